#pragma once
// Device memory helpers over the C ABI, with the reference's throw-on-error convention
// (CudaHelpers.hpp:5-23): a failing call prints "FMX error: <what> | <detail>" to stderr and
// throws std::runtime_error.
#include <cstddef>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "fmx_abi.h"

namespace fmx {

inline void check(int rc, const char *what) {
    if (rc != FMX_OK) {
        const std::string msg = std::string("FMX error: ") + what + " | " + fmx_last_error();
        std::fprintf(stderr, "%s\n", msg.c_str());
        throw std::runtime_error(msg);
    }
}

inline void *mallocBytes(std::size_t bytes) {
    void *p = nullptr;
    check(fmx_alloc(&p, bytes), "fmx_alloc");
    return p;
}
inline void freeBytes(void *p) {
    if (p) fmx_free(p);
}
inline void memcpyHtoD(void *d, const void *h, std::size_t bytes) {
    check(fmx_memcpy_h2d(d, h, bytes, nullptr), "fmx_memcpy_h2d");
    check(fmx_stream_sync(nullptr), "fmx_stream_sync");
}
inline void memcpyDtoH(void *h, const void *d, std::size_t bytes) {
    check(fmx_memcpy_d2h(h, d, bytes, nullptr), "fmx_memcpy_d2h");
    check(fmx_stream_sync(nullptr), "fmx_stream_sync");
}
inline void memcpyD2D(void *dst, const void *src, std::size_t bytes) {
    check(fmx_memcpy_d2d(dst, src, bytes, nullptr), "fmx_memcpy_d2d");
}
inline void memsetBytes(void *d, int value, std::size_t bytes) {
    check(fmx_memset(d, value, bytes, nullptr), "fmx_memset");
}
// last launch error + synchronise (cudaCheckThrow, CudaHelpers.cu:16-20)
inline void checkThrow(const char *what) { check(fmx_device_sync(), what); }

inline fmx_grid gridOf(int Nx, int Ny, int Nz, bool px, bool py, bool pz) {
    fmx_grid g;
    g.nx = Nx;
    g.ny = Ny;
    g.nz = Nz;
    g.periodic_x = px;
    g.periodic_y = py;
    g.periodic_z = pz;
    return g;
}

}  // namespace fmx
