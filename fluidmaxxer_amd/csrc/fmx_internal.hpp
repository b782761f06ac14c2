// fmx_internal.hpp -- error handling and launch helpers shared by the ABI files.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <exception>

#include "fmx_device.hpp"

namespace fmx {

// Thread-local last-error message (fmx_last_error); defined in runtime.cpp.
void set_error_msg(const char *msg);
int set_error(int code, const char *fmt, ...);
// Reference behaviour of cudaCheckThrow (CudaHelpers.cu:16-20) when enabled.
bool sync_check_enabled();

inline hipStream_t as_stream(void *s) { return static_cast<hipStream_t>(s); }

inline i64 grid_cells(const fmx_grid &g) { return i64(g.nx) * g.ny * g.nz; }
inline bool grid_ok(const fmx_grid &g) { return g.nx > 0 && g.ny > 0 && g.nz > 0 && g.ny < 65536 * 4 && g.nz < 65536; }
inline unsigned blocks_for(i64 n, int block = 256) { return unsigned((n + block - 1) / block); }

inline int check_hip(hipError_t e, const char *what) {
    if (e != hipSuccess) return set_error(FMX_ERR_HIP, "HIP error: %s | %s", what, hipGetErrorString(e));
    return FMX_OK;
}

// Launch-error check after every kernel; with sync-check on, also synchronise.
inline int after_launch(const char *what) {
    int rc = check_hip(hipGetLastError(), what);
    if (rc) return rc;
    if (sync_check_enabled()) return check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
    return FMX_OK;
}

inline int unsupported_prec(const char *fn, int prec) {
    return set_error(FMX_ERR_UNSUPPORTED, "%s: precision %d not supported here (FMX_F32 or FMX_F64)", fn, prec);
}

}  // namespace fmx

#define FMX_TRY_BEGIN try {
#define FMX_TRY_END                                                          \
    }                                                                        \
    catch (const std::exception &e) {                                        \
        return ::fmx::set_error(FMX_ERR_ARG, "host exception: %s", e.what()); \
    }                                                                        \
    catch (...) {                                                            \
        return ::fmx::set_error(FMX_ERR_ARG, "unknown host exception");      \
    }
