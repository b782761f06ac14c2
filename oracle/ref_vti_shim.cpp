// ref_vti_shim.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Exposes the reference's own header-only host code -- readVTI
// (include/vti/VTIReader.hpp:35-160) and buildXPlaneMask
// (include/core/helpers/helpers.hpp:7-14) -- through a C entry point, so the
// test suite can check this repository's VTI reader and mask builder against the
// reference implementation itself.  The headers are compiled where they lie in
// the reference tree (oracle/Makefile, target `ref`); nothing from them is copied
// here.  Output goes to oracle/_ref/ only.
#include <cstdint>
#include <cstring>
#include <string>

#include "core/helpers/helpers.hpp"
#include "vti/VTIReader.hpp"

extern "C" {

// Reads `path` with the reference readVTI.  dims[0..2] <- Nx, Ny, Nz.  Copies at
// most `cap` mask bytes into `mask` (nullable to query the size).  Returns 0 on
// success, 1 when readVTI threw (message in `err`, truncated to errlen).
int ref_read_vti(const char *path, double threshold, int *dims, std::uint8_t *mask, long long cap, char *err,
                 int errlen) {
    try {
        VTIResult r = readVTI(std::string(path), threshold);
        dims[0] = r.Nx;
        dims[1] = r.Ny;
        dims[2] = r.Nz;
        if (mask) {
            const long long n = static_cast<long long>(r.mask.size()) < cap ? static_cast<long long>(r.mask.size()) : cap;
            std::memcpy(mask, r.mask.data(), static_cast<std::size_t>(n));
        }
        return 0;
    } catch (const std::exception &e) {
        if (err && errlen > 0) {
            std::strncpy(err, e.what(), static_cast<std::size_t>(errlen - 1));
            err[errlen - 1] = 0;
        }
        return 1;
    }
}

// buildXPlaneMask from the reference helpers.
void ref_build_x_plane_mask(int Nx, int Ny, int Nz, int xPlane, std::uint8_t *out) {
    std::vector<std::uint8_t> m = buildXPlaneMask(Nx, Ny, Nz, xPlane);
    std::memcpy(out, m.data(), m.size());
}

}  // extern "C"
