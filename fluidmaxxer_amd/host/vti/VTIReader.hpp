#pragma once
// readVTI (VTIReader.hpp:14-160 of the reference) over fmx_vti_read_mask: ASCII VTK ImageData,
// dimensions from WholeExtent (else Extent), mask = first DataArray value > threshold
// (3-component arrays by magnitude).  Throws std::runtime_error with the reader's message.
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "fmx_abi.h"

struct VTIResult {
    int Nx = 0;
    int Ny = 0;
    int Nz = 0;
    std::vector<std::uint8_t> mask;
};

inline VTIResult readVTI(const std::string &filename, double threshold = 0.5) {
    VTIResult r;
    int32_t dims[3] = {0, 0, 0};
    if (fmx_vti_read_mask(filename.c_str(), threshold, dims, nullptr, 0) != FMX_OK)
        throw std::runtime_error(fmx_last_error());
    r.Nx = dims[0];
    r.Ny = dims[1];
    r.Nz = dims[2];
    r.mask.assign(std::size_t(r.Nx) * r.Ny * r.Nz, 0);
    if (fmx_vti_read_mask(filename.c_str(), threshold, dims, r.mask.data(), std::int64_t(r.mask.size())) != FMX_OK)
        throw std::runtime_error(fmx_last_error());
    return r;
}

inline void writeVTIMask(const std::string &filename, int Nx, int Ny, int Nz, const std::vector<std::uint8_t> &mask,
                         const char *name = "solid") {
    if (fmx_vti_write_mask(filename.c_str(), Nx, Ny, Nz, mask.data(), name) != FMX_OK)
        throw std::runtime_error(fmx_last_error());
}
