#pragma once
// buildXPlaneMask and hasFlag (helpers.hpp:7-22 of the reference).
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "fmx_abi.h"

inline std::vector<std::uint8_t> buildXPlaneMask(int Nx, int Ny, int Nz, int xPlane) {
    std::vector<std::uint8_t> m(std::size_t(Nx) * Ny * Nz, 0);
    if (fmx_build_x_plane_mask(Nx, Ny, Nz, xPlane, m.data()) != FMX_OK) throw std::runtime_error(fmx_last_error());
    return m;
}

inline bool hasFlag(int argc, char **argv, const std::string &flag) {
    for (int i = 1; i < argc; ++i)
        if (flag == argv[i]) return true;
    return false;
}
