#pragma once
// Zero-gradient outlet (ADNeumannOutlet.hpp:13-38, ADNeumannOutlet.cu:5-53 of the reference),
// PostStreaming on g_new: mask cells on x = Nx-1 copy all 7 populations from x = Nx-2.
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "core/DeviceIndexList.hpp"
#include "core/interfaces/ADInterfaces.hpp"
#include "core/lattice/ADLattice.hpp"

template <typename Descriptor> class ADNeumannOutlet : public ADBoundaryCondition<Descriptor> {
public:
    ADNeumannOutlet(const LBMParams &p, std::vector<std::uint8_t> mask) : params(p), h_mask(std::move(mask)) {
        if (std::int64_t(h_mask.size()) != std::int64_t(p.Nx) * p.Ny * p.Nz)
            throw std::runtime_error("ADNeumannOutlet mask mismatch");
        std::vector<std::int64_t> cells, src;
        for (std::int64_t c : fmx::maskCells(h_mask))
            if (c % p.Nx == p.Nx - 1 && p.Nx > 1) {
                cells.push_back(c);
                src.push_back(c - 1);
            }
        d_cells.assign(cells);
        d_src.assign(src);
    }
    void apply(ADLattice<Descriptor> &lat, BoundaryPhase phase, int) override {
        if (phase != BoundaryPhase::PostStreaming || d_cells.size() == 0) return;
        fmx::check(fmx_ad_bc_neumann(fmx::kPrec, lat.d_g_new_ptr(), d_cells.data(), d_src.data(), d_cells.size(),
                                     lat.size64(), nullptr),
                   "fmx_ad_bc_neumann");
    }
    const std::vector<std::uint8_t> &mask() const { return h_mask; }

private:
    LBMParams params{};
    std::vector<std::uint8_t> h_mask;
    fmx::DeviceIndexList d_cells, d_src;
};
