#pragma once
// Post-streaming bounce-back on a mask (BounceBack.hpp:11-44, BounceBack.cu:12-53 of the
// reference) as a true swap of each opposite pair; the reference's in-place loop
// (BounceBack.cu:29-30) overwrites each pair with one value, which is not reproduced.
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "core/DeviceIndexList.hpp"
#include "core/interfaces/NSInterfaces.hpp"
#include "core/lattice/NSLattice.hpp"

template <typename Descriptor> class BounceBack : public NSBoundaryCondition<Descriptor> {
public:
    BounceBack(const LBMParams &p, std::vector<std::uint8_t> mask) : params(p), h_mask(std::move(mask)) {
        if (std::int64_t(h_mask.size()) != std::int64_t(p.Nx) * p.Ny * p.Nz)
            throw std::runtime_error("BounceBack mask mismatch");
        d_cells.assign(fmx::maskCells(h_mask));
    }
    void apply(Lattice<Descriptor> &lat, BoundaryPhase phase, int) override {
        if (phase != BoundaryPhase::PostStreaming || d_cells.size() == 0) return;
        fmx::check(fmx_ns_bc_bounceback(fmx::kPrec, lat.d_f_new_ptr(), d_cells.data(), d_cells.size(), lat.size64(),
                                        nullptr),
                   "fmx_ns_bc_bounceback");
    }

private:
    LBMParams params{};
    std::vector<std::uint8_t> h_mask;
    fmx::DeviceIndexList d_cells;
};
