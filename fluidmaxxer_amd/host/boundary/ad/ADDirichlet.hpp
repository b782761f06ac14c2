#pragma once
// Constant-concentration inlet (ADDirichlet.hpp:14-42, ADDirichlet.cu:11-53 of the reference),
// PostStreaming on g_new: g <- phi * w on the mask.
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "core/DeviceIndexList.hpp"
#include "core/interfaces/ADInterfaces.hpp"
#include "core/lattice/ADLattice.hpp"

template <typename Descriptor> class ADDirichlet : public ADBoundaryCondition<Descriptor> {
public:
    ADDirichlet(const LBMParams &p, std::vector<std::uint8_t> mask, DATA_TYPE phi)
        : params(p), h_mask(std::move(mask)), phi0(phi) {
        if (std::int64_t(h_mask.size()) != std::int64_t(p.Nx) * p.Ny * p.Nz)
            throw std::runtime_error("ADDirichlet mask mismatch");
        d_cells.assign(fmx::maskCells(h_mask));
    }
    void apply(ADLattice<Descriptor> &lat, BoundaryPhase phase, int) override {
        if (phase != BoundaryPhase::PostStreaming || d_cells.size() == 0) return;
        fmx::check(fmx_ad_bc_dirichlet(fmx::kPrec, lat.d_g_new_ptr(), d_cells.data(), d_cells.size(), double(phi0),
                                       lat.size64(), nullptr),
                   "fmx_ad_bc_dirichlet");
    }
    const std::vector<std::uint8_t> &mask() const { return h_mask; }
    DATA_TYPE value() const { return phi0; }

private:
    LBMParams params{};
    std::vector<std::uint8_t> h_mask;
    fmx::DeviceIndexList d_cells;
    DATA_TYPE phi0 = DATA_TYPE(0);
};
