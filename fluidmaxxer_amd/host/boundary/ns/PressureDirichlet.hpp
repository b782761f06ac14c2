#pragma once
// Constant-density outlet (PressureDirichlet.hpp:12-28, PressureDirichlet.cu:26-145 of the
// reference), PostStreaming on f_new: mask cells with x > 0 get feq(rho_out, u of the x-1
// neighbour) in the form w rho (1 + 3eu + 4.5eu^2 - 1.5u^2).
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "core/DeviceIndexList.hpp"
#include "core/interfaces/NSInterfaces.hpp"
#include "core/lattice/NSLattice.hpp"

template <typename Descriptor> class PressureDirichlet : public NSBoundaryCondition<Descriptor> {
public:
    PressureDirichlet(const LBMParams &p, std::vector<std::uint8_t> mask, DATA_TYPE rho_out_)
        : params(p), mask_h(std::move(mask)), rho_out(rho_out_) {
        if (std::int64_t(mask_h.size()) != std::int64_t(p.Nx) * p.Ny * p.Nz)
            throw std::runtime_error("PressureDirichlet mask mismatch");
        std::vector<std::int64_t> cells, nei;
        for (std::int64_t c : fmx::maskCells(mask_h))
            if (c % p.Nx > 0) {  // PressureDirichlet.cu:50
                cells.push_back(c);
                nei.push_back(c - 1);
            }
        d_cells.assign(cells);
        d_nei.assign(nei);
    }

    void apply(Lattice<Descriptor> &lat, BoundaryPhase phase, int) override {
        if (phase != BoundaryPhase::PostStreaming || d_cells.size() == 0) return;
        fmx::check(fmx_ns_bc_pressure(fmx::kPrec, lat.d_f_new_ptr(), d_cells.data(), d_nei.data(), d_cells.size(),
                                      double(rho_out), lat.size64(), nullptr),
                   "fmx_ns_bc_pressure");
    }

    const std::vector<std::uint8_t> &mask() const { return mask_h; }
    DATA_TYPE rhoOut() const { return rho_out; }

private:
    LBMParams params{};
    std::vector<std::uint8_t> mask_h;
    fmx::DeviceIndexList d_cells, d_nei;
    DATA_TYPE rho_out = DATA_TYPE(1);
};
