#pragma once
// Equilibrium velocity inlet (VelocityDirichlet.hpp:11-36, VelocityDirichlet.cu:54-130 of the
// reference), PreCollision on f: f <- feq(params.rho0, u0) on the mask.
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "core/DeviceIndexList.hpp"
#include "core/interfaces/NSInterfaces.hpp"
#include "core/lattice/NSLattice.hpp"

template <typename Descriptor> class VelocityDirichletEq : public NSBoundaryCondition<Descriptor> {
public:
    VelocityDirichletEq(const LBMParams &p, std::vector<std::uint8_t> mask, DATA_TYPE ux, DATA_TYPE uy, DATA_TYPE uz)
        : params(p), h_mask(std::move(mask)), u0x(ux), u0y(uy), u0z(uz) {
        if (std::int64_t(h_mask.size()) != std::int64_t(p.Nx) * p.Ny * p.Nz)
            throw std::runtime_error("VelocityDirichletEq mask mismatch");
        d_cells.assign(fmx::maskCells(h_mask));
    }

    void apply(Lattice<Descriptor> &lat, BoundaryPhase phase, int) override {
        if (phase != BoundaryPhase::PreCollision || d_cells.size() == 0) return;
        fmx::check(fmx_ns_bc_velocity_eq(fmx::kPrec, lat.d_f_ptr(), d_cells.data(), d_cells.size(), double(params.rho0),
                                         double(u0x), double(u0y), double(u0z), lat.size64(), nullptr),
                   "fmx_ns_bc_velocity_eq");
    }

    const std::vector<std::uint8_t> &mask() const { return h_mask; }
    DATA_TYPE ux() const { return u0x; }
    DATA_TYPE uy() const { return u0y; }
    DATA_TYPE uz() const { return u0z; }

private:
    LBMParams params{};
    std::vector<std::uint8_t> h_mask;
    fmx::DeviceIndexList d_cells;
    DATA_TYPE u0x = DATA_TYPE(0), u0y = DATA_TYPE(0), u0z = DATA_TYPE(0);
};
