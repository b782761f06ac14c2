#pragma once
// Porous BGK (PorousBGKCollision.hpp, PorousBGKCollision.cu:43-197 of the reference):
// u <- clamp(porosity, 0, 1) * j / rho before the equilibrium; no obstacle handling.
#include "core/DeviceHelpers.hpp"
#include "core/interfaces/NSInterfaces.hpp"
#include "core/lattice/NSLattice.hpp"

template <typename Descriptor> class PorousBGKCollision : public NSCollisionOperator<Descriptor> {
public:
    explicit PorousBGKCollision(const LBMParams &p) : params(p) {}
    void initEq(Lattice<Descriptor> &lat) override {
        static_assert(Descriptor::Q == 19, "PorousBGKCollision is implemented for D3Q19");
        fmx_grid g = lat.grid();
        fmx::check(fmx_ns_init_eq(fmx::kPrec, lat.d_f_ptr(), &g, double(lat.P().rho0), nullptr), "fmx_ns_init_eq");
    }
    void collideStream(Lattice<Descriptor> &lat) override {
        fmx_grid g = lat.grid();
        fmx::check(fmx_ns_collide_stream(fmx::kPrec, FMX_NS_POROUS, lat.d_f_ptr(), lat.d_f_new_ptr(), nullptr,
                                         lat.d_porosity_ptr(), &g, double(lat.P().tau), double(lat.P().rho0), nullptr),
                   "fmx_ns_collide_stream");
    }

private:
    LBMParams params{};
};
