#pragma once
// Partial bounce-back BGK (no reference counterpart): Walsh, Burwinkle & Saar, Computers &
// Geosciences 35 (2009) 1186, f_out = (1 - n_s) f^c + n_s f_opp with the per-cell solid
// fraction n_s in [0, 1] taken from the lattice's porosity buffer (setPorosityField).
#include "core/DeviceHelpers.hpp"
#include "core/interfaces/NSInterfaces.hpp"
#include "core/lattice/NSLattice.hpp"

template <typename Descriptor> class PartialBounceBackCollision : public NSCollisionOperator<Descriptor> {
public:
    explicit PartialBounceBackCollision(const LBMParams &p) : params(p) {}
    void initEq(Lattice<Descriptor> &lat) override {
        static_assert(Descriptor::Q == 19, "PartialBounceBackCollision is implemented for D3Q19");
        fmx_grid g = lat.grid();
        fmx::check(fmx_ns_init_eq(fmx::kPrec, lat.d_f_ptr(), &g, double(lat.P().rho0), nullptr), "fmx_ns_init_eq");
    }
    void collideStream(Lattice<Descriptor> &lat) override {
        fmx_grid g = lat.grid();
        fmx::check(fmx_ns_collide_stream(fmx::kPrec, FMX_NS_PBB, lat.d_f_ptr(), lat.d_f_new_ptr(), nullptr,
                                         lat.d_porosity_ptr(), &g, double(lat.P().tau), double(lat.P().rho0), nullptr),
                   "fmx_ns_collide_stream");
    }

private:
    LBMParams params{};
};
