#pragma once
// BGK with obstacle bounce-back (BGKCollision.hpp, BGKCollision.cu:44-223 of the reference):
// solid cells reverse their populations on site, fluid cells pulling from a solid take their
// own opposite population (halfway bounce-back), tau clamped to >= 0.5001.
#include "core/DeviceHelpers.hpp"
#include "core/interfaces/NSInterfaces.hpp"
#include "core/lattice/NSLattice.hpp"

template <typename Descriptor> class BGKCollision : public NSCollisionOperator<Descriptor> {
public:
    explicit BGKCollision(const LBMParams &p) : params(p) {}
    void initEq(Lattice<Descriptor> &lat) override {
        static_assert(Descriptor::Q == 19, "BGKCollision is implemented for D3Q19");
        fmx_grid g = lat.grid();
        fmx::check(fmx_ns_init_eq(fmx::kPrec, lat.d_f_ptr(), &g, double(lat.P().rho0), nullptr), "fmx_ns_init_eq");
    }
    void collideStream(Lattice<Descriptor> &lat) override {
        fmx_grid g = lat.grid();
        fmx::check(fmx_ns_collide_stream(fmx::kPrec, FMX_NS_BGK, lat.d_f_ptr(), lat.d_f_new_ptr(), lat.d_obstacle_ptr(),
                                         nullptr, &g, double(lat.P().tau), double(lat.P().rho0), nullptr),
                   "fmx_ns_collide_stream");
    }

private:
    LBMParams params{};
};
