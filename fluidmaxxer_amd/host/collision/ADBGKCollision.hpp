#pragma once
// D3Q7 BGK with pull streaming and a per-cell source (ADBGKCollision.hpp,
// ADBGKCollision.cu:29-168 of the reference).
#include "core/DeviceHelpers.hpp"
#include "core/interfaces/ADInterfaces.hpp"
#include "core/lattice/ADLattice.hpp"

template <typename Descriptor> class ADBGKCollision : public ADCollisionOperator<Descriptor> {
public:
    explicit ADBGKCollision(const LBMParams &p) : params(p) {}
    void initEq(ADLattice<Descriptor> &lat) override {
        static_assert(Descriptor::Q == 7, "ADBGKCollision is implemented for D3Q7");
        fmx_grid g = lat.grid();
        fmx::check(fmx_ad_init_eq(fmx::kPrec, lat.d_g_ptr(), &g, double(lat.P().phi0), nullptr), "fmx_ad_init_eq");
    }
    void collideStream(ADLattice<Descriptor> &lat) override {
        fmx_grid g = lat.grid();
        fmx::check(fmx_ad_collide_stream(fmx::kPrec, lat.d_g_ptr(), lat.d_g_new_ptr(), lat.d_phi_ptr(), lat.d_tau_ptr(),
                                         lat.d_ux_adv_ptr(), lat.d_uy_adv_ptr(), lat.d_uz_adv_ptr(), lat.d_Sphi_ptr(),
                                         &g, double(lat.P().minConcentration), nullptr),
                   "fmx_ad_collide_stream");
    }

private:
    LBMParams params{};
};
