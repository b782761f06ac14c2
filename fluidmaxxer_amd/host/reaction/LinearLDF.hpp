#pragma once
// Henry isotherm with a linear-driving-force uptake (LinearLDF.hpp:8-45, LinearLDF.cu:13-75 of
// the reference): c = max(phi, m) - m, q_eq = K c, dq = k (q_eq - q), q += dq, S = -dq on the
// region, S = 0 elsewhere.  `phi` is the AD lattice's concentration from the previous step.
#include <cstdint>

#include "core/DeviceHelpers.hpp"
#include "core/Params.hpp"
#include "core/lattice/ADLattice.hpp"
#include "reaction/AdsorbedField.hpp"

// the reference's launcher signature (LinearLDF.hpp:8-17)
inline void launch_henry_ldf_source_d3q7(int nCells, const std::uint8_t *d_regionMask, const DATA_TYPE *d_phi,
                                         DATA_TYPE *d_Sphi, DATA_TYPE *d_q, DATA_TYPE K, DATA_TYPE kLDF,
                                         DATA_TYPE phiMinClamp) {
    fmx::check(fmx_ldf_linear(fmx::kPrec, nCells, d_regionMask, d_phi, d_Sphi, d_q, double(K), double(kLDF),
                              double(phiMinClamp), nullptr),
               "k_henry_ldf_source");
}

template <typename ADDesc> class LinearLDF {
public:
    explicit LinearLDF(const LBMParams &p) : p_(p) {}
    void computeSource(ADLattice<ADDesc> &ad, AdsorbedField &q, const std::uint8_t *d_regionMask) {
        fmx::check(fmx_ldf_linear(fmx::kPrec, ad.size64(), d_regionMask, ad.d_phi_ptr(), ad.d_Sphi_ptr(), q.d_q(),
                                  double(p_.KHenry), double(p_.kA), double(p_.minConcentration), nullptr),
                   "k_henry_ldf_source");
    }
    const LBMParams &P() const { return p_; }

private:
    LBMParams p_;
};
