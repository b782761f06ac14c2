#pragma once
// Langmuir isotherm with a linear-driving-force uptake (LangmuirLDF.hpp:7-34,
// LangmuirLDF.cu:7-59 of the reference): q_eq = q_max b c / (1 + b c).
#include <cstdint>

#include "core/DeviceHelpers.hpp"
#include "core/Params.hpp"
#include "core/lattice/ADLattice.hpp"
#include "reaction/AdsorbedField.hpp"

// the reference's launcher signature (LangmuirLDF.hpp:7-14)
inline void launch_langmuir_ldf_source_d3q7(int nCells, const std::uint8_t *d_regionMask, const DATA_TYPE *d_phi,
                                            DATA_TYPE *d_Sphi, DATA_TYPE *d_q, DATA_TYPE qMax, DATA_TYPE b,
                                            DATA_TYPE kLDF, DATA_TYPE phiMinClamp) {
    fmx::check(fmx_ldf_langmuir(fmx::kPrec, nCells, d_regionMask, d_phi, d_Sphi, d_q, double(qMax), double(b),
                                double(kLDF), double(phiMinClamp), nullptr),
               "k_langmuir_ldf_source");
}

template <typename ADDesc> class LangmuirLDF {
public:
    explicit LangmuirLDF(const LBMParams &p) : p_(p) {}
    void computeSource(ADLattice<ADDesc> &ad, AdsorbedField &q, const std::uint8_t *d_regionMask) {
        fmx::check(fmx_ldf_langmuir(fmx::kPrec, ad.size64(), d_regionMask, ad.d_phi_ptr(), ad.d_Sphi_ptr(), q.d_q(),
                                    double(p_.qMax), double(p_.bLangmuir), double(p_.kLDF),
                                    double(p_.minConcentration), nullptr),
                   "k_langmuir_ldf_source");
    }
    const LBMParams &P() const { return p_; }

private:
    LBMParams p_;
};
