// CPU checks of the host-side physical-unit layer: UnitConverter, ADUnitConverter, main.cpp's
// correlations and the VTI writer (no device needed).
//
//   host_units <ascii.vti> <binary.vti> <pvd>
//
// Prints one JSON object with the converter outputs for main.cpp's parameter set and writes a
// 5x4x3 ImageData file twice (ASCII and raw appended binary) with a scalar and a 3-vector field,
// plus a two-entry .pvd collection.
#include <cstdio>
#include <string>
#include <vector>

#include "boundary/ad/ADDirichletTanhPulse.hpp"
#include "core/descriptors/D3Q7.hpp"
#include "core/io/ADUnitConverter.hpp"
#include "core/io/Correlations.hpp"
#include "core/io/UnitConverter.hpp"
#include "vti/VTIWriter.hpp"

int main(int argc, char **argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: host_units <ascii.vti> <binary.vti> <pvd>\n");
        return 2;
    }
    // main.cpp:157-193 parameter set
    const DATA_TYPE dp = DATA_TYPE(0.0002), Pe = DATA_TYPE(40), nu = DATA_TYPE(1.0e-6), rho = DATA_TYPE(1000.0);
    const DATA_TYPE epsBed = DATA_TYPE(0.26), epsP = DATA_TYPE(0.67), Dm = DATA_TYPE(2e-09);
    const DATA_TYPE Deff = computeEffectiveDiffusivity(Dm, epsP);
    const DATA_TYPE us = calculateInletVelocityFromPeclet(Pe, Dm, dp, epsBed);
    const DATA_TYPE ui = calculateInterstitialVelocity(us, epsBed);
    UnitConverter uc(dp, ui, nu, rho, 15, 64, 32, 32, 0.75);
    const DATA_TYPE kf = calculatekFilm(DATA_TYPE(uc.reynolds()), nu, Dm, epsP, dp);
    const DATA_TYPE kldf = calculatekLDF(epsBed, dp, kf, DATA_TYPE(uc.dt_phys));
    ADUnitConverter<D3Q7> ad(5, 4, 3, uc.dx_phys, uc.dt_phys);

    std::vector<std::uint8_t> mask(std::size_t(5) * 4 * 3);
    for (std::size_t i = 0; i < mask.size(); ++i) mask[i] = (i * 7 + 3) % 5 < 2;
    const std::vector<DATA_TYPE> tau = ad.buildTauField(mask, Dm, Deff);

    std::printf("{\"Deff\": %.17g, \"u_superficial\": %.17g, \"u_interstitial\": %.17g, \"dx\": %.17g, \"dt\": %.17g, "
                "\"nu_lbm\": %.17g, \"u_lbm\": %.17g, \"Re\": %.17g, \"Ma\": %.17g, \"Kn\": %.17g, \"kfilm\": %.17g, "
                "\"kldf\": %.17g, \"tau_fluid\": %.17g, \"tau_solid\": %.17g, \"tau_field0\": %.17g, \"tau_field_mask\": %.17g, "
                "\"len_p2l\": %.17g, \"time_l2p\": %.17g, \"visc_p2l\": %.17g, \"dens_l2p\": %.17g, \"vel_l2p\": %.17g}\n",
                double(Deff), double(us), double(ui), uc.dx_phys, uc.dt_phys, uc.nu_lbm, uc.u_lbm, uc.reynolds(),
                uc.mach(), uc.knudsen(), double(kf), double(kldf), ad.tauFromDiffusivityPhys(Dm),
                ad.tauFromDiffusivityPhys(Deff), double(tau[0]), double(tau[1]), uc.lengthPhysToLattice(1e-3),
                uc.timeLatticeToPhys(1000.0), uc.viscosityPhysToLattice(1e-6), uc.densityLatticeToPhys(1.02),
                uc.velocityLatticeToPhys(0.01));

    // ADDirichletTanhPulse::valueAt across both edges of a pulse (base 0.1, amp 0.9, t_on 300, t_off 900, k 40)
    LBMParams pp{};
    pp.Nx = 5;
    pp.Ny = 4;
    pp.Nz = 3;
    const ADDirichletTanhPulse<D3Q7> pulse(pp, std::vector<std::uint8_t>(60, 0), DATA_TYPE(0.1), DATA_TYPE(0.9), 300, 900,
                                           DATA_TYPE(40));
    std::printf("{\"pulse\": [");
    const int pulse_steps[] = {0, 1, 150, 260, 299, 300, 301, 340, 600, 880, 900, 901, 950, 1200, 5000};
    for (int i = 0; i < 15; ++i) std::printf("%s[%d, %.17g]", i ? ", " : "", pulse_steps[i], double(pulse.valueAt(pulse_steps[i])));
    std::printf("]}\n");

    std::vector<fmx_vti::Field> fields;
    fields.push_back({"mask", 1, [&](std::int64_t c, int) { return DATA_TYPE(mask[std::size_t(c)]); }});
    fields.push_back({"vec", 3, [&](std::int64_t c, int k) { return DATA_TYPE(0.25) * DATA_TYPE(c % 7) + DATA_TYPE(k); }});
    fmx_vti::writeImageData(argv[1], 5, 4, 3, uc.dx_phys, fields, false);
    fmx_vti::writeImageData(argv[2], 5, 4, 3, uc.dx_phys, fields, true);
    fmx_vti::appendToPVD(argv[3], argv[1], 0.0);
    fmx_vti::appendToPVD(argv[3], argv[2], 0.5);
    return 0;
}
