// lbm_coupled -- the reference's main program (main.cpp:101-357) on this library: a physical-unit
// driver that reads geometry.vti, derives every lattice parameter from the physical ones (Peclet
// inlet velocity, Mackie-Meares particle diffusivity, film mass transfer, LDF rate), sets up the
// porous NS lattice with a velocity inlet and pressure outlet and the AD lattice with a per-cell
// relaxation time, a Dirichlet (or tanh-pulse) inlet and a Neumann outlet, and runs the coupled
// step with the linear isotherm on the obstacle region, logging the outlet breakthrough curve
// (OutletCSV) and writing VTI/PVD fields in physical units.
//
// Flags as main.cpp: --coupling --adsorb --pulse --vti.  Additions for running it at other sizes
// and in tests: --geometry FILE (default geometry.vti), --out DIR (default results), --steps N
// (default: maxPhysT in lattice steps), --every N (field output interval, default 10000), --csv-every N
// (default 10000), --binary (appended binary VTI arrays).  The lattice parameters are printed as
// one JSON line ("params: {...}").
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "fluidmaxxer.hpp"

namespace {

std::string argValue(int argc, char **argv, const std::string &flag, const std::string &def) {
    for (int i = 1; i + 1 < argc; ++i)
        if (flag == argv[i]) return argv[i + 1];
    return def;
}

}  // namespace

int main(int argc, char **argv) {
    try {
        using Descriptor_NS = D3Q19;
        using Descriptor_AD = D3Q7;
        const bool enableCoupling = hasFlag(argc, argv, "--coupling");
        const bool enablePulse = hasFlag(argc, argv, "--pulse");
        const bool enableAdsorb = hasFlag(argc, argv, "--adsorb");
        const bool writeVTI = hasFlag(argc, argv, "--vti");
        const bool binaryVTI = hasFlag(argc, argv, "--binary");
        const std::string geometryFile = argValue(argc, argv, "--geometry", "geometry.vti");
        const std::string outDir = argValue(argc, argv, "--out", "results");
        std::printf("[main] Starting Coupled LBM Simulation\n[main] Coupling: %s | Inlet pulse: %s | Adsorption: %s | "
                    "writeVTI: %s\n", enableCoupling ? "ON" : "OFF", enablePulse ? "ON" : "OFF",
                    enableAdsorb ? "ON" : "OFF", writeVTI ? "ON" : "OFF");
        uploadD3Q19Constants();
        uploadD3Q7Constants();

        // geometry (main.cpp:142-146)
        VTIResult geom = readVTI(geometryFile, 0.5);
        LBMParams p{};
        p.Nx = geom.Nx;
        p.Ny = geom.Ny;
        p.Nz = geom.Nz;
        p.periodicX = false;
        p.periodicY = true;
        p.periodicZ = true;

        // physical parameter set (main.cpp:157-179)
        const int resolution = 15;
        const DATA_TYPE particleDiameter = DATA_TYPE(0.0002);
        const DATA_TYPE Peclet = DATA_TYPE(40);
        const DATA_TYPE nu_phys = DATA_TYPE(1.0e-6);
        const DATA_TYPE rho_phys = DATA_TYPE(1000.0);
        const DATA_TYPE maxPhysT = DATA_TYPE(80);
        const DATA_TYPE injectionTimePhys = DATA_TYPE(0.5);
        const DATA_TYPE geometryPorosity = DATA_TYPE(0.26);
        const DATA_TYPE particlePorosity = DATA_TYPE(0.67);
        const DATA_TYPE D_fluid = DATA_TYPE(2e-09);
        const DATA_TYPE D_solid = computeEffectiveDiffusivity(D_fluid, particlePorosity);
        const DATA_TYPE inletVelocity = calculateInletVelocityFromPeclet(Peclet, D_fluid, particleDiameter, geometryPorosity);
        const DATA_TYPE charPhysVelocity = calculateInterstitialVelocity(inletVelocity, geometryPorosity);
        std::printf("[config] Deff = %g for eps = %g; inlet velocity %g m/s (Pe %g), interstitial %g m/s\n",
                    double(D_solid), double(particlePorosity), double(inletVelocity), double(Peclet),
                    double(charPhysVelocity));

        UnitConverter uc(particleDiameter, charPhysVelocity, nu_phys, rho_phys, resolution, p.Nx, p.Ny, p.Nz, 0.75);
        uc.printInfo();
        const DATA_TYPE qMax = DATA_TYPE(1.0);
        const DATA_TYPE KHenry = DATA_TYPE(2);
        const DATA_TYPE kA = calculatekFilm(DATA_TYPE(uc.reynolds()), nu_phys, D_fluid, particlePorosity, particleDiameter);
        const DATA_TYPE k_ldf_LU = calculatekLDF(geometryPorosity, particleDiameter, kA, DATA_TYPE(uc.dt_phys));
        p.enableAdsorption = enableAdsorb;
        p.qMax = qMax;
        p.KHenry = KHenry;
        p.kA = k_ldf_LU;
        p.rho0 = static_cast<DATA_TYPE>(uc.rho_lbm);
        p.tau = static_cast<DATA_TYPE>(uc.tau_lbm);

        // lattices and operators (main.cpp:209-219)
        Lattice<Descriptor_NS> nsLat(p);
        ADLattice<Descriptor_AD> adLat(p);
        nsLat.setCollisionOperator(std::make_shared<PorousBGKCollision<Descriptor_NS>>(p));
        adLat.setCollisionOperator(std::make_shared<ADBGKCollision<Descriptor_AD>>(p));

        // boundaries (main.cpp:222-287)
        const auto inletMask = buildXPlaneMask(p.Nx, p.Ny, p.Nz, 0);
        const auto outletMask = buildXPlaneMask(p.Nx, p.Ny, p.Nz, p.Nx - 1);
        nsLat.setObstacleMask(geom.mask);
        nsLat.setPorosityFromMask(geom.mask, particlePorosity, DATA_TYPE(1.0));
        const DATA_TYPE u_in = static_cast<DATA_TYPE>(uc.velocityPhysToLattice(static_cast<double>(inletVelocity)));
        const DATA_TYPE rho_out = p.rho0;
        nsLat.addBoundary<VelocityDirichletNEQ<Descriptor_NS>>(p, inletMask, u_in, DATA_TYPE(0), DATA_TYPE(0));
        nsLat.addBoundary<PressureDirichlet<Descriptor_NS>>(p, outletMask, rho_out);

        ADUnitConverter<Descriptor_AD> adUc(p.Nx, p.Ny, p.Nz, uc.dx_phys, uc.dt_phys);
        const std::vector<DATA_TYPE> tauField = adUc.buildTauField(geom.mask, D_fluid, D_solid);
        adLat.setTauField(tauField);
        AdsorbedField qField(p);
        LinearLDF<Descriptor_AD> isotherm(p);
        int iTmax = 0, t_on = 0, t_off = 0;
        DATA_TYPE k_steps = 0;
        if (enablePulse) {
            iTmax = static_cast<int>(uc.timePhysToLattice(injectionTimePhys));
            t_on = static_cast<int>(DATA_TYPE(0.10) * iTmax);
            t_off = static_cast<int>(DATA_TYPE(0.90) * iTmax);
            k_steps = DATA_TYPE(0.05) * DATA_TYPE(iTmax);
            std::printf("[main] Pulse enabled, iTmax=%d t_on=%d t_off=%d k=%.0f\n", iTmax, t_on, t_off, double(k_steps));
            adLat.addBoundary<ADDirichletTanhPulse<Descriptor_AD>>(p, inletMask, DATA_TYPE(0.0), DATA_TYPE(1.0), t_on,
                                                                    t_off, k_steps);
        } else {
            adLat.addBoundary<ADDirichlet<Descriptor_AD>>(p, inletMask, DATA_TYPE(1.0));
        }
        adLat.addBoundary<ADNeumannOutlet<Descriptor_AD>>(p, outletMask);

        nsLat.initEquilibrium();
        adLat.initEquilibrium();

        // the coupled step: one fused launch per step where the operators allow it (main.cpp:321-336)
        std::unique_ptr<CoupledStep<Descriptor_NS, Descriptor_AD>> cs;
        if (enableAdsorb)
            cs = std::make_unique<CoupledStep<Descriptor_NS, Descriptor_AD>>(nsLat, adLat, qField, isotherm,
                                                                             nsLat.obstacleMaskHost(), enableCoupling);
        else
            cs = std::make_unique<CoupledStep<Descriptor_NS, Descriptor_AD>>(nsLat, adLat, qField, enableCoupling);

        // output (main.cpp:290-311)
        std::filesystem::create_directories(outDir);
        VTIWriterPhys<Descriptor_NS> w(uc);
        w.setOutputPattern(outDir + "/step_", 6);
        w.setPVDFile(outDir + "/out.pvd");
        w.setBinary(binaryVTI);
        w.registerRhoPhys(nsLat);
        w.registerVelocityPhys(nsLat);
        w.registerSpeedMagPhys(nsLat);
        w.registerObstacle(nsLat);
        w.registerConcentration(adLat, "concentration");
        w.registerScalarField("adfsorbed amount", qField.host());
        const std::string csvPath = outDir + "/outlet_concentration.csv";
        OutletCSV<ADLattice<Descriptor_AD>, UnitConverter> outletCsv(csvPath.c_str(), uc, outletMask, adLat);

        const long defaultSteps = static_cast<long>(uc.timePhysToLattice(maxPhysT));
        const long steps = std::atol(argValue(argc, argv, "--steps", std::to_string(defaultSteps)).c_str());
        const long outputEvery = std::max(1L, std::atol(argValue(argc, argv, "--every", "10000").c_str()));
        const long csvEvery = std::max(1L, std::atol(argValue(argc, argv, "--csv-every", "10000").c_str()));
        std::printf("params: {\"Nx\": %d, \"Ny\": %d, \"Nz\": %d, \"tau\": %.17g, \"rho0\": %.17g, \"u_in\": %.17g, "
                    "\"tau_fluid\": %.17g, \"tau_solid\": %.17g, \"porosity_solid\": %.17g, \"KHenry\": %.17g, "
                    "\"k_ldf\": %.17g, \"steps\": %ld, \"dt\": %.17g, \"dx\": %.17g, \"fused\": %s, "
                    "\"pulse\": [%d, %d, %.17g]}\n",
                    p.Nx, p.Ny, p.Nz, double(p.tau), double(p.rho0), double(u_in), adUc.tauFromDiffusivityPhys(D_fluid),
                    adUc.tauFromDiffusivityPhys(D_solid), double(particlePorosity), double(KHenry), double(k_ldf_LU),
                    steps, uc.dt_phys, uc.dx_phys, cs->fused() ? "true" : "false", t_on, t_off, double(k_steps));
        std::fflush(stdout);

        if (writeVTI) w.writeStep(nsLat, uc.timeLatticeToPhys(0), 0);
        outletCsv.log(0);
        // main loop (main.cpp:324-355), fused steps batched between log / output events
        long s = 1;
        while (s <= steps) {
            const long nextCsv = (s + csvEvery - 1) / csvEvery * csvEvery;
            const long nextOut = (s + outputEvery - 1) / outputEvery * outputEvery;
            const long stop = std::min({steps, nextCsv, nextOut});
            cs->run(int(s), int(stop - s + 1));
            s = stop + 1;
            if (stop % csvEvery == 0) outletCsv.log(int(stop));
            if (stop % outputEvery == 0) {
                if (enableAdsorb) qField.download();
                std::printf("[timer] Step %ld/%ld | physT = %g (%.2f%%)\n", stop, steps, uc.timeLatticeToPhys(double(stop)),
                            100.0 * double(stop) / double(steps));
                if (writeVTI) {
                    cs->finish();
                    w.writeStep(nsLat, uc.timeLatticeToPhys(double(stop)), int(stop));
                }
                outletCsv.flush();
            }
        }
        cs->finish();
        outletCsv.flush();
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "lbm_coupled: %s\n", e.what());
        return 1;
    }
}
