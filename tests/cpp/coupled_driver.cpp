// coupled_driver.cpp -- the reference main loop (main.cpp:101-357) written against this
// repository's C++ host API: geometry from a VTI file, BGK NS lattice with bounce-back
// obstacles, x = 0 velocity inlet (equilibrium or NEQ) and x = Nx-1 pressure outlet, D3Q7
// AD lattice with a Dirichlet inlet and Neumann outlet, linear or Langmuir LDF on the
// sphere region.  Runs either the operator sequence or CoupledStep and writes the fields.
//
//   coupled_driver <solid.vti> <region.vti> <out_dir> <steps> <fused|operators> key=value...
//   keys: tau u_in tau_ad K kA qmax b kldf inlet=eq|neq iso=linear|langmuir
//         adinlet=dirichlet|pulse (ADDirichlet C = 1, or ADDirichletTanhPulse with
//         pulse=base,amp,t_on,t_off,k), split=N (fused: run steps 1..N, then the rest, in two calls)
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "fluidmaxxer.hpp"

template <class T> static void dump(const std::string &path, const std::vector<T> &v) {
    std::ofstream out(path, std::ios::binary);
    out.write(reinterpret_cast<const char *>(v.data()), std::streamsize(v.size() * sizeof(T)));
}

int main(int argc, char **argv) {
    if (argc < 6) {
        std::fprintf(stderr, "usage: %s solid.vti region.vti out_dir steps fused|operators [key=value...]\n", argv[0]);
        return 2;
    }
    std::map<std::string, std::string> kv = {{"tau", "0.8"}, {"u_in", "0.05"}, {"tau_ad", "0.7"}, {"K", "0.5"},
                                             {"kA", "0.05"}, {"qmax", "2"},    {"b", "5"},        {"kldf", "0.02"},
                                             {"inlet", "eq"}, {"iso", "linear"}, {"adinlet", "dirichlet"},
                                             {"pulse", "0.1,0.9,20,70,6"}, {"split", "0"}};
    for (int i = 6; i < argc; ++i) {
        std::string a = argv[i];
        const auto eq = a.find('=');
        if (eq != std::string::npos) kv[a.substr(0, eq)] = a.substr(eq + 1);
    }
    auto num = [&](const char *k) { return DATA_TYPE(std::atof(kv[k].c_str())); };
    const std::string out_dir = argv[3];
    const int steps = std::atoi(argv[4]);
    const bool fusedMode = std::string(argv[5]) == "fused";

    VTIResult geom = readVTI(argv[1], 0.5);
    VTIResult reg = readVTI(argv[2], 0.5);
    LBMParams p{};
    p.Nx = geom.Nx;
    p.Ny = geom.Ny;
    p.Nz = geom.Nz;
    p.periodicX = false;
    p.periodicY = true;
    p.periodicZ = true;
    p.rho0 = 1;
    p.tau = num("tau");
    p.tau_ad = num("tau_ad");
    p.phi0 = 0;
    p.KHenry = num("K");
    p.kA = num("kA");
    p.qMax = num("qmax");
    p.bLangmuir = num("b");
    p.kLDF = num("kldf");

    Lattice<D3Q19> ns(p);
    ADLattice<D3Q7> ad(p);
    ns.setCollisionOperator(std::make_shared<BGKCollision<D3Q19>>(p));
    ad.setCollisionOperator(std::make_shared<ADBGKCollision<D3Q7>>(p));
    auto inlet = buildXPlaneMask(p.Nx, p.Ny, p.Nz, 0);
    auto outlet = buildXPlaneMask(p.Nx, p.Ny, p.Nz, p.Nx - 1);
    ns.setObstacleMask(geom.mask);
    if (kv["inlet"] == "neq") ns.addBoundary<VelocityDirichletNEQ<D3Q19>>(p, inlet, num("u_in"), DATA_TYPE(0), DATA_TYPE(0));
    else ns.addBoundary<VelocityDirichletEq<D3Q19>>(p, inlet, num("u_in"), DATA_TYPE(0), DATA_TYPE(0));
    ns.addBoundary<PressureDirichlet<D3Q19>>(p, outlet, p.rho0);
    ad.setTauField(std::vector<DATA_TYPE>(geom.mask.size(), p.tau_ad));
    if (kv["adinlet"] == "pulse") {
        double base = 0, amp = 0, k = 1;
        int t_on = 0, t_off = 0;
        if (std::sscanf(kv["pulse"].c_str(), "%lf,%lf,%d,%d,%lf", &base, &amp, &t_on, &t_off, &k) != 5) {
            std::fprintf(stderr, "pulse=base,amp,t_on,t_off,k expected\n");
            return 2;
        }
        ad.addBoundary<ADDirichletTanhPulse<D3Q7>>(p, inlet, DATA_TYPE(base), DATA_TYPE(amp), t_on, t_off, DATA_TYPE(k));
    } else {
        ad.addBoundary<ADDirichlet<D3Q7>>(p, inlet, DATA_TYPE(1.0));
    }
    ad.addBoundary<ADNeumannOutlet<D3Q7>>(p, outlet);
    AdsorbedField q(p);
    LinearLDF<D3Q7> lin(p);
    LangmuirLDF<D3Q7> lang(p);
    ns.initEquilibrium();
    ad.initEquilibrium();

    std::vector<double> c_out;
    bool wasFused = false;
    if (fusedMode) {
        std::unique_ptr<CoupledStep<D3Q19, D3Q7>> cs;
        if (kv["iso"] == "langmuir") cs.reset(new CoupledStep<D3Q19, D3Q7>(ns, ad, q, lang, reg.mask));
        else cs.reset(new CoupledStep<D3Q19, D3Q7>(ns, ad, q, lin, reg.mask));
        wasFused = cs->fused();
        if (!wasFused) std::fprintf(stderr, "not fused: %s\n", cs->whyNotFused());
        if (kv["compact"] == "1") cs->useCompactSolids();
        const int split = std::atoi(kv["split"].c_str());
        if (split > 0 && split < steps) {  // two calls: the second grows the cached per-step inlet values
            cs->run(1, split, true);
            cs->run(split + 1, steps - split, true);
        } else {
            cs->run(1, steps, true);
        }
        cs->finish();
        c_out = cs->outletMeans();
    } else {
        std::vector<std::uint8_t> regionMask = reg.mask;
        std::uint8_t *d_region = static_cast<std::uint8_t *>(fmx::mallocBytes(regionMask.size()));
        fmx::memcpyHtoD(d_region, regionMask.data(), regionMask.size());
        const std::size_t nyz = std::size_t(p.Ny) * p.Nz;
        for (int s = 1; s <= steps; ++s) {  // main.cpp:319-336
            ns.step(s);
            ns.computeMacroscopicDevice();
            ad.setAdvectionFromFluidDevice(ns);
            ad.zeroSource();
            if (kv["iso"] == "langmuir") lang.computeSource(ad, q, d_region);
            else lin.computeSource(ad, q, d_region);
            ad.step(s);
            ad.downloadPhi();
            double sum = 0.0;
            for (std::size_t k = 0; k < nyz; ++k) sum += double(ad.phiHost()[k * std::size_t(p.Nx) + std::size_t(p.Nx - 1)]);
            c_out.push_back(sum / double(nyz));
        }
        fmx::freeBytes(d_region);
    }
    ns.downloadMacroscopic();
    ad.downloadPhi();
    q.download();
    dump(out_dir + "/rho.bin", ns.rhoHost());
    dump(out_dir + "/ux.bin", ns.uxHost());
    dump(out_dir + "/uy.bin", ns.uyHost());
    dump(out_dir + "/uz.bin", ns.uzHost());
    dump(out_dir + "/phi.bin", ad.phiHost());
    dump(out_dir + "/q.bin", q.host());
    dump(out_dir + "/c_out.bin", c_out);
    std::ofstream(out_dir + "/info.txt") << "fused=" << (wasFused ? 1 : 0) << "\nsizeof=" << sizeof(DATA_TYPE) << "\n";
    std::printf("done: %d steps, %s\n", steps, wasFused ? "fused" : "operators");
    return 0;
}
