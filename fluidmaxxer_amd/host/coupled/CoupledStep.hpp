#pragma once
// CoupledStep: one coupled step of the reference main loop (main.cpp:321-336)
//
//     ns.step(s);
//     if (coupling) { ns.computeMacroscopicDevice(); ad.setAdvectionFromFluidDevice(ns); }
//     if (adsorb)   { ad.zeroSource(); isotherm.computeSource(ad, q, region); }
//     ad.step(s);
//
// over the lattices' own device buffers.  When the registered operators and boundary
// conditions are ones the fused kernel implements -- BGK / porous / partial bounce-back NS
// collision, ADBGK collision, an x = 0 velocity inlet (NEQ or equilibrium) and AD inlet
// (Dirichlet, tanh pulse or knot series), an x = Nx-1 pressure outlet with the AD Neumann
// outlet, non-periodic x -- each step is one kernel launch (fmx_coupled_run); otherwise the
// sequence above runs operator by operator.  Results are the same either way.
//
// Between fused steps the x = 0 plane of the NS populations holds the next step's inlet
// values; finish() writes back the reference state (call it before reading d_f_ptr()
// directly; step() in the operator path calls it itself).
#include <algorithm>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <vector>

#include "boundary/ad/ADDirichlet.hpp"
#include "boundary/ad/ADDirichletSeries.hpp"
#include "boundary/ad/ADDirichletTanhPulse.hpp"
#include "boundary/ad/ADNeumannOutlet.hpp"
#include "boundary/ns/PressureDirichlet.hpp"
#include "boundary/ns/VelocityDirichlet.hpp"
#include "boundary/ns/VelocityDirichletNEQ.hpp"
#include "collision/ADBGKCollision.hpp"
#include "collision/BGKCollision.hpp"
#include "collision/PartialBounceBackCollision.hpp"
#include "collision/PorousBGKCollision.hpp"
#include "core/DeviceHelpers.hpp"
#include "core/DeviceIndexList.hpp"
#include "core/lattice/ADLattice.hpp"
#include "core/lattice/NSLattice.hpp"
#include "reaction/AdsorbedField.hpp"
#include "reaction/LangmuirLDF.hpp"
#include "reaction/LinearLDF.hpp"

template <typename NSDesc, typename ADDesc> class CoupledStep {
public:
    // no sorption
    CoupledStep(Lattice<NSDesc> &ns, ADLattice<ADDesc> &ad, AdsorbedField &q, bool coupling = true)
        : ns_(ns), ad_(ad), q_(q), coupling_(coupling) {
        init({});
    }
    // Henry isotherm on `region` (host mask, e.g. the obstacle mask as in main.cpp:333)
    CoupledStep(Lattice<NSDesc> &ns, ADLattice<ADDesc> &ad, AdsorbedField &q, LinearLDF<ADDesc> &iso,
                const std::vector<std::uint8_t> &region, bool coupling = true)
        : ns_(ns), ad_(ad), q_(q), linear_(&iso), coupling_(coupling) {
        init(region);
    }
    // Langmuir isotherm on `region`
    CoupledStep(Lattice<NSDesc> &ns, ADLattice<ADDesc> &ad, AdsorbedField &q, LangmuirLDF<ADDesc> &iso,
                const std::vector<std::uint8_t> &region, bool coupling = true)
        : ns_(ns), ad_(ad), q_(q), langmuir_(&iso), coupling_(coupling) {
        init(region);
    }
    ~CoupledStep() {
        fmx::freeBytes(d_flags_);
        fmx::freeBytes(d_links_);
        fmx::freeBytes(d_inlet_raw_);
        fmx::freeBytes(d_phi_alt_);
        fmx::freeBytes(d_series_);
        fmx::freeBytes(d_acc_);
        fmx::freeBytes(d_mean_);
        fmx::freeBytes(d_ring_);
        fmx::freeBytes(d_compact_);
    }
    CoupledStep(const CoupledStep &) = delete;
    CoupledStep &operator=(const CoupledStep &) = delete;

    bool fused() const { return fused_; }
    const char *whyNotFused() const { return why_; }

    // No counterpart in the reference (whose BGK kernel updates every solid cell).  Fused BGK runs of
    // 16 steps or more leave the solid cells at rest out of the NS traffic: they are frozen at the
    // prologue (their state only alternates) and the other cells' NS populations are stepped in
    // row-compacted buffers, with the same results.  Takes effect from the next run().
    void useCompactSolids(bool on = true) {
        finish();
        compact_ = on && desc_.ns_op == FMX_NS_BGK;
        desc_.freeze_solids = compact_ ? 1 : 0;
        primed_ = false;  // the next run marks (or clears) the frozen cells
    }

    // one step
    void step(int s) { run(s, 1); }

    // steps s0 .. s0+n-1; with logOutlet the outlet means land in outletMeans()
    void run(int s0, int n, bool logOutlet = false) {
        if (n <= 0) return;
        if (!fused_) {
            for (int k = 0; k < n; ++k) {
                operatorStep(s0 + k);
                if (logOutlet) outlet_.push_back(hostOutletMean());
            }
            return;
        }
        if (series_) ensureSeries(s0 + n);  // before the prologue: a per-step inlet needs its series buffer
        if (!primed_) {
            fmx::check(fmx_coupled_prologue(&desc_, bufsFor(nullptr), 0, nullptr), "fmx_coupled_prologue");
            primed_ = true;
            if (compact_) sizeCompact();
        }
        double *acc = nullptr;
        if (logOutlet) acc = accFor(n);
        fmx::check(fmx_coupled_run(&desc_, bufsFor(acc), s0, n, 0, nullptr), "fmx_coupled_run");
        if (n & 1) {
            ns_.swapPopulations();
            ad_.swapPopulations();
            d_phi_alt_ = ad_.exchangePhi(d_phi_alt_);
        }
        if (logOutlet) {
            fmx::check(fmx_outlet_mean(d_acc_, n, std::int64_t(nyz_), d_mean_, nullptr), "fmx_outlet_mean");
            std::vector<double> h(static_cast<std::size_t>(n));
            fmx::memcpyDtoH(h.data(), d_mean_, sizeof(double) * std::size_t(n));
            outlet_.insert(outlet_.end(), h.begin(), h.end());
        }
    }

    // restore the reference state of the x = 0 plane in the NS populations
    void finish() {
        if (!fused_ || !primed_ || !desc_.ns_inlet) return;
        fmx::check(fmx_coupled_view_f(&desc_, bufsFor(nullptr), 0, ns_.d_f_new_ptr(), nullptr), "fmx_coupled_view_f");
        ns_.swapPopulations();
        primed_ = false;
    }

    const std::vector<double> &outletMeans() const { return outlet_; }

private:
    void init(const std::vector<std::uint8_t> &region) {
        const LBMParams &p = ns_.P();
        region_ = region;
        nyz_ = std::size_t(p.Ny) * p.Nz;
        if (!region_.empty() && std::int64_t(region_.size()) != ns_.size64())
            throw std::runtime_error("CoupledStep: region mask size mismatch");
        if (!region_.empty()) d_region_.reset(new RegionBuf(region_));
        analyse();
        if (!fused_) return;
        // flag byte: obstacle bit (BGK bounce-back) + sorption region bit
        const std::vector<std::uint8_t> &obst = ns_.obstacleMaskHost();
        std::vector<std::uint8_t> flags(obst.size(), 0);
        for (std::size_t i = 0; i < flags.size(); ++i)
            flags[i] = std::uint8_t((obst[i] ? FMX_FLAG_SOLID : 0) |
                                    (!region_.empty() && region_[i] ? FMX_FLAG_REGION : 0));
        d_flags_ = static_cast<std::uint8_t *>(fmx::mallocBytes(flags.size()));
        fmx::memcpyHtoD(d_flags_, flags.data(), flags.size());
        d_links_ = static_cast<std::uint32_t *>(fmx::mallocBytes(flags.size() * sizeof(std::uint32_t)));
        fmx::check(fmx_coupled_build_links(&desc_, d_flags_, d_links_, nullptr), "fmx_coupled_build_links");
        if (desc_.ns_inlet)
            d_inlet_raw_ = static_cast<DATA_TYPE *>(fmx::mallocBytes(sizeof(DATA_TYPE) * 19 * nyz_));
        d_phi_alt_ = static_cast<DATA_TYPE *>(fmx::mallocBytes(sizeof(DATA_TYPE) * std::size_t(ns_.size64())));
    }

    void analyse() {
        const LBMParams &p = ns_.P();
        fused_ = false;
        desc_ = fmx_coupled_desc{};
        desc_.nx = p.Nx;
        desc_.ny = p.Ny;
        desc_.nz = p.Nz;
        desc_.periodic_y = p.periodicY;
        desc_.periodic_z = p.periodicZ;
        desc_.prec = fmx::kPrec;
        desc_.rho0 = double(p.rho0);
        desc_.tau = double(p.tau);
        desc_.min_conc = double(p.minConcentration);
        desc_.coupling = coupling_;
        desc_.rho_out = 1.0;
        if (p.periodicX) return fail("periodic x");
        if (p.Nx < 3) return fail("Nx < 3");
        auto *col = ns_.collisionOperator().get();
        if (dynamic_cast<BGKCollision<NSDesc> *>(col)) desc_.ns_op = FMX_NS_BGK;
        else if (dynamic_cast<PorousBGKCollision<NSDesc> *>(col)) desc_.ns_op = FMX_NS_POROUS;
        else if (dynamic_cast<PartialBounceBackCollision<NSDesc> *>(col)) desc_.ns_op = FMX_NS_PBB;
        else return fail("NS collision operator not fusable");
        for (auto &bc : ns_.boundaryList()) {
            auto *b = bc.get();
            if (auto *v = dynamic_cast<VelocityDirichletNEQ<NSDesc> *>(b)) {
                if (desc_.ns_inlet || !fmx::isFullXPlane(v->mask(), p.Nx, p.Ny, p.Nz, 0)) return fail("NS inlet mask");
                desc_.ns_inlet = 1;
                setInletVelocity(v->ux(), v->uy(), v->uz());
            } else if (auto *e = dynamic_cast<VelocityDirichletEq<NSDesc> *>(b)) {
                if (desc_.ns_inlet || !fmx::isFullXPlane(e->mask(), p.Nx, p.Ny, p.Nz, 0)) return fail("NS inlet mask");
                desc_.ns_inlet = 2;
                setInletVelocity(e->ux(), e->uy(), e->uz());
            } else if (auto *o = dynamic_cast<PressureDirichlet<NSDesc> *>(b)) {
                if (desc_.ns_outlet || !fmx::isFullXPlane(o->mask(), p.Nx, p.Ny, p.Nz, p.Nx - 1))
                    return fail("NS outlet mask");
                desc_.ns_outlet = 1;
                desc_.rho_out = double(o->rhoOut());
            } else {
                return fail("NS boundary condition not fusable");
            }
        }
        for (auto &bc : ad_.boundaryList()) {
            auto *b = bc.get();
            if (auto *d = dynamic_cast<ADDirichlet<ADDesc> *>(b)) {
                if (desc_.ad_inlet || !fmx::isFullXPlane(d->mask(), p.Nx, p.Ny, p.Nz, 0)) return fail("AD inlet mask");
                desc_.ad_inlet = 1;
                desc_.phi_in = double(d->value());
            } else if (auto *t = dynamic_cast<ADDirichletTanhPulse<ADDesc> *>(b)) {
                if (desc_.ad_inlet || !fmx::isFullXPlane(t->mask(), p.Nx, p.Ny, p.Nz, 0)) return fail("AD inlet mask");
                desc_.ad_inlet = 2;
                series_ = [t](int s) { return double(t->valueAt(s)); };
            } else if (auto *k = dynamic_cast<ADDirichletSeries<ADDesc> *>(b)) {
                if (desc_.ad_inlet || !fmx::isFullXPlane(k->mask(), p.Nx, p.Ny, p.Nz, 0)) return fail("AD inlet mask");
                desc_.ad_inlet = 2;
                series_ = [k](int s) { return double(k->valueAt(s)); };
            } else if (auto *nm = dynamic_cast<ADNeumannOutlet<ADDesc> *>(b)) {
                if (desc_.ad_outlet || !fmx::isFullXPlane(nm->mask(), p.Nx, p.Ny, p.Nz, p.Nx - 1))
                    return fail("AD outlet mask");
                desc_.ad_outlet = 1;
            } else {
                return fail("AD boundary condition not fusable");
            }
        }
        if (desc_.ns_outlet && !desc_.ad_outlet) return fail("pressure outlet without the AD Neumann outlet");
        // uniform AD relaxation time -> scalar
        const std::vector<DATA_TYPE> &tau = ad_.tauFieldHost();
        bool uniform = !tau.empty();
        for (std::size_t i = 1; i < tau.size() && uniform; ++i) uniform = tau[i] == tau[0];
        desc_.tau_ad = uniform ? double(tau[0]) : 0.0;
        tau_field_ = uniform ? nullptr : ad_.d_tau_ptr();
        if (linear_) {
            desc_.isotherm = FMX_ISO_LINEAR;
            desc_.iso_a = double(linear_->P().KHenry);
            desc_.k_ldf = double(linear_->P().kA);
        } else if (langmuir_) {
            desc_.isotherm = FMX_ISO_LANGMUIR;
            desc_.iso_a = double(langmuir_->P().qMax);
            desc_.iso_b = double(langmuir_->P().bLangmuir);
            desc_.k_ldf = double(langmuir_->P().kLDF);
        }
        fused_ = true;
        why_ = "";
    }

    void setInletVelocity(DATA_TYPE ux, DATA_TYPE uy, DATA_TYPE uz) {
        desc_.u_in[0] = double(ux);
        desc_.u_in[1] = double(uy);
        desc_.u_in[2] = double(uz);
    }
    void fail(const char *why) {
        fused_ = false;
        why_ = why;
    }

    const fmx_coupled_bufs *bufsFor(double *acc) {
        bufs_ = fmx_coupled_bufs{};
        bufs_.f[0] = ns_.d_f_ptr();
        bufs_.f[1] = ns_.d_f_new_ptr();
        bufs_.g[0] = ad_.d_g_ptr();
        bufs_.g[1] = ad_.d_g_new_ptr();
        bufs_.q = q_.d_q();
        bufs_.phi[0] = ad_.d_phi_ptr();
        bufs_.phi[1] = d_phi_alt_;
        bufs_.links = d_links_;
        bufs_.field = desc_.ns_op == FMX_NS_BGK ? nullptr : ns_.d_porosity_ptr();
        bufs_.tau_field = tau_field_;
        bufs_.inlet_raw = d_inlet_raw_;
        bufs_.inlet_series = d_series_;
        bufs_.inlet_series_len = d_series_ ? series_cap_ : 0;
        bufs_.outlet_acc = acc;
        bufs_.outlet_acc_rows = acc ? acc_cap_ : 0;
        if (compact_ && d_compact_) {
            bufs_.compact_work = d_compact_;
            bufs_.compact_work_bytes = compact_bytes_;
        }
        if (acc) {  // breakthrough sums in a fixed order (the same curve on every run)
            if (!d_ring_)
                d_ring_ = static_cast<double *>(fmx::mallocBytes(sizeof(double) * FMX_OUTLET_RING_ROWS * nyz_));
            bufs_.outlet_ring = d_ring_;
        }
        return &bufs_;
    }

    // workspace of the row-compacted runs, sized for the cells the prologue left unfrozen
    void sizeCompact() {
        const std::size_t n = std::size_t(desc_.nx) * std::size_t(desc_.ny) * std::size_t(desc_.nz);
        std::vector<std::uint32_t> links(n);
        fmx::memcpyDtoH(links.data(), d_links_, sizeof(std::uint32_t) * n);
        std::int64_t active = 0;
        for (std::uint32_t w : links) active += (w & FMX_LINK_FROZEN) ? 0 : 1;
        const std::int64_t bytes = fmx_coupled_compact_bytes(&desc_, active);
        if (bytes < 0) throw std::runtime_error("FMX error: fmx_coupled_compact_bytes refused the configuration");
        if (bytes > compact_bytes_) {
            fmx::freeBytes(d_compact_);
            d_compact_ = fmx::mallocBytes(std::size_t(bytes));
            compact_bytes_ = bytes;
        }
    }

    void ensureSeries(int last_step) {
        if (last_step < series_cap_) return;
        const int cap = std::max(last_step + 1, 2 * series_cap_ + 1024);
        std::vector<double> h(static_cast<std::size_t>(cap));
        for (int s = 0; s < cap; ++s) h[std::size_t(s)] = series_(s);
        fmx::freeBytes(d_series_);
        d_series_ = static_cast<double *>(fmx::mallocBytes(sizeof(double) * h.size()));
        fmx::memcpyHtoD(d_series_, h.data(), sizeof(double) * h.size());
        series_cap_ = cap;
    }

    double *accFor(int n) {
        const std::size_t rows = std::size_t(n) * FMX_OUTLET_SLOTS * FMX_OUTLET_STRIDE;
        if (n > acc_cap_) {
            fmx::freeBytes(d_acc_);
            fmx::freeBytes(d_mean_);
            d_acc_ = static_cast<double *>(fmx::mallocBytes(sizeof(double) * rows));
            d_mean_ = static_cast<double *>(fmx::mallocBytes(sizeof(double) * std::size_t(n)));
            acc_cap_ = n;
        }
        fmx::check(fmx_fill_zero(d_acc_, sizeof(double) * rows, nullptr), "fmx_fill_zero");
        return d_acc_;
    }

    void operatorStep(int s) {
        finish();
        ns_.step(s);
        if (coupling_) {
            ns_.computeMacroscopicDevice();
            ad_.setAdvectionFromFluidDevice(ns_);
        }
        if (linear_ || langmuir_) {
            ad_.zeroSource();
            const std::uint8_t *reg = d_region_ ? d_region_->ptr : nullptr;
            if (linear_) linear_->computeSource(ad_, q_, reg);
            else langmuir_->computeSource(ad_, q_, reg);
        }
        ad_.step(s);
    }

    double hostOutletMean() {
        ad_.downloadPhi();
        const auto &phi = ad_.phiHost();
        const int nx = ns_.P().Nx;
        double sum = 0.0;
        for (std::size_t k = 0; k < nyz_; ++k) sum += double(phi[k * std::size_t(nx) + std::size_t(nx - 1)]);
        return sum / double(nyz_);
    }

    struct RegionBuf {
        std::uint8_t *ptr = nullptr;
        explicit RegionBuf(const std::vector<std::uint8_t> &h) {
            ptr = static_cast<std::uint8_t *>(fmx::mallocBytes(h.size()));
            fmx::memcpyHtoD(ptr, h.data(), h.size());
        }
        ~RegionBuf() { fmx::freeBytes(ptr); }
    };

    Lattice<NSDesc> &ns_;
    ADLattice<ADDesc> &ad_;
    AdsorbedField &q_;
    LinearLDF<ADDesc> *linear_ = nullptr;
    LangmuirLDF<ADDesc> *langmuir_ = nullptr;
    bool coupling_ = true;
    std::vector<std::uint8_t> region_;
    std::unique_ptr<RegionBuf> d_region_;
    std::size_t nyz_ = 0;
    bool fused_ = false, primed_ = false, compact_ = false;
    void *d_compact_ = nullptr;
    std::int64_t compact_bytes_ = 0;
    const char *why_ = "not analysed";
    fmx_coupled_desc desc_{};
    fmx_coupled_bufs bufs_{};
    std::uint8_t *d_flags_ = nullptr;
    std::uint32_t *d_links_ = nullptr;
    DATA_TYPE *d_inlet_raw_ = nullptr;
    DATA_TYPE *d_phi_alt_ = nullptr;
    const DATA_TYPE *tau_field_ = nullptr;
    std::function<double(int)> series_;
    double *d_series_ = nullptr;
    int series_cap_ = 0;
    double *d_acc_ = nullptr, *d_mean_ = nullptr, *d_ring_ = nullptr;
    int acc_cap_ = 0;
    std::vector<double> outlet_;
};
