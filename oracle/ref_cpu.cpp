// ref_cpu.cpp -- C API over the CPU oracle (ref_cpu.hpp).
//
// TEST INFRASTRUCTURE ONLY (see ref_cpu.hpp).  The coupled "system" below strings
// the restated operators together in the reference's step order:
//   main.cpp:319-336   NS.step -> [moments + velocity copy] -> [zeroSource + LDF] -> AD.step
//   NSLattice.hpp:188-209  PreCollision BCs -> (D2D copy) -> collideStream -> PostStreaming BCs -> swap
//   ADLattice.hpp:144-162  PreCollision BCs -> collideStream -> PostStreaming BCs -> swap
// The D2D copy f -> f_new (NSLattice.hpp:198) is omitted: every collide writes all cells.
//
// orc_set_step_mutation(m != 0) plants one ordering mistake in the step sequence, so the
// sequence pins of the test suite can show that they fail on it (mutation testing of the
// checks; m = 0, the default, is the reference order):
//   1  the LDF reads the current populations' concentration instead of the previous step's phi
//   2  the AD lattice is advected by the moments of the populations before NS.step
//   3  the AD lattice is advected by the moments taken before the PostStreaming BCs
//   4  the NEQ velocity inlet is applied after streaming (on f_new) instead of before collision
//   5  the pressure outlet is applied before collision (on f) instead of after streaming
#ifdef _OPENMP
#include <omp.h>
#endif
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "ref_cpu.hpp"

extern "C" {

typedef struct {
    int kind;                 // NS: 1 VelocityDirichletNEQ, 2 VelocityDirichletEq, 3 PressureDirichlet, 4 BounceBack
                              // AD: 1 ADDirichlet, 2 ADDirichletTanhPulse, 3 ADNeumannOutlet, 4 inlet series
    const std::uint8_t *mask; // N bytes
    double v[4];              // velocity: ux,uy,uz | pressure: rho_out | dirichlet: phi | tanh: base,amp,k
    int i0, i1;               // tanh: t_on, t_off
    const double *knots;      // series knots
    int n_knots;
    double dt_knot;
} orc_bc;

typedef struct {
    int Nx, Ny, Nz;
    int px, py, pz;
    int ns_op;                     // 0 BGKCollision, 1 PorousBGKCollision, 2 partial bounce-back
    double rho0, tau;
    const std::uint8_t *obstacle;  // nullable
    const double *porosity;        // nullable; porous: porosity; pbb: solid fraction n_s
    double phi0, tau_ad, minC;
    const double *tau_field;       // nullable: tau_ad everywhere
    int n_ns_bc;
    const orc_bc *ns_bc;
    int n_ad_bc;
    const orc_bc *ad_bc;
    int coupling, adsorb, isotherm; // isotherm: 1 linear (Henry), 2 Langmuir
    const std::uint8_t *region;
    double KHenry, kA, qMax, bLangmuir, kLDF;
} orc_config;

}  // extern "C"

namespace {

using orc::i64;

struct BCHost {
    orc_bc d;
    std::vector<std::uint8_t> mask;
    std::vector<i64> bnd, nei;      // NEQ lists
    std::vector<double> knots;
};

struct SysBase {
    virtual ~SysBase() = default;
    virtual void step(int s) = 0;
    virtual void get(int field, double *out) = 0;
    virtual void set(int field, const double *in) = 0;
    virtual void outlet(const std::uint8_t *mask, double &avg, double &mn, double &mx) = 0;
    virtual i64 n() const = 0;
    virtual void set_mutation(int m) = 0;
};

template <class T> struct System : SysBase {
    orc::Grid g;
    orc_config c;
    std::vector<T> f, fn, gg, gn, phi, S, q, uxa, uya, uza, tau, por, rho_m, ux_m, uy_m, uz_m;
    std::vector<T> mux, muy, muz;  // moments taken by a planted ordering mistake
    int mutation = 0;
    std::vector<std::uint8_t> obst, region;
    std::vector<BCHost> nsbc, adbc;

    explicit System(const orc_config &cfg) : c(cfg) {
        if (c.Nx <= 0 || c.Ny <= 0 || c.Nz <= 0) throw std::runtime_error("NSLattice: invalid dimensions");
        g = orc::Grid{c.Nx, c.Ny, c.Nz, c.px != 0, c.py != 0, c.pz != 0};
        const i64 N = g.n();
        f.assign(N * 19, T(0));
        fn.assign(N * 19, T(0));
        gg.assign(N * 7, T(0));
        gn.assign(N * 7, T(0));
        phi.assign(N, T(0));   // ADLattice.hpp:49 memset
        S.assign(N, T(0));
        q.assign(N, T(0));     // AdsorbedField.hpp:19-20
        uxa.assign(N, T(0));
        uya.assign(N, T(0));
        uza.assign(N, T(0));
        tau.assign(N, T(c.tau_ad));
        if (c.tau_field)
            for (i64 i = 0; i < N; ++i) tau[i] = T(c.tau_field[i]);
        if (c.porosity) {
            por.resize(N);
            for (i64 i = 0; i < N; ++i) por[i] = T(c.porosity[i]);
        }
        if (c.obstacle) obst.assign(c.obstacle, c.obstacle + N);
        if (c.region) region.assign(c.region, c.region + N);
        for (int k = 0; k < c.n_ns_bc; ++k) {
            BCHost b;
            b.d = c.ns_bc[k];
            b.mask.assign(b.d.mask, b.d.mask + N);
            b.d.mask = nullptr;
            if (b.d.kind == 1) orc::neq_build_lists(b.mask.data(), g, b.bnd, b.nei);
            nsbc.push_back(std::move(b));
        }
        for (int k = 0; k < c.n_ad_bc; ++k) {
            BCHost b;
            b.d = c.ad_bc[k];
            b.mask.assign(b.d.mask, b.d.mask + N);
            b.d.mask = nullptr;
            if (b.d.knots) b.knots.assign(b.d.knots, b.d.knots + b.d.n_knots);
            b.d.knots = nullptr;
            adbc.push_back(std::move(b));
        }
        c.obstacle = nullptr;
        c.region = nullptr;
        c.porosity = nullptr;
        c.tau_field = nullptr;
        c.ns_bc = nullptr;
        c.ad_bc = nullptr;
        // initEquilibrium (NSLattice.hpp:177-185, ADLattice.hpp:134-141)
        if (c.ns_op == 0) orc::ns_init_eq_bgk(f.data(), g, T(c.rho0));
        else orc::ns_init_eq_porous(f.data(), g, T(c.rho0));
        fn = f;
        orc::ad_init_eq(gg.data(), g, T(c.phi0));
        gn = gg;
    }

    i64 n() const override { return g.n(); }
    void set_mutation(int m) override { mutation = m; }

    void ns_step(int s) {
        for (auto &b : nsbc) {  // PreCollision
            if (b.d.kind == 1 && !b.bnd.empty() && mutation != 4)
                orc::ns_bc_velocity_neq(f.data(), b.bnd, b.nei, T(b.d.v[0]), T(b.d.v[1]), T(b.d.v[2]));
            else if (b.d.kind == 2)
                orc::ns_bc_velocity_eq(f.data(), b.mask.data(), g, T(c.rho0), T(b.d.v[0]), T(b.d.v[1]), T(b.d.v[2]));
            else if (b.d.kind == 3 && mutation == 5)
                orc::ns_bc_pressure(f.data(), b.mask.data(), g, T(b.d.v[0]));
        }
        const T omega = orc::ns_omega(T(c.tau));
        if (c.ns_op == 0)
            orc::ns_collide_bgk(f.data(), fn.data(), obst.empty() ? nullptr : obst.data(), g, omega, T(c.rho0));
        else if (c.ns_op == 1)
            orc::ns_collide_porous(f.data(), fn.data(), por.empty() ? nullptr : por.data(), g, omega, T(c.rho0));
        else
            orc::ns_collide_pbb(f.data(), fn.data(), por.empty() ? nullptr : por.data(), g, omega, T(c.rho0));
        if (mutation == 3) moments_of(fn, mux, muy, muz);
        for (auto &b : nsbc) {  // PostStreaming
            if (b.d.kind == 3 && mutation != 5) orc::ns_bc_pressure(fn.data(), b.mask.data(), g, T(b.d.v[0]));
            else if (b.d.kind == 4) orc::ns_bc_bounceback(fn.data(), b.mask.data(), g);
            else if (b.d.kind == 1 && !b.bnd.empty() && mutation == 4)
                orc::ns_bc_velocity_neq(fn.data(), b.bnd, b.nei, T(b.d.v[0]), T(b.d.v[1]), T(b.d.v[2]));
        }
        std::swap(f, fn);
        (void)s;
    }

    void ad_step(int s) {
        for (auto &b : adbc) {  // PreCollision
            if (b.d.kind == 2) {
                const T v = orc::tanh_pulse_value<T>(s, T(b.d.v[0]), T(b.d.v[1]), b.d.i0, b.d.i1, T(b.d.v[2]));
                orc::ad_bc_dirichlet(gg.data(), b.mask.data(), g, v);
            } else if (b.d.kind == 4) {
                const T v = orc::series_value<T>(s, b.knots.data(), int(b.knots.size()), b.d.dt_knot);
                orc::ad_bc_dirichlet(gg.data(), b.mask.data(), g, v);
            }
        }
        orc::ad_collide(gg.data(), gn.data(), phi.data(), tau.data(), uxa.data(), uya.data(), uza.data(), S.data(),
                        g, T(c.minC));
        for (auto &b : adbc) {  // PostStreaming
            if (b.d.kind == 1) orc::ad_bc_dirichlet(gn.data(), b.mask.data(), g, T(b.d.v[0]));
            else if (b.d.kind == 3) orc::ad_bc_neumann(gn.data(), b.mask.data(), g);
        }
        std::swap(gg, gn);
    }

    void step(int s) override {
        if (mutation == 2) moments_of(f, mux, muy, muz);
        ns_step(s);
        if (c.coupling) {  // computeMacroscopicDevice + setAdvectionFromFluidDevice
            ensure_moments();
            uxa = ux_m;
            uya = uy_m;
            uza = uz_m;
            if (mutation == 2 || mutation == 3) {
                uxa = mux;
                uya = muy;
                uza = muz;
            }
        }
        if (c.adsorb) {  // zeroSource + computeSource(region)
            std::fill(S.begin(), S.end(), T(0));
            const T *phi_in = phi.data();
            std::vector<T> phi_now;
            if (mutation == 1) {  // the current populations' (clamped) concentration, no lag
                phi_now.assign(g.n(), T(0));
                for (i64 i = 0; i < g.n(); ++i) {
                    T sum = T(0);
                    for (int k = 0; k < 7; ++k) sum += gg[i * 7 + k];
                    phi_now[i] = sum < T(c.minC) ? T(c.minC) : sum;
                }
                phi_in = phi_now.data();
            }
            if (c.isotherm == 1)
                orc::ldf_linear(g.n(), region.data(), phi_in, S.data(), q.data(), T(c.KHenry), T(c.kA), T(c.minC));
            else if (c.isotherm == 2)
                orc::ldf_langmuir(g.n(), region.data(), phi_in, S.data(), q.data(), T(c.qMax), T(c.bLangmuir),
                                  T(c.kLDF), T(c.minC));
        }
        ad_step(s);
    }

    void moments_of(const std::vector<T> &pop, std::vector<T> &ux, std::vector<T> &uy, std::vector<T> &uz) {
        const i64 N = g.n();
        std::vector<T> rho(N);
        ux.resize(N);
        uy.resize(N);
        uz.resize(N);
        orc::ns_moments(pop.data(), rho.data(), ux.data(), uy.data(), uz.data(), g);
    }

    void ensure_moments() {
        const i64 N = g.n();
        rho_m.resize(N);
        ux_m.resize(N);
        uy_m.resize(N);
        uz_m.resize(N);
        orc::ns_moments(f.data(), rho_m.data(), ux_m.data(), uy_m.data(), uz_m.data(), g);
    }

    std::vector<T> *field(int id) {
        switch (id) {
            case 0: return &f;
            case 1: return &gg;
            case 2: return &phi;
            case 3: return &q;
            case 4: return &S;
            case 5: return &uxa;
            case 6: return &uya;
            case 7: return &uza;
            case 8: ensure_moments(); return &rho_m;
            case 9: ensure_moments(); return &ux_m;
            case 10: ensure_moments(); return &uy_m;
            case 11: ensure_moments(); return &uz_m;
            default: throw std::runtime_error("orc: unknown field id");
        }
    }
    void get(int id, double *out) override {
        auto *v = field(id);
        for (std::size_t i = 0; i < v->size(); ++i) out[i] = double((*v)[i]);
    }
    void set(int id, const double *in) override {
        if (id >= 8) throw std::runtime_error("orc: moment fields are read-only");
        auto *v = field(id);
        for (std::size_t i = 0; i < v->size(); ++i) (*v)[i] = T(in[i]);
    }
    void outlet(const std::uint8_t *mask, double &avg, double &mn, double &mx) override {
        orc::outlet_stats(phi.data(), mask, g.n(), avg, mn, mx);
    }
};

thread_local std::string g_err;

template <class F> int guarded(F &&fn) {
    try {
        fn();
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        return 1;
    }
}

orc::Grid make_grid(int Nx, int Ny, int Nz, int px, int py, int pz) {
    return orc::Grid{Nx, Ny, Nz, px != 0, py != 0, pz != 0};
}

template <class T> std::vector<T> to_t(const double *p, i64 n) {
    std::vector<T> v(p ? n : 0);
    for (i64 i = 0; p && i < n; ++i) v[i] = T(p[i]);
    return v;
}
template <class T> void from_t(const std::vector<T> &v, double *p) {
    for (std::size_t i = 0; i < v.size(); ++i) p[i] = double(v[i]);
}

}  // namespace

extern "C" {

const char *orc_last_error(void) { return g_err.c_str(); }

// Threads of the per-cell loops (n <= 0: the OpenMP default the process started with; 1 in a build
// without OpenMP); results do not depend on it.  Returns the count now in effect.
int orc_set_threads(int n) {
#ifdef _OPENMP
    static const int initial = omp_get_max_threads();
    omp_set_num_threads(n > 0 ? n : initial);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

void *orc_create(const orc_config *cfg, int prec) {
    SysBase *s = nullptr;
    int rc = guarded([&] {
        if (prec == 4) s = new System<float>(*cfg);
        else if (prec == 8) s = new System<double>(*cfg);
        else throw std::runtime_error("orc_create: prec must be 4 or 8");
    });
    return rc ? nullptr : s;
}

void orc_destroy(void *h) { delete static_cast<SysBase *>(h); }

// Plants ordering mistake m (1..5, listed at the top of this file) in the system's step sequence;
// 0 restores the reference order.  For the mutation tests of the sequence pins only.
int orc_set_step_mutation(void *h, int m) {
    return guarded([&] {
        if (m < 0 || m > 5) throw std::runtime_error("orc_set_step_mutation: unknown mutation");
        static_cast<SysBase *>(h)->set_mutation(m);
    });
}

// Runs steps s0 .. s0+nsteps-1.  When `outlet_mask` is non-null, the outlet
// statistics after every step go to avg/mn/mx (nullable each).  `seconds` (nullable)
// receives the wall time of the step loop alone.
int orc_run(void *h, int s0, int nsteps, const std::uint8_t *outlet_mask, double *avg, double *mn, double *mx,
            double *seconds) {
    return guarded([&] {
        auto *s = static_cast<SysBase *>(h);
        double t = 0.0;
        for (int k = 0; k < nsteps; ++k) {
            const auto t0 = std::chrono::steady_clock::now();
            s->step(s0 + k);
            t += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (outlet_mask) {
                double a, lo, hi;
                s->outlet(outlet_mask, a, lo, hi);
                if (avg) avg[k] = a;
                if (mn) mn[k] = lo;
                if (mx) mx[k] = hi;
            }
        }
        if (seconds) *seconds = t;
    });
}

int orc_get(void *h, int field, double *out) {
    return guarded([&] { static_cast<SysBase *>(h)->get(field, out); });
}
int orc_set(void *h, int field, const double *in) {
    return guarded([&] { static_cast<SysBase *>(h)->set(field, in); });
}

// ------------------------------------------------ single-operator entry points
// Inputs and outputs are float64 arrays in the reference AoS layout; for prec 4
// they are rounded to float on entry and widened on exit.

int orc_op_ns_collide(int prec, int ns_op, int Nx, int Ny, int Nz, int px, int py, int pz, const double *f,
                      double *f_new, const std::uint8_t *obst, const double *field, double tau, double rho0) {
    return guarded([&] {
        const auto g = make_grid(Nx, Ny, Nz, px, py, pz);
        auto run = [&](auto zero) {
            using T = decltype(zero);
            auto fi = to_t<T>(f, g.n() * 19);
            std::vector<T> fo(g.n() * 19);
            auto fl = to_t<T>(field, g.n());
            const T om = orc::ns_omega(T(tau));
            if (ns_op == 0) orc::ns_collide_bgk(fi.data(), fo.data(), obst, g, om, T(rho0));
            else if (ns_op == 1) orc::ns_collide_porous(fi.data(), fo.data(), field ? fl.data() : nullptr, g, om, T(rho0));
            else orc::ns_collide_pbb(fi.data(), fo.data(), field ? fl.data() : nullptr, g, om, T(rho0));
            from_t(fo, f_new);
        };
        if (prec == 4) run(0.0f);
        else run(0.0);
    });
}

int orc_op_ns_moments(int prec, int Nx, int Ny, int Nz, const double *f, double *rho, double *ux, double *uy,
                      double *uz) {
    return guarded([&] {
        const auto g = make_grid(Nx, Ny, Nz, 0, 0, 0);
        auto run = [&](auto zero) {
            using T = decltype(zero);
            auto fi = to_t<T>(f, g.n() * 19);
            std::vector<T> r(g.n()), a(g.n()), b(g.n()), c(g.n());
            orc::ns_moments(fi.data(), r.data(), a.data(), b.data(), c.data(), g);
            from_t(r, rho);
            from_t(a, ux);
            from_t(b, uy);
            from_t(c, uz);
        };
        if (prec == 4) run(0.0f);
        else run(0.0);
    });
}

// kind: 1 NEQ velocity (PreCollision), 2 Eq velocity (PreCollision; a = rho0),
// 3 pressure (a = rho_out), 4 bounce-back swap.  f is updated in place.
int orc_op_ns_bc(int prec, int kind, int Nx, int Ny, int Nz, double *f, const std::uint8_t *mask, double ux,
                 double uy, double uz, double a) {
    return guarded([&] {
        const auto g = make_grid(Nx, Ny, Nz, 0, 0, 0);
        auto run = [&](auto zero) {
            using T = decltype(zero);
            auto fi = to_t<T>(f, g.n() * 19);
            if (kind == 1) {
                std::vector<i64> bnd, nei;
                orc::neq_build_lists(mask, g, bnd, nei);
                orc::ns_bc_velocity_neq(fi.data(), bnd, nei, T(ux), T(uy), T(uz));
            } else if (kind == 2) {
                orc::ns_bc_velocity_eq(fi.data(), mask, g, T(a), T(ux), T(uy), T(uz));
            } else if (kind == 3) {
                orc::ns_bc_pressure(fi.data(), mask, g, T(a));
            } else if (kind == 4) {
                orc::ns_bc_bounceback(fi.data(), mask, g);
            } else {
                throw std::runtime_error("orc_op_ns_bc: unknown kind");
            }
            from_t(fi, f);
        };
        if (prec == 4) run(0.0f);
        else run(0.0);
    });
}

int orc_op_ad_collide(int prec, int Nx, int Ny, int Nz, int px, int py, int pz, const double *g_in, double *g_out,
                      double *phi, const double *tau, const double *ux, const double *uy, const double *uz,
                      const double *S, double minC) {
    return guarded([&] {
        const auto gr = make_grid(Nx, Ny, Nz, px, py, pz);
        const i64 N = gr.n();
        auto run = [&](auto zero) {
            using T = decltype(zero);
            auto gi = to_t<T>(g_in, N * 7);
            std::vector<T> go(N * 7), ph(N);
            auto t = to_t<T>(tau, N);
            auto a = to_t<T>(ux, N), b = to_t<T>(uy, N), c = to_t<T>(uz, N), s = to_t<T>(S, N);
            orc::ad_collide(gi.data(), go.data(), ph.data(), t.data(), a.data(), b.data(), c.data(),
                            S ? s.data() : nullptr, gr, T(minC));
            from_t(go, g_out);
            from_t(ph, phi);
        };
        if (prec == 4) run(0.0f);
        else run(0.0);
    });
}

// kind: 1 Dirichlet (g <- a * w on the mask), 3 Neumann outlet copy.  In place.
int orc_op_ad_bc(int prec, int kind, int Nx, int Ny, int Nz, double *g, const std::uint8_t *mask, double a) {
    return guarded([&] {
        const auto gr = make_grid(Nx, Ny, Nz, 0, 0, 0);
        auto run = [&](auto zero) {
            using T = decltype(zero);
            auto gi = to_t<T>(g, gr.n() * 7);
            if (kind == 1) orc::ad_bc_dirichlet(gi.data(), mask, gr, T(a));
            else if (kind == 3) orc::ad_bc_neumann(gi.data(), mask, gr);
            else throw std::runtime_error("orc_op_ad_bc: unknown kind");
            from_t(gi, g);
        };
        if (prec == 4) run(0.0f);
        else run(0.0);
    });
}

// isotherm 1: linear (a = K); 2: Langmuir (a = qMax, b = b).  q and S updated in place.
int orc_op_ldf(int prec, int isotherm, long long n, const std::uint8_t *region, const double *phi, double *S,
               double *q, double a, double b, double kLDF, double minC) {
    return guarded([&] {
        auto run = [&](auto zero) {
            using T = decltype(zero);
            auto ph = to_t<T>(phi, n), s = to_t<T>(S, n), qq = to_t<T>(q, n);
            if (isotherm == 1) orc::ldf_linear(n, region, ph.data(), s.data(), qq.data(), T(a), T(kLDF), T(minC));
            else orc::ldf_langmuir(n, region, ph.data(), s.data(), qq.data(), T(a), T(b), T(kLDF), T(minC));
            from_t(s, S);
            from_t(qq, q);
        };
        if (prec == 4) run(0.0f);
        else run(0.0);
    });
}

double orc_tanh_pulse(int prec, int step, double base, double amp, int t_on, int t_off, double k) {
    if (prec == 4) return double(orc::tanh_pulse_value<float>(step, float(base), float(amp), t_on, t_off, float(k)));
    return orc::tanh_pulse_value<double>(step, base, amp, t_on, t_off, k);
}

double orc_series_value(int prec, int step, const double *knots, int n_knots, double dt_knot) {
    if (prec == 4) return double(orc::series_value<float>(step, knots, n_knots, dt_knot));
    return orc::series_value<double>(step, knots, n_knots, dt_knot);
}

}  // extern "C"
