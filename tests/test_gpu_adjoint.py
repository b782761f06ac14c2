"""Discrete adjoint of the fused coupled step: gradients of an outlet breakthrough-curve
loss L = sum_s (C_out(s) - C*(s))^2 with respect to the Langmuir constants (q_max, K_L),
the per-cell partial-bounce-back fraction / porosity and the per-step inlet
concentration, checked against

  * central finite differences of the GPU forward run in FP64 (observed 1e-10 .. 1e-7 relative,
    each bound about 3x its observation; the FP64 fused step is bit-identical to the FP64 oracle),
  * central finite differences of the CPU oracle (the restatement of the reference
    update) in FP64 (observed 4e-10),
  * the FP64 adjoint, for the FP32 adjoint (6e-6 on the small pack, 2e-3 at C4 scale over 12000
    steps: the FP32 forward run's own drift from FP64 -- reverse steps taken in FP32 on the FP64
    run's states agree with the FP64 adjoint to 1e-7 .. 1e-5),
  * itself under a different checkpoint interval (relative 1e-12),
  * the FP32-checkpoint adjoint, for an FP32 run with 16-bit checkpoints and FP32 recompute,
  * the FP64 adjoint, for the FP32 reverse sweep of an FP64 run (FP32 checkpoints, FP32 recompute
    and adjoint steps: the forward curve is the FP64 run's, and the FP32 drift is confined to one
    segment),
and used through the torch.autograd.Function wrapper.
"""
import dataclasses

import numpy as np
import pytest

from fluidmaxxer_amd import workloads
from fluidmaxxer_amd.adjoint import AdjointRun, CoupledBreakthrough, segment_slots
from fluidmaxxer_amd.coupled import CoupledConfig
from fluidmaxxer_amd.geometry import sphere_pack
from oracle import ref_cpu as R
from tests.gpu_util import host, say
from tests.lattice import oracle_setup

pytestmark = pytest.mark.gpu

NSTEPS, K = 400, 64


def adj_config(op="pbb", series=True, closed=False, seed=21, dims=(16, 10, 10), variant=None):
    """A small sphere pack whose concentration front reaches the outlet within NSTEPS."""
    nx, ny, nz = dims
    pack = sphere_pack(nx, ny, nz, seed, 1.5, 2.5, 0.55, overlap=False, clearance=3)
    cfg = CoupledConfig(nx, ny, nz, solid=pack.solid, region=pack.sphere, tau=0.7, u_in=(0.08, 0.0, 0.0),
                        tau_ad=0.8, isotherm="langmuir", iso_a=2.0, iso_b=5.0, k_ldf=0.05, ns_inlet="eq")
    rng = np.random.default_rng(seed)
    if op == "pbb":
        cfg.ns_op = "pbb"
        cfg.field = 0.05 + 0.9 * pack.solid + rng.uniform(-0.03, 0.03, cfg.n)   # strictly inside (0, 1)
        cfg.solid = np.zeros_like(pack.solid)
    elif op == "porous":
        cfg.ns_op = "porous"
        cfg.field = np.where(pack.solid == 1, 0.6, 0.95) + rng.uniform(-0.03, 0.03, cfg.n)
    if series:
        cfg.ad_inlet = "series"
        cfg.inlet_series = 0.2 + 0.8 * workloads.series_per_step(workloads.series_knots(9, seed), NSTEPS)
    if variant == "no_ns_inlet":     # flow driven by a low outlet density; keep-local pulls at x = 0
        cfg.ns_inlet = None
        cfg.rho_out = 0.97
    elif variant == "no_outlets":    # keep-local pulls at x = nx-1 for both lattices
        cfg.ns_outlet = False
        cfg.ad_outlet = False
    elif variant == "neq_inlet":     # main.cpp's VelocityDirichletNEQ inlet
        cfg.ns_inlet = "neq"
    if closed:
        cfg.ns_inlet = None
        cfg.ns_outlet = False
        cfg.ad_outlet = False
        cfg.ad_inlet = "none"
        cfg.inlet_series = None
        cfg.phi0 = 0.4
    return cfg


class Problem:
    """Loss, adjoint gradients and finite differences for one configuration."""

    def __init__(self, cfg, prec="f64", nsteps=NSTEPS, k=K, target=None, ckpt=None, seg_slots=None):
        say(f"adjoint problem prec={prec} ckpt={ckpt or prec} op={cfg.ns_op} dims={(cfg.nx, cfg.ny, cfg.nz)} "
            f"steps={nsteps} k={k} seg_slots={seg_slots}")
        self.cfg = cfg
        self.run = AdjointRun(cfg, nsteps, k, prec, ckpt=ckpt, seg_slots=seg_slots)
        self.base = dict(iso_a=cfg.iso_a, iso_b=cfg.iso_b,
                         field=None if cfg.field is None else np.asarray(cfg.field, np.float64).copy(),
                         series=None if cfg.inlet_series is None else np.asarray(cfg.inlet_series, np.float64).copy())
        c = host(self.run.forward())
        assert np.all(np.isfinite(c))
        self.c0 = c
        self.target = target if target is not None else 0.85 * c + 0.01 * np.sin(np.arange(nsteps) / 17.0)

    def loss(self, **over):
        p = dict(self.base)
        p.update(over)
        self.run.set_params(**{k: v for k, v in p.items() if v is not None})
        c = host(self.run.forward())
        return float(np.sum((c - self.target) ** 2))

    def grads(self):
        self.run.set_params(**{k: v for k, v in self.base.items() if v is not None})
        c = host(self.run.forward())
        g = self.run.backward(2.0 * (c - self.target))
        return {k: (host(v) if hasattr(v, "detach") else v) for k, v in g.items()}

    def fd_scalar(self, name, h):
        v = self.base[name]
        return (self.loss(**{name: v + h}) - self.loss(**{name: v - h})) / (2 * h)

    def fd_dir(self, name, direction, h):
        v = self.base[name]
        return (self.loss(**{name: v + h * direction}) - self.loss(**{name: v - h * direction})) / (2 * h)


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.fixture(scope="module")
def pbb_problem():
    return Problem(adj_config("pbb"))


def test_front_reaches_outlet(pbb_problem):
    # the loss must see a real breakthrough curve, not the clamped floor
    assert pbb_problem.c0[-1] > 0.02, pbb_problem.c0[-1]


def test_iso_grads_match_fd_pbb(pbb_problem):
    g = pbb_problem.grads()
    for name in ("iso_a", "iso_b"):
        fd = pbb_problem.fd_scalar(name, 1e-5 * pbb_problem.base[name])
        say(f"  d L/d {name}: adjoint {g[name]:.12e} fd {fd:.12e} rel {rel(g[name], fd):.2e}")
        assert rel(g[name], fd) < 1.5e-9  # observed 4.3e-10


def test_field_grads_match_fd_pbb(pbb_problem):
    g = pbb_problem.grads()
    gf = g["field"]
    rng = np.random.default_rng(3)
    v = rng.standard_normal(pbb_problem.cfg.n)
    fd = pbb_problem.fd_dir("field", v, 1e-6)
    ad = float(gf @ v)
    say(f"  <dL/dfield, v>: adjoint {ad:.12e} fd {fd:.12e} rel {rel(ad, fd):.2e}")
    assert rel(ad, fd) < 3e-8  # observed 9.5e-9
    for c in np.argsort(-np.abs(gf))[:3]:
        e = np.zeros(pbb_problem.cfg.n)
        e[c] = 1.0
        fdc = pbb_problem.fd_dir("field", e, 1e-6)
        say(f"  d L/d field[{c}]: adjoint {gf[c]:.12e} fd {fdc:.12e} rel {rel(gf[c], fdc):.2e}")
        assert rel(gf[c], fdc) < 5e-9  # observed <= 1.6e-9


def test_series_grads_match_fd(pbb_problem):
    g = pbb_problem.grads()
    gs = g["series"]
    assert gs[0] == 0.0 and np.count_nonzero(gs[1:]) > NSTEPS // 2
    rng = np.random.default_rng(4)
    v = rng.standard_normal(gs.size)
    fd = pbb_problem.fd_dir("series", v, 1e-6)
    ad = float(gs @ v)
    say(f"  <dL/dseries, v>: adjoint {ad:.12e} fd {fd:.12e} rel {rel(ad, fd):.2e}")
    assert rel(ad, fd) < 1.5e-9  # observed 4.9e-10
    for s in np.argsort(-np.abs(gs))[:3]:
        e = np.zeros(gs.size)
        e[s] = 1.0
        fds = pbb_problem.fd_dir("series", e, 1e-5)
        say(f"  d L/d series[{s}]: adjoint {gs[s]:.12e} fd {fds:.12e} rel {rel(gs[s], fds):.2e}")
        assert rel(gs[s], fds) < 1.5e-9, (s, gs[s], fds)  # observed <= 4.6e-10


@pytest.mark.parametrize("op", ["bgk", "porous"])
def test_grads_match_fd_other_operators(op):
    # BGK: bounce-back solids and the ADDirichlet inlet; porous: the porosity gradient
    pr = Problem(adj_config(op, series=(op == "porous")))
    g = pr.grads()
    for name in ("iso_a", "iso_b"):
        fd = pr.fd_scalar(name, 1e-5 * pr.base[name])
        say(f"  {op} d L/d {name}: adjoint {g[name]:.12e} fd {fd:.12e} rel {rel(g[name], fd):.2e}")
        assert rel(g[name], fd) < 1.5e-9  # observed <= 3.7e-10
    if op == "porous":
        v = np.random.default_rng(5).standard_normal(pr.cfg.n)
        fd = pr.fd_dir("field", v, 1e-6)
        ad = float(g["field"] @ v)
        say(f"  porous <dL/dfield, v>: adjoint {ad:.12e} fd {fd:.12e} rel {rel(ad, fd):.2e}")
        assert rel(ad, fd) < 1e-9  # observed 2.5e-10


def test_grads_match_fd_closed_box():
    # no inlet/outlet conditions: the fluid stays at rest (so the field has no effect) and every
    # x = 0 / x = nx-1 slot of g has the keep-local second reader
    pr = Problem(adj_config("pbb", closed=True), target=None)
    pr.target = pr.c0 * 1.05
    g = pr.grads()
    for name in ("iso_a", "iso_b"):
        fd = pr.fd_scalar(name, 1e-5 * pr.base[name])
        say(f"  closed d L/d {name}: adjoint {g[name]:.12e} fd {fd:.12e} rel {rel(g[name], fd):.2e}")
        assert rel(g[name], fd) < 5e-10  # observed 1.3e-10
    assert np.max(np.abs(g["field"])) < 1e-12 * abs(g["iso_a"])


@pytest.mark.parametrize("variant", ["no_ns_inlet", "no_outlets"])
def test_grads_match_fd_open_ends(variant):
    # the x = 0 (no NS inlet) or x = nx-1 (no outlets) slots with two readers, with flow
    pr = Problem(adj_config("pbb", variant=variant))
    g = pr.grads()
    for name in ("iso_a", "iso_b"):
        fd = pr.fd_scalar(name, 1e-5 * pr.base[name])
        say(f"  {variant} d L/d {name}: adjoint {g[name]:.12e} fd {fd:.12e} rel {rel(g[name], fd):.2e}")
        assert rel(g[name], fd) < 1.5e-9  # observed <= 4.9e-10
    for name, seed in (("field", 6), ("series", 8)):
        v = np.random.default_rng(seed).standard_normal(g[name].size)
        fd = pr.fd_dir(name, v, 1e-6)
        ad = float(g[name] @ v)
        say(f"  {variant} <dL/d{name}, v>: adjoint {ad:.12e} fd {fd:.12e} rel {rel(ad, fd):.2e}")
        assert rel(ad, fd) < 3e-7  # observed <= 9.4e-8 (no NS inlet, field)


@pytest.mark.parametrize("op", ["pbb", "bgk"])
def test_grads_match_fd_neq_inlet(op):
    """main.cpp's VelocityDirichletNEQ inlet (VelocityDirichletNEQ.cu:81-95): the x = 0 populations are
    functions of the (1, y, z) cell's post-collision state, so their adjoint flows back into it
    (through f, its density and its velocity).  FP64 adjoint against FP64 central differences, for the
    isotherm constants and random directions in the field and the inlet series; and the gradients are
    not those of the equilibrium inlet (the NEQ terms are live)."""
    pr = Problem(adj_config(op, variant="neq_inlet"))
    g = pr.grads()
    for name in ("iso_a", "iso_b"):
        fd = pr.fd_scalar(name, 1e-5 * pr.base[name])
        say(f"  neq {op} d L/d {name}: adjoint {g[name]:.12e} fd {fd:.12e} rel {rel(g[name], fd):.2e}")
        assert rel(g[name], fd) < 5e-10  # observed <= 1.1e-10
    for name, seed in (("field", 6), ("series", 8)):
        if g[name] is None:
            continue
        v = np.random.default_rng(seed).standard_normal(g[name].size)
        fd = pr.fd_dir(name, v, 1e-6)
        ad = float(g[name] @ v)
        say(f"  neq {op} <dL/d{name}, v>: adjoint {ad:.12e} fd {fd:.12e} rel {rel(ad, fd):.2e}")
        assert rel(ad, fd) < 3e-8  # observed <= 8.4e-9
    eq = Problem(adj_config(op), target=pr.target).grads()
    assert rel(eq["iso_a"], g["iso_a"]) > 1e-4


def test_grads_match_oracle_fd(pbb_problem):
    """Central differences of the CPU oracle (FP64 restatement of the reference update)."""
    cfg = pbb_problem.cfg
    outlet = (np.arange(cfg.n) % cfg.nx == cfg.nx - 1).astype(np.uint8)
    target = pbb_problem.target

    def oracle_loss(**over):
        c = dataclasses.replace(cfg, **over)
        sysm = R.System(oracle_setup(c, R), 8)
        cr, _, _, _ = sysm.run(NSTEPS, outlet)
        sysm.close()
        return float(np.sum((np.asarray(cr) - target) ** 2))

    g = pbb_problem.grads()
    for name, attr in (("iso_a", "iso_a"), ("iso_b", "iso_b")):
        v0 = getattr(cfg, attr)
        h = 1e-5 * v0
        fd = (oracle_loss(**{attr: v0 + h}) - oracle_loss(**{attr: v0 - h})) / (2 * h)
        say(f"  oracle fd d L/d {name}: adjoint {g[name]:.12e} oracle {fd:.12e} rel {rel(g[name], fd):.2e}")
        assert rel(g[name], fd) < 1.5e-9  # observed 4.2e-10
    c = int(np.argmax(np.abs(g["field"])))
    f0 = np.asarray(cfg.field, np.float64)
    fp, fm = f0.copy(), f0.copy()
    fp[c] += 1e-6
    fm[c] -= 1e-6
    fd = (oracle_loss(field=fp) - oracle_loss(field=fm)) / 2e-6
    say(f"  oracle fd d L/d field[{c}]: adjoint {g['field'][c]:.12e} oracle {fd:.12e}")
    assert rel(g["field"][c], fd) < 6e-9  # observed 2.0e-9


def test_fp32_adjoint_matches_fp64(pbb_problem):
    g64 = pbb_problem.grads()
    p32 = Problem(adj_config("pbb"), prec="f32", target=pbb_problem.target)
    g32 = p32.grads()
    for name in ("iso_a", "iso_b"):
        say(f"  f32 vs f64 d L/d {name}: {g32[name]:.8e} vs {g64[name]:.8e} rel {rel(g32[name], g64[name]):.2e}")
        assert rel(g32[name], g64[name]) < 1.5e-5  # observed 6.2e-6
    for name in ("field", "series"):
        e = np.linalg.norm(g32[name] - g64[name]) / np.linalg.norm(g64[name])
        say(f"  f32 vs f64 {name}: relative l2 {e:.2e}")
        assert e < 2e-5  # observed 8.8e-6 (field), 6.2e-6 (series)


def test_fp32_sweep_of_fp64_run(pbb_problem):
    """An FP64 run with FP32 checkpoints: its forward curve is the FP64 run's to the bit, and its
    gradients (FP32 recompute from the rounded checkpoints, FP32 adjoint steps) agree with the FP64
    adjoint far inside the 1e-4 of an FP64 finite difference."""
    g64 = pbb_problem.grads()
    pm = Problem(adj_config("pbb"), prec="f64", ckpt="f32", target=pbb_problem.target)
    assert np.array_equal(pm.c0, pbb_problem.c0)
    gm = pm.grads()
    for name in ("iso_a", "iso_b"):
        say(f"  f64 run, f32 sweep vs f64 d L/d {name}: rel {rel(gm[name], g64[name]):.2e}")
        assert rel(gm[name], g64[name]) < 7e-7, name  # observed 2.7e-7, 3.2e-7
    for name, bound in (("field", 6e-6), ("series", 1.5e-6)):  # observed 2.6e-6, 6.4e-7
        e = rel_l2(gm[name], g64[name])
        say(f"  f64 run, f32 sweep vs f64 {name}: relative l2 {e:.2e}")
        assert e < bound, name


def test_fp32_checkpoint_of_fp64_run_round_trip(pbb_problem):
    """fmx_checkpoint_save of an FP64 state at FP32 rounds each value to nearest; fmx_checkpoint_load
    widens it back exactly."""
    import torch

    from fluidmaxxer_amd import abi
    run = AdjointRun(adj_config("pbb"), 50, 50, "f64", ckpt="f32")
    run.forward()
    sim = run.sim
    cur = sim.cur
    before = [t.clone() for t in (sim.f[cur], sim.g[cur], sim.q, sim.phis[cur])]
    abi.call("fmx_checkpoint_save", sim._desc_p(), sim._bufs_p(), cur, run.ckpt.data_ptr(), abi.F32, 0, sim._s())
    for t in (sim.f[cur], sim.g[cur], sim.q, sim.phis[cur]):
        t.fill_(-1.0)
    abi.call("fmx_checkpoint_load", sim._desc_p(), sim._bufs_p(), cur, run.ckpt.data_ptr(), abi.F32, 0, sim._s())
    after = (sim.f[cur], sim.g[cur], sim.q, sim.phis[cur])
    for name, a, b in zip("fgq" + "p", before, after):
        assert torch.equal(b, a.to(torch.float32).to(torch.float64)), name


@pytest.mark.parametrize("k", [1, 7, NSTEPS])
def test_checkpoint_interval_invariance(pbb_problem, k):
    g = pbb_problem.grads()
    pr = Problem(adj_config("pbb"), k=k, target=pbb_problem.target)
    gk = pr.grads()
    for name in ("iso_a", "iso_b"):
        assert rel(gk[name], g[name]) < 1e-12, (k, name)
    for name in ("field", "series"):
        assert np.max(np.abs(gk[name] - g[name])) <= 1e-12 * np.max(np.abs(g[name])), (k, name)


def test_autograd_function(pbb_problem):
    import torch

    cfg = pbb_problem.cfg
    run = AdjointRun(cfg, NSTEPS, K, "f64")
    qmax = torch.tensor(cfg.iso_a, dtype=torch.float64, requires_grad=True)
    kl = torch.tensor(cfg.iso_b, dtype=torch.float64, requires_grad=True)
    field = torch.tensor(np.asarray(cfg.field), dtype=torch.float64, device="cuda", requires_grad=True)
    series = torch.tensor(np.asarray(cfg.inlet_series), dtype=torch.float64, device="cuda", requires_grad=True)
    target = torch.tensor(pbb_problem.target, device="cuda")
    c = CoupledBreakthrough.apply(run, qmax, kl, field, series)
    loss = ((c - target) ** 2).sum()
    loss.backward()
    g = pbb_problem.grads()
    assert rel(qmax.grad.item(), g["iso_a"]) < 1e-12
    assert rel(kl.grad.item(), g["iso_b"]) < 1e-12
    assert np.allclose(host(field.grad), g["field"], rtol=1e-12, atol=1e-300)
    assert np.allclose(host(series.grad), g["series"], rtol=1e-12, atol=1e-300)
    assert abs(loss.item() - pbb_problem.loss()) <= 1e-12 * loss.item()


def test_autograd_rejects_a_stale_run(pbb_problem):
    """Two nodes of one graph on the same AdjointRun: the second forward overwrites the first node's
    checkpoints and parameters, so the first node's backward must raise instead of returning the
    second run's gradients; the second node's own backward still works."""
    import torch

    from fluidmaxxer_amd import abi

    cfg = pbb_problem.cfg
    run = AdjointRun(cfg, 64, 16, "f64")
    kl = torch.tensor(cfg.iso_b, dtype=torch.float64)
    qa = torch.tensor(cfg.iso_a, dtype=torch.float64, requires_grad=True)
    qb = torch.tensor(cfg.iso_a * 1.1, dtype=torch.float64, requires_grad=True)
    c1 = CoupledBreakthrough.apply(run, qa, kl, None, None)
    c2 = CoupledBreakthrough.apply(run, qb, kl, None, None)
    c2.sum().backward()
    assert qb.grad is not None and torch.isfinite(qb.grad)
    with pytest.raises(abi.FmxError, match="belongs to forward run"):
        c1.sum().backward()


def c4_config(steps):
    """C4 geometry (128x64x64 overlapping sphere pack, partial bounce-back, Langmuir) with two
    departures that give the loss something smooth to differentiate:

    * C4's +-0.05 solid-fraction noise on the fluid cells acts as a uniform drag (mean u_x ~ 0.0015
      against u_in = 0.03), transport is diffusive and C_out stays at the clamp floor (~1e-8) for
      16000 steps; the fluid cells are set to 0 here (spheres and walls keep theirs), and the front
      reaches the outlet after about 6000 steps;
    * the minimum-concentration clamp (1e-8) is set to 0: at the front tip and in depleted sorbing
      cells the clamp's kink makes central differences drift with the step (h vs 4h differed by
      3e-3), so they would not be a reference there."""
    cfg = workloads.build("C4", steps=steps)
    cfg.field = np.where(cfg.field > 0.5, cfg.field, 0.0)
    cfg.min_conc = 0.0
    return cfg


@pytest.mark.parametrize("prec,ckpt,slots", [("f64", None, 14), ("f64", None, 20), ("f32", "f16", 15),
                                             ("f64", "f32", 14)])
def test_two_level_segments_bit_identical(pbb_problem, prec, ckpt, slots):
    """A segment store of about 2 sqrt(k) slots (sub-checkpoints within each segment, one more forward
    pass) gives the one-level sweep's gradients to the bit: the recomputed states are the same and
    every sum over the grid runs in a fixed order."""
    one = Problem(adj_config("pbb"), prec=prec, target=pbb_problem.target, ckpt=ckpt)
    g1 = one.grads()
    two = Problem(adj_config("pbb"), prec=prec, target=pbb_problem.target, ckpt=ckpt)
    two.run = AdjointRun(two.cfg, NSTEPS, K, prec, ckpt=ckpt, seg_slots=slots)
    say(f"  two-level: {slots} segment slots for k = {K} (one level: {one.run.seg_slots})")
    g2 = two.grads()
    assert g1["iso_a"] == g2["iso_a"] and g1["iso_b"] == g2["iso_b"]
    assert np.array_equal(g1["field"], g2["field"])
    assert np.array_equal(g1["series"], g2["series"])


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_gradients_repeat_bitwise(pbb_problem, prec):
    """Two backward sweeps over the same forward give the same gradients to the bit (the parameter and
    inlet-series sums add per-block partials in block order, with no atomics)."""
    pr = Problem(adj_config("pbb"), prec=prec, target=pbb_problem.target)
    g1 = pr.grads()
    g2 = Problem(adj_config("pbb"), prec=prec, target=pbb_problem.target).grads()
    assert g1["iso_a"] == g2["iso_a"] and g1["iso_b"] == g2["iso_b"]
    for name in ("field", "series"):
        assert np.array_equal(g1[name], g2[name]), name


@pytest.mark.parametrize("env", [{"FMX_WIDE": "1"}, {"FMX_BLOCK": "128x2"}, {"FMX_BLOCK": "64x1", "FMX_XCD_REMAP": "0"}])
def test_adjoint_launch_variants(pbb_problem, monkeypatch, env):
    """The adjoint kernel with 64-bit offsets and under other launch shapes: the per-cell gradients
    agree to the bit; the sums over the grid group cells by launch block, so the isotherm and
    inlet-series gradients agree to 1e-12 (to the bit with 64-bit offsets, whose blocks are the same)."""
    g = pbb_problem.grads()
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    say(f"  adjoint launch variant {env}")
    g2 = Problem(adj_config("pbb"), target=pbb_problem.target).grads()
    assert np.array_equal(g["field"], g2["field"])
    if "FMX_WIDE" in env:
        assert g["iso_a"] == g2["iso_a"] and g["iso_b"] == g2["iso_b"]
        assert np.array_equal(g["series"], g2["series"])
    else:
        for name in ("iso_a", "iso_b"):
            assert rel(g2[name], g[name]) < 1e-12, name
        assert np.max(np.abs(g2["series"] - g["series"])) <= 1e-12 * np.max(np.abs(g["series"]))


@pytest.mark.parametrize("prec", ["f64", "f32", "f16"])
def test_adjoint_step_matches_sweep(pbb_problem, prec):
    """Reverse steps taken one at a time (fmx_checkpoint_load of every stored state, then
    fmx_adjoint_step, last step first) give fmx_adjoint_backward's gradients to the bit."""
    import ctypes
    import torch

    from fluidmaxxer_amd import abi
    pr = Problem(adj_config("pbb"), prec=prec, k=1, target=pbb_problem.target)
    g = pr.grads()
    run, sim = pr.run, pr.run.sim
    w = 2.0 * (host(run.forward()) - pr.target)  # every state before steps 1 .. NSTEPS in ckpt slots 0 ..
    ow = torch.zeros(NSTEPS + 1, dtype=torch.float64, device="cuda")
    ow[1:] = torch.as_tensor(w, device="cuda")  # indexed by the step itself
    gi = torch.zeros(2, dtype=torch.float64, device="cuda")
    gf = torch.zeros_like(run.grad_field)
    gs = torch.zeros_like(run.grad_series)
    a = abi.AdjointBufs()
    a.lam_f[0], a.lam_f[1] = run.lam_f[0].data_ptr(), run.lam_f[1].data_ptr()
    a.lam_g[0], a.lam_g[1] = run.lam_g[0].data_ptr(), run.lam_g[1].data_ptr()
    a.lam_q, a.lam_phi = run.lam_q.data_ptr(), run.lam_phi.data_ptr()
    a.workspace, a.workspace_bytes = run.workspace.data_ptr(), run.workspace.numel()
    a.outlet_weight, a.grad_iso, a.grad_field, a.grad_series = ow.data_ptr(), gi.data_ptr(), gf.data_ptr(), gs.data_ptr()
    a.outlet_weight_len = ow.numel()
    say(f"  adjoint one step at a time, prec={prec}")
    lam = 0
    for j in range(NSTEPS - 1, -1, -1):
        abi.call("fmx_checkpoint_load", sim._desc_p(), sim._bufs_p(), 0, run.ckpt.data_ptr(), run.ckpt_prec, j,
                 sim._s())
        abi.call("fmx_adjoint_step", sim._desc_p(), sim._bufs_p(), 0, 1 + j, lam, int(j == NSTEPS - 1),
                 ctypes.byref(a), sim._s())
        lam ^= 1
    gi, gf, gs = host(gi), host(gf), host(gs)
    assert gi[0] == g["iso_a"] and gi[1] == g["iso_b"]
    assert np.array_equal(gf, g["field"])
    assert np.array_equal(gs, g["series"])


def test_c4_recovers_isotherm_constants():
    """The C4 inverse problem as specified: the target curve is the forward run with q_max x 1.1 and
    K_L x 0.9, so the minimum of L = sum_s (C_out(s) - C*(s))^2 is known (0, at those constants).
    L-BFGS on (log q_max, log K_L), through the torch.autograd.Function and the FP32 adjoint
    (checkpoints every 200 steps), recovers both constants."""
    import math

    import torch

    steps = 12000
    cfg = c4_config(steps)
    truth = (cfg.iso_a * 1.1, cfg.iso_b * 0.9)
    trun = AdjointRun(dataclasses.replace(cfg, iso_a=truth[0], iso_b=truth[1]), steps, 200, "f32")
    target = trun.forward().clone()
    del trun
    run = AdjointRun(cfg, steps, 200, "f32")
    theta = torch.tensor([math.log(cfg.iso_a), math.log(cfg.iso_b)], dtype=torch.float64, requires_grad=True)
    opt = torch.optim.LBFGS([theta], lr=1.0, max_iter=25, line_search_fn="strong_wolfe", tolerance_grad=1e-14,
                            tolerance_change=1e-20)
    scale = float((target ** 2).sum())  # C_out stays below ~1e-3 here: the loss relative to the curve
    hist = []

    def closure():
        opt.zero_grad()
        c = CoupledBreakthrough.apply(run, theta[0].exp(), theta[1].exp(), None, None)
        loss = ((c - target) ** 2).sum() / scale
        loss.backward()
        hist.append(float(loss))
        return loss

    opt.step(closure)
    q_max, k_l = (float(v) for v in theta.detach().exp())
    say(f"  C4 recovery: {len(hist)} evaluations, loss {hist[0]:.3e} -> {min(hist):.3e}; q_max {q_max:.5f} "
        f"(truth {truth[0]:.5f}), K_L {k_l:.5f} (truth {truth[1]:.5f})")
    assert min(hist) < 1e-8 * hist[0], hist
    assert abs(q_max / truth[0] - 1) < 1e-3 and abs(k_l / truth[1] - 1) < 1e-3, (q_max, k_l, truth)


def test_c4_differentiation_fp64_vs_fd():
    """C4-scale FP64 adjoint (12000 steps, checkpoint every 200; target = the forward run with
    q_max x 1.1 and K_L x 0.9) against FP64 central differences of the GPU forward run."""
    steps = 12000
    cfg = c4_config(steps)
    tcfg = dataclasses.replace(cfg, iso_a=cfg.iso_a * 1.1, iso_b=cfg.iso_b * 0.9)
    trun = AdjointRun(tcfg, steps, 200, "f64")
    target = host(trun.forward())
    del trun
    pr = Problem(cfg, prec="f64", nsteps=steps, k=200, target=target)
    g = pr.grads()
    say(f"  C4 f64: c_out at steps 4000/8000/12000 = {[float('%.4e' % pr.c0[s - 1]) for s in (4000, 8000, 12000)]} "
        f"target[-1]={target[-1]:.4e} dL/dq_max={g['iso_a']:.6e} dL/dK_L={g['iso_b']:.6e}")
    for name in ("iso_a", "iso_b"):
        fd = pr.fd_scalar(name, 1e-5 * pr.base[name])
        say(f"    d L/d {name}: adjoint {g[name]:.10e} fd {fd:.10e} rel {rel(g[name], fd):.2e}")
        assert rel(g[name], fd) < 1e-7  # observed 3.1e-8
    f0 = cfg.field
    # cells at least 1e-3 inside (0, 1), so that no perturbation crosses the clamp at 0 or 1
    # (C4's sphere cells lie in (0.95, 1]; a step of a few 1e-6 would push some across 1)
    inner = (f0 > 1e-3) & (f0 < 1 - 1e-3)
    v = np.random.default_rng(7).standard_normal(cfg.n) * inner
    fd = pr.fd_dir("field", v, 1e-6)
    fd4 = pr.fd_dir("field", v, 4e-6)
    ad = float(g["field"] @ v)
    say(f"    <dL/dfield, v> (1e-3 < field < 1 - 1e-3): adjoint {ad:.10e} fd(h) {fd:.10e} fd(4h) {fd4:.10e} "
        f"rel {rel(ad, fd):.2e}")
    # central differences at h = 1e-6 carry ~1e-6 truncation here (their h / 4h difference), so the
    # bound is 5e-6 (observed 2.2e-6)
    assert rel(ad, fd) < 5e-6
    # fluid cells sit at field = 0, where the gradient passed through is the one-sided (upward) one
    # (forward differences at h and 2h, extrapolated: 2 D(h) - D(2h) cancels the O(h) term, which the
    # drag of partial bounce-back in every fluid cell makes large)
    w = np.abs(np.random.default_rng(8).standard_normal(cfg.n)) * (f0 == 0)
    h = 1e-7
    l0 = pr.loss()
    d1 = (pr.loss(field=f0 + h * w) - l0) / h
    d2 = (pr.loss(field=f0 + 2 * h * w) - l0) / (2 * h)
    fd1 = 2 * d1 - d2
    ad1 = float(g["field"] @ w)
    say(f"    <dL/dfield, w> (field = 0, upward): adjoint {ad1:.10e} one-sided fd {fd1:.10e} (D(h) {d1:.6e}) "
        f"rel {rel(ad1, fd1):.2e}")
    assert rel(ad1, fd1) < 5e-6  # observed 1.8e-6


def test_c4_fp32_adjoint_vs_fp64_adjoint():
    """C4-scale FP32 adjoint against the FP64 adjoint, loss against 0.85 x the FP64 curve (residuals
    of 15% of the curve and of one sign, so neither the FP32 forward error nor cancellation between
    steps dominates the gradient); and the FP64 run with FP32 checkpoints and an FP32 reverse sweep,
    whose FP32 drift is restarted from the FP64 trajectory at every checkpoint."""
    steps = 12000
    cfg = c4_config(steps)
    p64 = Problem(cfg, prec="f64", nsteps=steps, k=200)
    p64.target = 0.85 * p64.c0
    g64 = p64.grads()
    p32 = Problem(cfg, prec="f32", nsteps=steps, k=200, target=p64.target)
    g32 = p32.grads()
    for name in ("iso_a", "iso_b"):
        say(f"  C4 f32 vs f64 d L/d {name}: {g32[name]:.8e} vs {g64[name]:.8e} rel {rel(g32[name], g64[name]):.2e}")
        assert rel(g32[name], g64[name]) < 4e-3  # observed 1.96e-3: the FP32 forward run's own drift
    e = np.linalg.norm(g32["field"] - g64["field"]) / np.linalg.norm(g64["field"])
    say(f"  C4 f32 vs f64 field gradient: relative l2 {e:.2e}")
    assert e < 4e-3  # observed 1.92e-3
    del p32
    pm = Problem(cfg, prec="f64", nsteps=steps, k=200, target=p64.target, ckpt="f32")
    assert np.array_equal(pm.c0, p64.c0)
    gm = pm.grads()
    for name in ("iso_a", "iso_b"):
        say(f"  C4 f64 run + f32 sweep vs f64 d L/d {name}: rel {rel(gm[name], g64[name]):.2e}")
        assert rel(gm[name], g64[name]) < 4e-5  # observed 1.5e-5, 1.7e-5
    e = rel_l2(gm["field"], g64["field"])
    say(f"  C4 f64 run + f32 sweep vs f64 field gradient: relative l2 {e:.2e}")
    assert e < 6e-5  # observed 2.5e-5


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


# observed: FP16 1.04e-3 .. 1.32e-3, BF16 1.19e-2 .. 1.56e-2
@pytest.mark.parametrize("prec,bound_iso,bound_vec", [("f16", 3e-3, 3e-3), ("bf16", 3e-2, 3.5e-2)])
def test_half_state_adjoint_against_fp64(pbb_problem, prec, bound_iso, bound_vec):
    """16-bit deviation-stored checkpoints and segment states (FP32 arithmetic in the recompute and
    the adjoint) against the FP64 adjoint; the errors are reported."""
    g64 = pbb_problem.grads()
    ph = Problem(adj_config("pbb"), prec=prec, target=pbb_problem.target)
    gh = ph.grads()
    errs = {name: rel(gh[name], g64[name]) for name in ("iso_a", "iso_b")}
    errs.update({name: rel_l2(gh[name], g64[name]) for name in ("field", "series")})
    say(f"  {prec} states vs f64 adjoint:", {k: f"{v:.2e}" for k, v in errs.items()})
    assert errs["iso_a"] < bound_iso and errs["iso_b"] < bound_iso, errs
    assert errs["field"] < bound_vec and errs["series"] < bound_vec, errs


def interpolation_matrix(n_knots, steps):
    """d(per-step inlet value)/d(knot) of workloads.series_per_step (linear interpolation)."""
    dt = steps / (n_knots - 1)
    s = np.arange(steps + 1) / dt
    B = np.zeros((steps + 1, n_knots))
    i0 = np.minimum(np.floor(s).astype(int), n_knots - 2)
    t = s - i0
    B[np.arange(steps + 1), i0] = 1 - t
    B[np.arange(steps + 1), i0 + 1] = t
    return B


@pytest.mark.parametrize("ckpt,k", [("f16", K), ("bf16", K), ("f16", 1)])
def test_half_checkpoints_of_fp32_run(pbb_problem, ckpt, k):
    """An FP32 run whose checkpoints are stored as 16-bit deviations and decoded for an FP32 recompute:
    the forward curve is the FP32 run's own, bit for bit, and the gradients stay within the 16-bit
    rounding of the checkpointed states of the FP32-checkpoint adjoint (k = 1 rounds every state)."""
    p32 = Problem(adj_config("pbb"), prec="f32", target=pbb_problem.target)
    g32 = p32.grads()
    ph = Problem(adj_config("pbb"), prec="f32", k=k, target=pbb_problem.target, ckpt=ckpt)
    assert np.array_equal(ph.c0, p32.c0)
    gh = ph.grads()
    g64 = pbb_problem.grads()
    errs = {name: rel(gh[name], g32[name]) for name in ("iso_a", "iso_b")}
    errs.update({name: rel_l2(gh[name], g32[name]) for name in ("field", "series")})
    e64 = {name: rel(gh[name], g64[name]) for name in ("iso_a", "iso_b")}
    e64.update({name: rel_l2(gh[name], g64[name]) for name in ("field", "series")})
    say(f"  f32 run, {ckpt} checkpoints (k={k}) vs f32 checkpoints:", {n: f"{v:.2e}" for n, v in errs.items()})
    say(f"  f32 run, {ckpt} checkpoints (k={k}) vs f64 adjoint:", {n: f"{v:.2e}" for n, v in e64.items()})
    # observed: FP16 k = 64 6.8e-6, BF16 k = 64 5.8e-5, FP16 k = 1 1.3e-4
    bound = 3e-4 if k == 1 else (1.5e-4 if ckpt == "bf16" else 2e-5)
    assert max(errs.values()) < bound, errs


def test_c5_half_checkpoints_against_fp64():
    """C5 physics (overlapping sphere pack at porosity 0.4, partial bounce-back, Langmuir, 200-knot inlet
    series), 20000 steps, on a 48x64x64 box (the FP64 reference adjoint's store stays at a few GB): the
    FP32 run with FP16-stored checkpoints every 500 steps and FP32 recompute against the FP64 adjoint
    (checkpoints every 200); gradients with respect to the 200 knots (chain rule through the per-step
    interpolation) and the per-cell solid fraction, errors reported.  As for C4, the fluid cells'
    solid fraction and the concentration clamp are set to 0.  With the equilibrium inlet holding the
    inlet density at rho0 the flow through this dense pack is weak and C_out stays small (~1e-4), so
    the loss is taken against 0.85 x the FP64 curve."""
    steps = 20000
    cfg = workloads.build("C5", dims=(48, 64, 64), steps=steps)
    cfg.field = np.where(cfg.field > 0.5, cfg.field, 0.0)
    cfg.min_conc = 0.0
    knots = cfg.meta["knots"]
    B = interpolation_matrix(len(knots), steps)
    assert np.allclose(B @ knots, cfg.inlet_series)
    p64 = Problem(cfg, prec="f64", nsteps=steps, k=200)
    say(f"  C5 48x64x64 c_out at 5000/10000/15000/20000: "
        f"{[float('%.3e' % p64.c0[s - 1]) for s in (5000, 10000, 15000, 20000)]}")
    assert p64.c0[-1] > 1e-6
    p64.target = 0.85 * p64.c0
    g64 = p64.grads()
    p16 = Problem(cfg, prec="f32", nsteps=steps, k=500, target=p64.target, ckpt="f16")
    g16 = p16.grads()
    k64, k16 = B.T @ g64["series"], B.T @ g16["series"]
    errs = dict(knots=rel_l2(k16, k64), field=rel_l2(g16["field"], g64["field"]), c_out=rel_l2(p16.c0, p64.c0),
                iso_a=rel(g16["iso_a"], g64["iso_a"]), iso_b=rel(g16["iso_b"], g64["iso_b"]))
    say(f"  C5 f32 run + f16 checkpoints vs f64 adjoint: {({k: f'{v:.2e}' for k, v in errs.items()})}, "
        f"|dL/dknots| {np.linalg.norm(k64):.3e}")
    # observed: c_out 3.6e-5, knots 2.7e-4, field 2.8e-4, iso 2.6e-4
    assert errs["c_out"] < 1e-4 and errs["knots"] < 6e-4 and errs["field"] < 6e-4, errs
    assert errs["iso_a"] < 6e-4 and errs["iso_b"] < 6e-4, errs
    del p16
    # the FP64 run with FP32 checkpoints every 500 steps and an FP32 reverse sweep
    pm = Problem(cfg, prec="f64", nsteps=steps, k=500, target=p64.target, ckpt="f32")
    assert np.array_equal(pm.c0, p64.c0)
    gm = pm.grads()
    km = B.T @ gm["series"]
    em = dict(knots=rel_l2(km, k64), field=rel_l2(gm["field"], g64["field"]), iso_a=rel(gm["iso_a"], g64["iso_a"]),
              iso_b=rel(gm["iso_b"], g64["iso_b"]))
    say(f"  C5 f64 run + f32 sweep vs f64 adjoint: {({k: f'{v:.2e}' for k, v in em.items()})}")
    assert max(em.values()) < 1e-4, em  # observed: knots 6.7e-5, field 5.8e-5, iso 1.0e-5


def test_c5_full_size_fp16_checkpoints_against_fp64():
    """C5 at its own size: the 256x128x128 overlapping sphere pack at porosity 0.4 (seed 5), partial
    bounce-back, Langmuir, the 200-knot inlet series, 20000 steps; the FP32 run with FP16-stored
    checkpoints every 500 steps and FP32 recompute against the FP64 adjoint (both with two-level
    segment stores, so the FP64 store is about 80 GB), and against the FP32 run with FP32 checkpoints.
    As in the reduced-size test the fluid cells' solid fraction and the concentration clamp are 0.

    The concentration front has NOT reached the outlet by step 20000 (C_out ~ 1e-15, asserted): the loss
    sum (C_out - 0.85 C_out^FP64)^2 weighs only the front's far tail, fifteen orders of magnitude below
    the inlet value.  There the FP32 forward run itself is about 0.6% off the FP64 curve and its
    gradients about 4.5% off the FP64 adjoint's, with FP32 and FP16 checkpoints alike, so the FP16
    checkpoints' own effect is measured against the FP32-checkpoint run (the forward curves are the same
    bits).  The FP64 run with FP32 checkpoints every 500 steps and an FP32 reverse sweep restarts the
    FP32 recompute from the FP64 trajectory at every checkpoint; its gradients are compared with the
    FP64 adjoint directly.  Bounds: about 2x the measured errors."""
    import torch

    steps, k = 20000, 500
    cfg = workloads.build("C5", steps=steps)
    assert (cfg.nx, cfg.ny, cfg.nz) == (256, 128, 128) and len(cfg.meta["knots"]) == 200
    cfg.field = np.where(cfg.field > 0.5, cfg.field, 0.0)
    cfg.min_conc = 0.0
    knots = cfg.meta["knots"]
    B = interpolation_matrix(len(knots), steps)
    from fluidmaxxer_amd import abi

    def grads(prec, ckpt, target):
        seg = segment_slots(k, prec, ckpt, two_level=True)
        p = Problem(cfg, prec=prec, nsteps=steps, k=k, target=target, ckpt=ckpt, seg_slots=seg)
        if target is None:
            p.target = 0.85 * p.c0
        g = p.grads()
        out = dict(c=p.c0, target=p.target, knots=B.T @ g["series"], field=g["field"], iso_a=g["iso_a"],
                   iso_b=g["iso_b"])
        del p, g
        torch.cuda.empty_cache()
        return out

    g64 = grads("f64", None, None)
    say(f"  C5 256x128x128 FP64 c_out at 5000/10000/20000: {[float('%.3e' % g64['c'][s - 1]) for s in (5000, 10000, 20000)]}")
    assert g64["c"][-1] < 1e-10  # the front is still inside the pack
    g16 = grads("f32", "f16", g64["target"])
    g32 = grads("f32", None, g64["target"])
    gm = grads("f64", "f32", g64["target"])

    def errs(a, b):
        return dict(knots=rel_l2(a["knots"], b["knots"]), field=rel_l2(a["field"], b["field"]),
                    c_out=rel_l2(a["c"], b["c"]), iso_a=rel(a["iso_a"], b["iso_a"]), iso_b=rel(a["iso_b"], b["iso_b"]))

    e16_64, e32_64, e16_32, em_64 = errs(g16, g64), errs(g32, g64), errs(g16, g32), errs(gm, g64)
    for name, e in (("f32 run + f16 ckpt vs f64", e16_64), ("f32 run + f32 ckpt vs f64", e32_64),
                    ("f16 ckpt vs f32 ckpt", e16_32), ("f64 run + f32 sweep vs f64", em_64)):
        say(f"  C5 {name}: {({kk: f'{v:.2e}' for kk, v in e.items()})}")
    assert np.array_equal(g16["c"], g32["c"])  # 16-bit checkpoints do not touch the forward run
    assert np.array_equal(gm["c"], g64["c"])  # nor do FP32 checkpoints of the FP64 run
    # observed: knots 7.7e-4, field 1.2e-3, iso 1.1e-3 (the FP32 runs: 4.5e-2)
    assert max(em_64[n] for n in ("knots", "field", "iso_a", "iso_b")) < 2.5e-3, em_64
    assert e16_64["c_out"] < 1.2e-2
    assert max(e16_64[n] for n in ("knots", "field", "iso_a", "iso_b")) < 9e-2, e16_64
    # observed (AD checkpoint values relative to their concentration): knots 4.8e-4, field 1.6e-3, iso 9.9e-5
    assert e16_32["knots"] < 1e-3 and e16_32["field"] < 3.5e-3, e16_32
    assert e16_32["iso_a"] < 2.5e-4 and e16_32["iso_b"] < 2.5e-4, e16_32
