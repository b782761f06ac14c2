"""Row-compacted NS storage (fmx_coupled_bufs::compact_work): the frozen solid cells are left out of
the NS planes a run steps, the other cells are stepped in row-compacted buffers and copied back.  The
state after every run must be the one the box-layout kernel leaves, to the bit (FP32 / FP64; the
16-bit kernels are compiled with FMA contraction, so there the two instantiations agree to the
storage rounding), across run boundaries, odd and even step counts, every boundary combination the
fused step folds in, box widths that are not whole waves and widths whose boundary cells fall in two
waves."""
import numpy as np
import pytest

from fluidmaxxer_amd import abi, workloads
from fluidmaxxer_amd.coupled import CoupledSim
from tests.gpu_util import host, say
from tests.lattice import W19, W7, rel_err
from tests.test_gpu_coupled import small_config

pytestmark = pytest.mark.gpu


def mixed_state(cfg, seed):
    """+-1% perturbation of the equilibria on every fluid cell and on one solid cell in five (those
    carry momentum and stay on the full path); the other solid cells stay at rest and freeze."""
    rng = np.random.default_rng(seed)
    n = cfg.n
    solid = np.asarray(cfg.solid).astype(bool)
    keep = solid & (rng.random(n) < 0.8)
    pert = np.where(keep, 0.0, 0.01)
    f = np.outer(W19, np.full(n, cfg.rho0)) * (1 + pert * rng.standard_normal((19, n)))
    g = np.outer(W7, np.full(n, max(cfg.phi0, 0.05))) * (1 + 0.01 * rng.standard_normal((7, n)))
    return f.astype(np.float32).astype(np.float64), g.astype(np.float32).astype(np.float64)


def frozen_count(sim):
    return int((sim.links & abi.LINK_FROZEN).ne(0).sum().item())


def wave_straddling_pairs(cfg, sim):
    """Rows whose inlet pair (x = 0, 1) or outlet pair (x = nx-2, nx-1) falls in two waves of the split
    NS kernel (the first cell of the pair in lane 63 of the compact order): the boundary lanes of such
    pairs cannot take their neighbour's values by a lane shuffle, so the neighbour writes them."""
    frozen = (sim.links.cpu().numpy().astype(np.int64) & abi.LINK_FROZEN) != 0
    k = np.cumsum(~frozen) - 1
    x = np.arange(cfg.n) % cfg.nx
    return int(((x == 0) & (k % 64 == 63)).sum()), int(((x == cfg.nx - 2) & (k % 64 == 63)).sum())


def run_both(cfg, prec, runs, seed=5, mass_every=0):
    out = []
    for compact in (False, True):
        say(f"compact={compact} prec={prec} dims={(cfg.nx, cfg.ny, cfg.nz)} inlet={cfg.ns_inlet} runs={runs}")
        sim = CoupledSim(cfg, prec=prec, compact=compact)
        f, g = mixed_state(cfg, seed)
        sim.load(f=f, g=g)
        curves = [host(sim.run(k, mass_every=mass_every)) for k in runs]
        sim.synchronize()
        out.append(dict(sim=sim, f=sim.f[sim.cur].clone(), fo=sim.f[1 - sim.cur].clone(), g=sim.g[sim.cur].clone(),
                        phi=sim.phi.clone(), q=sim.q.clone(), view=sim.view_f().clone(),
                        raw=None if sim.inlet_raw is None else sim.inlet_raw.clone(), c=np.concatenate(curves)))
    return out


def assert_same(dense, comp, names=("f", "g", "phi", "q", "view", "raw")):
    """The state (f[cur] and the rest) to the bit, and the frozen cells' second phase in f[1 - cur] (the
    other cells' values there are scratch, not state: the compact run does not write them)."""
    import torch
    for k in names:
        if dense[k] is None:
            continue
        assert torch.equal(dense[k], comp[k]), k
    assert np.array_equal(dense["c"], comp["c"]), "c_out"  # the outlet ring sums in a fixed order
    fz = (comp["sim"].links & abi.LINK_FROZEN).ne(0).repeat(19)
    assert torch.equal(dense["fo"][fz], comp["fo"][fz]), "frozen cells' second phase"


VARIANTS = {
    "neq_24": ("bgk", (24, 12, 10)),
    "eq_inlet_130": ("eq_inlet", (130, 6, 5)),
    "series_96": ("langmuir_series", (96, 8, 6)),
    "walls_64": ("closed_walls", (64, 9, 7)),
    "plain_48": ("no_inlet_outlet", (48, 7, 5)),
    "split_boundary_65": ("bgk", (65, 6, 5)),
    "wide_256": ("eq_inlet", (256, 8, 6)),
}


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("case", sorted(VARIANTS))
def test_compact_run_bit_identical(prec, case):
    variant, dims = VARIANTS[case]
    cfg = small_config(variant, dims=dims)
    dense, comp = run_both(cfg, prec, (37, 1, 22))
    nfz = frozen_count(comp["sim"])
    pairs = wave_straddling_pairs(cfg, comp["sim"])
    say(f"  frozen cells {nfz} of {cfg.n}; workspace {comp['sim']._compact_ws.numel()} B; "
        f"inlet / outlet pairs across two waves {pairs}")
    assert nfz > 0 and frozen_count(dense["sim"]) == 0
    if case == "neq_24":  # both kinds of boundary pair split between two waves are exercised
        assert pairs[0] > 0 and pairs[1] > 0
    assert_same(dense, comp)


@pytest.mark.parametrize("prec", ["f32", "f64", "f16"])
@pytest.mark.parametrize("case", ["neq_24", "split_boundary_65", "wide_256"])
def test_compact_fused_kernel(prec, case, monkeypatch):
    """FMX_SPLIT=0: compact runs take the fused compact kernel (one launch per step, k_coupled_rc)
    instead of the split step; the same state as the box layout (FP16: to its rounding)."""
    import torch
    monkeypatch.setenv("FMX_SPLIT", "0")
    variant, dims = VARIANTS[case]
    cfg = small_config(variant, dims=dims)
    dense, comp = run_both(cfg, prec, (37, 22))
    assert frozen_count(comp["sim"]) > 0
    if prec == "f16":
        same = {k: bool(torch.equal(dense[k], comp[k])) for k in ("f", "g", "phi", "q")}
        errs = {k: rel_err(host(comp[k]), host(dense[k])) for k in ("view", "phi", "q")}
        say("  bit-identical:", same, " fused compact vs box layout:", {k: f"{v:.1e}" for k, v in errs.items()})
        assert all(v < 2e-3 for v in errs.values()), errs
    else:
        assert_same(dense, comp)


@pytest.mark.parametrize("prec", ["f16", "bf16"])
@pytest.mark.parametrize("case", ["neq_24", "series_96", "wide_256"])
def test_compact_run_16bit(prec, case):
    variant, dims = VARIANTS[case]
    cfg = small_config(variant, dims=dims)
    dense, comp = run_both(cfg, prec, (30, 21))
    assert frozen_count(comp["sim"]) > 0
    import torch
    same = {k: bool(torch.equal(dense[k], comp[k])) for k in ("f", "g", "phi", "q")}
    errs = {k: rel_err(host(comp[k]), host(dense[k])) for k in ("view", "phi", "q")}
    errs["c"] = rel_err(comp["c"], dense["c"])
    say("  bit-identical:", same, " compact vs box layout:", {k: f"{v:.1e}" for k, v in errs.items()})
    assert all(v < 2e-3 for v in errs.values()), errs


def test_compact_c1_against_box_layout_long_run():
    """C1 (the 64x32x32 pack at porosity 0.6 from rest, NEQ inlet): every solid cell on 2 <= x <= nx-3
    freezes, and 600 steps in three runs leave the box-layout state to the bit."""
    cfg = workloads.build("C1")
    out = []
    for compact in (False, True):
        say(f"C1 compact={compact}")
        sim = CoupledSim(cfg, prec="f32", compact=compact)
        c = np.concatenate([host(sim.run(k)) for k in (200, 201, 199)])
        sim.synchronize()
        out.append(dict(sim=sim, f=sim.f[sim.cur].clone(), fo=sim.f[1 - sim.cur].clone(), g=sim.g[sim.cur].clone(),
                        phi=sim.phi.clone(), q=sim.q.clone(), view=sim.view_f().clone(), raw=sim.inlet_raw.clone(),
                        c=c))
    x = np.arange(cfg.n) % cfg.nx
    expect = int(((np.asarray(cfg.solid) == 1) & (x >= 2) & (x <= cfg.nx - 3)).sum())
    assert frozen_count(out[1]["sim"]) == expect
    assert_same(out[0], out[1])


def test_compact_with_mass_log_runs_in_the_box_layout():
    """A run that logs the mass diagnostic reads the state between steps, so it ignores the compact
    workspace; the result is the same either way."""
    cfg = small_config("bgk", dims=(64, 8, 6))
    dense, comp = run_both(cfg, "f32", (20,), mass_every=5)
    assert_same(dense, comp)
    import torch
    assert torch.equal(dense["sim"].mass_log, comp["sim"].mass_log)


def test_compact_workspace_that_does_not_fit_runs_the_box_layout():
    """A workspace sized for fewer active cells than the link words now have (a state that froze fewer
    cells) leaves the run in the box layout, with the same result."""
    cfg = small_config("bgk", dims=(64, 8, 6))
    dense, comp = run_both(cfg, "f32", (5,))
    active = cfg.n - frozen_count(comp["sim"])
    small = int(abi.lib().fmx_coupled_compact_bytes(comp["sim"]._desc_p(), active - 200))
    assert 0 < small < comp["sim"]._compact_ws.numel()
    comp["sim"]._compact_ws = comp["sim"]._compact_ws[:small]
    import torch
    for sim in (dense["sim"], comp["sim"]):
        sim.run(40)
    for a, b in ((dense["sim"].f[dense["sim"].cur], comp["sim"].f[comp["sim"].cur]),
                 (dense["sim"].phi, comp["sim"].phi), (dense["sim"].q, comp["sim"].q)):
        assert torch.equal(a, b)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_compact_after_checkpoint_restart(prec):
    """fmx_checkpoint_load marks the frozen cells again from the restored state (a checkpoint holds one
    phase), so compact runs after a restart go on to the same bits as a box-layout run."""
    import torch
    cfg = small_config("langmuir_series", dims=(96, 8, 6))
    dense, comp = run_both(cfg, prec, (20,))
    cp = abi.PREC_NAMES[prec]
    for o in (dense, comp):
        sim = o["sim"]
        nb = int(abi.lib().fmx_adjoint_ckpt_bytes(sim._desc_p(), cp))
        o["store"] = torch.empty(nb, dtype=torch.uint8, device="cuda")
        o["cur0"], o["step0"] = sim.cur, sim.step_index
        abi.call("fmx_checkpoint_save", sim._desc_p(), sim._bufs_p(), sim.cur, o["store"].data_ptr(), cp, 0, sim._s())
        sim.run(25)
        abi.call("fmx_checkpoint_load", sim._desc_p(), sim._bufs_p(), o["cur0"], o["store"].data_ptr(), cp, 0, sim._s())
        sim.cur, sim.step_index = o["cur0"], o["step0"]
    nfz = frozen_count(comp["sim"])
    assert nfz > 0 and frozen_count(dense["sim"]) == 0
    for o in (dense, comp):
        o["c"] = host(o["sim"].run(33))
    assert np.array_equal(dense["c"], comp["c"])
    for k in ("f", "g"):
        a, b = getattr(dense["sim"], k)[dense["sim"].cur], getattr(comp["sim"], k)[comp["sim"].cur]
        assert torch.equal(a, b), k
    assert torch.equal(dense["sim"].phi, comp["sim"].phi) and torch.equal(dense["sim"].q, comp["sim"].q)
