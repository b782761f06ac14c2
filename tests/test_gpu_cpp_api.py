"""The C++ host API (fluidmaxxer_amd/host, the reference's class names and signatures)
driving a main.cpp-style coupled run, through both the operator sequence and
CoupledStep's fused launch, against the oracle.  The driver is built by
`python -m fluidmaxxer_amd.build` (tests/cpp/_build/coupled_driver_{float,double})."""
import os
import subprocess

import numpy as np
import pytest

from fluidmaxxer_amd import geometry
from fluidmaxxer_amd.coupled import CoupledConfig
from oracle import ref_cpu as R
from tests.gpu_util import say
from tests.lattice import oracle_setup, rel_err

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(__file__), "cpp", "_build", "coupled_driver_{}")


def _config(inlet, iso):
    nx, ny, nz = 28, 12, 10
    pack = geometry.sphere_pack(nx, ny, nz, 21, 2.0, 3.0, 0.6, overlap=False, clearance=4)
    cfg = CoupledConfig(nx, ny, nz, solid=pack.solid, region=pack.sphere, tau=0.8, u_in=(0.05, 0, 0), tau_ad=0.7,
                        ns_inlet=inlet, isotherm=iso)
    if iso == "linear":
        cfg.iso_a, cfg.k_ldf = 0.5, 0.05
    else:
        cfg.iso_a, cfg.iso_b, cfg.k_ldf = 2.0, 5.0, 0.02
    return cfg


def _run_driver(tmp_path, cfg, dtype, mode, steps, inlet, iso, extra=()):
    exe = DRIVER.format(dtype)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} not built (python -m fluidmaxxer_amd.build)")
    solid = str(tmp_path / "solid.vti")
    region = str(tmp_path / "region.vti")
    geometry.write_vti(solid, cfg.nx, cfg.ny, cfg.nz, cfg.solid)
    geometry.write_vti(region, cfg.nx, cfg.ny, cfg.nz, cfg.region, "region")
    out = tmp_path / f"{dtype}_{mode}_{len(list(tmp_path.iterdir()))}"
    out.mkdir()
    say(f"cpp driver {dtype} {mode} inlet={inlet} iso={iso} steps={steps} {' '.join(extra)}")
    r = subprocess.run([exe, solid, region, str(out), str(steps), mode, f"inlet={inlet}", f"iso={iso}",
                        "kA=0.05", "kldf=0.02", *extra], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    npdt = np.float32 if dtype == "float" else np.float64
    res = {k: np.fromfile(out / f"{k}.bin", dtype=npdt).astype(np.float64) for k in ("rho", "ux", "uy", "uz", "phi", "q")}
    res["c_out"] = np.fromfile(out / "c_out.bin", dtype=np.float64)
    info = (out / "info.txt").read_text()
    return res, info


@pytest.mark.parametrize("dtype,oprec,tol", [("float", 4, 1e-5), ("double", 8, 1e-11)])
@pytest.mark.parametrize("inlet,iso", [("eq", "linear"), ("neq", "langmuir")])
def test_cpp_api_matches_oracle(tmp_path, dtype, oprec, tol, inlet, iso):
    cfg = _config(inlet, iso)
    steps = 120
    orc = R.System(oracle_setup(cfg, R), oprec)
    outlet = (np.arange(cfg.n) % cfg.nx == cfg.nx - 1).astype(np.uint8)
    c_ref, _, _, _ = orc.run(steps, outlet)
    ref = dict(rho=orc.get(R.RHO), ux=orc.get(R.UX), uy=orc.get(R.UY), uz=orc.get(R.UZ), phi=orc.get(R.PHI),
               q=orc.get(R.Q), c_out=c_ref)
    for mode in ("operators", "fused"):
        got, info = _run_driver(tmp_path, cfg, dtype, mode, steps, inlet, iso)
        assert ("fused=1" in info) == (mode == "fused"), info
        errs = {k: rel_err(got[k], ref[k]) for k in ref}
        say(f"  {mode}:", {k: f"{v:.2e}" for k, v in errs.items()})
        assert all(v <= tol for v in errs.values()), (mode, errs)


@pytest.mark.parametrize("dtype", ["float", "double"])
def test_cpp_api_compact_solids_same_bits(tmp_path, dtype):
    """CoupledStep::useCompactSolids(): the fused runs leave the solid cells at rest out of the NS
    traffic (row-compacted planes) and give the same fields and outlet curve, to the bit, as the fused
    runs without it; split into two runs (each at least 16 steps, so both are compact)."""
    cfg = _config("neq", "langmuir")
    runs = {}
    for extra in ((), ("compact=1",)):
        got, info = _run_driver(tmp_path, cfg, dtype, "fused", 120, "neq", "langmuir", ("split=50", *extra))
        assert "fused=1" in info, info
        runs[extra] = got
    for k, v in runs[()].items():
        assert np.array_equal(v, runs[("compact=1",)][k]), k


PULSE = (0.1, 0.9, 300, 900, 40.0)  # base, amp, t_on, t_off, k: the run straddles both edges


@pytest.mark.parametrize("dtype,oprec,tol", [("float", 4, 1e-5), ("double", 8, 1e-11)])
def test_cpp_api_tanh_pulse_matches_oracle(tmp_path, dtype, oprec, tol):
    """ADDirichletTanhPulse (ADDirichletTanhPulse.cu:66-100, wired as in main.cpp:254-271) through the
    operator path (pre-collision g <- phi_in(s) w on the inlet plane) and through CoupledStep's fused
    route (a per-step inlet series cached on the device from valueAt(s)), the fused run split into
    two calls at step 1000 so that the second one grows the cache (the first reserves 1024 steps),
    against the oracle's tanh-pulse boundary over 1100 steps rising at t_on = 300 and falling at
    t_off = 900."""
    cfg = _config("eq", "linear")
    steps = 1100
    s = oracle_setup(cfg, R)
    inlet = (np.arange(cfg.n) % cfg.nx == 0).astype(np.uint8)
    s.ad_bc = [R.BC(R.AD_TANH, inlet, (PULSE[0], PULSE[1], PULSE[4]), i0=PULSE[2], i1=PULSE[3])] + \
        [b for b in s.ad_bc if b.kind != R.AD_DIRICHLET]
    orc = R.System(s, oprec)
    outlet = (np.arange(cfg.n) % cfg.nx == cfg.nx - 1).astype(np.uint8)
    c_ref, _, _, _ = orc.run(steps, outlet)
    ref = dict(rho=orc.get(R.RHO), ux=orc.get(R.UX), uy=orc.get(R.UY), uz=orc.get(R.UZ), phi=orc.get(R.PHI),
               q=orc.get(R.Q), c_out=c_ref)
    pulse = "pulse=" + ",".join(str(v) for v in PULSE)
    for mode, extra in (("operators", ()), ("fused", ("split=1000",)), ("fused", ())):
        got, info = _run_driver(tmp_path, cfg, dtype, mode, steps, "eq", "linear", ("adinlet=pulse", pulse, *extra))
        assert ("fused=1" in info) == (mode == "fused"), info
        errs = {k: rel_err(got[k], ref[k]) for k in ref}
        say(f"  {mode} {extra}:", {k: f"{v:.2e}" for k, v in errs.items()})
        assert all(v <= tol for v in errs.values()), (mode, extra, errs)
    # the inlet plane follows the pulse: 200 steps after t_off its concentration is back near base
    assert ref["phi"][inlet == 1].max() < PULSE[0] + 0.05
