"""The main.cpp-equivalent driver (fluidmaxxer_amd/host/apps/lbm_coupled.cpp) end to end on a small
sphere-pack geometry: it reads geometry.vti, derives the lattice parameters from main.cpp's physical
set, runs the coupled step through the C++ API and writes the outlet CSV and VTI/PVD fields.

Checked against the oracle (FP64 restatement of the reference's step sequence) configured with the
parameters the driver reports: the outlet CSV (6 significant digits, as OutletCSV prints them) and the
concentration field of the last VTI file.
"""
import json
import os
import re
import subprocess

import numpy as np
import pytest

from fluidmaxxer_amd import geometry
from oracle import ref_cpu as R
from tests.gpu_util import say
from tests.lattice import x_plane

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "fluidmaxxer_amd", "host", "apps", "_build", "lbm_coupled_double")
STEPS, EVERY, CSV_EVERY = 300, 150, 50


def vti_array(path, name):
    text = open(path).read()
    m = re.search(r'<DataArray[^>]*Name="%s"[^>]*format="ascii">(.*?)</DataArray>' % re.escape(name), text, re.S)
    assert m, (path, name)
    return np.array(m.group(1).split(), np.float64)


@pytest.fixture(scope="module", params=["dirichlet", "pulse"])
def driver_run(tmp_path_factory, request):
    if not os.path.exists(APP):
        pytest.skip("lbm_coupled not built (python -m fluidmaxxer_amd.build)")
    d = tmp_path_factory.mktemp("driver")
    nx, ny, nz = 48, 20, 20
    pack = geometry.sphere_pack(nx, ny, nz, 12, 2.0, 3.0, 0.7, overlap=False, clearance=4)
    geo = str(d / "geometry.vti")
    geometry.write_vti(geo, nx, ny, nz, pack.sphere)
    out = str(d / "results")
    cmd = [APP, "--coupling", "--adsorb", "--vti", "--geometry", geo, "--out", out, "--steps", str(STEPS),
           "--every", str(EVERY), "--csv-every", str(CSV_EVERY)]
    if request.param == "pulse":  # main.cpp:254-271: ADDirichletTanhPulse(0, 1, 0.1 iTmax, 0.9 iTmax, 0.05 iTmax)
        cmd.append("--pulse")
    say("lbm_coupled", " ".join(cmd[1:]))
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    line = [l for l in r.stdout.splitlines() if l.startswith("params: ")][-1]
    params = json.loads(line[len("params: "):])
    return dict(dir=out, params=params, mask=pack.sphere, dims=(nx, ny, nz), stdout=r.stdout, inlet=request.param)


def test_driver_outputs(driver_run):
    out, prm = driver_run["dir"], driver_run["params"]
    assert prm["fused"] is True
    assert prm["steps"] == STEPS
    rows = open(os.path.join(out, "outlet_concentration.csv")).read().strip().splitlines()
    assert rows[0] == "step,time_s,phi_out_avg,phi_out_min,phi_out_max"
    steps = [int(r.split(",")[0]) for r in rows[1:]]
    assert steps == list(range(0, STEPS + 1, CSV_EVERY))
    for s in (0, EVERY, 2 * EVERY):
        assert os.path.exists(os.path.join(out, f"step_{s:06d}.vti"))
    pvd = open(os.path.join(out, "out.pvd")).read()
    assert pvd.count("<DataSet") == 3
    dims, m = geometry.read_vti(os.path.join(out, f"step_{STEPS:06d}.vti"), 0.5)
    assert dims == driver_run["dims"]


def test_driver_matches_oracle(driver_run):
    prm, mask = driver_run["params"], driver_run["mask"]
    nx, ny, nz = driver_run["dims"]
    inlet, outlet = x_plane(nx, ny, nz, 0), x_plane(nx, ny, nz, nx - 1)
    porosity = np.where(mask == 1, prm["porosity_solid"], 1.0)
    if driver_run["inlet"] == "pulse":
        t_on, t_off, k = prm["pulse"]
        assert 0 < t_on < t_off and k > 0
        ad_inlet = R.BC(R.AD_TANH, inlet, (0.0, 1.0, k), i0=t_on, i1=t_off)
    else:
        ad_inlet = R.BC(R.AD_DIRICHLET, inlet, (1.0,))
    tau_field = np.where(mask == 1, prm["tau_solid"], prm["tau_fluid"])
    s = R.Setup(nx, ny, nz, periodic=(False, True, True), ns_op=R.POROUS, rho0=prm["rho0"], tau=prm["tau"],
                obstacle=None, porosity=porosity, phi0=0.0, tau_ad=prm["tau_fluid"], minC=1e-8, tau_field=tau_field,
                ns_bc=[R.BC(R.NS_NEQ, inlet, (prm["u_in"], 0.0, 0.0)), R.BC(R.NS_PRESSURE, outlet, (prm["rho0"],))],
                ad_bc=[ad_inlet, R.BC(R.AD_NEUMANN, outlet)], coupling=True, adsorb=True,
                isotherm=R.LINEAR, region=mask)
    s.KHenry, s.kA = prm["KHenry"], prm["k_ldf"]
    orc = R.System(s, 8)
    avg, mn, mx, _ = orc.run(STEPS, outlet)
    rows = [r.split(",") for r in open(os.path.join(driver_run["dir"], "outlet_concentration.csv")).read().split()[1:]]
    for r in rows[1:]:
        k = int(r[0]) - 1
        got = np.array([float(v) for v in r[2:5]])
        ref = np.array([avg[k], mn[k], mx[k]])
        assert np.allclose(got, ref, rtol=2e-5, atol=1e-12), (r[0], got, ref)
    phi = vti_array(os.path.join(driver_run["dir"], f"step_{STEPS:06d}.vti"), "concentration")
    ref_phi = orc.get(R.PHI)
    say(f"  driver vs oracle concentration: max |diff| {np.max(np.abs(phi - ref_phi)):.2e}, max {ref_phi.max():.3e}")
    assert np.max(np.abs(phi - ref_phi)) <= 1e-12 + 1e-12 * np.max(np.abs(ref_phi))
