"""The host-side physical-unit layer (no device): UnitConverter, ADUnitConverter and main.cpp's
correlations against their closed forms, and the VTI writer read back by this repository's reader
(ASCII and raw appended binary) and by the reference's own readVTI (ASCII; built from the
reference headers into oracle/_ref, skipped where that build is absent).

The closed forms (main.cpp:43-99, UnitConverter.hpp:66-88, ADUnitConverter.hpp:47-60):
  D_eff = eps/(2-eps)^2 D_m;  u_s = Pe D/d_p eps;  u_i = u_s/eps;  dx = L/res;  nu_lbm = (tau-1/2)/3;
  dt = nu_lbm dx^2/nu;  u_lbm = u_i dt/dx;  Re = u_i L/nu;  Ma = u_lbm/sqrt(1/3);  Kn = Ma/Re;
  Sh = 1.09 eps_p (eps_p Re)^0.33 Sc^0.33 with Sc = nu/D;  k_film = Sh D/d_p;
  k_ldf = 3(1-eps)/(d_p/2) k_film dt;  tau_AD = D dt/dx^2 / (1/4) + 1/2.
"""
import json
import math
import os
import subprocess

import numpy as np
import pytest

from fluidmaxxer_amd import geometry
from oracle import ref_cpu as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "tests", "cpp", "_build")


def run_units(dt, tmp_path):
    exe = os.path.join(BUILD, f"host_units_{dt}")
    if not os.path.exists(exe):
        pytest.skip("tests/cpp programs not built (python -m fluidmaxxer_amd.build)")
    paths = [str(tmp_path / n) for n in ("a.vti", "b.vti", "out.pvd")]
    r = subprocess.run([exe, *paths], capture_output=True, text=True, check=True)
    out = {}
    for line in r.stdout.splitlines():
        if line.startswith("{"):
            out.update(json.loads(line))
    return out, paths


def closed_forms():
    dp, Pe, nu, Dm = 2e-4, 40.0, 1e-6, 2e-9
    eps, epsp, res, tau = 0.26, 0.67, 15, 0.75
    Deff = epsp / (2 - epsp) ** 2 * Dm
    us = Pe * Dm / dp * eps
    ui = us / eps
    dx = dp / res
    nu_l = (tau - 0.5) / 3.0
    dt = nu_l * dx * dx / nu
    u_l = ui * dt / dx
    Re = ui * dp / nu
    Ma = u_l / math.sqrt(1 / 3)
    Sc = nu / Dm
    Sh = 1.09 * epsp * (epsp * Re) ** 0.33 * Sc ** 0.33
    kf = Sh * Dm / dp
    kldf = 3 * (1 - eps) / (dp / 2) * kf * dt
    tau_ad = lambda D: D * dt / dx ** 2 / 0.25 + 0.5  # noqa: E731
    return dict(Deff=Deff, u_superficial=us, u_interstitial=ui, dx=dx, dt=dt, nu_lbm=nu_l, u_lbm=u_l, Re=Re, Ma=Ma,
                Kn=Ma / Re, kfilm=kf, kldf=kldf, tau_fluid=tau_ad(Dm), tau_solid=tau_ad(Deff), tau_field0=tau_ad(Dm),
                tau_field_mask=tau_ad(Deff), len_p2l=1e-3 / dx, time_l2p=1000 * dt, visc_p2l=1e-6 * dt / dx ** 2,
                dens_l2p=1.02 * 1000.0, vel_l2p=0.01 * dx / dt)


@pytest.mark.parametrize("dt,tol", [("double", 1e-12), ("float", 2e-6)])
def test_unit_converters_and_correlations(dt, tol, tmp_path):
    got, _ = run_units(dt, tmp_path)
    want = closed_forms()
    bad = {k: (got[k], v) for k, v in want.items() if not abs(got[k] - v) <= tol * abs(v)}
    assert not bad, bad
    # main.cpp's scenario sits at tau_AD barely above 1/2 (D = 2e-9 m^2/s at dx = 13.3 um)
    assert 0.5 < got["tau_solid"] < got["tau_fluid"] < 0.501


def test_vti_writer_round_trip(tmp_path):
    _, (ascii_path, binary_path, pvd) = run_units("double", tmp_path)
    mask = np.array([(i * 7 + 3) % 5 < 2 for i in range(60)], np.uint8)
    for path in (ascii_path, binary_path):
        dims, m = geometry.read_vti(path, 0.5)
        assert dims == (5, 4, 3), path
        assert np.array_equal(m, mask), path
    if R.ref_vti_available():
        rdims, rm = R.ref_read_vti(ascii_path, 0.5)
        assert rdims == (5, 4, 3) and np.array_equal(rm, mask)
    text = open(pvd).read()
    assert text.count("<DataSet") == 2 and 'file="a.vti"' in text and 'timestep="0.5"' in text
    assert text.rstrip().endswith("</VTKFile>")


def test_vti_reader_binary_forms(tmp_path):
    """Inline base64 binary and appended raw arrays of several types, read by fmx_vti_read_mask."""
    import base64
    import struct

    vals = np.array([0.1, 0.9, 0.4, 0.6, 1.2, 0.0, 0.51, 0.49], np.float32)
    want = (vals > 0.5).astype(np.uint8)
    head = '<VTKFile type="ImageData" byte_order="LittleEndian"{ht}><ImageData WholeExtent="0 1 0 1 0 1"><PointData>'
    # inline base64, UInt32 header
    blob = struct.pack("<I", vals.nbytes) + vals.tobytes()
    f1 = tmp_path / "b64.vti"
    f1.write_text(head.format(ht="") + '<DataArray type="Float32" format="binary">'
                  + base64.b64encode(blob).decode() + "</DataArray></PointData></ImageData></VTKFile>")
    # appended raw, UInt64 header, Float64 and UInt8 arrays (the first array is read)
    d64 = vals.astype(np.float64)
    f2 = tmp_path / "app.vti"
    with open(f2, "wb") as fh:
        fh.write((head.format(ht=' header_type="UInt64"') + '<DataArray type="Float64" format="appended" offset="0"/>'
                  '</PointData></ImageData><AppendedData encoding="raw">_').encode())
        fh.write(struct.pack("<Q", d64.nbytes) + d64.tobytes())
        fh.write(b"</AppendedData></VTKFile>")
    u8 = (vals * 10).astype(np.uint8)
    f3 = tmp_path / "u8.vti"
    with open(f3, "wb") as fh:
        fh.write((head.format(ht="") + '<DataArray type="UInt8" format="appended" offset="0"/>'
                  '</PointData></ImageData><AppendedData encoding="raw">_').encode())
        fh.write(struct.pack("<I", u8.nbytes) + u8.tobytes())
        fh.write(b"</AppendedData></VTKFile>")
    for path, thr, expect in ((f1, 0.5, want), (f2, 0.5, want), (f3, 4.5, (u8 > 4.5).astype(np.uint8))):
        dims, m = geometry.read_vti(str(path), thr)
        assert dims == (2, 2, 2)
        assert np.array_equal(m, expect), path


def test_fp16_decode_all_patterns(tmp_path):
    """The FP16 decode of the 16-bit population storage (fmx_half.hpp, compiled for the host) against
    IEEE binary16 for all 65,536 bit patterns: bit-exact for every finite half (subnormals, +-0, the
    largest 65504), +-Inf decoded as +-Inf and every NaN as a NaN (not as a finite 2^16 (1 + m))."""
    exe = os.path.join(BUILD, "half_units")
    if not os.path.exists(exe):
        pytest.skip("tests/cpp programs not built (python -m fluidmaxxer_amd.build)")
    out = tmp_path / "half.f32"
    subprocess.run([exe, str(out)], check=True)
    got = np.fromfile(out, dtype=np.float32)
    ref = np.arange(65536, dtype=np.uint32).astype(np.uint16).view(np.float16).astype(np.float32)
    finite = np.isfinite(ref)
    assert finite.sum() == 63488
    assert np.array_equal(got[finite].view(np.uint32), ref[finite].view(np.uint32))
    assert np.array_equal(np.isnan(got), np.isnan(ref)) and np.array_equal(np.isinf(got), np.isinf(ref))
    assert got[0x7c00] == np.inf and got[0xfc00] == -np.inf and got[0x7bff] == 65504.0
    assert np.array_equal(np.signbit(got[~np.isnan(ref)]), np.signbit(ref[~np.isnan(ref)]))


@pytest.mark.parametrize("dt,prec", [("double", 8), ("float", 4)])
def test_tanh_pulse_value_matches_oracle(dt, prec, tmp_path):
    """ADDirichletTanhPulse::valueAt (the host class the operator path and CoupledStep's fused series
    use) against the oracle's restatement of ADDirichletTanhPulse.cu:72-83 -- the pulse in double, cast
    to DATA_TYPE, times amp, plus base -- at steps across both edges, to the bit; and the pulse's
    closed-form landmarks: base + amp/2 at t_on and t_off, base + amp on the plateau."""
    got, _ = run_units(dt, tmp_path)
    base, amp, t_on, t_off, k = 0.1, 0.9, 300, 900, 40.0
    for step, v in got["pulse"]:
        assert v == R.tanh_pulse(prec, step, base, amp, t_on, t_off, k), step
    vals = dict((s, v) for s, v in got["pulse"])
    tol = 1e-15 if prec == 8 else 1e-7
    assert abs(vals[300] - (base + amp * 0.5 * (0.0 - math.tanh(-600 / k)))) < tol
    assert abs(vals[900] - (base + amp * 0.5 * (math.tanh(600 / k) - 0.0))) < tol
    assert abs(vals[600] - (base + amp)) < 1e-6 and abs(vals[5000] - base) < 1e-6
