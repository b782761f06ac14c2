"""CPU checks of the C ABI library: it loads, exports everything the header
declares, validates descriptors, and its host-side geometry / VTI code matches
the reference's own readVTI and buildXPlaneMask (built into oracle/_ref from the
reference headers; skipped where that build is absent)."""
import ctypes as C
import os

import numpy as np
import pytest

from fluidmaxxer_amd import abi, geometry
from fluidmaxxer_amd.coupled import CoupledConfig
from oracle import ref_cpu as R


def test_library_exports_every_header_symbol():
    L = abi.lib()
    names = abi.header_functions()
    assert len(names) >= 45
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(abi._SIGS), set(names) ^ set(abi._SIGS)
    assert L.fmx_abi_version() == 5 == abi.ABI_VERSION


def test_validate_rejects_bad_descriptors():
    n = 4 * 3 * 3
    cfg = CoupledConfig(4, 3, 3, solid=np.zeros(n, np.uint8), region=np.zeros(n, np.uint8))
    d = cfg.desc(abi.F32)
    b = abi.CoupledBufs()
    for i in range(2):
        b.f[i] = b.g[i] = 0x1000
    b.q = b.links = b.inlet_raw = 0x1000
    b.phi[0] = b.phi[1] = 0x1000
    assert abi.lib().fmx_coupled_validate(C.byref(d), C.byref(b)) == 0
    d.nx = 2
    assert abi.lib().fmx_coupled_validate(C.byref(d), C.byref(b)) == abi.lib().fmx_coupled_validate(C.byref(d), C.byref(b)) != 0
    assert b"nx >= 3" in abi.lib().fmx_last_error()
    d.nx = 4
    d.ad_outlet = 0
    assert abi.lib().fmx_coupled_validate(C.byref(d), C.byref(b)) == 3
    d.ad_outlet = 1
    b.inlet_raw = None
    assert abi.lib().fmx_coupled_validate(C.byref(d), C.byref(b)) == 1
    with pytest.raises(abi.FmxError):
        abi.call("fmx_coupled_run", C.byref(d), C.byref(b), 1, 1, 0, None)


def test_compact_workspace_size_and_validation():
    """fmx_coupled_compact_bytes: a row table of 16 bytes per (row, 64-cell segment), 256-byte aligned,
    the scan's chunk sums, a 256-byte header and an 8-byte region mask per (row, segment), then two
    compact copies of the 19 NS planes of the
    active cells rounded up to 128 cells, and for the split step an 8-byte box index and link word
    and a velocity (three values of the compute precision) per active cell; out-of-range arguments give -1.  A workspace below the
    table's size is refused by validation (no GPU needed)."""
    L = abi.lib()
    n = 130 * 6 * 5
    cfg = CoupledConfig(130, 6, 5, solid=np.zeros(n, np.uint8), region=np.zeros(n, np.uint8))
    ents = 6 * 5 * 3  # rows x ceil(130 / 64)
    head = (ents * 16 + 255) // 256 * 256 + 256 + 256 + (ents * 8 + 255) // 256 * 256  # table, chunk sums
    # (one chunk), header, sorption-region masks
    for prec, per in ((abi.F32, 2 * 19 * 4 + 8 + 12), (abi.F64, 2 * 19 * 8 + 8 + 24), (abi.F16, 2 * 19 * 2 + 8 + 12),
                      (abi.BF16, 2 * 19 * 2 + 8 + 12)):
        d = cfg.desc(prec)
        for active in (0, 1, 128, 129, n):
            want = head + per * ((active + 127) // 128 * 128)
            assert L.fmx_coupled_compact_bytes(C.byref(d), active) == want, (prec, active)
        assert L.fmx_coupled_compact_bytes(C.byref(d), -1) == -1
        assert L.fmx_coupled_compact_bytes(C.byref(d), n + 1) == -1
    # FMX_SPLIT=0 (the fused compact kernel): the NS planes only
    os.environ["FMX_SPLIT"] = "0"
    try:
        for prec, per in ((abi.F64, 2 * 19 * 8), (abi.F32, 2 * 19 * 4), (abi.F16, 2 * 19 * 2)):
            d = cfg.desc(prec)
            assert L.fmx_coupled_compact_bytes(C.byref(d), 200) == head + per * 256, prec
    finally:
        del os.environ["FMX_SPLIT"]
    d = cfg.desc(abi.F32)
    b = abi.CoupledBufs()
    for i in range(2):
        b.f[i] = b.g[i] = b.phi[i] = 0x1000
    b.q = b.links = b.inlet_raw = 0x1000
    assert L.fmx_coupled_validate(C.byref(d), C.byref(b)) == abi.OK
    b.compact_work, b.compact_work_bytes = 0x1000, L.fmx_coupled_compact_bytes(C.byref(d), 1) - 1
    assert L.fmx_coupled_validate(C.byref(d), C.byref(b)) == abi.ERR_ARG
    assert b"compact_work_bytes" in L.fmx_last_error()
    b.compact_work_bytes += 1
    assert L.fmx_coupled_validate(C.byref(d), C.byref(b)) == abi.OK
    d.freeze_solids = 2
    assert L.fmx_coupled_validate(C.byref(d), C.byref(b)) == abi.ERR_ARG


def test_recenter_refuses_a_far_reference_before_any_launch():
    """fmx_coupled_recenter moves the 16-bit density reference only within a factor of two of the current
    one (the shift h += f_ref - target is then exact in FP32) and never to a non-finite value: FMX_ERR_ARG
    before anything is launched (these calls run without a GPU).  A missing output pointer is refused too."""
    n = 8 * 4 * 4
    cfg = CoupledConfig(8, 4, 4, solid=np.zeros(n, np.uint8), region=np.zeros(n, np.uint8))
    L = abi.lib()
    b = abi.CoupledBufs()
    for i in range(2):
        b.f[i] = b.g[i] = b.phi[i] = 0x1000
    b.q = b.links = b.inlet_raw = 0x1000
    out = C.c_double(0.0)
    for prec in (abi.F16, abi.BF16):
        d = cfg.desc(prec)
        assert d.f_ref == 1.0
        for bad in (2.5, 0.4, float("inf")):
            assert L.fmx_coupled_recenter(C.byref(d), C.byref(b), 0, bad, C.byref(out), None, 0, None) == abi.ERR_ARG
        assert b"non-finite" in L.fmx_last_error()
        assert L.fmx_coupled_recenter(C.byref(d), C.byref(b), 0, 1.05, None, None, 0, None) == abi.ERR_ARG
        # the mean needs the reduction's workspace
        assert L.fmx_coupled_recenter(C.byref(d), C.byref(b), 0, float("nan"), C.byref(out), None, 0, None) == abi.ERR_ARG
        assert b"fmx_coupled_recenter_workspace_bytes" in L.fmx_last_error()
    d = cfg.desc(abi.F32)
    # FP32 storage: nothing to re-encode, the target is returned as given
    assert L.fmx_coupled_recenter(C.byref(d), C.byref(b), 0, 3.0, C.byref(out), None, 0, None) == abi.OK
    assert out.value == 3.0
    assert L.fmx_coupled_recenter_workspace_bytes(C.byref(d)) == (L.fmx_reduce_mass_workspace_bytes(n) + 32)


def test_short_buffers_are_refused_before_any_launch():
    """ABI 2 carries the lengths of the buffers a run indexes by step: a run over steps step0 ..
    step0+nsteps-1 needs step0+nsteps inlet-series entries and nsteps outlet rows, a reverse sweep
    nsteps loss weights (fmx_adjoint_step: step+1).  A short buffer is FMX_ERR_ARG, returned before
    anything is launched (these calls run without a GPU)."""
    n = 8 * 4 * 4
    cfg = CoupledConfig(8, 4, 4, solid=np.zeros(n, np.uint8), region=np.zeros(n, np.uint8), ad_inlet="series")
    L = abi.lib()
    d = cfg.desc(abi.F32)
    assert d.ad_inlet == 2
    b = abi.CoupledBufs()
    for i in range(2):
        b.f[i] = b.g[i] = b.phi[i] = 0x1000
    b.q = b.links = b.inlet_raw = b.inlet_series = 0x1000
    b.inlet_series_len = 10
    # steps 5 .. 14 read inlet_series[14]: 15 entries needed
    assert L.fmx_coupled_run(C.byref(d), C.byref(b), 5, 10, 0, None) == abi.ERR_ARG
    assert b"inlet_series_len is 10" in L.fmx_last_error()
    b.inlet_series_len = 15
    b.outlet_acc = 0x2000
    b.outlet_acc_rows = 9
    assert L.fmx_coupled_run(C.byref(d), C.byref(b), 5, 10, 0, None) == abi.ERR_ARG
    assert b"outlet_acc_rows is 9" in L.fmx_last_error()
    assert L.fmx_coupled_step(C.byref(d), C.byref(b), 15, 0, None) == abi.ERR_ARG  # reads inlet_series[15]
    rc = L.fmx_adjoint_forward(C.byref(d), C.byref(b), 5, 10, 4, 0, 0x3000, abi.F32, None)
    assert rc == abi.ERR_ARG and b"outlet_acc_rows" in L.fmx_last_error()
    b.outlet_acc_rows = 10
    b.inlet_series_len = 14
    rc = L.fmx_adjoint_forward(C.byref(d), C.byref(b), 5, 10, 4, 0, 0x3000, abi.F32, None)
    assert rc == abi.ERR_ARG and b"inlet_series" in L.fmx_last_error()
    b.inlet_series_len = 15
    a = abi.AdjointBufs()
    for i in range(2):
        a.lam_f[i] = a.lam_g[i] = 0x1000
    a.lam_q = a.lam_phi = a.grad_iso = a.workspace = 0x1000
    a.workspace_bytes = 1 << 40
    a.outlet_weight, a.outlet_weight_len = 0x4000, 9
    rc = L.fmx_adjoint_backward(C.byref(d), C.byref(b), 5, 10, 4, 0x3000, abi.F32, 0x5000, 4, C.byref(a), None)
    assert rc == abi.ERR_ARG and b"outlet_weight has 9 entries, 10 needed" in L.fmx_last_error()
    a.outlet_weight_len = 10
    b.inlet_series_len = 14
    rc = L.fmx_adjoint_backward(C.byref(d), C.byref(b), 5, 10, 4, 0x3000, abi.F32, 0x5000, 4, C.byref(a), None)
    assert rc == abi.ERR_ARG and b"inlet_series" in L.fmx_last_error()
    b.inlet_series_len = 15
    a.outlet_weight_len = 12
    rc = L.fmx_adjoint_step(C.byref(d), C.byref(b), 0, 12, 0, 1, C.byref(a), None)  # reads outlet_weight[12]
    assert rc == abi.ERR_ARG and b"outlet_weight has 12 entries, 13 needed" in L.fmx_last_error()


def test_adjoint_state_bytes_and_argument_checks():
    n = 8 * 4 * 4
    cfg = CoupledConfig(8, 4, 4, solid=np.zeros(n, np.uint8), region=np.zeros(n, np.uint8))
    L = abi.lib()

    def al(b):
        return (b + 255) // 256 * 256

    for prec, s, c in ((abi.F32, 4, 4), (abi.F64, 8, 8), (abi.F16, 2, 4), (abi.BF16, 2, 4)):
        d = cfg.desc(prec)
        assert L.fmx_adjoint_state_bytes(C.byref(d)) == al(19 * n * s) + al(7 * n * s) + 2 * al(n * c)
    b = abi.CoupledBufs()
    for i in range(2):
        b.f[i] = b.g[i] = b.phi[i] = 0x1000
    b.q = b.links = b.inlet_raw = 0x1000
    a = abi.AdjointBufs()
    for i in range(2):
        a.lam_f[i] = a.lam_g[i] = 0x1000
    a.lam_q = a.lam_phi = a.grad_iso = 0x1000
    a.workspace = 0x1000
    a.workspace_bytes = 1 << 40

    def backward(desc, k=5):
        return L.fmx_adjoint_backward(C.byref(desc), C.byref(b), 1, 10, k, 0x1000, desc.prec, 0x1000, k,
                                      C.byref(a), None)

    # 16-bit checkpoints only for an FP32 run
    d = cfg.desc(abi.F32)
    for cp, s in ((abi.F32, 4), (abi.F16, 2), (abi.BF16, 2)):
        assert L.fmx_adjoint_ckpt_bytes(C.byref(d), cp) == al(19 * n * s) + al(7 * n * s) + 2 * al(n * 4)
    assert L.fmx_adjoint_ckpt_bytes(C.byref(d), abi.F64) == -1
    d64 = cfg.desc(abi.F64)
    assert L.fmx_adjoint_ckpt_bytes(C.byref(d64), abi.F16) == -1
    assert L.fmx_adjoint_ckpt_bytes(C.byref(d64), abi.F64) == L.fmx_adjoint_state_bytes(C.byref(d64))
    # FP32 checkpoints of an FP64 run: FP32 state slots; its FP32 sweep needs the FP32 field copies
    assert L.fmx_adjoint_ckpt_bytes(C.byref(d64), abi.F32) == L.fmx_adjoint_state_bytes(C.byref(d))
    assert L.fmx_adjoint_ckpt_bytes(C.byref(d), abi.F64) == -1
    b.field = 0x1000
    rc = L.fmx_adjoint_backward(C.byref(d64), C.byref(b), 1, 10, 5, 0x1000, abi.F32, 0x1000, 4, C.byref(a), None)
    assert rc == abi.ERR_ARG and b"field_rev" in L.fmx_last_error()
    b.field = None
    rc = L.fmx_adjoint_backward(C.byref(d64), C.byref(b), 1, 10, 5, 0x1000, abi.F16, 0x1000, 5, C.byref(a), None)
    assert rc != 0 and b"checkpoint precision" in L.fmx_last_error()
    # a 16-bit checkpoint store needs a segment store even for k = 1
    rc = L.fmx_adjoint_backward(C.byref(d), C.byref(b), 1, 10, 1, 0x1000, abi.F16, None, 0, C.byref(a), None)
    assert rc != 0 and b"segment store" in L.fmx_last_error()
    # segment slots: one level k - 1 (k with 16-bit checkpoints); two levels about 2 sqrt(k)
    assert L.fmx_adjoint_seg_slots(64, 0, 0) == 63 and L.fmx_adjoint_seg_slots(64, 1, 0) == 64
    assert L.fmx_adjoint_seg_slots(64, 0, 1) == 14 and L.fmx_adjoint_seg_slots(64, 1, 1) == 15
    assert L.fmx_adjoint_seg_slots(500, 0, 1) == 43 and L.fmx_adjoint_seg_slots(1, 0, 1) == 0
    assert L.fmx_adjoint_seg_slots(0, 0, 0) == -1
    # workspace: 2 doubles per launch block, 66 with a per-step inlet series; a short one is refused
    ws = L.fmx_adjoint_workspace_bytes(C.byref(d))
    assert ws > 0 and ws % 16 == 0
    d.ad_inlet = 2
    assert L.fmx_adjoint_workspace_bytes(C.byref(d)) == ws // 16 * (66 * 8)
    d.ad_inlet = cfg.desc(abi.F32).ad_inlet
    a.workspace_bytes = ws - 8
    rc = L.fmx_adjoint_backward(C.byref(d), C.byref(b), 1, 10, 5, 0x1000, abi.F32, 0x1000, 5, C.byref(a), None)
    assert rc != 0 and b"workspace" in L.fmx_last_error()
    a.workspace_bytes = 1 << 40
    rc = L.fmx_adjoint_backward(C.byref(d), C.byref(b), 1, 200, 64, 0x1000, abi.F32, 0x1000, 13, C.byref(a), None)
    assert rc != 0 and b"at least 14 segment slots" in L.fmx_last_error()
    assert backward(d, k=0) != 0 and b"interval" in L.fmx_last_error()
    a.grad_series = 0x1000
    assert backward(d) != 0 and b"grad_series" in L.fmx_last_error()
    a.grad_series = None
    a.grad_iso = None
    assert backward(d) != 0 and b"adjoint buffer" in L.fmx_last_error()


def test_sphere_pack_deterministic_and_on_target():
    a = geometry.sphere_pack(64, 32, 32, 1, 4, 4, 0.6, overlap=False)
    b = geometry.sphere_pack(64, 32, 32, 1, 4, 4, 0.6, overlap=False)
    c = geometry.sphere_pack(64, 32, 32, 2, 4, 4, 0.6, overlap=False)
    assert np.array_equal(a.solid, b.solid) and not np.array_equal(a.solid, c.solid)
    assert abs(a.porosity - 0.6) < 0.01
    assert np.isclose(a.porosity, 1 - a.solid.mean())
    s = a.solid.reshape(32, 32, 64)
    assert s[0].all() and s[-1].all() and s[:, 0].all() and s[:, -1].all()   # z and y walls
    sph = a.sphere.reshape(32, 32, 64)
    assert not sph[:, :, :2].any() and not sph[:, :, -2:].any()                # inlet/outlet planes stay clear
    assert (a.solid >= a.sphere).all()
    d = geometry.sphere_pack(96, 48, 48, 2, 4, 8, 0.4, overlap=True)
    assert abs(d.porosity - 0.4) < 0.01


def test_x_plane_mask_matches_reference():
    m = geometry.x_plane_mask(7, 5, 3, 6)
    assert m.sum() == 15 and m.reshape(3, 5, 7)[:, :, 6].all()
    if not R.ref_vti_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    assert np.array_equal(m, R.ref_build_x_plane_mask(7, 5, 3, 6))


def _write(path, text):
    with open(path, "w") as fh:
        fh.write(text)


def test_vti_reader_matches_reference(tmp_path):
    if not R.ref_vti_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    p = geometry.sphere_pack(20, 10, 12, 7, 2, 3, 0.5)
    f1 = str(tmp_path / "scalar.vti")
    geometry.write_vti(f1, 20, 10, 12, p.solid)
    cases = [f1]
    rng = np.random.default_rng(0)
    vals = rng.random(4 * 3 * 2)
    body = " ".join(f"{v:.6f}" for v in vals)
    f2 = str(tmp_path / "float.vti")
    _write(f2, f'<VTKFile><ImageData WholeExtent="0 3 0 2 0 1"><Piece Extent="0 3 0 2 0 1"><PointData>'
               f'<DataArray type="Float32" Name="a" format="ascii">{body}</DataArray></PointData></Piece>'
               f'</ImageData></VTKFile>')
    cases.append(f2)
    v3 = rng.random(3 * 24) * 0.6
    f3 = str(tmp_path / "vec.vti")
    _write(f3, '<VTKFile><ImageData WholeExtent="1 4 0 2 0 1"><PointData><DataArray NumberOfComponents="3">'
               + " ".join(f"{v:.5f}" for v in v3) + "</DataArray></PointData></ImageData></VTKFile>")
    cases.append(f3)
    f4 = str(tmp_path / "short.vti")
    _write(f4, '<VTKFile><ImageData><Piece Extent="0 3 0 2 0 1"><DataArray>1 0 1 0.7 0.2 end 1 1'
               "</DataArray></Piece></ImageData></VTKFile>")
    cases.append(f4)
    for path in cases:
        dims, mine = geometry.read_vti(path, 0.5)
        rdims, ref = R.ref_read_vti(path, 0.5)
        assert dims == rdims, path
        assert np.array_equal(mine, ref), path


def test_vti_reader_errors(tmp_path):
    f = str(tmp_path / "bad.vti")
    _write(f, "<VTKFile><ImageData></ImageData></VTKFile>")
    with pytest.raises(abi.FmxError, match="Extent"):
        geometry.read_vti(f)
    _write(f, '<VTKFile><ImageData WholeExtent="0 1 0 1 0 1"></ImageData></VTKFile>')
    with pytest.raises(abi.FmxError, match="DataArray"):
        geometry.read_vti(f)
    with pytest.raises(abi.FmxError, match="could not open"):
        geometry.read_vti(str(tmp_path / "missing.vti"))


def test_header_is_plain_c_and_links(tmp_path):
    """include/fmx_abi.h is C99 (any FFI can read it) and a C program calling the fused path, the
    checkpoint and adjoint entry points links against libfmx_hip.so (compile and link only)."""
    import subprocess

    src = tmp_path / "use_abi.c"
    src.write_text(r'''
#include <stdio.h>
#include "fmx_abi.h"
int main(void) {
    fmx_coupled_desc d = {0};
    fmx_coupled_bufs b = {0};
    fmx_adjoint_bufs a = {0};
    d.nx = 8; d.ny = 4; d.nz = 4; d.prec = FMX_F32; d.ns_op = FMX_NS_BGK;
    if (fmx_coupled_validate(&d, &b) == 0) return 1;
    printf("%s %lld %lld\n", fmx_last_error(), (long long)fmx_adjoint_workspace_bytes(&d),
           (long long)FMX_OUTLET_RING_ROWS);
    (void)fmx_coupled_step; (void)fmx_checkpoint_save; (void)fmx_checkpoint_load; (void)fmx_adjoint_step;
    (void)fmx_reduce_outlet_mean; (void)fmx_ad_bc_dirichlet_series; (void)a;
    return 0;
}
''')
    exe = tmp_path / "use_abi"
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(abi.ROOT, "include"), str(src),
                        "-o", str(exe), "-L", os.path.dirname(abi.LIB_PATH), "-lfmx_hip"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_ctypes_structs_match_the_header_layout(tmp_path):
    """Every struct the bindings mirror has the header's size and field offsets: a C program
    compiled against include/fmx_abi.h prints offsetof / sizeof for each ctypes field."""
    import subprocess

    structs = {"fmx_grid": abi.Grid, "fmx_coupled_desc": abi.CoupledDesc, "fmx_coupled_bufs": abi.CoupledBufs,
               "fmx_adjoint_bufs": abi.AdjointBufs}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "fmx_abi.h"', 'int main(void) {']
    for cname, py in structs.items():
        lines.append(f'    printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'    printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ['    return 0;', '}']
    src, exe = tmp_path / "layout.c", tmp_path / "layout"
    src.write_text("\n".join(lines) + "\n")
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(abi.ROOT, "include"), str(src),
                        "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        cname, fname, v = line.split()
        got[(cname, fname)] = int(v)
    for cname, py in structs.items():
        assert got[(cname, "sizeof")] == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
