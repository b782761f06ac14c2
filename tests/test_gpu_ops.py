"""Each gfx950 operator kernel against its oracle restatement, one application on
seeded random inputs (perturbed equilibria, random masks), FP32 and FP64.

Tolerances: FP32 <= 2e-6 and FP64 <= 1e-13 relative L-infinity error (the
kernels and the oracle evaluate the same expressions; the difference is FMA
contraction and division rounding on the GPU)."""
import ctypes as C

import numpy as np
import pytest

from fluidmaxxer_amd import abi
from oracle import ref_cpu as R
from tests.gpu_util import dev, host, say, stream, sync, torch
from tests.lattice import W19, W7, aos_to_soa, coords, rel_err, soa_to_aos, x_plane

pytestmark = pytest.mark.gpu

PRECS = [(abi.F32, 4, 2e-6), (abi.F64, 8, 1e-13)]


def _tdt(prec):
    t = torch()
    return t.float32 if prec == abi.F32 else t.float64


def _rand_f(n, seed, amp=0.05):
    rng = np.random.default_rng(seed)
    return np.tile(W19, n) * (1 + amp * rng.standard_normal(n * 19))


@pytest.mark.parametrize("prec,oprec,tol", PRECS)
@pytest.mark.parametrize("op", [abi.NS_BGK, abi.NS_POROUS, abi.NS_PBB])
@pytest.mark.parametrize("periodic", [(False, True, True), (True, False, True), (False, False, False)])
def test_ns_collide_stream(prec, oprec, tol, op, periodic):
    dims = (13, 9, 7)
    n = 13 * 9 * 7
    say(f"ns_collide op={op} prec={prec} periodic={periodic} dims={dims}")
    rng = np.random.default_rng(op * 10 + prec)
    f = _rand_f(n, 1 + op)
    obst = (rng.random(n) < 0.3).astype(np.uint8)
    field = rng.uniform(-0.1, 1.1, n)  # exercises both clamps
    tdt = _tdt(prec)
    fin = dev(aos_to_soa(f, 19).ravel(), tdt)
    fout = torch().empty_like(fin)
    ob = dev(obst, torch().uint8) if op == abi.NS_BGK else None
    fl = dev(field, tdt) if op != abi.NS_BGK else None
    g = abi.grid(*dims, periodic)
    abi.call("fmx_ns_collide_stream", prec, op, fin.data_ptr(), fout.data_ptr(), ob.data_ptr() if ob is not None else None,
             fl.data_ptr() if fl is not None else None, C.byref(g), 0.62, 1.0, stream())
    sync()
    ref = R.op_ns_collide(oprec, {abi.NS_BGK: R.BGK, abi.NS_POROUS: R.POROUS, abi.NS_PBB: R.PBB}[op], dims, periodic, f,
                          obstacle=obst if op == abi.NS_BGK else None, field=field if op != abi.NS_BGK else None,
                          tau=0.62, rho0=1.0)
    assert rel_err(soa_to_aos(host(fout), 19), ref) < tol


@pytest.mark.parametrize("prec,oprec,tol", PRECS)
def test_ns_collide_density_fallback_and_tau_clamp(prec, oprec, tol):
    """rho <= 0 takes the rho0 / u = 0 branch; tau below 0.5001 is clamped."""
    dims = (6, 5, 4)
    n = 120
    say(f"ns_collide fallback prec={prec}")
    f = _rand_f(n, 3)
    f[: 19 * 10] = -np.abs(f[: 19 * 10])  # negative density in the first cells
    tdt = _tdt(prec)
    for op, rop in ((abi.NS_BGK, R.BGK), (abi.NS_POROUS, R.POROUS)):
        fin = dev(aos_to_soa(f, 19).ravel(), tdt)
        fout = torch().empty_like(fin)
        g = abi.grid(*dims, (True, True, True))
        abi.call("fmx_ns_collide_stream", prec, op, fin.data_ptr(), fout.data_ptr(), None, None, C.byref(g), 0.3, 1.2,
                 stream())
        sync()
        ref = R.op_ns_collide(oprec, rop, dims, (True, True, True), f, tau=0.3, rho0=1.2)
        assert rel_err(soa_to_aos(host(fout), 19), ref) < 10 * tol


@pytest.mark.parametrize("prec,oprec,tol", PRECS)
def test_ns_moments(prec, oprec, tol):
    dims = (11, 6, 5)
    n = 330
    say(f"ns_moments prec={prec}")
    f = _rand_f(n, 4)
    f[:19] = 0.0  # zero density -> u = 0 branch
    tdt = _tdt(prec)
    fin = dev(aos_to_soa(f, 19).ravel(), tdt)
    outs = [torch().empty(n, dtype=tdt, device="cuda") for _ in range(4)]
    g = abi.grid(*dims)
    abi.call("fmx_ns_moments", prec, fin.data_ptr(), *(o.data_ptr() for o in outs), C.byref(g), stream())
    sync()
    ref = R.op_ns_moments(oprec, dims, f)
    for o, r in zip(outs, ref):
        assert rel_err(host(o), r) < tol


def _lists_neq(mask, dims):
    nx, ny, nz = dims
    b = np.nonzero(mask)[0].astype(np.int64)
    x = b % nx
    step = np.where(x == 0, 1, -1)
    return b, b + step


@pytest.mark.parametrize("prec,oprec,tol", PRECS)
@pytest.mark.parametrize("plane", [0, -1])
def test_ns_velocity_neq(prec, oprec, tol, plane):
    dims = (9, 6, 5)
    n = 270
    say(f"ns_bc_neq prec={prec} plane={plane}")
    f = _rand_f(n, 5)
    mask = x_plane(*dims, 0 if plane == 0 else dims[0] - 1)
    bnd, nei = _lists_neq(mask, dims)
    tdt = _tdt(prec)
    fd = dev(aos_to_soa(f, 19).ravel(), tdt)
    bd, nd = dev(bnd, torch().int64), dev(nei, torch().int64)
    abi.call("fmx_ns_bc_velocity_neq", prec, fd.data_ptr(), bd.data_ptr(), nd.data_ptr(), len(bnd), 0.05, 0.01, -0.02, n,
             stream())
    sync()
    ref = R.op_ns_bc(oprec, R.NS_NEQ, dims, f, mask, u=(0.05, 0.01, -0.02))
    assert rel_err(soa_to_aos(host(fd), 19), ref) < tol


@pytest.mark.parametrize("prec,oprec,tol", PRECS)
def test_ns_velocity_eq_pressure_bounceback(prec, oprec, tol):
    dims = (9, 6, 5)
    n = 270
    say(f"ns_bc eq/pressure/bounceback prec={prec}")
    f = _rand_f(n, 6)
    tdt = _tdt(prec)
    inlet, outlet = x_plane(*dims, 0), x_plane(*dims, 8)
    cells_in = dev(np.nonzero(inlet)[0].astype(np.int64), torch().int64)
    # Eq velocity
    fd = dev(aos_to_soa(f, 19).ravel(), tdt)
    abi.call("fmx_ns_bc_velocity_eq", prec, fd.data_ptr(), cells_in.data_ptr(), cells_in.numel(), 1.1, 0.04, 0.0, 0.01,
             n, stream())
    sync()
    assert rel_err(soa_to_aos(host(fd), 19), R.op_ns_bc(oprec, R.NS_EQ, dims, f, inlet, u=(0.04, 0, 0.01), a=1.1)) < tol
    # pressure
    oc = np.nonzero(outlet)[0].astype(np.int64)
    fd = dev(aos_to_soa(f, 19).ravel(), tdt)
    c1, c2 = dev(oc, torch().int64), dev(oc - 1, torch().int64)
    abi.call("fmx_ns_bc_pressure", prec, fd.data_ptr(), c1.data_ptr(), c2.data_ptr(), len(oc), 1.02, n, stream())
    sync()
    assert rel_err(soa_to_aos(host(fd), 19), R.op_ns_bc(oprec, R.NS_PRESSURE, dims, f, outlet, a=1.02)) < tol
    # bounce-back swap on a random mask
    rng = np.random.default_rng(7)
    m = (rng.random(n) < 0.4).astype(np.uint8)
    fd = dev(aos_to_soa(f, 19).ravel(), tdt)
    cm = dev(np.nonzero(m)[0].astype(np.int64), torch().int64)
    abi.call("fmx_ns_bc_bounceback", prec, fd.data_ptr(), cm.data_ptr(), cm.numel(), n, stream())
    sync()
    assert rel_err(soa_to_aos(host(fd), 19), R.op_ns_bc(oprec, R.NS_BOUNCEBACK, dims, f, m)) < tol


@pytest.mark.parametrize("prec,oprec,tol", PRECS)
@pytest.mark.parametrize("periodic", [(False, True, True), (True, True, False)])
def test_ad_collide_stream(prec, oprec, tol, periodic):
    dims = (12, 7, 6)
    n = 504
    say(f"ad_collide prec={prec} periodic={periodic}")
    rng = np.random.default_rng(8)
    phi = rng.uniform(0, 1, n)
    phi[:5] = 1e-12  # clamp branch
    g = soa_to_aos(np.outer(W7, phi) * (1 + 0.05 * rng.standard_normal((7, n))), 7)
    tau = rng.uniform(0.52, 1.2, n)
    u = [rng.uniform(-0.05, 0.05, n) for _ in range(3)]
    S = rng.uniform(-0.01, 0.01, n)
    tdt = _tdt(prec)
    gd = dev(aos_to_soa(g, 7).ravel(), tdt)
    go = torch().empty_like(gd)
    ph = torch().empty(n, dtype=tdt, device="cuda")
    td, ud, sd = dev(tau, tdt), [dev(x, tdt) for x in u], dev(S, tdt)
    gr = abi.grid(*dims, periodic)
    abi.call("fmx_ad_collide_stream", prec, gd.data_ptr(), go.data_ptr(), ph.data_ptr(), td.data_ptr(),
             *(x.data_ptr() for x in ud), sd.data_ptr(), C.byref(gr), 1e-8, stream())
    sync()
    rg, rphi = R.op_ad_collide(oprec, dims, periodic, g, tau, *u, S, minC=1e-8)
    assert rel_err(soa_to_aos(host(go), 7), rg) < tol
    assert rel_err(host(ph), rphi) < tol


@pytest.mark.parametrize("prec,oprec,tol", PRECS)
def test_ad_dirichlet_neumann(prec, oprec, tol):
    dims = (8, 5, 4)
    n = 160
    say(f"ad_bc prec={prec}")
    rng = np.random.default_rng(9)
    g = rng.random(n * 7)
    tdt = _tdt(prec)
    inlet, outlet = x_plane(*dims, 0), x_plane(*dims, 7)
    gd = dev(aos_to_soa(g, 7).ravel(), tdt)
    ci = dev(np.nonzero(inlet)[0].astype(np.int64), torch().int64)
    abi.call("fmx_ad_bc_dirichlet", prec, gd.data_ptr(), ci.data_ptr(), ci.numel(), 0.8, n, stream())
    sync()
    assert rel_err(soa_to_aos(host(gd), 7), R.op_ad_bc(oprec, R.AD_DIRICHLET, dims, g, inlet, a=0.8)) < tol
    oc = np.nonzero(outlet)[0].astype(np.int64)
    gd = dev(aos_to_soa(g, 7).ravel(), tdt)
    c1, c2 = dev(oc, torch().int64), dev(oc - 1, torch().int64)
    abi.call("fmx_ad_bc_neumann", prec, gd.data_ptr(), c1.data_ptr(), c2.data_ptr(), len(oc), n, stream())
    sync()
    assert rel_err(soa_to_aos(host(gd), 7), R.op_ad_bc(oprec, R.AD_NEUMANN, dims, g, outlet)) < tol


@pytest.mark.parametrize("prec,oprec,tol", PRECS)
@pytest.mark.parametrize("iso", [abi.ISO_LINEAR, abi.ISO_LANGMUIR])
def test_ldf(prec, oprec, tol, iso):
    n = 3000
    say(f"ldf prec={prec} iso={iso}")
    rng = np.random.default_rng(10 + iso)
    region = (rng.random(n) < 0.6).astype(np.uint8)
    phi = rng.uniform(-0.1, 1.0, n)
    q = rng.uniform(0, 1, n)
    S = rng.uniform(-1, 1, n)
    tdt = _tdt(prec)
    rd, pd, sd, qd = dev(region, torch().uint8), dev(phi, tdt), dev(S, tdt), dev(q, tdt)
    if iso == abi.ISO_LINEAR:
        abi.call("fmx_ldf_linear", prec, n, rd.data_ptr(), pd.data_ptr(), sd.data_ptr(), qd.data_ptr(), 0.5, 0.01, 1e-8,
                 stream())
        rs, rq = R.op_ldf(oprec, R.LINEAR, region, phi, S, q, 0.5, 0.0, 0.01)
    else:
        abi.call("fmx_ldf_langmuir", prec, n, rd.data_ptr(), pd.data_ptr(), sd.data_ptr(), qd.data_ptr(), 2.0, 5.0,
                 0.02, 1e-8, stream())
        rs, rq = R.op_ldf(oprec, R.LANGMUIR, region, phi, S, q, 2.0, 5.0, 0.02)
    sync()
    assert rel_err(host(sd), rs) < tol and rel_err(host(qd), rq) < tol


@pytest.mark.parametrize("prec", [abi.F32, abi.F64])
def test_init_pack_unpack(prec):
    dims = (7, 5, 3)
    n = 105
    say(f"init/pack prec={prec}")
    tdt = _tdt(prec)
    f = torch().empty(19 * n, dtype=tdt, device="cuda")
    g = abi.grid(*dims)
    abi.call("fmx_ns_init_eq", prec, f.data_ptr(), C.byref(g), 1.3, stream())
    gg = torch().empty(7 * n, dtype=tdt, device="cuda")
    abi.call("fmx_ad_init_eq", prec, gg.data_ptr(), C.byref(g), 0.25, stream())
    aos = torch().empty_like(f)
    abi.call("fmx_pack_aos", prec, f.data_ptr(), aos.data_ptr(), 19, n, stream())
    back = torch().empty_like(f)
    abi.call("fmx_unpack_aos", prec, aos.data_ptr(), back.data_ptr(), 19, n, stream())
    sync()
    w = W19.astype(np.float32) if prec == abi.F32 else W19
    assert np.allclose(host(aos), np.tile(w * (np.float32(1.3) if prec == abi.F32 else 1.3), n), rtol=1e-7)
    assert np.array_equal(host(back), host(f))
    assert np.allclose(aos_to_soa(soa_to_aos(host(gg), 7), 7).sum(0), 0.25, rtol=1e-7)


@pytest.mark.parametrize("prec,oprec,tol", PRECS)
def test_ad_dirichlet_series_and_outlet_stats(prec, oprec, tol):
    """The inlet value read on the device from series[step] (ADDirichletTanhPulse with the pulse
    tabulated per step), and OutletCSV's mean / min / max reduced on the device in FP64."""
    dims = (8, 5, 4)
    n = 160
    say(f"ad_bc series / outlet stats prec={prec}")
    rng = np.random.default_rng(23)
    g = rng.random(n * 7)
    tdt = _tdt(prec)
    inlet = x_plane(*dims, 0)
    series = rng.uniform(0.0, 1.0, 12)
    ci = dev(np.nonzero(inlet)[0].astype(np.int64), torch().int64)
    sd = dev(series, torch().float64)
    gd = dev(aos_to_soa(g, 7).ravel(), tdt)
    abi.call("fmx_ad_bc_dirichlet_series", prec, gd.data_ptr(), ci.data_ptr(), ci.numel(), sd.data_ptr(), 7, n,
             stream())
    gc = dev(aos_to_soa(g, 7).ravel(), tdt)
    abi.call("fmx_ad_bc_dirichlet", prec, gc.data_ptr(), ci.data_ptr(), ci.numel(), float(series[7]), n, stream())
    sync()
    assert np.array_equal(host(gd), host(gc))
    assert rel_err(soa_to_aos(host(gd), 7), R.op_ad_bc(oprec, R.AD_DIRICHLET, dims, g, inlet, a=series[7])) < tol
    for count in (len(np.nonzero(x_plane(*dims, 7))[0]), 3000):
        phi = rng.uniform(0.0, 1.0, max(n, count))
        cells = np.nonzero(x_plane(*dims, 7))[0] if count < n else rng.permutation(phi.size)[:count]
        pd, cd = dev(phi, tdt), dev(cells.astype(np.int64), torch().int64)
        out = torch().zeros(3, dtype=torch().float64, device="cuda")
        abi.call("fmx_reduce_outlet_mean", prec, pd.data_ptr(), cd.data_ptr(), len(cells), out.data_ptr(), stream())
        sync()
        vals = host(pd).astype(np.float64)[cells]
        o = host(out)
        assert abs(o[0] - vals.mean()) <= 1e-14 * abs(vals.mean()), (o[0], vals.mean())
        assert o[1] == vals.min() and o[2] == vals.max()


def test_sync_check_mode():
    """fmx_set_sync_check(1) restores the reference's synchronise-and-check after every launch
    (cudaCheckThrow, CudaHelpers.cu:16-20): an operator and a short fused run give the same results
    as without it, and the mode switches back off."""
    from fluidmaxxer_amd.coupled import CoupledSim
    from tests.test_gpu_coupled import small_config

    say("sync-check mode: fused run and ns_collide with FMX sync check on and off")
    L = abi.lib()
    cfg = small_config("bgk")
    results = []
    dims = (13, 9, 7)
    n = 13 * 9 * 7
    f = _rand_f(n, 5)
    for on in (1, 0):
        assert L.fmx_set_sync_check(on) == abi.OK
        sim = CoupledSim(cfg, prec="f32")
        c = host(sim.run(30))
        fin = dev(aos_to_soa(f, 19).ravel(), torch().float32)
        fout = torch().empty_like(fin)
        g = abi.grid(*dims, (False, True, True))
        abi.call("fmx_ns_collide_stream", abi.F32, abi.NS_PBB, fin.data_ptr(), fout.data_ptr(), None, None, C.byref(g),
                 0.62, 1.0, stream())
        sync()
        results.append((c, host(sim.phi), host(fout)))
    for a, b in zip(*results):
        assert np.array_equal(a, b)


def test_memset_and_fill_zero_bytes():
    """fmx_memset sets every byte to value & 0xff (cudaMemset's contract) and fmx_fill_zero to 0, at
    every start alignment and length around the 4-byte words the fill kernel writes; the bytes on either
    side stay untouched."""
    t = torch()
    say("fmx_memset / fmx_fill_zero over offsets 0..7 and lengths 0..70, 4099")
    buf = t.empty(8192, dtype=t.uint8, device="cuda")
    for off in range(8):
        for length in list(range(0, 71)) + [4099]:
            for value, fn in ((0x1A5, "fmx_memset"), (0, "fmx_fill_zero")):
                buf.fill_(0x33)
                ptr = buf.data_ptr() + off
                if fn == "fmx_memset":
                    abi.call("fmx_memset", ptr, value, length, stream())
                else:
                    abi.call("fmx_fill_zero", ptr, length, stream())
                sync()
                h = host(buf)
                want = np.full(8192, 0x33, np.uint8)
                want[off:off + length] = value & 0xFF
                assert np.array_equal(h, want), (fn, off, length)


def test_memset_in_a_captured_graph():
    """fmx_memset and fmx_fill_zero captured into a HIP graph (they launch kernels, not memset nodes):
    three replays, each after eager work that overwrites the buffers, each checked."""
    t = torch()
    say("fmx_memset / fmx_fill_zero under torch.cuda.graph capture, 3 replays")
    a = t.empty(1 << 20, dtype=t.uint8, device="cuda")
    b = t.empty((1 << 20) + 3, dtype=t.uint8, device="cuda")
    a.fill_(1)
    b.fill_(1)
    sync()
    g = t.cuda.CUDAGraph()
    with t.cuda.graph(g):
        s = t.cuda.current_stream().cuda_stream
        abi.call("fmx_memset", a.data_ptr() + 1, 0x7E, a.numel() - 2, s)
        abi.call("fmx_fill_zero", b.data_ptr(), b.numel(), s)
    for rep in range(3):
        a.fill_(rep + 2)
        b.fill_(rep + 2)
        t.cuda.synchronize()
        g.replay()
        t.cuda.synchronize()
        ha, hb = host(a), host(b)
        assert ha[0] == rep + 2 and ha[-1] == rep + 2 and np.all(ha[1:-1] == 0x7E), rep
        assert not hb.any(), rep
