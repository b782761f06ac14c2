"""ctypes bindings for the CPU oracle (oracle/ref_cpu.cpp).

TEST INFRASTRUCTURE ONLY: the oracle is a serial C++ restatement of the
reference's coupled D3Q19 + D3Q7 update (see oracle/ref_cpu.hpp for the
file:line map).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, and only as the checker or as the timed
CPU baseline -- never as part of the product path.

Arrays cross the boundary as float64 numpy arrays in the reference AoS layout
(f[cell*19 + q], g[cell*7 + q]); for prec=4 the oracle rounds them to float32 on
entry.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libref_cpu.so")
_REF_VTI_PATH = os.path.join(_HERE, "_ref", "libref_vti.so")

# field ids understood by orc_get / orc_set
F, G, PHI, Q, S, UXA, UYA, UZA, RHO, UX, UY, UZ = range(12)

# boundary kinds (NS)
NS_NEQ, NS_EQ, NS_PRESSURE, NS_BOUNCEBACK = 1, 2, 3, 4
# boundary kinds (AD)
AD_DIRICHLET, AD_TANH, AD_NEUMANN, AD_SERIES = 1, 2, 3, 4
# collision operators
BGK, POROUS, PBB = 0, 1, 2
# isotherms
LINEAR, LANGMUIR = 1, 2


class _BC(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("mask", C.POINTER(C.c_uint8)),
        ("v", C.c_double * 4),
        ("i0", C.c_int),
        ("i1", C.c_int),
        ("knots", C.POINTER(C.c_double)),
        ("n_knots", C.c_int),
        ("dt_knot", C.c_double),
    ]


class _Config(C.Structure):
    _fields_ = [
        ("Nx", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int),
        ("px", C.c_int), ("py", C.c_int), ("pz", C.c_int),
        ("ns_op", C.c_int),
        ("rho0", C.c_double), ("tau", C.c_double),
        ("obstacle", C.POINTER(C.c_uint8)),
        ("porosity", C.POINTER(C.c_double)),
        ("phi0", C.c_double), ("tau_ad", C.c_double), ("minC", C.c_double),
        ("tau_field", C.POINTER(C.c_double)),
        ("n_ns_bc", C.c_int), ("ns_bc", C.POINTER(_BC)),
        ("n_ad_bc", C.c_int), ("ad_bc", C.POINTER(_BC)),
        ("coupling", C.c_int), ("adsorb", C.c_int), ("isotherm", C.c_int),
        ("region", C.POINTER(C.c_uint8)),
        ("KHenry", C.c_double), ("kA", C.c_double), ("qMax", C.c_double),
        ("bLangmuir", C.c_double), ("kLDF", C.c_double),
    ]


_lib = None


def build() -> None:
    """Compile the oracle with its Makefile (g++ only, a few seconds)."""
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp, dp, u8p = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_uint8)
        L.orc_last_error.restype = C.c_char_p
        L.orc_set_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.POINTER(_Config), C.c_int]
        L.orc_destroy.argtypes = [vp]
        L.orc_set_step_mutation.argtypes = [vp, C.c_int]
        L.orc_run.argtypes = [vp, C.c_int, C.c_int, u8p, dp, dp, dp, dp]
        L.orc_get.argtypes = [vp, C.c_int, dp]
        L.orc_set.argtypes = [vp, C.c_int, dp]
        L.orc_op_ns_collide.argtypes = [C.c_int] * 8 + [dp, dp, u8p, dp, C.c_double, C.c_double]
        L.orc_op_ns_moments.argtypes = [C.c_int] * 4 + [dp] * 5
        L.orc_op_ns_bc.argtypes = [C.c_int] * 5 + [dp, u8p] + [C.c_double] * 4
        L.orc_op_ad_collide.argtypes = [C.c_int] * 7 + [dp] * 8 + [C.c_double]
        L.orc_op_ad_bc.argtypes = [C.c_int] * 5 + [dp, u8p, C.c_double]
        L.orc_op_ldf.argtypes = [C.c_int, C.c_int, C.c_longlong, u8p, dp, dp, dp] + [C.c_double] * 4
        L.orc_tanh_pulse.restype = C.c_double
        L.orc_tanh_pulse.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_double]
        L.orc_series_value.restype = C.c_double
        L.orc_series_value.argtypes = [C.c_int, C.c_int, dp, C.c_int, C.c_double]
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())


def _dp(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _u8(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def _f64(a) -> Optional[np.ndarray]:
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _mask(a) -> Optional[np.ndarray]:
    return None if a is None else np.ascontiguousarray(a, dtype=np.uint8)


# ----------------------------------------------------------------- system API
@dataclass
class BC:
    kind: int
    mask: np.ndarray
    v: tuple = (0.0, 0.0, 0.0, 0.0)
    i0: int = 0
    i1: int = 0
    knots: Optional[np.ndarray] = None
    dt_knot: float = 1.0


@dataclass
class Setup:
    """One coupled configuration, in the vocabulary of the reference's LBMParams."""
    Nx: int
    Ny: int
    Nz: int
    periodic: tuple = (False, True, True)
    ns_op: int = BGK
    rho0: float = 1.0
    tau: float = 0.8
    obstacle: Optional[np.ndarray] = None
    porosity: Optional[np.ndarray] = None
    phi0: float = 0.0
    tau_ad: float = 0.7
    minC: float = 1e-8
    tau_field: Optional[np.ndarray] = None
    ns_bc: List[BC] = field(default_factory=list)
    ad_bc: List[BC] = field(default_factory=list)
    coupling: bool = True
    adsorb: bool = True
    isotherm: int = LINEAR
    region: Optional[np.ndarray] = None
    KHenry: float = 0.0
    kA: float = 0.0
    qMax: float = 0.0
    bLangmuir: float = 0.0
    kLDF: float = 0.0

    @property
    def n(self) -> int:
        return self.Nx * self.Ny * self.Nz


class System:
    """The coupled reference update on the CPU, T = float (prec=4) or double (prec=8)."""

    def __init__(self, setup: Setup, prec: int = 8):
        self.setup = setup
        self.prec = prec
        self._keep = []
        s = setup

        def keep(a):
            self._keep.append(a)
            return a

        def bcs(lst):
            arr = (_BC * max(1, len(lst)))()
            for i, b in enumerate(lst):
                arr[i].kind = b.kind
                arr[i].mask = _u8(keep(_mask(b.mask)))
                for j in range(4):
                    arr[i].v[j] = float(b.v[j]) if j < len(b.v) else 0.0
                arr[i].i0, arr[i].i1 = int(b.i0), int(b.i1)
                if b.knots is not None:
                    arr[i].knots = _dp(keep(_f64(b.knots)))
                    arr[i].n_knots = len(b.knots)
                arr[i].dt_knot = float(b.dt_knot)
            return keep(arr)

        cfg = _Config()
        cfg.Nx, cfg.Ny, cfg.Nz = s.Nx, s.Ny, s.Nz
        cfg.px, cfg.py, cfg.pz = (int(bool(p)) for p in s.periodic)
        cfg.ns_op = s.ns_op
        cfg.rho0, cfg.tau = s.rho0, s.tau
        cfg.obstacle = _u8(keep(_mask(s.obstacle)))
        cfg.porosity = _dp(keep(_f64(s.porosity)))
        cfg.phi0, cfg.tau_ad, cfg.minC = s.phi0, s.tau_ad, s.minC
        cfg.tau_field = _dp(keep(_f64(s.tau_field)))
        cfg.n_ns_bc, cfg.ns_bc = len(s.ns_bc), bcs(s.ns_bc)
        cfg.n_ad_bc, cfg.ad_bc = len(s.ad_bc), bcs(s.ad_bc)
        cfg.coupling, cfg.adsorb, cfg.isotherm = int(s.coupling), int(s.adsorb), s.isotherm
        cfg.region = _u8(keep(_mask(s.region if s.region is not None else np.zeros(s.n, np.uint8))))
        cfg.KHenry, cfg.kA, cfg.qMax, cfg.bLangmuir, cfg.kLDF = s.KHenry, s.kA, s.qMax, s.bLangmuir, s.kLDF
        self._cfg = cfg
        h = lib().orc_create(C.byref(cfg), prec)
        if not h:
            raise RuntimeError(lib().orc_last_error().decode())
        self._h = C.c_void_p(h)
        self.step_index = 0

    def close(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def run(self, nsteps: int, outlet_mask: Optional[np.ndarray] = None):
        """Advance nsteps; returns (avg, min, max, seconds) of the outlet phi after each step."""
        avg = np.zeros(nsteps) if outlet_mask is not None else None
        mn = np.zeros(nsteps) if outlet_mask is not None else None
        mx = np.zeros(nsteps) if outlet_mask is not None else None
        sec = C.c_double(0.0)
        m = _mask(outlet_mask)
        _check(lib().orc_run(self._h, self.step_index + 1, nsteps, _u8(m), _dp(avg), _dp(mn), _dp(mx),
                             C.byref(sec)))
        self.step_index += nsteps
        return avg, mn, mx, sec.value

    def set_step_mutation(self, m: int) -> None:
        """Plant ordering mistake m in the step sequence (0: the reference order); for the mutation
        tests of the sequence pins only (the list is at the top of oracle/ref_cpu.cpp)."""
        _check(lib().orc_set_step_mutation(self._h, int(m)))

    def get(self, fid: int) -> np.ndarray:
        n = self.setup.n * (19 if fid == F else 7 if fid == G else 1)
        out = np.empty(n, np.float64)
        _check(lib().orc_get(self._h, fid, _dp(out)))
        return out

    def set(self, fid: int, values) -> None:
        v = _f64(values).ravel()
        _check(lib().orc_set(self._h, fid, _dp(v)))


# ------------------------------------------------------- single operators
def op_ns_collide(prec, ns_op, dims, periodic, f, obstacle=None, field=None, tau=0.8, rho0=1.0):
    Nx, Ny, Nz = dims
    f = _f64(f)
    out = np.empty_like(f)
    _check(lib().orc_op_ns_collide(prec, ns_op, Nx, Ny, Nz, *(int(bool(p)) for p in periodic), _dp(f), _dp(out),
                                   _u8(_mask(obstacle)), _dp(_f64(field)), tau, rho0))
    return out


def op_ns_moments(prec, dims, f):
    Nx, Ny, Nz = dims
    n = Nx * Ny * Nz
    r, a, b, c = (np.empty(n) for _ in range(4))
    _check(lib().orc_op_ns_moments(prec, Nx, Ny, Nz, _dp(_f64(f)), _dp(r), _dp(a), _dp(b), _dp(c)))
    return r, a, b, c


def op_ns_bc(prec, kind, dims, f, mask, u=(0.0, 0.0, 0.0), a=0.0):
    Nx, Ny, Nz = dims
    f = _f64(f).copy()
    _check(lib().orc_op_ns_bc(prec, kind, Nx, Ny, Nz, _dp(f), _u8(_mask(mask)), u[0], u[1], u[2], a))
    return f


def op_ad_collide(prec, dims, periodic, g, tau, ux, uy, uz, S, minC=1e-8):
    Nx, Ny, Nz = dims
    g = _f64(g)
    out = np.empty_like(g)
    phi = np.empty(Nx * Ny * Nz)
    _check(lib().orc_op_ad_collide(prec, Nx, Ny, Nz, *(int(bool(p)) for p in periodic), _dp(g), _dp(out), _dp(phi),
                                   _dp(_f64(tau)), _dp(_f64(ux)), _dp(_f64(uy)), _dp(_f64(uz)), _dp(_f64(S)), minC))
    return out, phi


def op_ad_bc(prec, kind, dims, g, mask, a=0.0):
    Nx, Ny, Nz = dims
    g = _f64(g).copy()
    _check(lib().orc_op_ad_bc(prec, kind, Nx, Ny, Nz, _dp(g), _u8(_mask(mask)), a))
    return g


def op_ldf(prec, isotherm, region, phi, S, q, a, b, kLDF, minC=1e-8):
    S = _f64(S).copy()
    q = _f64(q).copy()
    r = _mask(region)
    _check(lib().orc_op_ldf(prec, isotherm, len(q), _u8(r), _dp(_f64(phi)), _dp(S), _dp(q), a, b, kLDF, minC))
    return S, q


def tanh_pulse(prec, step, base, amp, t_on, t_off, k):
    return lib().orc_tanh_pulse(prec, step, base, amp, t_on, t_off, k)


def series_value(prec, step, knots, dt_knot):
    k = _f64(knots)
    return lib().orc_series_value(prec, step, _dp(k), len(k), dt_knot)


# ------------------------------------------------------ reference VTI reader
def ref_vti_available() -> bool:
    return os.path.exists(_REF_VTI_PATH)


def ref_read_vti(path: str, threshold: float = 0.5):
    """The reference's own readVTI (VTIReader.hpp:35-160), built into oracle/_ref/."""
    L = C.CDLL(_REF_VTI_PATH)
    dims = (C.c_int * 3)()
    err = C.create_string_buffer(512)
    rc = L.ref_read_vti(path.encode(), C.c_double(threshold), dims, None, C.c_longlong(0), err, 512)
    if rc:
        raise RuntimeError(err.value.decode())
    n = dims[0] * dims[1] * dims[2]
    mask = np.zeros(n, np.uint8)
    rc = L.ref_read_vti(path.encode(), C.c_double(threshold), dims, mask.ctypes.data_as(C.c_void_p),
                        C.c_longlong(n), err, 512)
    if rc:
        raise RuntimeError(err.value.decode())
    return (dims[0], dims[1], dims[2]), mask


def ref_build_x_plane_mask(Nx, Ny, Nz, x):
    L = C.CDLL(_REF_VTI_PATH)
    out = np.zeros(Nx * Ny * Nz, np.uint8)
    L.ref_build_x_plane_mask(Nx, Ny, Nz, x, out.ctypes.data_as(C.c_void_p))
    return out
