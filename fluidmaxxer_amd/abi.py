"""ctypes binding of libfmx_hip.so (include/fmx_abi.h).

The library is loaded from fluidmaxxer_amd/lib/ and never rebuilt here: a
missing library is an error (build with `python -m fluidmaxxer_amd.build`).
Every wrapped call raises FmxError with the library's message on a non-zero
status.
"""
from __future__ import annotations

import ctypes as C
import os
import re

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "lib", "libfmx_hip.so")
HEADER = os.path.join(ROOT, "include", "fmx_abi.h")

F32, F64, F16, BF16 = 0, 1, 2, 3
NS_BGK, NS_POROUS, NS_PBB = 0, 1, 2
ISO_NONE, ISO_LINEAR, ISO_LANGMUIR = 0, 1, 2
FLAG_SOLID, FLAG_REGION = 1, 2
LINK_FROZEN = 1 << 21  # FMX_LINK_FROZEN
OUTLET_SLOTS, OUTLET_STRIDE, OUTLET_RING_ROWS = 64, 16, 64
MASS_COLS = 4
ABI_VERSION = 5  # FMX_ABI_VERSION of the header these structs mirror
OK, ERR_ARG, ERR_HIP, ERR_UNSUPPORTED = 0, 1, 2, 3

PREC_NAMES = {"f32": F32, "f64": F64, "f16": F16, "bf16": BF16}


class FmxError(RuntimeError):
    pass


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("periodic_x", C.c_int32), ("periodic_y", C.c_int32), ("periodic_z", C.c_int32)]


class CoupledDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("periodic_y", C.c_int32), ("periodic_z", C.c_int32),
        ("prec", C.c_int32), ("ns_op", C.c_int32),
        ("ns_inlet", C.c_int32), ("ns_outlet", C.c_int32),
        ("ad_inlet", C.c_int32), ("ad_outlet", C.c_int32),
        ("isotherm", C.c_int32), ("coupling", C.c_int32),
        ("rho0", C.c_double), ("tau", C.c_double),
        ("u_in", C.c_double * 3),
        ("rho_out", C.c_double), ("phi_in", C.c_double), ("tau_ad", C.c_double), ("min_conc", C.c_double),
        ("iso_a", C.c_double), ("iso_b", C.c_double), ("k_ldf", C.c_double),
        ("f_ref", C.c_double),
        ("freeze_solids", C.c_int32), ("reserved0", C.c_int32),
    ]


class CoupledBufs(C.Structure):
    _fields_ = [
        ("f", C.c_void_p * 2), ("g", C.c_void_p * 2),
        ("q", C.c_void_p), ("phi", C.c_void_p * 2),
        ("links", C.c_void_p), ("field", C.c_void_p), ("tau_field", C.c_void_p),
        ("inlet_raw", C.c_void_p), ("inlet_series", C.c_void_p), ("outlet_acc", C.c_void_p),
        ("outlet_ring", C.c_void_p), ("inlet_series_len", C.c_int64), ("outlet_acc_rows", C.c_int64),
        ("mass_log", C.c_void_p), ("mass_workspace", C.c_void_p), ("mass_log_rows", C.c_int64),
        ("mass_workspace_bytes", C.c_int64), ("mass_every", C.c_int64),
        ("compact_work", C.c_void_p), ("compact_work_bytes", C.c_int64),
    ]


class AdjointBufs(C.Structure):
    _fields_ = [
        ("lam_f", C.c_void_p * 2), ("lam_g", C.c_void_p * 2),
        ("lam_q", C.c_void_p), ("lam_phi", C.c_void_p),
        ("workspace", C.c_void_p), ("workspace_bytes", C.c_int64), ("outlet_weight", C.c_void_p),
        ("outlet_weight_len", C.c_int64),
        ("grad_iso", C.c_void_p), ("grad_field", C.c_void_p), ("grad_series", C.c_void_p),
        ("field_rev", C.c_void_p), ("tau_field_rev", C.c_void_p),
    ]


_vp, _i, _i64, _d, _sz = C.c_void_p, C.c_int, C.c_int64, C.c_double, C.c_size_t
_DP, _BP, _GP = C.POINTER(CoupledDesc), C.POINTER(CoupledBufs), C.POINTER(Grid)

_SIGS = {
    "fmx_abi_version": ([], C.c_int),
    "fmx_last_error": ([], C.c_char_p),
    "fmx_device_count": ([C.POINTER(C.c_int)], _i),
    "fmx_device_info": ([_i, C.c_char_p, _i, C.POINTER(C.c_int64), C.POINTER(C.c_int), C.c_char_p, _i], _i),
    "fmx_set_device": ([_i], _i),
    "fmx_alloc": ([C.POINTER(_vp), _sz], _i),
    "fmx_free": ([_vp], _i),
    "fmx_memcpy_h2d": ([_vp, _vp, _sz, _vp], _i),
    "fmx_memcpy_d2h": ([_vp, _vp, _sz, _vp], _i),
    "fmx_memcpy_d2d": ([_vp, _vp, _sz, _vp], _i),
    "fmx_memset": ([_vp, _i, _sz, _vp], _i),
    "fmx_fill_zero": ([_vp, _sz, _vp], _i),
    "fmx_stream_create": ([C.POINTER(_vp)], _i),
    "fmx_stream_destroy": ([_vp], _i),
    "fmx_stream_sync": ([_vp], _i),
    "fmx_device_sync": ([], _i),
    "fmx_set_sync_check": ([_i], _i),
    "fmx_event_create": ([C.POINTER(_vp)], _i),
    "fmx_event_destroy": ([_vp], _i),
    "fmx_event_record": ([_vp, _vp], _i),
    "fmx_event_elapsed_ms": ([_vp, _vp, C.POINTER(C.c_float)], _i),
    "fmx_pack_aos": ([_i, _vp, _vp, _i, _i64, _vp], _i),
    "fmx_unpack_aos": ([_i, _vp, _vp, _i, _i64, _vp], _i),
    "fmx_ns_init_eq": ([_i, _vp, _GP, _d, _vp], _i),
    "fmx_ns_collide_stream": ([_i, _i, _vp, _vp, _vp, _vp, _GP, _d, _d, _vp], _i),
    "fmx_ns_moments": ([_i, _vp, _vp, _vp, _vp, _vp, _GP, _vp], _i),
    "fmx_ns_bc_velocity_neq": ([_i, _vp, _vp, _vp, _i64, _d, _d, _d, _i64, _vp], _i),
    "fmx_ns_bc_velocity_eq": ([_i, _vp, _vp, _i64, _d, _d, _d, _d, _i64, _vp], _i),
    "fmx_ns_bc_pressure": ([_i, _vp, _vp, _vp, _i64, _d, _i64, _vp], _i),
    "fmx_ns_bc_bounceback": ([_i, _vp, _vp, _i64, _i64, _vp], _i),
    "fmx_ad_init_eq": ([_i, _vp, _GP, _d, _vp], _i),
    "fmx_ad_collide_stream": ([_i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _GP, _d, _vp], _i),
    "fmx_ad_bc_dirichlet": ([_i, _vp, _vp, _i64, _d, _i64, _vp], _i),
    "fmx_ad_bc_neumann": ([_i, _vp, _vp, _vp, _i64, _i64, _vp], _i),
    "fmx_ad_bc_dirichlet_series": ([_i, _vp, _vp, _i64, _vp, _i64, _i64, _vp], _i),
    "fmx_ldf_linear": ([_i, _i64, _vp, _vp, _vp, _vp, _d, _d, _d, _vp], _i),
    "fmx_ldf_langmuir": ([_i, _i64, _vp, _vp, _vp, _vp, _d, _d, _d, _d, _vp], _i),
    "fmx_coupled_validate": ([_DP, _BP], _i),
    "fmx_coupled_build_links": ([_DP, _vp, _vp, _vp], _i),
    "fmx_coupled_init": ([_DP, _BP, _i, _d, _vp], _i),
    "fmx_coupled_load": ([_DP, _BP, _i, _vp, _vp, _vp, _vp, _vp], _i),
    "fmx_coupled_prologue": ([_DP, _BP, _i, _vp], _i),
    "fmx_coupled_run": ([_DP, _BP, _i64, C.c_int32, _i, _vp], _i),
    "fmx_coupled_step": ([_DP, _BP, _i64, _i, _vp], _i),
    "fmx_coupled_view_f": ([_DP, _BP, _i, _vp, _vp], _i),
    "fmx_coupled_view_g": ([_DP, _BP, _i, _vp, _vp], _i),
    "fmx_outlet_mean": ([_vp, _i64, _i64, _vp, _vp], _i),
    "fmx_coupled_recenter_workspace_bytes": ([_DP], C.c_int64),
    "fmx_coupled_compact_bytes": ([_DP, _i64], C.c_int64),
    "fmx_coupled_recenter": ([_DP, _BP, _i, _d, C.POINTER(C.c_double), _vp, _i64, _vp], _i),
    "fmx_adjoint_state_bytes": ([_DP], C.c_int64),
    "fmx_adjoint_ckpt_bytes": ([_DP, _i], C.c_int64),
    "fmx_adjoint_forward": ([_DP, _BP, _i64, C.c_int32, C.c_int32, _i, _vp, _i, _vp], _i),
    "fmx_adjoint_seg_slots": ([C.c_int32, _i, _i], C.c_int64),
    "fmx_adjoint_workspace_bytes": ([_DP], C.c_int64),
    "fmx_checkpoint_save": ([_DP, _BP, _i, _vp, _i, _i64, _vp], _i),
    "fmx_checkpoint_load": ([_DP, _BP, _i, _vp, _i, _i64, _vp], _i),
    "fmx_adjoint_step": ([_DP, _BP, _i, _i64, _i, _i, C.POINTER(AdjointBufs), _vp], _i),
    "fmx_adjoint_backward": ([_DP, _BP, _i64, C.c_int32, C.c_int32, _vp, _i, _vp, _i64, C.POINTER(AdjointBufs), _vp],
                             _i),
    "fmx_reduce_mass_workspace_bytes": ([_i64], _i64),
    "fmx_reduce_mass": ([_i, _vp, _vp, _vp, _i64, _d, _vp, _vp, _i64, _vp], _i),
    "fmx_reduce_mass_static": ([_i, _vp, _vp, _vp, _i64, _d, _vp, _vp], _i),
    "fmx_reduce_outlet_mean": ([_i, _vp, _vp, _i64, _vp, _vp], _i),
    "fmx_stream_probe": ([_i, _vp, _vp, _sz, _vp], _i),
    "fmx_geom_sphere_pack": ([C.c_int32, C.c_int32, C.c_int32, C.c_uint64, _d, _d, _d, C.c_int32, C.c_int32,
                              C.c_int64, _vp, _vp, C.POINTER(C.c_double)], _i),
    "fmx_build_x_plane_mask": ([C.c_int32, C.c_int32, C.c_int32, C.c_int32, _vp], _i),
    "fmx_vti_write_mask": ([C.c_char_p, C.c_int32, C.c_int32, C.c_int32, _vp, C.c_char_p], _i),
    "fmx_vti_read_mask": ([C.c_char_p, _d, C.POINTER(C.c_int32), _vp, C.c_int64], _i),
}

_lib = None


def header_functions() -> list:
    """Names of every function include/fmx_abi.h declares."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fmx_\w+)\s*\(", text)))


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FmxError(f"{LIB_PATH} is missing: build it with `python -m fluidmaxxer_amd.build`")
        # torch's HIP runtime first: its libamdhip64 carries the soname this library links against, so
        # loaded first it serves both.  Loaded after this library, torch brings a second HIP / HSA
        # runtime into the process and this library's first launch finds no device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        # the structs above mirror one layout of include/fmx_abi.h: a library built from another
        # header would read them wrongly
        if L.fmx_abi_version() != ABI_VERSION:
            raise FmxError(f"{LIB_PATH} implements ABI version {L.fmx_abi_version()}, these bindings version "
                           f"{ABI_VERSION}: rebuild it with `python -m fluidmaxxer_amd.build`")
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().fmx_last_error()
        raise FmxError(msg.decode() if msg else f"fmx status {rc}")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))


def grid(nx, ny, nz, periodic=(False, True, True)) -> Grid:
    return Grid(nx, ny, nz, *(int(bool(p)) for p in periodic))
