"""Fused coupled step on one GPU: configuration and driver.

`CoupledConfig` holds one scenario in the vocabulary of the reference's
LBMParams (Params.hpp:6-37) and main.cpp's set-up (main.cpp:137-279): a D3Q19
NS lattice with an operator (BGK with obstacle bounce-back, porous BGK, or
partial bounce-back), a velocity inlet on x = 0 and a pressure outlet on
x = nx-1, a D3Q7 AD lattice with a Dirichlet (or per-step series) inlet and a
zero-gradient outlet, and a linear or Langmuir LDF sorption source on a region.

`CoupledSim` owns the device state (torch tensors used as plain device
memory) and advances it with fmx_coupled_run, one kernel launch per step.
"""
from __future__ import annotations

from dataclasses import dataclass, field as dc_field
from typing import Optional

import numpy as np

from . import abi

_NS_OPS = {"bgk": abi.NS_BGK, "porous": abi.NS_POROUS, "pbb": abi.NS_PBB}
_ISO = {None: abi.ISO_NONE, "none": abi.ISO_NONE, "linear": abi.ISO_LINEAR, "langmuir": abi.ISO_LANGMUIR}


@dataclass
class CoupledConfig:
    nx: int
    ny: int
    nz: int
    solid: np.ndarray                      # uint8[N]; BGK bounce-back cells
    region: np.ndarray                     # uint8[N]; sorption cells
    periodic_y: bool = True                # main.cpp:143-145
    periodic_z: bool = True
    ns_op: str = "bgk"
    rho0: float = 1.0
    tau: float = 0.8
    u_in: tuple = (0.05, 0.0, 0.0)
    ns_inlet: Optional[str] = "eq"         # "eq": VelocityDirichletEq, "neq": VelocityDirichletNEQ, on x = 0
    ns_outlet: bool = True                 # PressureDirichlet on x = nx-1
    rho_out: float = 1.0
    field: Optional[np.ndarray] = None     # porosity (porous) / solid fraction (pbb), float[N]
    ad_inlet: str = "dirichlet"            # "dirichlet" | "series" | "none"
    phi_in: float = 1.0
    inlet_series: Optional[np.ndarray] = None  # ad_inlet == "series": value of step s at [s]
    ad_outlet: bool = True                 # ADNeumannOutlet on x = nx-1
    tau_ad: float = 0.7
    tau_field: Optional[np.ndarray] = None
    phi0: float = 0.0
    min_conc: float = 1e-8
    isotherm: Optional[str] = "linear"
    iso_a: float = 0.5                     # linear: KHenry; Langmuir: qMax
    iso_b: float = 0.0                     # Langmuir: bLangmuir (K_L)
    k_ldf: float = 0.01                    # linear: kA; Langmuir: kLDF
    coupling: bool = True
    f_ref: float = 1.0                     # half storage: f = w * (f_ref + h)
    meta: dict = dc_field(default_factory=dict)

    @property
    def n(self) -> int:
        return self.nx * self.ny * self.nz

    def flags(self) -> np.ndarray:
        return (np.asarray(self.solid, np.uint8) * abi.FLAG_SOLID
                | np.asarray(self.region, np.uint8) * abi.FLAG_REGION).astype(np.uint8)

    def desc(self, prec: int) -> abi.CoupledDesc:
        d = abi.CoupledDesc()
        d.nx, d.ny, d.nz = self.nx, self.ny, self.nz
        d.periodic_y, d.periodic_z = int(self.periodic_y), int(self.periodic_z)
        d.prec = prec
        d.ns_op = _NS_OPS[self.ns_op]
        d.ns_inlet = {None: 0, "none": 0, "neq": 1, "eq": 2}[self.ns_inlet]
        d.ns_outlet = int(self.ns_outlet)
        d.ad_inlet = {"none": 0, None: 0, "dirichlet": 1, "series": 2}[self.ad_inlet]
        d.ad_outlet = int(self.ad_outlet)
        d.isotherm = _ISO[self.isotherm]
        d.coupling = int(self.coupling)
        d.rho0, d.tau = self.rho0, self.tau
        for i in range(3):
            d.u_in[i] = float(self.u_in[i])
        d.rho_out, d.phi_in, d.tau_ad, d.min_conc = self.rho_out, self.phi_in, self.tau_ad, self.min_conc
        d.iso_a, d.iso_b, d.k_ldf = self.iso_a, self.iso_b, self.k_ldf
        d.f_ref = self.f_ref
        return d


def _torch():
    import torch
    return torch


class CoupledSim:
    """Device state of one coupled configuration and its stepper.

    prec: "f32" | "f64" | "f16" | "bf16" (population storage; scalar state is
    float64 for f64 and float32 otherwise).
    compact: BGK only -- solid cells at rest are frozen at the prologue and runs of 16 steps or more
    step the NS populations of the other cells in row-compacted buffers (a workspace sized for them),
    so the frozen ones cost no population traffic, each step being an NS launch over those cells and an
    AD launch over the box; results are the same to the bit (FP32 / FP64; 16-bit storage to its
    rounding, the kernels being compiled with FMA contraction).  Default (None): on (512^3 sphere pack:
    FP32 1.28x the box layout, FP64 1.31x, FP16 / BF16 1.07-1.13x).
    """

    def __init__(self, cfg: CoupledConfig, prec: str = "f32", device: str = "cuda:0", compact: bool = None):
        torch = _torch()
        if not torch.cuda.is_available():
            raise abi.FmxError("CoupledSim needs a GPU")
        self.cfg = cfg
        self.prec_name = prec
        self.prec = abi.PREC_NAMES[prec]
        self.device = torch.device(device)
        n = cfg.n
        sdt = {abi.F32: torch.float32, abi.F64: torch.float64, abi.F16: torch.float16,
               abi.BF16: torch.bfloat16}[self.prec]
        self.cdt = torch.float64 if self.prec == abi.F64 else torch.float32
        dev = self.device
        self.f = [torch.empty(19 * n, dtype=sdt, device=dev) for _ in range(2)]
        self.g = [torch.empty(7 * n, dtype=sdt, device=dev) for _ in range(2)]
        self.q = torch.zeros(n, dtype=self.cdt, device=dev)
        self.phis = [torch.zeros(n, dtype=self.cdt, device=dev) for _ in range(2)]
        self.flags = torch.from_numpy(cfg.flags()).to(dev)
        self.links = torch.empty(n, dtype=torch.int32, device=dev)
        self.field = None if cfg.field is None else torch.as_tensor(cfg.field, dtype=self.cdt).to(dev)
        self.tau_field = None if cfg.tau_field is None else torch.as_tensor(cfg.tau_field, dtype=self.cdt).to(dev)
        has_inlet = cfg.ns_inlet not in (None, "none")
        self.inlet_raw = torch.zeros(19 * cfg.ny * cfg.nz, dtype=self.cdt, device=dev) if has_inlet else None
        self.inlet_series = None
        if cfg.ad_inlet == "series":
            if cfg.inlet_series is None:
                raise abi.FmxError("ad_inlet='series' needs inlet_series")
            self.inlet_series = torch.as_tensor(np.asarray(cfg.inlet_series, np.float64)).to(dev)
        self._desc = cfg.desc(self.prec)
        if compact is None:
            compact = True
        self.compact = bool(compact) and cfg.ns_op == "bgk"
        self._desc.freeze_solids = int(self.compact)
        self._compact_ws = None
        self.cur = 0
        self.step_index = 0
        self._acc = None
        self._ring = None
        self._mass_ws = None
        self._recenter_ws = None
        self.mass_log = None   # (rows, 4) float64 device tensor of the last run's periodic sums, or None
        self.stream = torch.cuda.current_stream(dev)
        abi.call("fmx_coupled_build_links", self._desc_p(), self.flags.data_ptr(), self.links.data_ptr(), self._s())
        abi.call("fmx_coupled_init", self._desc_p(), self._bufs_p(), self.cur, float(cfg.phi0), self._s())
        self._size_compact()

    def _size_compact(self):
        """(Re)sizes the compact workspace for the cells the last prologue left unfrozen."""
        if not self.compact:
            return
        torch = _torch()
        active = int((self.links & abi.LINK_FROZEN).eq(0).sum().item())
        nb = int(abi.lib().fmx_coupled_compact_bytes(self._desc_p(), active))
        if nb < 0:
            raise abi.FmxError("fmx_coupled_compact_bytes refused the configuration")
        if self._compact_ws is None or self._compact_ws.numel() < nb:
            self._compact_ws = None
            self._compact_ws = torch.empty(nb, dtype=torch.uint8, device=self.device)

    @property
    def phi(self):
        """Concentration of the current state (the last step's clamped pre-collision sum)."""
        return self.phis[self.cur]

    # ------------------------------------------------------------ plumbing
    def _s(self):
        return self.stream.cuda_stream

    def _desc_p(self):
        return abi.C.byref(self._desc)

    def _mass_workspace(self):
        if self._mass_ws is None:
            nb = int(abi.lib().fmx_reduce_mass_workspace_bytes(self.cfg.n))
            self._mass_ws = _torch().empty(nb, dtype=_torch().uint8, device=self.device)
        return self._mass_ws

    def _bufs(self, acc=None, mass=None, mass_every=0) -> abi.CoupledBufs:
        b = abi.CoupledBufs()
        b.f[0], b.f[1] = self.f[0].data_ptr(), self.f[1].data_ptr()
        b.g[0], b.g[1] = self.g[0].data_ptr(), self.g[1].data_ptr()
        b.q, b.links = self.q.data_ptr(), self.links.data_ptr()
        b.phi[0], b.phi[1] = self.phis[0].data_ptr(), self.phis[1].data_ptr()
        b.field = self.field.data_ptr() if self.field is not None else None
        b.tau_field = self.tau_field.data_ptr() if self.tau_field is not None else None
        b.inlet_raw = self.inlet_raw.data_ptr() if self.inlet_raw is not None else None
        b.inlet_series = self.inlet_series.data_ptr() if self.inlet_series is not None else None
        b.inlet_series_len = self.inlet_series.numel() if self.inlet_series is not None else 0
        b.outlet_acc = acc.data_ptr() if acc is not None else None
        b.outlet_acc_rows = acc.numel() // (abi.OUTLET_SLOTS * abi.OUTLET_STRIDE) if acc is not None else 0
        if acc is not None:  # outlet sums in a fixed order (repeatable to the bit)
            if self._ring is None:
                self._ring = _torch().empty(abi.OUTLET_RING_ROWS * self.cfg.ny * self.cfg.nz,
                                            dtype=_torch().float64, device=self.device)
            b.outlet_ring = self._ring.data_ptr()
        if mass is not None:
            ws = self._mass_workspace()
            b.mass_log, b.mass_log_rows, b.mass_every = mass.data_ptr(), mass.shape[0], int(mass_every)
            b.mass_workspace, b.mass_workspace_bytes = ws.data_ptr(), ws.numel()
        if self._compact_ws is not None:
            b.compact_work, b.compact_work_bytes = self._compact_ws.data_ptr(), self._compact_ws.numel()
        return b

    def _bufs_p(self, acc=None, mass=None, mass_every=0):
        self._b = self._bufs(acc, mass, mass_every)
        return abi.C.byref(self._b)

    # ------------------------------------------------------------- stepping
    def outlet_acc(self, nsteps: int):
        """Zeroed FP64 accumulator rows for `nsteps` steps of outlet sums."""
        torch = _torch()
        need = nsteps * abi.OUTLET_SLOTS * abi.OUTLET_STRIDE
        if self._acc is None or self._acc.numel() < need:
            self._acc = torch.empty(need, dtype=torch.float64, device=self.device)
        acc = self._acc[:need]
        abi.call("fmx_fill_zero", acc.data_ptr(), need * 8, self._s())
        return acc

    def run(self, nsteps: int, log_outlet: bool = True, acc=None, mass_every: int = 0):
        """Advance `nsteps` steps; returns the per-step outlet mean (float64 device tensor) or None.
        With mass_every > 0, self.mass_log holds one row (sum phi, sum q, sum of stored f, non-finite
        count) after every mass_every-th step of this run."""
        torch = _torch()
        if self.inlet_series is not None and self.step_index + nsteps >= self.inlet_series.numel():
            raise abi.FmxError("inlet_series is shorter than the requested steps")
        if log_outlet and acc is None:
            acc = self.outlet_acc(nsteps)
        mass = None
        if mass_every > 0:
            mass = torch.zeros((max(nsteps // mass_every, 1), abi.MASS_COLS), dtype=torch.float64, device=self.device)
        self.mass_log = mass
        abi.call("fmx_coupled_run", self._desc_p(), self._bufs_p(acc if log_outlet else None, mass, mass_every),
                 self.step_index + 1, nsteps, self.cur, self._s())
        self.cur = (self.cur + nsteps) & 1
        self.step_index += nsteps
        if not log_outlet:
            return None
        mean = torch.empty(nsteps, dtype=torch.float64, device=self.device)
        abi.call("fmx_outlet_mean", acc.data_ptr(), nsteps, self.cfg.ny * self.cfg.nz, mean.data_ptr(), self._s())
        return mean

    def launch_only(self, nsteps: int, acc=None):
        """Steps without any host-side bookkeeping beyond the ping-pong index (for timing)."""
        abi.call("fmx_coupled_run", self._desc_p(), self._bufs_p(acc), self.step_index + 1, nsteps, self.cur, self._s())
        self.cur = (self.cur + nsteps) & 1
        self.step_index += nsteps

    # ------------------------------------------------------------ read-back
    def view_f(self):
        """Reference-view NS populations, SoA [19, N], compute precision."""
        torch = _torch()
        out = torch.empty(19 * self.cfg.n, dtype=self.cdt, device=self.device)
        abi.call("fmx_coupled_view_f", self._desc_p(), self._bufs_p(), self.cur, out.data_ptr(), self._s())
        return out.view(19, self.cfg.n)

    def view_g(self):
        torch = _torch()
        out = torch.empty(7 * self.cfg.n, dtype=self.cdt, device=self.device)
        abi.call("fmx_coupled_view_g", self._desc_p(), self._bufs_p(), self.cur, out.data_ptr(), self._s())
        return out.view(7, self.cfg.n)

    def moments(self):
        """rho, ux, uy, uz of the reference-view populations (k_compute_macros_d3q19 order)."""
        torch = _torch()
        f = self.view_f()
        n = self.cfg.n
        outs = [torch.empty(n, dtype=self.cdt, device=self.device) for _ in range(4)]
        g = abi.grid(self.cfg.nx, self.cfg.ny, self.cfg.nz)
        prec = abi.F64 if self.cdt == torch.float64 else abi.F32
        abi.call("fmx_ns_moments", prec, f.data_ptr(), *(o.data_ptr() for o in outs), abi.C.byref(g), self._s())
        return outs

    def load(self, f=None, g=None, q=None, phi=None):
        """Load a state given as compute-precision SoA arrays ([19,N], [7,N], [N], [N]); re-applies the
        inlet prologue."""
        torch = _torch()
        keep = []

        def dev(a):
            if a is None:
                return None
            if isinstance(a, torch.Tensor):  # device tensors stay on the device
                t = a.reshape(-1).to(device=self.device, dtype=self.cdt).contiguous()
            else:
                t = torch.as_tensor(np.asarray(a)).to(device=self.device, dtype=self.cdt).contiguous()
            keep.append(t)
            return t.data_ptr()

        abi.call("fmx_coupled_load", self._desc_p(), self._bufs_p(), self.cur, dev(f), dev(g), dev(q), dev(phi),
                 self._s())
        self.stream.synchronize()
        self._size_compact()

    def mass(self):
        """FP64 (sum phi, sum q, sum rho of the stored f, non-finite count) over all cells."""
        torch = _torch()
        out = torch.zeros(abi.MASS_COLS, dtype=torch.float64, device=self.device)
        ws = self._mass_workspace()
        abi.call("fmx_reduce_mass", self.prec, self.phi.data_ptr(), self.q.data_ptr(), self.f[self.cur].data_ptr(),
                 self.cfg.n, float(self._desc.f_ref), out.data_ptr(), ws.data_ptr(), ws.numel(), self._s())
        return out

    @property
    def f_ref(self) -> float:
        """Density reference of the stored NS populations (16-bit storage: f = w (f_ref + h))."""
        return float(self._desc.f_ref)

    def recenter(self, target: float = None) -> float:
        """Moves the density reference of FP16 / BF16 NS storage to `target`, or by default to the mean
        density of the non-solid cells, re-encoding the stored populations (fmx_coupled_recenter); with
        FP32 / FP64 storage the state is unchanged.  Returns the new reference.  Called every ~1000 steps
        while a pressure drop builds up, it keeps the stored deviations centred on the flow."""
        torch = _torch()
        nb = int(abi.lib().fmx_coupled_recenter_workspace_bytes(self._desc_p()))
        if self._recenter_ws is None or self._recenter_ws.numel() < nb:
            self._recenter_ws = torch.empty(nb, dtype=torch.uint8, device=self.device)
        out = abi.C.c_double(0.0)
        tgt = float("nan") if target is None else float(target)
        abi.call("fmx_coupled_recenter", self._desc_p(), self._bufs_p(), self.cur, tgt, abi.C.byref(out),
                 self._recenter_ws.data_ptr(), self._recenter_ws.numel(), self._s())
        self._desc.f_ref = out.value
        return out.value

    def synchronize(self):
        self.stream.synchronize()
