"""Breakthrough-curve gradients through the discrete adjoint of the fused coupled step.

`AdjointRun` owns one coupled configuration on the GPU, a checkpoint store and the
adjoint buffers.  `forward()` runs the configuration from its initial state for
`nsteps` steps (fmx_adjoint_forward, a checkpoint every `k` steps) and returns the
outlet-averaged concentration C_out(s), s = 1..nsteps.  Checkpoints keep the storage
precision, or, for an FP32 run, `ckpt="f16"` / `"bf16"` stores them as 16-bit
deviations and the reverse sweep recomputes each segment in FP32 from them, or, for an
FP64 run, `ckpt="f32"` rounds them to FP32 and the whole reverse sweep runs in FP32 (the
forward curve stays the FP64 run's; the gradients are the FP32 update's along the FP64
trajectory, restarted from it at every checkpoint).  `backward(w)` runs the
reverse sweep (fmx_adjoint_backward) for a loss with dL/dC_out(s) = w[s-1] and
returns the gradients with respect to the isotherm constants (Langmuir q_max and
K_L, or the linear K_Henry), the per-cell field (partial-bounce-back solid fraction
or porosity) and the per-step inlet concentration series.

`CoupledBreakthrough` wraps the pair as a torch.autograd.Function:

    run = AdjointRun(cfg, nsteps=4000, k=200, prec="f32")
    c = CoupledBreakthrough.apply(run, q_max, k_l, field, series)   # tensors
    loss = ((c - target) ** 2).sum(); loss.backward()

Both passes run entirely on the GPU through the C ABI; nothing here computes a
population update.
"""
from __future__ import annotations

import numpy as np

from . import abi
from .coupled import CoupledConfig, CoupledSim


def _torch():
    import torch
    return torch


def decoded_checkpoints(prec: str, ckpt: str = None) -> bool:
    """True for 16-bit checkpoints of an FP32 run, which the reverse sweep decodes into a segment slot
    (the `recast` argument of fmx_adjoint_seg_slots)."""
    return prec == "f32" and ckpt not in (None, "f32")


def segment_slots(k: int, prec: str, ckpt: str = None, two_level: bool = False) -> int:
    """Segment slots a reverse sweep over segments of k steps needs (fmx_adjoint_seg_slots)."""
    return int(abi.lib().fmx_adjoint_seg_slots(int(k), int(decoded_checkpoints(prec, ckpt)), int(two_level)))


class AdjointRun:
    def __init__(self, cfg: CoupledConfig, nsteps: int, k: int, prec: str = "f64", device: str = "cuda:0",
                 ckpt: str = None, seg_slots: int = None):
        torch = _torch()
        if prec not in ("f32", "f64", "f16", "bf16"):
            raise abi.FmxError(f"unknown precision {prec}")
        ckpt = prec if ckpt is None else ckpt
        if ckpt != prec and not ((prec == "f32" and ckpt in ("f16", "bf16")) or (prec == "f64" and ckpt == "f32")):
            raise abi.FmxError(f"checkpoints of a {prec} run are stored in {prec} (or f16 / bf16 for f32, f32 for "
                               f"f64)")
        if cfg.ad_inlet == "series" and (cfg.inlet_series is None or len(cfg.inlet_series) < nsteps + 1):
            raise abi.FmxError("inlet_series must hold steps 0..nsteps")
        self.cfg, self.nsteps, self.k = cfg, int(nsteps), int(k)
        self.sim = CoupledSim(cfg, prec=prec, device=device, compact=False)  # every cell on the recorded path
        self.device = self.sim.device
        self.cdt = self.sim.cdt
        # an FP64 run with FP32 checkpoints sweeps back in FP32: its adjoint state and segment slots
        # are FP32, and so are the copies of the per-cell fields the FP32 kernels read
        self.mixed = prec == "f64" and ckpt == "f32"
        self.rdt = torch.float32 if self.mixed else self.cdt
        n = cfg.n
        d = self.sim._desc
        self.state_bytes = int(abi.lib().fmx_adjoint_state_bytes(abi.C.byref(d)))
        self.ckpt_prec = abi.PREC_NAMES[ckpt]
        self.ckpt_bytes = int(abi.lib().fmx_adjoint_ckpt_bytes(abi.C.byref(d), self.ckpt_prec))
        nseg = (self.nsteps + self.k - 1) // self.k
        kk = min(self.k, self.nsteps)
        # segment store: one level by default; fewer slots (at least the two-level minimum, about
        # 2 sqrt(k)) make the reverse sweep keep sub-checkpoints within each segment
        single = segment_slots(kk, prec, ckpt)
        self.seg_slots = single if seg_slots is None else int(seg_slots)
        least = segment_slots(kk, prec, ckpt, two_level=True)
        if self.seg_slots < least:
            raise abi.FmxError(f"segments of {kk} steps need at least {least} segment slots")
        u8 = torch.uint8
        self.ckpt = torch.empty(nseg * self.ckpt_bytes, dtype=u8, device=self.device)
        slot_bytes = self.ckpt_bytes if self.mixed else self.state_bytes
        self.seg = torch.empty(max(self.seg_slots, 1) * slot_bytes, dtype=u8, device=self.device)
        self.lam_f = [torch.empty(19 * n, dtype=self.rdt, device=self.device) for _ in range(2)]
        self.lam_g = [torch.empty(7 * n, dtype=self.rdt, device=self.device) for _ in range(2)]
        self.lam_q = torch.empty(n, dtype=self.rdt, device=self.device)
        self.lam_phi = torch.empty(n, dtype=self.rdt, device=self.device)
        self.field_rev = self.tau_field_rev = None  # FP32 copies, made at backward()
        self.workspace = None  # sized at backward() (it follows the launch shape)
        self.grad_iso = torch.zeros(2, dtype=torch.float64, device=self.device)
        self.grad_field = torch.zeros(n, dtype=torch.float64, device=self.device) if self.sim.field is not None else None
        self.grad_series = (torch.zeros(self.sim.inlet_series.numel(), dtype=torch.float64, device=self.device)
                            if self.sim.inlet_series is not None else None)
        self.acc = None
        self.forward_done = False
        # counts forward() calls: a reverse sweep belongs to the forward run that wrote the checkpoints
        self.generation = 0

    # ------------------------------------------------------------ parameters
    def set_params(self, iso_a=None, iso_b=None, field=None, series=None):
        """Update the differentiable inputs in place (device tensors or arrays)."""
        torch = _torch()
        d = self.sim._desc
        if iso_a is not None:
            d.iso_a = float(iso_a)
        if iso_b is not None:
            d.iso_b = float(iso_b)
        if field is not None:
            if self.sim.field is None:
                raise abi.FmxError("this configuration has no field")
            self.sim.field.copy_(torch.as_tensor(field).to(self.device, self.cdt).reshape(-1))
        if series is not None:
            if self.sim.inlet_series is None:
                raise abi.FmxError("this configuration has no inlet series")
            self.sim.inlet_series.copy_(torch.as_tensor(series).to(self.device, torch.float64).reshape(-1))
        self.forward_done = False

    # ------------------------------------------------------------ passes
    def forward(self):
        """C_out(s) for s = 1..nsteps (float64 device tensor), from the configuration's initial state."""
        torch = _torch()
        sim = self.sim
        sim.cur, sim.step_index = 0, 0
        abi.call("fmx_coupled_init", sim._desc_p(), sim._bufs_p(), 0, float(self.cfg.phi0), sim._s())
        self.acc = sim.outlet_acc(self.nsteps)
        abi.call("fmx_adjoint_forward", sim._desc_p(), sim._bufs_p(self.acc), 1, self.nsteps, self.k, 0,
                 self.ckpt.data_ptr(), self.ckpt_prec, sim._s())
        sim.cur = self.nsteps & 1
        sim.step_index = self.nsteps
        self.generation += 1
        mean = torch.empty(self.nsteps, dtype=torch.float64, device=self.device)
        abi.call("fmx_outlet_mean", self.acc.data_ptr(), self.nsteps, self.cfg.ny * self.cfg.nz, mean.data_ptr(),
                 sim._s())
        self.forward_done = True
        return mean

    def backward(self, weight, generation=None):
        """Gradients of a loss with dL/dC_out(s) = weight[s-1]: dict of iso_a, iso_b (floats),
        field (per cell, or None) and series (indexed like inlet_series, or None).  `generation`, when
        given, is the value of self.generation right after the forward run these gradients belong to:
        a later forward() has overwritten that run's checkpoints and parameters, which is an error."""
        torch = _torch()
        if not self.forward_done:
            raise abi.FmxError("backward() needs a forward() with the current parameters")
        if generation is not None and generation != self.generation:
            raise abi.FmxError(f"this reverse sweep belongs to forward run {generation}, but the AdjointRun has "
                               f"since run forward {self.generation - generation} more time(s) and holds that run's "
                               f"checkpoints; use one AdjointRun per node of the autograd graph")
        w = torch.as_tensor(weight).to(self.device, torch.float64).contiguous()
        if w.numel() != self.nsteps:
            raise abi.FmxError(f"weight must have {self.nsteps} entries")
        a = abi.AdjointBufs()
        a.lam_f[0], a.lam_f[1] = self.lam_f[0].data_ptr(), self.lam_f[1].data_ptr()
        a.lam_g[0], a.lam_g[1] = self.lam_g[0].data_ptr(), self.lam_g[1].data_ptr()
        a.lam_q, a.lam_phi = self.lam_q.data_ptr(), self.lam_phi.data_ptr()
        sim = self.sim
        wb = int(abi.lib().fmx_adjoint_workspace_bytes(sim._desc_p()))
        if wb < 0:
            raise abi.FmxError("fmx_adjoint_workspace_bytes: bad descriptor")
        if self.workspace is None or self.workspace.numel() < wb:
            self.workspace = torch.empty(max(wb, 1), dtype=torch.uint8, device=self.device)
        a.workspace, a.workspace_bytes = self.workspace.data_ptr(), self.workspace.numel()
        a.outlet_weight = w.data_ptr()
        a.outlet_weight_len = w.numel()
        a.grad_iso = self.grad_iso.data_ptr()
        a.grad_field = self.grad_field.data_ptr() if self.grad_field is not None else None
        a.grad_series = self.grad_series.data_ptr() if self.grad_series is not None else None
        if self.mixed:
            if sim.field is not None:
                self.field_rev = sim.field.to(torch.float32)
                a.field_rev = self.field_rev.data_ptr()
            if sim.tau_field is not None:
                self.tau_field_rev = sim.tau_field.to(torch.float32)
                a.tau_field_rev = self.tau_field_rev.data_ptr()
        abi.call("fmx_adjoint_backward", sim._desc_p(), sim._bufs_p(), 1, self.nsteps, self.k, self.ckpt.data_ptr(),
                 self.ckpt_prec, self.seg.data_ptr(), self.seg_slots, abi.C.byref(a), sim._s())
        sim.synchronize()
        gi = self.grad_iso.cpu().numpy()
        return {
            "iso_a": float(gi[0]),
            "iso_b": float(gi[1]),
            "field": self.grad_field.clone() if self.grad_field is not None else None,
            "series": self.grad_series.clone() if self.grad_series is not None else None,
        }


def _autograd_function():
    torch = _torch()

    class CoupledBreakthrough(torch.autograd.Function):
        """C_out(s), s = 1..nsteps, as a function of (iso_a, iso_b, field, series).

        iso_a / iso_b: 0-d tensors; field: per-cell tensor or None; series: per-step tensor
        (index = step) or None.  Gradients flow to every tensor input that requires them.
        """

        @staticmethod
        def forward(ctx, run, iso_a, iso_b, field=None, series=None):
            run.set_params(iso_a=float(iso_a), iso_b=float(iso_b),
                           field=None if field is None else field.detach(),
                           series=None if series is None else series.detach())
            ctx.run = run
            ctx.meta = (iso_a.dtype, iso_a.device, None if field is None else (field.dtype, field.device),
                        None if series is None else (series.dtype, series.device))
            out = run.forward()
            ctx.generation = run.generation
            return out

        @staticmethod
        def backward(ctx, grad_c):
            run = ctx.run
            g = run.backward(grad_c.detach().to(torch.float64), generation=ctx.generation)
            adt, adev, fmeta, smeta = ctx.meta
            ga = torch.tensor(g["iso_a"], dtype=adt, device=adev)
            gb = torch.tensor(g["iso_b"], dtype=adt, device=adev)
            gf = None if fmeta is None else g["field"].to(fmeta[1], fmeta[0])
            gs = None if smeta is None else g["series"].to(smeta[1], smeta[0])
            return None, ga, gb, gf, gs

    return CoupledBreakthrough


class _Lazy:
    _fn = None

    def __getattr__(self, name):
        if _Lazy._fn is None:
            _Lazy._fn = _autograd_function()
        return getattr(_Lazy._fn, name)


# torch.autograd.Function subclass, created on first use so that importing this module needs no torch
CoupledBreakthrough = _Lazy()


def breakthrough_loss_weights(c_out, target) -> np.ndarray:
    """dL/dC_out for L = sum_s (C_out(s) - C*(s))^2."""
    return 2.0 * (np.asarray(c_out, np.float64) - np.asarray(target, np.float64))
