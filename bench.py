"""Benchmark of the fused coupled D3Q19 + D3Q7 step on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload C3] [--prec f32]

Default workload: C3, the 512^3 overlapping sphere pack (porosity 0.4, 134M
cells) with the Langmuir LDF, FP32 storage -- the largest single-GPU
configuration -- with the library's default NS layout for it (row-compacted NS
planes that leave the frozen solid cells out; --layout box for the box layout).
Prints ONE JSON line: MLUPS over all cells (solids included),
the HBM roofline of the fused kernel (algorithmic bytes per cell update x cells
/ launch time measured with HIP events on the launch stream, against 8 TB/s,
plus the HBM bytes per launch from the committed rocprofv3 counter summary of
the same kernel and workload), and the CPU oracle (restatement of the
reference update) timed serially on a bounded sample of the same workload.
Extra keys: the same box with FP16 / BF16 storage, the 256x128x128 pack (C2),
and the adjoint/forward time ratio of the checkpointed gradient on the
differentiation workload (C4, 128x64x64 partial bounce-back, 4000 steps,
checkpoint every 200): headed by the FP64 run with FP32 checkpoints and an FP32
reverse sweep, with the all-FP32 and all-FP64 gradients beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per coupled cell update: 19 + 7 populations read and written once,
# adsorbed load and lagged concentration read and written once, one flag byte
B_ALG = {"f32": 225, "f64": 449, "f16": 121, "bf16": 121}
B_FIELD = {"f32": 4, "f64": 8, "f16": 4, "bf16": 4}  # per-cell porosity / solid-fraction read
B_NS = {"f32": 152, "f64": 304, "f16": 76, "bf16": 76}  # of B_ALG: the 19 NS populations read and written


def split_step():
    """Whether compact runs take the split step (the library's default; FMX_SPLIT=0 turns it off)."""
    return os.environ.get("FMX_SPLIT", "1").strip() not in ("0",)


def compact_bytes(sim, prec):
    """Algorithmic bytes per cell update of a run in row-compacted NS planes: the NS populations of the
    active cells only (the frozen solid cells' are never moved), the rest as B_ALG, and with the split
    step each active cell's 8-byte record read and its velocity (three values of the compute precision)
    written and read; and the active fraction."""
    from fluidmaxxer_amd import abi
    n = sim.cfg.n
    active = int((sim.links & abi.LINK_FROZEN).eq(0).sum().item())
    a = active / n
    extra = a * (8 + 2 * 3 * (8 if prec == "f64" else 4)) if split_step() else 0.0
    return B_ALG[prec] - B_NS[prec] * (1.0 - a) + extra, a


def cpu_info():
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(cfg, prec, budget_s):
    """The CPU oracle (restatement of the reference update) on a bounded sample of the workload: serial
    (the baseline), then the same system with its per-cell loops on all the threads OpenMP is given.
    The first serial step is timed; more steps are timed only while they fit in `budget_s` (a 512^3
    step alone takes 15-35 s on one core)."""
    from oracle import ref_cpu as R
    from tests.lattice import oracle_setup

    oprec = 8 if prec == "f64" else 4
    model, ncpu = cpu_info()
    R.lib().orc_set_threads(1)
    sysm = R.System(oracle_setup(cfg, R), oprec)
    out = []
    try:
        _, _, _, t1 = sysm.run(1)
        steps, t = 1, t1
        if 2 * t1 < budget_s:  # small boxes: the first step was a warm-up, time a window of steps
            steps = int(max(1, min(200, budget_s / max(t1, 1e-6))))
            _, _, _, t = sysm.run(steps)
        timed = [(1, steps, t)]
        used = int(R.lib().orc_set_threads(0))  # OpenMP default: OMP_NUM_THREADS, else every logical CPU
        if used > 1:
            _, _, _, tp = sysm.run(1)
            psteps = int(max(1, min(200, budget_s / 2 / max(tp, 1e-6))))
            _, _, _, tp = sysm.run(psteps)
            timed.append((used, psteps, tp))
    finally:
        sysm.close()
        R.lib().orc_set_threads(0)
    for used, steps, t in timed:
        mode = "serial" if used == 1 else f"OpenMP, {used} threads"
        out.append({
            "value": round(cfg.n * steps / t / 1e6, 3),
            "unit": "MLUPS",
            "cores": used,
            "kind": "port",
            "sample": f"{cfg.meta['workload']} {cfg.nx}x{cfg.ny}x{cfg.nz} (the whole box), {'FP64' if oprec == 8 else 'FP32'} "
                      f"oracle ({mode}), {steps} step{'s' if steps > 1 else ''} timed ({t:.1f} s), host '{model}' "
                      f"({ncpu} logical CPUs)",
        })
    return out[0], (out[1] if len(out) > 1 else None)


def stream_probe(gib=2, reps=10):
    """Achievable HBM bandwidth of plain streaming kernels on this GPU (the practical ceiling)."""
    import torch

    from fluidmaxxer_amd import abi
    n = gib * (1 << 30) // 4
    a = torch.empty(n, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    out = {}

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps / 1e3

    nbytes = n * 4
    out["copy_GBps"] = round(2 * nbytes / timed(lambda: b.copy_(a)) / 1e9, 1)
    out["runtime_copy_GBps"] = round(2 * nbytes / timed(
        lambda: abi.call("fmx_memcpy_d2d", b.data_ptr(), a.data_ptr(), nbytes, torch.cuda.current_stream().cuda_stream)) / 1e9, 1)
    s = torch.cuda.current_stream().cuda_stream
    out["stream_copy_GBps"] = round(2 * nbytes / timed(
        lambda: abi.call("fmx_stream_probe", 0, b.data_ptr(), a.data_ptr(), nbytes, s)) / 1e9, 1)
    sink = torch.zeros(1, dtype=torch.int32, device="cuda")
    out["stream_read_GBps"] = round(nbytes / timed(
        lambda: abi.call("fmx_stream_probe", 1, sink.data_ptr(), a.data_ptr(), nbytes, s)) / 1e9, 1)
    out["torch_sum_read_GBps"] = round(nbytes / timed(lambda: a.sum()) / 1e9, 1)
    for mode, esz in ((2, 2), (3, 4), (4, 8)):
        moved = 2 * (nbytes // (26 * esz)) * 26 * esz
        out[f"plane_copy_{8 * esz}bit_GBps"] = round(moved / timed(
            lambda: abi.call("fmx_stream_probe", mode, b.data_ptr(), a.data_ptr(), nbytes, s)) / 1e9, 1)
    out["write_GBps"] = round(nbytes / timed(
        lambda: abi.call("fmx_fill_zero", b.data_ptr(), nbytes, torch.cuda.current_stream().cuda_stream)) / 1e9, 1)
    out["bytes"] = nbytes
    return {"stream_probe": out}


def adjoint_timing(workload="C4", prec="f32", k=200, steps=None, ckpt=None, two_level=False):
    """Time of the checkpointed reverse sweep (recompute + adjoint steps) against the plain forward
    run of the same steps, on a differentiation workload (default C4: 128x64x64 partial bounce-back;
    C5 at 256x128x128 with 16-bit checkpoints every 500 steps and a two-level segment store)."""
    import torch

    from fluidmaxxer_amd import abi, workloads
    from fluidmaxxer_amd.adjoint import AdjointRun, segment_slots

    nsteps = steps or workloads.SPECS[workload]["steps"]
    cfg = workloads.build(workload, steps=nsteps)
    seg_slots = None
    if two_level:
        seg_slots = segment_slots(k, prec, ckpt, two_level=True)
    run = AdjointRun(cfg, nsteps, k, prec, ckpt=ckpt, seg_slots=seg_slots)
    sim = run.sim
    ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731

    def timed(fn):
        torch.cuda.synchronize()
        e0, e1 = ev(), ev()
        e0.record(sim.stream)
        out = fn()
        e1.record(sim.stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1), out

    # plain forward: the same steps with the outlet reduction, no checkpoints
    def plain():
        abi.call("fmx_coupled_init", sim._desc_p(), sim._bufs_p(), 0, float(cfg.phi0), sim._s())
        sim.cur, sim.step_index = 0, 0
        acc = sim.outlet_acc(nsteps)
        sim.launch_only(nsteps, acc=acc)

    plain()  # warm-up
    t_plain, _ = timed(plain)
    t_fwd, c = timed(run.forward)
    w = 2.0 * c  # loss sum_s C_out(s)^2
    run.backward(w)  # warm-up
    t_fwd, c = timed(run.forward)
    t_bwd, _ = timed(lambda: run.backward(w))
    res = {
        "workload": f"{workload}: {cfg.nx}x{cfg.ny}x{cfg.nz} {cfg.ns_op} {cfg.isotherm}, {nsteps} steps, "
                    f"checkpoint every {k} ({ckpt or prec}), {prec} run, {run.seg_slots} segment slots",
        "run_precision": prec,
        "sweep_precision": "f64" if (prec == "f64" and ckpt in (None, "f64")) else "f32",
        "store_gib": round((run.ckpt.numel() + run.seg.numel()) / 2 ** 30, 2),
        "forward_ms": round(t_plain, 3),
        "forward_with_checkpoints_ms": round(t_fwd, 3),
        "backward_ms": round(t_bwd, 3),
        # the gradient's whole cost (checkpointed forward + reverse sweep) over the plain forward run
        "ratio": round((t_fwd + t_bwd) / t_plain, 3),
        "reverse_sweep_ratio": round(t_bwd / t_plain, 3),
    }
    del run
    torch.cuda.empty_cache()
    return res


def adjoint_c5_accuracy(steps=None, k=500):
    """C5 at its full 256x128x128: the FP32 run with FP16-stored checkpoints every k steps and FP32
    recompute against the FP64 adjoint (both with two-level segment stores, so the FP64 store fits in
    about 80 GB); relative l2 errors of dL/d(200 knots), dL/d(solid fraction), the isotherm gradients
    and C_out.  As in the reduced-size test, the fluid cells' solid fraction and the concentration clamp
    are 0 (the specified +-0.05 noise stops the flow), and the loss is sum (C_out - 0.85 C_out^FP64)^2."""
    import numpy as np
    import torch

    from fluidmaxxer_amd import abi, workloads
    from fluidmaxxer_amd.adjoint import AdjointRun, segment_slots

    nsteps = steps or workloads.SPECS["C5"]["steps"]
    cfg = workloads.build("C5", steps=nsteps)
    cfg.field = np.where(cfg.field > 0.5, cfg.field, 0.0)
    cfg.min_conc = 0.0
    knots = np.asarray(cfg.meta["knots"])
    # d(per-step inlet value)/d(knot) of the per-step linear interpolation
    s = np.arange(nsteps + 1) * (len(knots) - 1) / nsteps
    i0 = np.minimum(np.floor(s).astype(int), len(knots) - 2)
    B = np.zeros((nsteps + 1, len(knots)))
    B[np.arange(nsteps + 1), i0] = 1 - (s - i0)
    B[np.arange(nsteps + 1), i0 + 1] = s - i0

    def gradients(prec, ckpt, target_of):
        seg = segment_slots(k, prec, ckpt, two_level=True)
        run = AdjointRun(cfg, nsteps, k, prec, ckpt=ckpt, seg_slots=seg)
        c = run.forward().cpu().numpy()
        target = target_of(c)
        t0 = time.perf_counter()
        g = run.backward(2.0 * (c - target))
        torch.cuda.synchronize()
        out = {"c": c, "target": target, "seconds": time.perf_counter() - t0, "iso_a": g["iso_a"],
               "iso_b": g["iso_b"], "field": g["field"].cpu().numpy(), "knots": B.T @ g["series"].cpu().numpy()}
        del run
        torch.cuda.empty_cache()
        return out

    def rel_l2(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    def errors(a, b):
        return {"knots": rel_l2(a["knots"], b["knots"]), "field": rel_l2(a["field"], b["field"]),
                "c_out": rel_l2(a["c"], b["c"]), "iso_a": abs(a["iso_a"] - b["iso_a"]) / abs(b["iso_a"]),
                "iso_b": abs(a["iso_b"] - b["iso_b"]) / abs(b["iso_b"])}

    g64 = gradients("f64", None, lambda c: 0.85 * c)
    g16 = gradients("f32", "f16", lambda c: g64["target"])
    g32 = gradients("f32", None, lambda c: g64["target"])
    gm = gradients("f64", "f32", lambda c: g64["target"])
    return {
        "workload": f"C5: {cfg.nx}x{cfg.ny}x{cfg.nz} pbb langmuir, {nsteps} steps; FP32 run, FP16 checkpoints every {k}, "
                    f"FP32 recompute, against the FP64 adjoint (two-level stores)",
        "c_out_last": float(g64["c"][-1]),
        "rel_l2": errors(g16, g64),
        "rel_l2_fp32_ckpt_vs_fp64": errors(g32, g64),
        "rel_l2_fp16_ckpt_vs_fp32_ckpt": errors(g16, g32),
        "rel_l2_fp64_run_fp32_sweep_vs_fp64": errors(gm, g64),
        "backward_seconds": {"f64": round(g64["seconds"], 2), "f32_f16ckpt": round(g16["seconds"], 2),
                             "f32": round(g32["seconds"], 2), "f64_run_f32_sweep": round(gm["seconds"], 2)},
    }


def precision_study(workload="C3", steps=None, every=1000, precs=("f64", "f32", "f16", "bf16", "f16r", "bf16r")):
    """The workload run for `steps` steps (C3: 512^3, 20000) in each storage precision, against the FP64
    run: per-field errors after the last step (rho, u of the populations as the reference holds them, C,
    q), the outlet breakthrough curve C_out(t) of every step, the breakthrough surface C(x, t) -- the
    plane-averaged concentration of every x-plane every `every` steps, which carries the front before it
    reaches the outlet -- and the FP64 sums of C and q (fmx_reduce_mass) at the same times.  Errors:
    max |a - b| / max |b| (and the relative l2 norm for the fields).  A trailing "r" (f16r, bf16r): the
    16-bit storage with its density reference moved to the mean density of the flow every `every` steps
    (CoupledSim.recenter); without it the reference stays at rho0."""
    import torch

    from fluidmaxxer_amd import workloads
    from fluidmaxxer_amd.coupled import CoupledSim

    nsteps = steps or workloads.SPECS[workload]["steps"]
    every = max(1, min(every, nsteps))
    nsteps -= nsteps % every
    cfg = workloads.build(workload, steps=nsteps + 1)
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    ref = None
    out = {"workload": f"{workload}: {nx}x{ny}x{nz} {cfg.ns_op} {cfg.isotherm}, porosity {cfg.meta['porosity']:.4f}, "
                       f"{nsteps} steps from the equilibrium start, C(x, t) every {every} steps"}

    def linf(a, b):
        return float((a - b).abs().max() / b.abs().max().clamp_min(1e-300))

    def l2(a, b):
        return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b).clamp_min(1e-300))

    for name in precs:
        prec, recentre = name.rstrip("r"), name.endswith("r")
        t0 = time.perf_counter()
        sim = CoupledSim(cfg, prec=prec)
        curve, surface, mass = [], [], []
        for blk in range(nsteps // every):
            curve.append(sim.run(every))  # the library's default layout (row-compacted, split step)
            mass.append(sim.mass())  # the FP64 sums of the state after the block (before any recentring)
            if recentre:
                sim.recenter()
            surface.append(sim.phi.view(nz, ny, nx).double().mean(dim=(0, 1)))
            print(f"[precision] {name} step {(blk + 1) * every}/{nsteps} C_out {float(curve[-1][-1]):.3e} "
                  f"front C(x={nx // 4}) {float(surface[-1][nx // 4]):.3e} ({time.perf_counter() - t0:.0f} s)",
                  flush=True)
        curve, surface, mass = torch.cat(curve), torch.stack(surface), torch.stack(mass)
        rho, ux, uy, uz = sim.moments()
        fields = {"rho": rho.double(), "ux": ux.double(), "uy": uy.double(), "uz": uz.double(),
                  "C": sim.phi.double().clone(), "q": sim.q.double().clone()}
        f_ref = sim.f_ref
        del rho, ux, uy, uz, sim
        torch.cuda.empty_cache()
        res = {"seconds": round(time.perf_counter() - t0, 1), "f_ref": round(f_ref, 6), "c_out_last": float(curve[-1]),
               "front_c_quarter": float(surface[-1][nx // 4]), "nonfinite": float(mass[:, 3].sum())}
        if ref is None:
            ref = dict(fields=fields, curve=curve, surface=surface, mass=mass)
        else:
            res["fields_linf"] = {k: round(linf(v, ref["fields"][k]), 9) for k, v in fields.items()}
            res["fields_l2"] = {k: round(l2(v, ref["fields"][k]), 9) for k, v in fields.items()}
            res["c_out_curve"] = linf(curve, ref["curve"])
            res["breakthrough_surface"] = linf(surface, ref["surface"])
            res["sum_C_plus_q"] = linf(mass[:, 0] + mass[:, 1], ref["mass"][:, 0] + ref["mass"][:, 1])
            del fields
        out[name] = res
        print(f"[precision] {name}: {json.dumps(res)}", flush=True)
    return out


def time_steps(sim, steps, warmup):
    """Warm-up steps, then exactly `steps` launches between two HIP events on the launch stream (with the
    per-step outlet reduction); returns ms per step."""
    import torch

    if warmup:
        sim.run(warmup, log_outlet=True)
    acc = sim.outlet_acc(steps)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(sim.stream)
    sim.launch_only(steps, acc=acc)
    e1.record(sim.stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps, acc


def storage_timing(cfg, precs, steps=100, warmup=20):
    """The same workload for each storage precision: step time, MLUPS and the fraction of the
    algorithmic roofline (121 B/cell for FP16 / BF16 storage, 449 for FP64).  A name with ":box"
    (e.g. "f32:box") times that storage in the box layout instead of the library's default."""
    import torch

    from fluidmaxxer_amd.coupled import CoupledSim

    out = {"workload": f"{cfg.meta['workload']}: {cfg.nx}x{cfg.ny}x{cfg.nz} {cfg.ns_op} {cfg.isotherm}, "
                       f"porosity {cfg.meta['porosity']:.4f}, {steps} steps after {warmup}"}
    for name in precs:
        prec, _, layout = name.partition(":")
        sim = CoupledSim(cfg, prec=prec, compact=False if layout == "box" else None)
        ms_step, acc = time_steps(sim, steps, warmup)
        b_alg = B_ALG[prec] + (B_FIELD[prec] if cfg.field is not None else 0)
        achieved = b_alg * cfg.n / (ms_step / 1e3) / 1e9
        out[name.replace(":", "_")] = {"mlups": round(cfg.n / (ms_step / 1e3) / 1e6, 1),
                                       "ms_per_step": round(ms_step, 4),
                                       "roofline_frac": round(achieved / HBM_PEAK_GBS, 4),
                                       "ns_layout": "row-compacted" if sim.compact else "box"}
        del sim, acc
        torch.cuda.empty_cache()
    return out


def traffic_from_profile(workload, prec, compact=False):
    """HBM bytes per launch of the fused kernel from the committed counter summary (FETCH_SIZE and
    WRITE_SIZE runs, scaled by a plain copy's reading) of this workload, storage and NS layout, and its
    path."""
    rel = os.path.join("profiles", f"traffic_{'rc_' if compact else ''}{workload.lower()}_{prec}.json")
    path = os.path.join(ROOT, rel)
    if not os.path.exists(path):
        return None, None
    with open(path) as fh:
        d = json.load(fh)
    return d.get("hbm_bytes_per_launch"), rel


def targets(line):
    """The 512^3 targets beside what this run measured; the storage error and the gradient accuracy are
    not measured by the timing run and name the committed record they come from."""
    out = {"fp32_roofline_frac": {"target": 0.70, "measured": line["roofline"]["frac"]}}
    if "compact_layout" in line["roofline"]:
        out["fp32_roofline_frac"]["measured_on_the_bytes_the_compact_layout_moves"] = \
            line["roofline"]["compact_layout"]["frac"]
    if "storage" in line:
        out["fp16_over_fp32_mlups"] = {"target": 1.6, "measured": line["storage"]["f16_over_f32"]}
        if "f16_over_f32_box" in line["storage"]:
            out["fp16_over_fp32_mlups"]["measured_with_both_in_the_box_layout"] = \
                line["storage"]["f16_over_f32_box"]
    rel = os.path.join("profiles", "precision_c3.json")
    if os.path.exists(os.path.join(ROOT, rel)):
        with open(os.path.join(ROOT, rel)) as fh:
            study = json.load(fh)
        for name, label in (("f16r", "recentred density reference"), ("f16", "density reference held at rho0")):
            if name in study:
                out[f"fp16_breakthrough_error_vs_fp64 ({label})"] = {
                    "target": 1e-3, "measured": study[name]["breakthrough_surface"],
                    "source": f"{rel} (bench.py --precision-study: C(x, t) over 20,000 steps); not this run"}
    if "adjoint" in line:
        out["gradient_cost_over_forward"] = {"target": 3.0, "measured": line["adjoint"]["ratio"],
                                             "run": "FP64 run, FP32 checkpoints and FP32 reverse sweep (C4)"}
    out["gradient_error_vs_fd"] = {"target": 1e-4,
                                   "measured": "FP64 adjoint 1e-10 .. 6e-8 against central differences; the FP64 "
                                               "run with the FP32 sweep 1.5e-5 .. 2.5e-5 against the FP64 adjoint",
                                   "source": "tests/test_gpu_adjoint.py; not this run"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="C3")
    ap.add_argument("--dims", default=None, help="override box, e.g. 256x256x256")
    ap.add_argument("--prec", default="f32", choices=["f32", "f64", "f16", "bf16"])
    ap.add_argument("--layout", default="default", choices=["default", "box", "compact"],
                    help="NS layout of the timed sim: the library's default (row-compacted), or forced")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="CPU baseline budget (one serial step is always timed)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-outlet", action="store_true", help="skip the per-step outlet reduction")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="first run one plain device-to-device copy of BYTES (HBM-counter calibration)")
    ap.add_argument("--stream-probe", action="store_true",
                    help="instead of the benchmark, measure plain streaming bandwidth (copy / read / write)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward timing on C4")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the adjoint/forward timings on C4 (FP64 run with FP32 sweep, FP32, FP64)")
    ap.add_argument("--no-storage-sweep", "--no-large-box", dest="no_storage_sweep", action="store_true",
                    help="skip the FP16 / BF16 storage timings of the same workload")
    ap.add_argument("--no-c2", action="store_true", help="skip the 256x128x128 (C2) FP32 timing")
    ap.add_argument("--adjoint-steps", type=int, default=None, metavar="N",
                    help="steps of the adjoint timing (default: the workload's own)")
    ap.add_argument("--adjoint-c5-accuracy", action="store_true",
                    help="only C5 at full size: FP16-checkpoint gradients against the FP64 adjoint")
    ap.add_argument("--precision-study", action="store_true",
                    help="only the storage-precision study: the workload (default C3, 512^3) for its own steps "
                         "(20000) in FP64, FP32, FP16 and BF16, errors against FP64")
    ap.add_argument("--adjoint-c5", action="store_true",
                    help="only the adjoint/forward timing on C5 as specified (256x128x128, 20000 steps, FP16 "
                         "checkpoints every 500 steps, FP32 recompute; one-level and two-level segment stores)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU only (the slab-decomposed multi-GPU path is a design, "
                                   "not built), so --gpus must be 1"}))
        sys.exit(2)

    import torch

    from fluidmaxxer_amd import workloads
    from fluidmaxxer_amd.coupled import CoupledSim

    if args.stream_probe:
        print(json.dumps(stream_probe()), flush=True)
        return

    if args.precision_study:
        print(json.dumps({"precision_study": precision_study(args.workload, args.steps)}), flush=True)
        return

    if args.adjoint_c5_accuracy:
        print(json.dumps({"adjoint_c5_accuracy": adjoint_c5_accuracy(args.adjoint_steps)}), flush=True)
        return

    if args.adjoint_c5:
        # one level: 499 FP32 segment slots (about 220 GiB); two levels: about 2 sqrt(500) slots
        for two_level in (False, True):
            print(json.dumps({"adjoint": adjoint_timing("C5", "f32", k=500, steps=args.adjoint_steps, ckpt="f16",
                                                        two_level=two_level)}), flush=True)
        # the FP64 run with FP32 checkpoints and an FP32 reverse sweep: one level (499 FP32 segment slots,
        # about 236 GiB with the checkpoints) and two levels
        for two_level in (False, True):
            print(json.dumps({"adjoint": adjoint_timing("C5", "f64", k=500, steps=args.adjoint_steps, ckpt="f32",
                                                        two_level=two_level)}), flush=True)
        return

    if args.adjoint_only:
        for prec, ckpt in (("f64", "f32"), ("f32", None), ("f64", None)):
            print(json.dumps({"adjoint": adjoint_timing(prec=prec, ckpt=ckpt, steps=args.adjoint_steps)}), flush=True)
        return

    if args.calibrate_copy:
        from fluidmaxxer_amd import abi
        src = torch.empty(args.calibrate_copy // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        abi.call("fmx_memcpy_d2d", dst.data_ptr(), src.data_ptr(), src.numel() * 4, None)
        torch.cuda.synchronize()
        del src, dst

    dims = tuple(int(v) for v in args.dims.split("x")) if args.dims else None
    spec_steps = workloads.SPECS[args.workload]["steps"]
    steps = args.steps if args.steps is not None else min(spec_steps, 200)
    cfg = workloads.build(args.workload, dims=dims, steps=max(steps + args.warmup + 1, spec_steps))
    n = cfg.n
    sim = CoupledSim(cfg, prec=args.prec, compact={"default": None, "box": False, "compact": True}[args.layout])
    log = not args.no_outlet

    # warm-up (untimed), then exactly `steps` steps between two events on the launch stream
    if args.warmup:
        sim.run(args.warmup, log_outlet=log)
    acc = sim.outlet_acc(steps) if log else None
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(sim.stream)
    sim.launch_only(steps, acc=acc)
    stop.record(sim.stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop)
    ms_step = ms / steps
    mlups = n * steps / (ms / 1e3) / 1e6
    fluid = float((cfg.solid == 0).mean()) if cfg.ns_op == "bgk" else float(np.mean(cfg.field < 0.5))
    b_alg = B_ALG[args.prec] + (B_FIELD[args.prec] if cfg.field is not None else 0)
    achieved = b_alg * n / (ms_step / 1e3) / 1e9
    c_out_last = None
    if log:
        from fluidmaxxer_amd import abi
        mean = torch.empty(steps, dtype=torch.float64, device=sim.device)
        abi.call("fmx_outlet_mean", acc.data_ptr(), steps, cfg.ny * cfg.nz, mean.data_ptr(), sim.stream.cuda_stream)
        torch.cuda.synchronize()
        c_out_last = float(mean[-1].item())

    compact = sim.compact
    b_moved, active = compact_bytes(sim, args.prec) if compact else (b_alg, 1.0)
    traffic, traffic_src = traffic_from_profile(args.workload, args.prec, compact) if dims is None else (None, None)
    line = {
        "metric": "Coupled D3Q19+D3Q7 MLUPS & %HBM roofline; adjoint/forward time ratio",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f32": "fp32", "f64": "fp64", "f16": "fp16-storage/fp32-compute",
                  "bf16": "bf16-storage/fp32-compute"}[args.prec],
        "data": "synthetic (seeded sphere-pack geometry, equilibrium start)",
        "config": {
            "workload": f"{args.workload}: {cfg.nx}x{cfg.ny}x{cfg.nz} coupled NS-AD, {cfg.ns_op} NS operator, "
                        f"{cfg.isotherm} LDF, porosity {cfg.meta['porosity']:.4f}, seed {cfg.meta['seed']}",
            "cells": n,
            "fluid_fraction": round(fluid, 4),
            "fluid_mlups": round(mlups * fluid, 1),
            "storage": args.prec,
            "outlet_reduction_per_step": log,
            "c_out_last_step": c_out_last,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": (f"{traffic_src}: rocprofv3 FETCH_SIZE and WRITE_SIZE runs of this kernel and workload, "
                               "scaled by a plain copy's reading; not measured in this run") if traffic_src else None,
            "bytes_per_cell_update": b_alg,
            "kernel": (("k_ns_split + k_ad_split (row-compacted NS planes: NS over the active cells, then AD over "
                        "the box; two launches per step" if split_step() else
                        "k_coupled_rc (row-compacted NS planes; one launch per step") +
                       ", plus one row-table build and one copy in and out per run)") if compact
                      else "k_coupled (one launch per step)",
        },
    }
    if compact:
        # the same step time over the bytes the row-compacted layout moves: the frozen solid cells' NS
        # populations (B_NS of each) are neither read nor written; the split step hands each active
        # cell's velocity over in memory
        line["roofline"]["compact_layout"] = {
            "active_fraction": round(active, 4),
            "bytes_per_cell_update": round(b_moved, 1),
            "achieved": round(b_moved * n / (ms_step / 1e3) / 1e9, 1),
            "frac": round(b_moved * n / (ms_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            # 225 B/cell counts every cell's NS populations, which this layout does not all move, so
            # the roofline fraction on it can pass 1; this one, on the bytes moved, cannot
            "note": "roofline.frac counts 225 B per cell update including the frozen solid cells' NS "
                    "populations, which are not moved; it can exceed 1. compact_layout.frac uses the bytes "
                    "moved.",
        }
    line["config"]["ns_layout"] = "row-compacted (frozen solid cells left out)" if compact else "box"
    del sim, acc
    torch.cuda.empty_cache()
    if not args.no_storage_sweep and args.prec == "f32":
        extra = ("f16", "bf16", "f32:box", "f16:box") if compact else ("f16", "bf16")
        st = storage_timing(cfg, extra, steps=steps, warmup=min(args.warmup, 20))
        st["f16_over_f32"] = round(st["f16"]["mlups"] / mlups, 3)
        st["bf16_over_f32"] = round(st["bf16"]["mlups"] / mlups, 3)
        if compact:
            # like for like in the box layout, and what the row-compacted layout gains in each storage
            st["f16_over_f32_box"] = round(st["f16_box"]["mlups"] / st["f32_box"]["mlups"], 3)
            st["f32_over_f32_box"] = round(mlups / st["f32_box"]["mlups"], 3)
            st["f16_over_f16_box"] = round(st["f16"]["mlups"] / st["f16_box"]["mlups"], 3)
        line["storage"] = st
    if not args.no_c2 and args.workload != "C2" and dims is None:
        c2 = workloads.build("C2", steps=max(steps, 100) + 21)
        line["c2"] = storage_timing(c2, ("f32", "f32:box", "f16"), steps=max(steps, 100), warmup=20)
        del c2
    if not args.no_adjoint and args.prec in ("f32", "f64"):
        # the FP64 run with an FP32 reverse sweep: FP64 forward curve, gradients within 1e-5 of the FP64
        # adjoint; ratio over the FP64 forward run
        line["adjoint"] = adjoint_timing(prec="f64", ckpt="f32")
        line["adjoint_single_precision"] = {p: adjoint_timing(prec=p) for p in ("f32", "f64")}
    if args.workload == "C3" and dims is None and args.prec == "f32":
        line["targets"] = targets(line)
        # the practical ceiling of this box, measured in the same process: boxes differ in what a plain
        # stream reaches, and the step's roofline fraction moves with it
        probe = stream_probe(gib=2, reps=5)["stream_probe"]
        line["same_box_streams"] = {
            "copy_GBps": probe["stream_copy_GBps"], "read_GBps": probe["stream_read_GBps"],
            "write_GBps": probe["write_GBps"],
            "step_achieved_over_copy": round(achieved / probe["stream_copy_GBps"], 3),
        }
        if compact:
            line["same_box_streams"]["step_moved_over_copy"] = round(
                line["roofline"]["compact_layout"]["achieved"] / probe["stream_copy_GBps"], 3)
    if not args.no_cpu_baseline:
        line["cpu_baseline"], all_cores = cpu_baseline(cfg, args.prec, args.cpu_seconds)
        if all_cores is not None:
            line["cpu_baseline_all_cores"] = all_cores
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
