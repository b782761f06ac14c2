"""The five benchmark/parity workloads, built from seeded sphere packs.

C1  64x32x32,   non-overlapping r=4 spheres, porosity ~0.6, BGK tau 0.8, u_in 0.05,
    AD tau 0.7, inlet C=1, linear isotherm K_d 0.5, k 0.01, 1000 steps
C2  256x128x128, overlapping r~U[4,8] spheres, porosity 0.4, tau 0.6, u_in 0.03,
    AD tau 0.55, Langmuir q_max 2, K_L 5, k 0.02, 10000 steps
C3  512^3, C2 physics, 20000 steps (precision study: f32 / f16 / bf16 / f64 storage)
C4  128x64x64, porosity 0.5, C2 physics with partial bounce-back solid fraction,
    4000 steps (differentiation)
C5  256x128x128, porosity 0.4, inlet concentration series of 200 knots, partial
    bounce-back, 20000 steps (differentiation + precision)

Inlet: VelocityDirichletEq (VelocityDirichlet.cu:54-84), which holds the inlet
density at rho0.  main.cpp's VelocityDirichletNEQ takes the density from the
neighbour cell (VelocityDirichletNEQ.cu:86-94); against these sphere packs that
feeds back -- the inflow rho_N u_in grows with the pressure it builds -- and the
inlet density diverges geometrically (C1: x10 in 1000 steps), so it is available
but not used by the workloads.

Walls: solid layers y in {0, ny-1} and z in {0, nz-1}; periodic y/z wrap as in
main.cpp:143-145 (the walls make it immaterial for the NS lattice).  Sorption
region: sphere voxels.  Porosity = fluid cells / all cells, walls counted as
solid.
"""
from __future__ import annotations

import numpy as np

from .coupled import CoupledConfig
from .geometry import sphere_pack

SPECS = {
    "C1": dict(dims=(64, 32, 32), seed=1, r=(4.0, 4.0), porosity=0.6, overlap=False, steps=1000,
               tau=0.8, u_in=0.05, tau_ad=0.7, isotherm="linear", iso_a=0.5, iso_b=0.0, k_ldf=0.01),
    "C2": dict(dims=(256, 128, 128), seed=2, r=(4.0, 8.0), porosity=0.4, overlap=True, steps=10000,
               tau=0.6, u_in=0.03, tau_ad=0.55, isotherm="langmuir", iso_a=2.0, iso_b=5.0, k_ldf=0.02),
    "C3": dict(dims=(512, 512, 512), seed=3, r=(4.0, 8.0), porosity=0.4, overlap=True, steps=20000,
               tau=0.6, u_in=0.03, tau_ad=0.55, isotherm="langmuir", iso_a=2.0, iso_b=5.0, k_ldf=0.02),
    "C4": dict(dims=(128, 64, 64), seed=4, r=(4.0, 8.0), porosity=0.5, overlap=True, steps=4000,
               tau=0.6, u_in=0.03, tau_ad=0.55, isotherm="langmuir", iso_a=2.0, iso_b=5.0, k_ldf=0.02,
               ns_op="pbb"),
    "C5": dict(dims=(256, 128, 128), seed=5, r=(4.0, 8.0), porosity=0.4, overlap=True, steps=20000,
               tau=0.6, u_in=0.03, tau_ad=0.55, isotherm="langmuir", iso_a=2.0, iso_b=5.0, k_ldf=0.02,
               ns_op="pbb", series_knots=200),
}


def pbb_fraction(sphere: np.ndarray, solid: np.ndarray, seed: int) -> np.ndarray:
    """Solid fraction for partial bounce-back: 1 in walls and spheres, 0 elsewhere,
    plus U[-0.05, 0.05] noise, clipped to [0, 1]."""
    rng = np.random.default_rng(seed)
    base = solid.astype(np.float64)
    return np.clip(base + rng.uniform(-0.05, 0.05, size=base.shape), 0.0, 1.0)


def series_knots(n_knots: int, seed: int) -> np.ndarray:
    """Knots U[0,1] smoothed by a 5-tap box filter (edges renormalised)."""
    rng = np.random.default_rng(seed)
    k = rng.uniform(0.0, 1.0, n_knots)
    box = np.ones(5)
    num = np.convolve(k, box, mode="same")
    den = np.convolve(np.ones(n_knots), box, mode="same")
    return num / den


def series_per_step(knots: np.ndarray, steps: int) -> np.ndarray:
    """Per-step inlet values (index = step, 0..steps) by linear interpolation,
    knot k at step k * steps / (n_knots - 1)."""
    dt = steps / (len(knots) - 1)
    s = np.arange(steps + 1) / dt
    return np.interp(s, np.arange(len(knots)), knots)


def build(name: str, dims=None, steps=None) -> CoupledConfig:
    """The named workload; `dims` overrides the box (same physics, same seed)."""
    spec = dict(SPECS[name])
    nx, ny, nz = dims or spec["dims"]
    pack = sphere_pack(nx, ny, nz, spec["seed"], spec["r"][0], spec["r"][1], spec["porosity"],
                       overlap=spec["overlap"])
    ns_op = spec.get("ns_op", "bgk")
    cfg = CoupledConfig(
        nx=nx, ny=ny, nz=nz, solid=pack.solid, region=pack.sphere, ns_op=ns_op,
        tau=spec["tau"], u_in=(spec["u_in"], 0.0, 0.0), tau_ad=spec["tau_ad"], ns_inlet="eq",
        isotherm=spec["isotherm"], iso_a=spec["iso_a"], iso_b=spec["iso_b"], k_ldf=spec["k_ldf"],
    )
    nsteps = steps if steps is not None else spec["steps"]
    if ns_op == "pbb":
        cfg.field = pbb_fraction(pack.sphere, pack.solid, spec["seed"] * 10)
        cfg.solid = np.zeros_like(pack.solid)
    if "series_knots" in spec:
        knots = series_knots(spec["series_knots"], spec["seed"])
        cfg.ad_inlet = "series"
        cfg.inlet_series = series_per_step(knots, nsteps)
        cfg.meta["knots"] = knots
    cfg.meta.update(workload=name, seed=spec["seed"], porosity=pack.porosity, steps=nsteps,
                    dims=(nx, ny, nz))
    return cfg
