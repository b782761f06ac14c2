"""Geometry inputs: seeded sphere packs, x-plane masks, ASCII VTI I/O.

Thin wrappers over the host functions of libfmx_hip.so (fmx_geom_sphere_pack,
fmx_build_x_plane_mask, fmx_vti_write_mask, fmx_vti_read_mask).  Masks are
uint8 numpy arrays of length nx*ny*nz in the reference cell order
x + nx*(y + ny*z).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import abi


@dataclass
class SpherePack:
    nx: int
    ny: int
    nz: int
    seed: int
    solid: np.ndarray     # walls + spheres
    sphere: np.ndarray    # sphere voxels only (the sorption region)
    porosity: float       # fluid cells / all cells

    @property
    def n(self) -> int:
        return self.nx * self.ny * self.nz


def sphere_pack(nx, ny, nz, seed, r_min, r_max, porosity, overlap=True, clearance=None,
                max_tries=200000) -> SpherePack:
    """Random sequential addition of spheres between solid y/z walls.

    Centres are drawn uniformly with x in [clearance, nx-1-clearance] (default
    clearance r_max + 2, keeping the inlet/outlet planes all-fluid) and y, z
    inside the walls; spheres are added until the fluid fraction reaches
    `porosity` or `max_tries` consecutive placements were rejected (only
    rejected when overlap=False).
    """
    if clearance is None:
        clearance = int(np.ceil(r_max)) + 2
    n = nx * ny * nz
    solid = np.zeros(n, np.uint8)
    sph = np.zeros(n, np.uint8)
    por = C.c_double(0.0)
    abi.call("fmx_geom_sphere_pack", nx, ny, nz, C.c_uint64(seed), float(r_min), float(r_max), float(porosity),
             int(bool(overlap)), int(clearance), int(max_tries), solid.ctypes.data, sph.ctypes.data, C.byref(por))
    return SpherePack(nx, ny, nz, seed, solid, sph, por.value)


def x_plane_mask(nx, ny, nz, x) -> np.ndarray:
    """buildXPlaneMask (helpers.hpp:7-14)."""
    m = np.zeros(nx * ny * nz, np.uint8)
    abi.call("fmx_build_x_plane_mask", nx, ny, nz, x, m.ctypes.data)
    return m


def write_vti(path: str, nx, ny, nz, mask: np.ndarray, name: str = "solid") -> None:
    m = np.ascontiguousarray(mask, np.uint8)
    abi.call("fmx_vti_write_mask", path.encode(), nx, ny, nz, m.ctypes.data, name.encode())


def read_vti(path: str, threshold: float = 0.5):
    """readVTI (VTIReader.hpp:35-160): returns ((nx, ny, nz), mask)."""
    dims = (C.c_int32 * 3)()
    abi.call("fmx_vti_read_mask", path.encode(), float(threshold), dims, None, 0)
    n = dims[0] * dims[1] * dims[2]
    m = np.zeros(n, np.uint8)
    abi.call("fmx_vti_read_mask", path.encode(), float(threshold), dims, m.ctypes.data, n)
    return (dims[0], dims[1], dims[2]), m
