"""Shared test helpers: lattice tables restated from the reference descriptors
(D3Q19.hpp:5-50, D3Q7.hpp:5-28) for the closed-form checks, and conversion of a
CoupledConfig into the oracle's set-up."""
import numpy as np

CX19 = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
CY19 = np.array([0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1])
CZ19 = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1])
W19 = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
OPP19 = np.array([0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15])
CX7 = np.array([0, 1, -1, 0, 0, 0, 0])
CY7 = np.array([0, 0, 0, 1, -1, 0, 0])
CZ7 = np.array([0, 0, 0, 0, 0, 1, -1])
W7 = np.array([0.25] + [0.125] * 6)


def coords(nx, ny, nz):
    c = np.arange(nx * ny * nz)
    return c % nx, (c // nx) % ny, c // (nx * ny)


def x_plane(nx, ny, nz, x):
    cx, _, _ = coords(nx, ny, nz)
    return (cx == x).astype(np.uint8)


def aos_to_soa(a, q):
    return np.ascontiguousarray(np.asarray(a).reshape(-1, q).T)


def soa_to_aos(a, q):
    return np.ascontiguousarray(np.asarray(a).reshape(q, -1).T).ravel()


def rel_err(a, b):
    """max |a - b| / max |b|  (relative L-infinity error, normalised by the field's scale)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = np.max(np.abs(b))
    if scale == 0:
        return float(np.max(np.abs(a)))
    return float(np.max(np.abs(a - b)) / scale)


def oracle_setup(cfg, R):
    """The oracle set-up (oracle.ref_cpu.Setup) equivalent to a CoupledConfig."""
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    inlet, outlet = x_plane(nx, ny, nz, 0), x_plane(nx, ny, nz, nx - 1)
    ns_bc = []
    if cfg.ns_inlet == "neq":
        ns_bc.append(R.BC(R.NS_NEQ, inlet, tuple(cfg.u_in)))
    elif cfg.ns_inlet == "eq":
        ns_bc.append(R.BC(R.NS_EQ, inlet, tuple(cfg.u_in)))
    if cfg.ns_outlet:
        ns_bc.append(R.BC(R.NS_PRESSURE, outlet, (cfg.rho_out,)))
    ad_bc = []
    if cfg.ad_inlet == "dirichlet":
        ad_bc.append(R.BC(R.AD_DIRICHLET, inlet, (cfg.phi_in,)))
    elif cfg.ad_inlet == "series":
        ad_bc.append(R.BC(R.AD_SERIES, inlet, knots=np.asarray(cfg.inlet_series, np.float64), dt_knot=1.0))
    if cfg.ad_outlet:
        ad_bc.append(R.BC(R.AD_NEUMANN, outlet))
    op = {"bgk": R.BGK, "porous": R.POROUS, "pbb": R.PBB}[cfg.ns_op]
    iso = {"linear": R.LINEAR, "langmuir": R.LANGMUIR, None: 0, "none": 0}[cfg.isotherm]
    s = R.Setup(nx, ny, nz, periodic=(False, cfg.periodic_y, cfg.periodic_z), ns_op=op, rho0=cfg.rho0, tau=cfg.tau,
                obstacle=cfg.solid if op == R.BGK else None, porosity=cfg.field, phi0=cfg.phi0, tau_ad=cfg.tau_ad,
                minC=cfg.min_conc, tau_field=cfg.tau_field, ns_bc=ns_bc, ad_bc=ad_bc, coupling=cfg.coupling,
                adsorb=iso != 0, isotherm=iso, region=cfg.region)
    if iso == R.LINEAR:
        s.KHenry, s.kA = cfg.iso_a, cfg.k_ldf
    elif iso == R.LANGMUIR:
        s.qMax, s.bLangmuir, s.kLDF = cfg.iso_a, cfg.iso_b, cfg.k_ldf
    return s
