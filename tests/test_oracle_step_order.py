"""Pins the oracle's step SEQUENCE (oracle/ref_cpu.cpp, System::step) without calling it for the
expected values.

Each check computes its expectation from first principles (the lattice tables, the conservation
of density and momentum by BGK collision, exact moment recursions) and is run twice: on the
oracle as committed, where it must hold, and on the oracle with the one ordering mistake it is
meant to catch planted by `System.set_step_mutation` (mutation testing), where it must fail.
The orders pinned are the reference's:

* the LDF of step s reads the concentration the AD collision wrote at step s-1
  (main.cpp:328-336, LinearLDF.hpp:38; ADLattice.hpp:49 starts it at 0);
* the AD lattice is advected by the moments of the NS populations after NS.step, that is after
  collision, streaming and the PostStreaming boundaries (NSLattice.hpp:188-209, Macros.cu:44-60,
  ADLattice.hpp:104-111);
* VelocityDirichletNEQ acts before collision, PressureDirichlet after streaming
  (NSLattice.hpp:194-205, VelocityDirichletNEQ.cu:101, PressureDirichlet.cu:127).
"""
import numpy as np
import pytest

from oracle import ref_cpu as R
from tests.lattice import CX19, CY19, CZ19, W19, W7, aos_to_soa, coords, soa_to_aos, x_plane

NO_LAG, VEL_PRE_STEP, VEL_PRE_BC, NEQ_POST_STREAM, PRESSURE_PRE_COLLISION = 1, 2, 3, 4, 5


def feq19(rho, u):
    """D3Q19 BGK equilibrium w rho (1 + 3 e.u + 4.5 (e.u)^2 - 1.5 u^2), one column per direction."""
    eu = CX19 * u[0] + CY19 * u[1] + CZ19 * u[2]
    u2 = u[0] ** 2 + u[1] ** 2 + u[2] ** 2
    return W19 * rho * (1 + 3 * eu + 4.5 * eu ** 2 - 1.5 * u2)


def moments(f_soa):
    rho = f_soa.sum(0)
    return rho, CX19 @ f_soa, CY19 @ f_soa, CZ19 @ f_soa


# ------------------------------------------------------------------ LDF lag
def _ldf_box(mutation):
    """Uniform periodic box at rest, every cell sorbing, concentration well above the clamp."""
    dims = (4, 3, 3)
    n = 36
    s = R.Setup(*dims, periodic=(True, True, True), ns_op=R.BGK, tau=0.8, phi0=0.8, tau_ad=0.7, minC=1e-8,
                coupling=True, adsorb=True, isotherm=R.LINEAR, region=np.ones(n, np.uint8), KHenry=0.5, kA=0.05)
    sysm = R.System(s, 8)
    sysm.set_step_mutation(mutation)
    return sysm, n


def _ldf_recursion(steps, lag, phi0=0.8, K=0.5, k=0.05, m=1e-8):
    """Per-cell state of the uniform box: Phi = sum g, P = the concentration buffer, q."""
    Phi, P, q = phi0, 0.0, 0.0
    for _ in range(steps):
        c = max(P if lag else Phi, m) - m
        qn = q + k * (K * c - q)
        S = -(qn - q)
        q = qn
        P = max(Phi, m)       # the AD collision writes the clamped pre-collision sum
        Phi = Phi + S         # uniform equilibrium populations: the cell sum gains the source
    return Phi, P, q


@pytest.mark.parametrize("mutation", [0, NO_LAG])
def test_ldf_reads_previous_step_concentration(mutation):
    """u = 0 and a uniform box reduce the coupled step to a scalar recursion.  With the one-step
    lag (LDF of step s reads the phi written by AD step s-1, 0 before the first step) the oracle's
    sum g, phi buffer and q follow it to rounding; the no-lag recursion differs by > 1e-3.
    Fails on mutation 1 (LDF reading the current populations' concentration)."""
    sysm, n = _ldf_box(mutation)
    steps = 25
    sysm.run(steps)
    Phi = aos_to_soa(sysm.get(R.G), 7).sum(0)
    P, q = sysm.get(R.PHI), sysm.get(R.Q)
    e_Phi, e_P, e_q = _ldf_recursion(steps, lag=True)
    n_Phi, _, n_q = _ldf_recursion(steps, lag=False)
    assert abs(n_q - e_q) > 1e-3 and abs(n_Phi - e_Phi) > 1e-3  # the check discriminates
    err = max(np.max(np.abs(Phi - e_Phi)), np.max(np.abs(P - e_P)), np.max(np.abs(q - e_q)))
    if mutation == 0:
        assert err < 1e-13
    else:
        assert err > 1e-3


# ---------------------------------------------------------- velocity hand-off
def _porous_drift(mutation, coupling=True):
    """Periodic porous box (uniform porosity s) started at feq(rho0, u0): BGK collision toward
    feq(rho, s u) keeps rho and multiplies j by r = 1 - omega (1 - s) each step, exactly."""
    dims = (80, 4, 4)
    n = 80 * 4 * 4
    s_por, tau, tau_ad, u0 = 0.8, 0.8, 0.7, 0.06
    s = R.Setup(*dims, periodic=(True, True, True), ns_op=R.POROUS, tau=tau, porosity=np.full(n, s_por),
                phi0=0.0, tau_ad=tau_ad, minC=1e-8, coupling=coupling, adsorb=False)
    sysm = R.System(s, 8)
    sysm.set_step_mutation(mutation)
    f = np.repeat(feq19(1.0, (u0, 0.0, 0.0))[:, None], n, axis=1)
    sysm.set(R.F, soa_to_aos(f, 19))
    x, _, _ = coords(*dims)
    # background 1 keeps phi above the clamp; populations move one cell per step, so in 12 steps
    # nothing but the blob's initial 1e-27 tail reaches the periodic seam 40 cells away
    blob = 1.0 + np.exp(-((x - 40.0) ** 2) / (2 * 2.5 ** 2))
    sysm.set(R.G, soa_to_aos(np.outer(W7, blob), 7))
    return sysm, x, (s_por, tau, tau_ad, u0)


def _centroid(phi, x):
    p = phi - 1.0
    return (p * x).sum() / p.sum()


def _centroid_drift(u_seq, om):
    """Exact first-moment recursion of the D3Q7 step (pull, then collide) for a spatially uniform
    velocity u_k at step k: the pull moves the centroid by J_k / M, the flux of the populations it
    pulls, and the collision sets J_{k+1} = (1 - om) J_k + om M u_k (J_0 = 0, equilibrium start)."""
    J, X = 0.0, 0.0
    for u in u_seq:
        X += J
        J = (1 - om) * J + om * u
    return X


@pytest.mark.parametrize("mutation", [0, VEL_PRE_STEP])
def test_advection_velocity_is_post_collision_moment(mutation):
    """The blob's centroid drifts as the recursion driven by u_s = u0 r^(s+1) -- the velocity of the
    populations after step s's collision -- and not by u0 r^s (the populations entering the step).
    Fails on mutation 2 (advection by the moments taken before NS.step)."""
    sysm, x, (s_por, tau, tau_ad, u0) = _porous_drift(mutation)
    c0 = _centroid(aos_to_soa(sysm.get(R.G), 7).sum(0), x)
    steps = 12
    sysm.run(steps)
    drift = _centroid(aos_to_soa(sysm.get(R.G), 7).sum(0), x) - c0
    r = 1 - (1 / tau) * (1 - s_por)
    post = _centroid_drift([u0 * r ** (k + 1) for k in range(steps)], 1 / tau_ad)
    pre = _centroid_drift([u0 * r ** k for k in range(steps)], 1 / tau_ad)
    assert abs(post - pre) > 1e-3  # the check discriminates
    # the NS state itself: uniform, rho kept, j scaled by r per step
    rho, jx, _, _ = moments(aos_to_soa(sysm.get(R.F), 19))
    assert np.allclose(rho, 1.0, atol=1e-14) and np.allclose(jx, u0 * r ** steps, atol=1e-14)
    if mutation == 0:
        assert abs(drift - post) < 1e-10
    else:
        assert abs(drift - post) > 1e-4


def test_advection_off_without_coupling():
    """coupling = false: the blob diffuses in place (centroid fixed), whatever the NS flow does."""
    sysm, x, _ = _porous_drift(0, coupling=False)
    c0 = _centroid(aos_to_soa(sysm.get(R.G), 7).sum(0), x)
    sysm.run(12)
    assert abs(_centroid(aos_to_soa(sysm.get(R.G), 7).sum(0), x) - c0) < 1e-12


def _channel(mutation, ns_inlet=R.NS_EQ, steps=6, seed=21):
    nx, ny, nz = 12, 5, 4
    n = nx * ny * nz
    inlet, outlet = x_plane(nx, ny, nz, 0), x_plane(nx, ny, nz, nx - 1)
    s = R.Setup(nx, ny, nz, periodic=(False, True, True), ns_op=R.BGK, tau=0.7,
                ns_bc=[R.BC(ns_inlet, inlet, (0.05, 0.01, -0.01)), R.BC(R.NS_PRESSURE, outlet, (1.02,))],
                tau_ad=0.7, coupling=True, adsorb=False)
    sysm = R.System(s, 8)
    sysm.set_step_mutation(mutation)
    rng = np.random.default_rng(seed)
    f = sysm.get(R.F)
    sysm.set(R.F, f * (1 + 0.05 * rng.standard_normal(f.size)))
    sysm.run(steps)
    return sysm, (nx, ny, nz)


@pytest.mark.parametrize("mutation", [0, VEL_PRE_BC])
def test_advection_velocity_after_poststreaming_bcs(mutation):
    """The pressure outlet leaves feq(rho_out, u_N) on x = nx-1, whose velocity is exactly u_N, the
    velocity of the x = nx-2 cell.  Taken after the PostStreaming BCs, the AD lattice's velocity on
    the outlet plane therefore equals the NS velocity one cell upstream (to rounding).
    Fails on mutation 3 (moments taken before the PostStreaming BCs)."""
    sysm, (nx, ny, nz) = _channel(mutation)
    x, _, _ = coords(nx, ny, nz)
    rho, jx, jy, jz = moments(aos_to_soa(sysm.get(R.F), 19))
    out, up = np.nonzero(x == nx - 1)[0], np.nonzero(x == nx - 2)[0]
    err = 0.0
    for a, j in ((R.UXA, jx), (R.UYA, jy), (R.UZA, jz)):
        err = max(err, np.max(np.abs(sysm.get(a)[out] - j[up] / rho[up])))
    if mutation == 0:
        assert err < 1e-14
    else:
        assert err > 1e-4


# ------------------------------------------------------------------ BC phases
def _half_space_moments(rho0, u, sel):
    """Density and momentum of a cell whose populations are feq(rho0, u) on the directions `sel` and
    rest (w rho0) on the others: from the equilibrium's moments, no collision involved."""
    fe = feq19(rho0, u)
    f = np.where(sel, fe, W19 * rho0)
    return f.sum(), np.array([CX19 @ f, CY19 @ f, CZ19 @ f])


@pytest.mark.parametrize("mutation", [0, NEQ_POST_STREAM])
def test_neq_inlet_acts_before_collision(mutation):
    """From rest, step 0: the NEQ inlet turns x = 0 into feq(rho0, u_BC) (its neighbour is at rest with
    rho0) BEFORE the pull, so x = 1 pulls feq(rho0, u_BC) along the directions with c_x = +1 and
    rest along the others; BGK keeps that cell's density and momentum, so after the step they are
    the half-space moments of feq(rho0, u_BC) -- nonzero.  After streaming instead, x = 1 would stay
    at rest.  Fails on mutation 4 (NEQ applied on f_new after streaming)."""
    nx, ny, nz = 6, 3, 3
    n = nx * ny * nz
    u_bc = (0.04, 0.01, 0.0)
    inlet = x_plane(nx, ny, nz, 0)
    s = R.Setup(nx, ny, nz, periodic=(False, True, True), ns_op=R.BGK, tau=0.8,
                ns_bc=[R.BC(R.NS_NEQ, inlet, u_bc)], coupling=False, adsorb=False)
    sysm = R.System(s, 8)
    sysm.set_step_mutation(mutation)
    sysm.run(1)
    x, _, _ = coords(nx, ny, nz)
    rho, jx, jy, jz = moments(aos_to_soa(sysm.get(R.F), 19))
    c1 = np.nonzero(x == 1)[0]
    e_rho, e_j = _half_space_moments(1.0, u_bc, CX19 == 1)
    assert abs(e_j[0]) > 1e-3  # the check discriminates from a cell left at rest
    err = max(np.max(np.abs(rho[c1] - e_rho)), np.max(np.abs(jx[c1] - e_j[0])), np.max(np.abs(jy[c1] - e_j[1])),
              np.max(np.abs(jz[c1] - e_j[2])))
    # the inlet cell itself: it pulls from x = 1 (rest) along c_x = -1 and keeps feq(rho0, u_BC) elsewhere
    c0 = np.nonzero(x == 0)[0]
    i_rho, i_j = _half_space_moments(1.0, u_bc, CX19 != -1)
    err = max(err, np.max(np.abs(rho[c0] - i_rho)), np.max(np.abs(jx[c0] - i_j[0])))
    if mutation == 0:
        assert err < 1e-14
    else:
        assert err > 1e-4


@pytest.mark.parametrize("mutation", [0, PRESSURE_PRE_COLLISION])
def test_pressure_outlet_reads_post_collision_neighbour(mutation):
    """After every step the outlet plane holds feq(rho_out, u_N) with u_N the velocity of the x = nx-2
    cell's post-collision, streamed populations: density exactly rho_out and every population equal
    to the equilibrium of the neighbour's current velocity.  Applied before collision it would be
    pulled over and relaxed again.  Fails on mutation 5 (pressure outlet on f before collision)."""
    sysm, (nx, ny, nz) = _channel(mutation, steps=5)
    x, _, _ = coords(nx, ny, nz)
    f = aos_to_soa(sysm.get(R.F), 19)
    rho, jx, jy, jz = moments(f)
    out, up = np.nonzero(x == nx - 1)[0], np.nonzero(x == nx - 2)[0]
    err = np.max(np.abs(rho[out] - 1.02))
    for k in range(len(out)):
        u = (jx[up[k]] / rho[up[k]], jy[up[k]] / rho[up[k]], jz[up[k]] / rho[up[k]])
        err = max(err, np.max(np.abs(f[:, out[k]] - feq19(1.02, u))))
    if mutation == 0:
        assert err < 1e-14
    else:
        assert err > 1e-4
