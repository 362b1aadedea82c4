"""CPU: pins of the oracle's transport fields (dynamical_fields.py:24-171):
the vortex velocities against the tracers they carry exactly, the divergent
flow against the derivative of its angle map, the tracer shapes against
their geometric definitions, and the random field against an independent
spherical-harmonic synthesis."""
import numpy as np
import pytest

from helpers import random_unit
from oracle import fields as ofl, geometry as g, sht as osh, meshes as om


def _tangential_derivative(f, t, X, U, h=1e-6):
    """d/ds f(t, normalise(X + s U)) at s = 0 (central difference)."""
    n = lambda Y: Y / np.linalg.norm(Y, axis=0)[None, :]
    return (f(t, n(X + h * U)) - f(t, n(X - h * U))) / (2 * h)


@pytest.mark.parametrize("vel,tracer", [("u_static_vortex", "static_vort_IC"),
                                        ("equator_vortex", "rotating_vort_IC")])
def test_vortex_tracers_solve_the_transport_equation(vel, tracer):
    """static_vort_IC(t, .) is the tracer the static vortex carries (and
    rotating_vort_IC the one of the equator vortex, in the Rot_x(pi/2)
    frame): d/dt c + u . grad c = 0, checked by central differences.  Pins
    omega_r against the velocity, the frame rotations and their signs."""
    X = random_unit(4000, 11).T
    V, C = getattr(ofl, vel), getattr(ofl, tracer)
    for t in (0.0, 0.3, 0.9):
        U = V(t, 0.0, X)
        assert np.abs((U * X).sum(axis=0)).max() < 1e-13  # tangent
        dt = 1e-6
        ct = (C(t + dt, X) - C(t - dt, X)) / (2 * dt)
        adv = _tangential_derivative(C, t, X, U)
        scale = np.abs(ct).max()
        assert scale > 1e-2
        assert np.abs(ct + adv).max() < 2e-7 * max(1.0, scale)


def test_rotating_vortex_is_the_rotated_equator_vortex():
    """u_rotating_vortex(t) - u_sbr = Rz(tau) E(Rz(tau)^T X), tau = 2 pi t,
    with E the equator vortex (Rot_x(pi/2) frame): the static vortex carried
    round by the frame rotation."""
    X = random_unit(3000, 12).T
    for t in (0.0, 0.17, 0.6):
        tau = 2 * np.pi * t
        Rz = np.array([[np.cos(tau), -np.sin(tau), 0], [np.sin(tau), np.cos(tau), 0], [0, 0, 1]])
        lhs = ofl.u_rotating_vortex(t, 0.0, X) - ofl.u_sbr(t, 0.0, X)
        rhs = Rz @ ofl.equator_vortex(t, 0.0, Rz.T @ X)
        assert np.abs(lhs - rhs).max() < 1e-12


def test_divergent_flow_is_the_derivative_of_its_angle_rates():
    """u_div gives (dphi/dt, dtheta/dt) = (u, v) in Cartesian components:
    equal to d/dh sphere2cart(phi + h u, theta + h v) at h = 0."""
    X = random_unit(3000, 13).T
    for t in (0.0, 0.4):
        U = ofl.u_div(t, 0.0, X)
        phi, theta = ofl.cart2sphere(X)
        u = -(np.sin(phi / 2) ** 2) * np.sin(2 * theta) * np.sin(theta) * np.cos(np.pi * t)
        v = -0.5 * np.sin(phi) * np.sin(theta) ** 3 * np.cos(np.pi * t)
        h = 1e-6
        fd = (np.array(g.sphere2cart(phi + h * u, theta + h * v)) -
              np.array(g.sphere2cart(phi - h * u, theta - h * v))) / (2 * h)
        assert np.abs(U - fd).max() < 1e-8
        assert np.abs((U * X).sum(axis=0)).max() < 1e-14


def test_cosine_bell_and_slotted_cylinder_geometry():
    """Bell distance = great-circle distance arccos(X . Xc); profile
    h0/2 (1 + cos(pi r / R)) scaled by 0.9 over 0.1: 1.0 at the centres,
    0.55 at half radius, 0.1 outside; the cylinders are 1.0 inside with a
    slot of width R/3 cut from the far side."""
    pc = [np.pi + np.pi / 6, np.pi - np.pi / 6]
    for k in (0, 1):
        Xc = np.array(g.sphere2cart(pc[k], np.pi / 2))
        for r, want in [(0.0, 1.0), (0.25, 0.55), (0.6, 0.1)]:
            # a point at distance r from the centre, along the equator
            phi, th = pc[k] + (r if k == 0 else -r), np.pi / 2
            X = np.array(g.sphere2cart(phi, th))
            assert abs(np.arccos(np.clip(X @ Xc, -1, 1)) - r) < 1e-12
            assert abs(float(ofl.cosine_bell(np.array([phi]), np.array([th]))[0]) - want) < 1e-12
    rng = np.random.default_rng(5)
    phi = rng.uniform(0, 2 * np.pi, 20000)
    th = np.arccos(rng.uniform(-1, 1, 20000))
    v = ofl.IC_SC(phi, th)
    assert set(np.unique(v).round(12)) <= {0.1, 1.0}
    X = np.array(g.sphere2cart(phi, th))
    R, e = 0.5, 1e-9
    for k in (0, 1):
        Xc = np.array(g.sphere2cart(pc[k], np.pi / 2))
        d = np.arccos(np.clip((X * Xc[:, None]).sum(axis=0), -1, 1))
        inside = d < R - 0.01
        strip = np.abs(phi - pc[k]) < R / 6 - e
        # the strip |phi - phi_c| < R/6 is cut out up to 5R/12 beyond the
        # equator: towards the south pole for the first cylinder (kept for
        # theta < pi/2 - 5R/12), towards the north for the second
        keep = th < np.pi / 2 - 5 * R / 12 - e if k == 0 else th > np.pi / 2 + 5 * R / 12 + e
        cut = th > np.pi / 2 - 5 * R / 12 + e if k == 0 else th < np.pi / 2 + 5 * R / 12 - e
        assert (v[inside & ~(np.abs(phi - pc[k]) < R / 6 + e)] == 1.0).all()
        assert (v[inside & strip & keep] == 1.0).all() and (inside & strip & keep).sum() > 5
        assert (v[inside & strip & cut] == 0.1).all() and (inside & strip & cut).sum() > 20
    far = np.ones_like(phi, bool)
    for k in (0, 1):
        Xc = np.array(g.sphere2cart(pc[k], np.pi / 2))
        far &= np.arccos(np.clip((X * Xc[:, None]).sum(axis=0), -1, 1)) > 0.51
    assert (v[far] == 0.1).all()


def test_random_initial_condition_against_synthesis():
    """The reference's seeded random field (33^2 uniform coefficients over
    l <= 32, m = -l .. l-1, Re Y_l^m from scipy) equals the oracle's MWSS
    synthesis (recurrence-based Legendre functions) of the same field's
    coefficients a_l0 = c_l0, a_lm = (c_lm + (-1)^m c_l,-m) / 2."""
    from cmm_s2.core.dynamical_fields import random_ic_table
    L = 40
    tab = random_ic_table()
    a = np.zeros((L, L), complex)
    a[:33, :33] = (tab[:, :, 0] + 1j * tab[:, :, 1]).T
    f = osh.inverse_lm(a, L)
    th, ph = om.mwss_thetas(L), om.mwss_phis(L)
    T, P = np.meshgrid(th, ph, indexing="ij")
    X = np.array(g.sphere2cart(P.ravel(), T.ravel()))
    ref = ofl.random_initial_condition(X).reshape(L + 1, 2 * L)
    assert np.abs(f - ref).max() < 1e-11 * np.abs(ref).max()
