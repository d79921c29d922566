"""Pin the oracle's Becke partition and the grid orientation by what the
mathematics fixes, independently of the oracle's own formulas
(DfGenerateGrid.cpp:1811-1876 calcMultiCenterWeight_Becke, :2043-2086 Ps_uij,
:1958-1968 getCovalentRadiiForBecke, :1715-1734 getSphericalGrids,
:1739-1809 getOMatrix).

For two atoms A, B the Becke weight of A is a function of the confocal
coordinate mu = (r_A - r_B) / R_AB alone, and equals 1/2 exactly on the cell
boundary.  Becke (J. Chem. Phys. 88, 2547 (1988), appendix) places that
boundary where r_A / r_B = chi = R_A / R_B, i.e. at mu = (chi - 1) / (chi + 1),
as long as the size adjustment |a_AB| stays below 1/2; when it is clamped to
a = -1/2 the boundary solves mu - (1 - mu^2) / 2 = 0, mu = sqrt(2) - 1.  The
radii are the Bragg-Slater values with H taken as 0.35 (Becke's choice).  At
a homonuclear midpoint three iterations of f(x) = x (3 - x^2) / 2 give the
slope dw/dmu = -(3/2)^3 / 2 = -27/16 (two would give -9/8), and with a size
adjustment the slope picks up d nu / d mu = 1 - 2 a mu.

Lebedev rules are invariant under the 48 signed permutations of the axes, so
the grid built with O Q (Q a signed permutation) is the grid built with O,
point for point up to order: the sign and order convention of the O
matrix's eigenvectors (LAPACK's dsyev is implementation-defined) does not
change any quadrature result.
"""
import itertools

import numpy as np
import pytest
from scipy.optimize import brentq

from oracle import oracle

LEB = [6, 14, 26, 38, 50, 74, 86, 110, 146, 170, 194, 230, 266, 302, 350, 434, 590, 770, 974, 1202, 1454, 1730]
# Bragg-Slater radii (Slater, J. Chem. Phys. 41, 3199 (1964)) with H = 0.35 as Becke uses it
RADIUS = {1: 0.35, 6: 0.70, 7: 0.65, 8: 0.60, 16: 1.00}


def _pair(za, zb, dist=2.3):
    Z = np.array([za, zb], dtype=np.int32)
    xyz = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, dist]])
    return Z, xyz


def _point_at_mu(mu, dist, x_off):
    """A point with r_A - r_B = mu dist at distance x_off from the axis."""
    f = lambda z: np.hypot(x_off, z) - np.hypot(x_off, z - dist) - mu * dist
    z = brentq(f, -50.0, 50.0, xtol=1e-15, rtol=1e-15)
    return np.array([x_off, 0.0, z])


def _w(Z, xyz, a, p, size_adjust=True):
    return oracle.becke_weight(Z, xyz, a, p, size_adjust)


def _mu_of(p, xyz):
    d = np.linalg.norm(xyz[1] - xyz[0])
    return (np.linalg.norm(p - xyz[0]) - np.linalg.norm(p - xyz[1])) / d


def _slope(Z, xyz, mu0, h=1e-5):
    d = np.linalg.norm(xyz[1] - xyz[0])
    # on the axis between the nuclei, mu = (2 z - d) / d
    zp, zm = (mu0 + h + 1.0) * d / 2.0, (mu0 - h + 1.0) * d / 2.0
    wp = _w(Z, xyz, 0, np.array([0.0, 0.0, zp]))
    wm = _w(Z, xyz, 0, np.array([0.0, 0.0, zm]))
    return (wp - wm) / (2.0 * h)


@pytest.mark.parametrize("za,zb", [(6, 1), (7, 1), (8, 1), (6, 8), (7, 8)], ids=["CH", "NH", "OH", "CO", "NO"])
def test_becke_size_adjusted_boundary(za, zb):
    chi = RADIUS[za] / RADIUS[zb]
    mu_b = (chi - 1.0) / (chi + 1.0)
    Z, xyz = _pair(za, zb)
    for x_off in (0.0, 0.4, 1.7):
        p = _point_at_mu(mu_b, 2.3, x_off)
        assert abs(_mu_of(p, xyz) - mu_b) < 1e-13
        assert abs(_w(Z, xyz, 0, p) - 0.5) < 1e-12, (za, zb, x_off)
        assert abs(_w(Z, xyz, 1, p) - 0.5) < 1e-12, (za, zb, x_off)
    # the unadjusted partition puts the boundary at the midplane instead
    assert abs(_w(Z, xyz, 0, np.array([0.0, 0.0, 1.15]), size_adjust=False) - 0.5) < 1e-14
    # slope at the boundary: -27/16 (1 - 2 a mu_b) with a = u / (u^2 - 1), u = mu_b
    a = mu_b / (mu_b * mu_b - 1.0)
    assert abs(a) < 0.5
    np.testing.assert_allclose(_slope(Z, xyz, mu_b), -27.0 / 16.0 * (1.0 - 2.0 * a * mu_b), rtol=1e-7)


def test_becke_clamped_size_adjustment():
    # S-H: chi = 1/0.35, u = 0.4815, u / (u^2 - 1) = -0.627 is clamped to -1/2,
    # so the boundary solves mu - (1 - mu^2)/2 = 0 and not mu = u
    chi = RADIUS[16] / RADIUS[1]
    u = (chi - 1.0) / (chi + 1.0)
    assert u / (u * u - 1.0) < -0.5
    mu_b = np.sqrt(2.0) - 1.0
    Z, xyz = _pair(16, 1)
    for x_off in (0.0, 0.9):
        p = _point_at_mu(mu_b, 2.3, x_off)
        assert abs(_w(Z, xyz, 0, p) - 0.5) < 1e-12
    assert abs(_w(Z, xyz, 0, _point_at_mu(u, 2.3, 0.0)) - 0.5) > 1e-3
    # d nu / d mu = 1 - 2 (-1/2) mu_b = sqrt(2)
    np.testing.assert_allclose(_slope(Z, xyz, mu_b), -27.0 / 16.0 * np.sqrt(2.0), rtol=1e-7)


@pytest.mark.parametrize("z", [1, 8])
def test_becke_homonuclear_midpoint_slope(z):
    Z, xyz = _pair(z, z, 1.9)
    assert abs(_w(Z, xyz, 0, np.array([0.3, -0.2, 0.95])) - 0.5) < 1e-14
    np.testing.assert_allclose(_slope(Z, xyz, 0.0), -27.0 / 16.0, rtol=1e-8)
    # away from the midpoint the weight still depends on mu alone
    for mu in (-0.6, 0.25, 0.8):
        w = [_w(Z, xyz, 0, _point_at_mu(mu, 1.9, x)) for x in (0.0, 0.5, 2.0)]
        assert max(w) - min(w) < 1e-13
        # w_A(mu) + w_A(-mu) = 1 for equal atoms
        assert abs(w[0] + _w(Z, xyz, 0, _point_at_mu(-mu, 1.9, 0.0)) - 1.0) < 1e-14


def test_becke_weight_depends_on_mu_only_heteronuclear():
    Z, xyz = _pair(6, 1)
    for mu in (-0.5, 0.1, 0.6):
        w = [_w(Z, xyz, 0, _point_at_mu(mu, 2.3, x)) for x in (0.0, 0.7, 3.0)]
        assert max(w) - min(w) < 1e-13


def _signed_permutations():
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1.0, -1.0), repeat=3):
            Q = np.zeros((3, 3))
            for r, c in enumerate(perm):
                Q[r, c] = signs[r]
            yield Q


def _same_point_set(a, wa, b, wb, tol=1e-12):
    ka = np.lexsort(np.round(a / tol).T[::-1])
    kb = np.lexsort(np.round(b / tol).T[::-1])
    return (np.abs(a[ka] - b[kb]).max() < 10 * tol) and (np.abs(wa[ka] - wb[kb]).max() < 10 * tol)


@pytest.mark.parametrize("order", LEB)
def test_lebedev_rules_are_octahedrally_invariant(order):
    x, w = oracle.lebedev(order)
    for Q in _signed_permutations():
        assert _same_point_set(x, w, x @ Q.T, w), (order, Q)


def test_grid_independent_of_o_matrix_sign_and_order():
    Z = np.array([8, 1, 1, 6], dtype=np.int32)
    xyz = np.array([[0.1, 0.0, -0.2], [1.5, 0.3, 0.9], [-1.4, 0.2, 1.0], [0.3, 2.4, -0.6]])
    O = oracle.o_matrix(Z, xyz)
    g0, a0 = oracle.generate_grid(Z, xyz, 20, 50, O=O)
    for Q in (np.diag([-1.0, 1.0, 1.0]), np.diag([1.0, -1.0, -1.0]), np.eye(3)[:, [2, 0, 1]],
              np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, -1.0]])):
        g1, a1 = oracle.generate_grid(Z, xyz, 20, 50, O=O @ Q)
        assert len(g1) == len(g0)
        for atom in range(len(Z)):
            s0, s1 = a0 == atom, a1 == atom
            assert _same_point_set(g0[s0, :3], g0[s0, 3], g1[s1, :3], g1[s1, 3], tol=1e-11)
