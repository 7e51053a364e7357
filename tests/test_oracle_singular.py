"""Pins the oracle's quadrature and closed-form singular integrals (CPU)."""
import csv
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import quadrature as Q
from oracle import singular as S


@pytest.mark.parametrize('order', [1, 3, 8, 16, 20, 32])
def test_gauss_legendre_matches_numpy(order):
    p, w = Q.gauss_legendre(order, 0.0, 1.0)
    ref_p, ref_w = np.polynomial.legendre.leggauss(order)
    np.testing.assert_allclose(p, (ref_p + 1) / 2, rtol=0, atol=2e-15)
    np.testing.assert_allclose(w, ref_w / 2, rtol=0, atol=2e-15)


@pytest.mark.parametrize('parts', [(1, 1), (3, 2)])
def test_reduced_rule_exact_for_polynomial_kernels(parts):
    """GenerateReducedRuleTest.m:19-35: the reduced rule against the tensor rule on both cells."""
    def kernel(ax, ay):
        return (ax - 0.3) ** 3 * (ay + 0.2) ** 2 + ax * ay

    rule = Q.reduced_rule(4, parts)
    nu, eta = np.meshgrid(rule['nodesX'], rule['nodesY'], indexing='ij')
    got = rule['weights'] @ kernel(nu.ravel(order='F'), eta.ravel(order='F'))
    p, w = Q.gauss_legendre(4, 0, 1)
    xi, xo, e1, eo = np.meshgrid(p, p, p, p, indexing='ij')
    pw = w[:, None, None, None] * w[None, :, None, None] * w[None, None, :, None] * w[None, None, None, :]
    v = pw * kernel(xi - xo, e1 - eo)
    expected = np.array([v.sum(), (xi * v).sum(), (xo * v).sum(), (xi * xo * v).sum(), (e1 * v).sum(),
                         (eo * v).sum(), (e1 * eo * v).sum()])
    np.testing.assert_allclose(got, expected, rtol=1e-13)


def test_rule_nodes_avoid_part_ends():
    rule = Q.reduced_rule(4, (3, 2))
    ends = np.arange(-3, 4) / 3
    assert rule['nodesX'].size == 24 and rule['nodesY'].size == 16 and rule['weights'].shape == (7, 384)
    assert np.min(np.abs(rule['nodesX'][:, None] - ends[None, :])) > 0


def test_remainder_matches_direct_difference():
    r, k = np.array([0.05, 0.3, 1, 2.5]), 2 * np.pi
    expected = np.exp(-1j * k * r) / r - (1 / r - k ** 2 * r / 2 + k ** 4 * r ** 3 / 24)
    np.testing.assert_allclose(S.evaluate_remainder(r, k), expected, rtol=1e-13)


def test_remainder_keeps_accuracy_at_small_distances():
    """EvaluateRemainderTest.m:13-20: the Taylor series."""
    r, k = np.array([1e-8, 1e-5, 1e-3]), 2 * np.pi
    series = -1j * k + 1j * k ** 3 * r ** 2 / 6 - 1j * k ** 5 * r ** 4 / 120 - k ** 6 * r ** 5 / 720
    np.testing.assert_allclose(S.evaluate_remainder(r, k), series, rtol=1e-14)


def test_power_moments_match_exact_values():
    """GeneratePowerMomentsTest.m:9-30 against tests/data/exactMoments.csv (30 digits, mpmath)."""
    with open(os.path.join(GOLDEN, 'exact_moments.csv')) as f:
        rows = list(csv.reader(f))[1:]
    offsets = [(0, 0), (1, 0), (0, 1), (1, 1)]
    assert len(rows) == 60
    for row in rows:
        a, b, power = float(row[0]), float(row[1]), int(row[2])
        column = offsets.index((int(row[3]), int(row[4])))
        expected = np.array([float(v) for v in row[5:12]])
        tol = 1e-12 if max(a, b) / min(a, b) > 4 else 1e-13
        np.testing.assert_allclose(S.power_moments(a, b, power)[:, column], expected, rtol=tol,
                                   err_msg='cell %g x %g power %d offset %s' % (a, b, power, row[3:5]))


def _self_cell(a, b):
    d = np.hypot(a, b)
    return 2 * (a ** 3 - a ** 2 * d + b ** 3 - b ** 2 * d + 3 * a * b ** 2 * np.log((a + d) / b)
                + 3 * a ** 2 * b * np.log((b + d) / a)) / (3 * a ** 2 * b ** 2)


def _self_equal(a, b):
    d = np.hypot(a, b)
    return (12 * a ** 5 - 12 * a ** 4 * d + a ** 2 * b ** 2 * (30 * b - 29 * d) + 2 * b ** 4 * (b - d)
            + 15 * a ** 3 * b * (4 * b * np.log((a + d) / b) + 3 * a * np.log((b + d) / a))) / (90 * a ** 4 * b ** 2)


def _self_different(a, b):
    d = np.hypot(a, b)
    return (18 * a ** 5 - 2 * b ** 5 - 18 * a ** 4 * d - a ** 2 * b ** 2 * d + 2 * b ** 4 * d
            + 30 * a ** 3 * b ** 2 * np.log((a + d) / b) + 45 * a ** 4 * b * np.log((b + d) / a)) / (90 * a ** 4 * b ** 2)


@pytest.mark.parametrize('a,b', [(0.1, 0.15), (0.3, 0.05)])
def test_self_terms_of_the_thesis(a, b):
    """GeneratePowerMomentsTest.m:66-82: the closed self terms of 1/R."""
    m = S.power_moments(a, b, -1)
    np.testing.assert_allclose(m[0, 0], _self_cell(a, b), rtol=1e-14)
    np.testing.assert_allclose(m[3, 0], _self_equal(a, b), rtol=1e-14)
    np.testing.assert_allclose(m[1, 0] - m[3, 0], _self_different(a, b), rtol=1e-14)
    np.testing.assert_allclose(m[6, 0], _self_equal(b, a), rtol=1e-14)
    np.testing.assert_allclose(m[4, 0] - m[6, 0], _self_different(b, a), rtol=1e-14)


@pytest.mark.parametrize('power', [-1, 1, 3])
def test_power_moments_transpose_and_homogeneity(power):
    m = S.power_moments(0.1, 0.25, power)
    t = S.power_moments(0.25, 0.1, power)
    np.testing.assert_array_equal(m[4:7], t[1:4][:, [0, 2, 1, 3]])
    np.testing.assert_allclose(m[0], t[0][[0, 2, 1, 3]], rtol=1e-14)
    np.testing.assert_allclose(S.power_moments(0.3, 0.75, power), 3.0 ** power * m, rtol=3e-14)


@pytest.mark.parametrize('power,m,n', [(p, m, n) for p in (-1, 1, 3) for m in range(4) for n in range(2)])
def test_monomial_closed_forms_against_quadrature(power, m, n):
    """Every closed form against 2D Gauss-Legendre on a rectangle away from the singular corner."""
    a0, a1, b0, b1 = 1.0, 1.7, 0.5, 1.3
    p, w = Q.gauss_legendre(30, 0, 1)
    u = a0 + (a1 - a0) * p
    v = b0 + (b1 - b0) * p
    uu, vv = np.meshgrid(u, v, indexing='ij')
    f = uu ** m * vv ** n * np.hypot(uu, vv) ** power
    expected = (a1 - a0) * (b1 - b0) * w @ f @ w
    G = lambda x, y: S.integrate_monomial(power, m, n, np.array([x]), np.array([y]))[0]
    got = G(a1, b1) - G(a0, b1) - G(a1, b0) + G(a0, b0)
    np.testing.assert_allclose(got, expected, rtol=1e-12)
