"""Pins of the oracle's generator accumulation and radar cross section against independent definitions (CPU).

* The generators magneticXX, magneticYY and electric are compared entry by entry with a brute-force integral of their
  definition (docs/nearfield/src/sections/02_setting.tex:42-58, eq. zpq and the generators A^xx, A^yy, Phi): the two
  rooftops are written out as piecewise-linear functions on their two cells and the 4D integral of
  Psi_0 . Psi_d exp(-ikR)/R is taken by a tensor Gauss-Legendre rule on every pair of cells. No moment, reduced rule or
  accumulation stencil is involved, so a wrong row, sign or shift in accumulate_moments (oracle/assembly.py) fails.
* accumulate_moments against the generator formula of the near-field note (Proposition "Generators from moments",
  02_setting.tex:124-137), written here offset by offset, on random moments added in chunks
  (tests/+unit/+assembler/+greensMatrix/+utils/AccumulateMomentsTest.m:5-60).
* The stored generators of reference.mat (an older, lower-order integration) as a coarse check.
* radar_cross_section against the closed form of tests/+unit/+operator/ComputeRadarCrossSectionTest.m:5-23.
"""
import numpy as np
import pytest

from oracle import assembly as A
from oracle import excitation as X

Z0 = 376.730313668   # src/+rooftopMoM/+assembler/Constants.m:11


# --- brute force of the definition ---------------------------------------------------------------------------------

def _gauss01(n):
    t, w = np.polynomial.legendre.leggauss(n)
    return (t + 1) / 2, w / 2


def _cell_pair(cell1, prof1, cell2, prof2, sizes, k, n):
    """Integral over cell1 x cell2 (face indices (i, j), cells [i a, (i+1) a] x [j b, (j+1) b]) of
    prof1(u1, v1) prof2(u2, v2) exp(-ikR)/R, with (u, v) the local coordinates in [0, 1]^2 of each cell: a tensor
    Gauss-Legendre rule of n points per dimension. The cells must not touch (the kernel is then smooth)."""
    a, b = sizes
    t, w = _gauss01(n)
    u1 = t[:, None, None, None]
    v1 = t[None, :, None, None]
    u2 = t[None, None, :, None]
    v2 = t[None, None, None, :]
    W = w[:, None, None, None] * w[None, :, None, None] * w[None, None, :, None] * w[None, None, None, :]
    R = np.hypot(a * ((cell1[0] + u1) - (cell2[0] + u2)), b * ((cell1[1] + v1) - (cell2[1] + v2)))
    f = prof1(u1, v1) * prof2(u2, v2) * np.exp(-1j * k * R) / R
    return (a * b) ** 2 * np.sum(W * f)


def _rooftop_pieces(direction, origin):
    """The two cells of the rooftop whose first face is `origin`, with its profile on each: rising (local coordinate)
    on the first face, falling (1 - local coordinate) on the second (02_setting.tex, Definition rooftop)."""
    if direction == 'x':
        return [((origin[0], origin[1]), lambda u, v: u + 0 * v),
                ((origin[0] + 1, origin[1]), lambda u, v: 1 - u + 0 * v)]
    return [((origin[0], origin[1]), lambda u, v: v + 0 * u),
            ((origin[0], origin[1] + 1), lambda u, v: 1 - v + 0 * u)]


def _a_by_definition(direction, offset, sizes, k, n=16):
    """A^xx(d) or A^yy(d): the integral of Psi_0 . Psi_d g over the rooftops' four cell pairs."""
    total = 0
    for c1, p1 in _rooftop_pieces(direction, (0, 0)):
        for c2, p2 in _rooftop_pieces(direction, offset):
            total += _cell_pair(c1, p1, c2, p2, sizes, k, n)
    return total


def _phi_by_definition(offset, sizes, k, n=16):
    one = lambda u, v: 1 + 0 * u * v
    return _cell_pair((0, 0), one, offset, one, sizes, k, n)


GRID = (9, 8, 0.1, 0.13)   # unequal face sizes, so that the two directions cannot stand in for each other
K = 2 * np.pi
# Offsets whose cell pairs are all separated by at least one face: the rule of _cell_pair is then exact to rounding.
XX_OFFSETS = [(1, 2), (2, 3), (3, 0), (4, 1), (6, 5), (0, 3), (7, 7), (3, 2)]
YY_OFFSETS = [(2, 1), (3, 2), (0, 3), (1, 5), (7, 4), (4, 0), (2, 6), (8, 3)]
E_OFFSETS = [(2, 0), (0, 2), (3, 4), (8, 7), (5, 1), (1, 6)]


def _check_against_definition(gen, tol):
    sizes = GRID[2:]
    worst = 0.0
    for key, offsets, value in (
            ('magneticXX', XX_OFFSETS, lambda d: 1j * Z0 * K / (4 * np.pi) * _a_by_definition('x', d, sizes, K)),
            ('magneticYY', YY_OFFSETS, lambda d: 1j * Z0 * K / (4 * np.pi) * _a_by_definition('y', d, sizes, K)),
            ('electric', E_OFFSETS, lambda d: 1j * Z0 / (4 * np.pi * K) * _phi_by_definition(d, sizes, K))):
        for d in offsets:
            expected = value(d)
            err = abs(gen[key][d] - expected) / abs(expected)
            worst = max(worst, err)
            if tol is not None:
                assert err < tol, (key, d, err)
    return worst


def test_generators_match_the_rooftop_definition():
    """Far and near (separated) generator entries of A.generate against the 4D integral of the definition, 1e-12."""
    gen = A.generate(A.Grid(*GRID), K, (16, 8))
    _check_against_definition(gen, 1e-12)


def test_definition_check_catches_a_swapped_moment_row(monkeypatch):
    """A planted mistake in the oracle (the e-u and e+u terms of accumulate_moments exchanged, i.e. M10 and M01
    swapped) is caught by the definition check: the pin is not a restatement of the oracle's formula."""
    swapped = A._COMBINATION.copy()
    swapped[[2, 3]] = swapped[[3, 2]]
    swapped[[5, 6]] = swapped[[6, 5]]
    monkeypatch.setattr(A, '_COMBINATION', swapped)
    gen = A.generate(A.Grid(*GRID), K, (16, 8))
    assert _check_against_definition(gen, None) > 1e-4


def test_stored_reference_generators_agree_coarsely(reference_data):
    """reference.mat's generators come from the package's older, lower-order integration: they agree with the oracle
    to about 1e-2 next to the singular cells and to 1e-4 where the faces are two or more apart."""
    for case in reference_data:
        grid = A.Grid(case['countX'], case['countY'], case['sizeX'], case['sizeY'])
        gen = A.generate(grid, case['wavenumber'], (16, 8))
        for key in ('magneticXX', 'magneticYY', 'electric'):
            stored = case[key]
            if stored is None or stored.size == 0:
                continue
            err = np.abs(gen[key] - stored) / np.abs(stored)
            assert np.max(err) < 2e-2, (key, np.max(err))
            far = np.add.outer(np.arange(stored.shape[0]), np.arange(stored.shape[1])) >= 3
            if np.any(far):
                assert np.max(err[far]) < 1e-4, (key, np.max(err[far]))


# --- accumulate_moments against the note's formula ------------------------------------------------------------------

def _generators_from_moments(moments, counts, scales):
    """The generators from the seven moments of every offset, entry by entry (02_setting.tex:124-137): at d = (p, q),
    A^xx/(ab)^2 = M11(d) + (M00 - M10 - M01 + M11)(d) + (M10 - M11)(d + ex) + (M01 - M11)(d - ex), with
    (M01 - M11)(-1, q) read as (M10 - M11)(1, q); A^yy the same with N and ey; electric = s1 M00."""
    cx, cy = counts
    M = moments.reshape(7, cx, cy, order='F')
    M00, M10, M01, M11, N10, N01, N11 = M
    e = scales[0] * M00
    mxx = np.zeros((cx - 1, cy), complex) if cx > 1 else None
    myy = np.zeros((cx, cy - 1), complex) if cy > 1 else None
    for p in range(cx - 1):
        for q in range(cy):
            back = (M01[p - 1, q] - M11[p - 1, q]) if p >= 1 else (M10[1, q] - M11[1, q])
            mxx[p, q] = scales[1] * (M11[p, q] + (M00[p, q] - M10[p, q] - M01[p, q] + M11[p, q])
                                     + (M10[p + 1, q] - M11[p + 1, q]) + back)
    for p in range(cx):
        for q in range(cy - 1):
            back = (N01[p, q - 1] - N11[p, q - 1]) if q >= 1 else (N10[p, 1] - N11[p, 1])
            myy[p, q] = scales[1] * (N11[p, q] + (M00[p, q] - N10[p, q] - N01[p, q] + N11[p, q])
                                     + (N10[p, q + 1] - N11[p, q + 1]) + back)
    return {'electric': e, 'magneticXX': mxx, 'magneticYY': myy}


@pytest.mark.parametrize('counts,chunks', [((5, 4), 'random'), ((5, 4), 'runs'), ((1, 6), 'one'), ((7, 1), 'one')])
def test_accumulate_moments_builds_the_generators(counts, chunks):
    """AccumulateMomentsTest.m:5-60: random moments of all offsets added in chunks (in random order, or in runs
    that end inside a column) give the generators of the note's formula; a one-face-wide grid has one direction."""
    rng = np.random.default_rng(hash(chunks) % 1000 + counts[0])
    cx, cy = counts
    count = cx * cy
    moments = rng.standard_normal((7, count)) + 1j * rng.standard_normal((7, count))
    idx = np.arange(count)
    offsets = np.vstack([idx % cx, idx // cx])
    parts = A.allocate_parts(A.Grid(cx, cy, 0.1, 0.1))
    if chunks == 'random':
        order = rng.permutation(count)
        pieces = [order[:7], order[7:]]
    elif chunks == 'runs':
        pieces = [idx[:7], idx[7:13], idx[13:]]
    else:
        pieces = [idx]
    scales = (2.0, 3.0)
    for piece in pieces:
        A.accumulate_moments(parts, moments[:, piece], offsets[:, piece], scales)
    expected = _generators_from_moments(moments, counts, scales)
    np.testing.assert_allclose(parts['electric'], expected['electric'], rtol=1e-15)
    for key in ('magneticXX', 'magneticYY'):
        if expected[key] is None:
            assert parts[key] is None
        else:
            np.testing.assert_allclose(parts[key], expected[key], rtol=1e-14, atol=0)


# --- radar cross section -------------------------------------------------------------------------------------------

def test_radar_cross_section_closed_form():
    """ComputeRadarCrossSectionTest.m:5-11: b.' I = 2i, k = 2, E = 1 gives (2 Z0)^2 4 / (4 pi)."""
    sigma = X.radar_cross_section(2.0, np.array([1, 1j]), np.array([1j, 1]), 1.0)
    expected = (2 * Z0) ** 2 * 4 / (4 * np.pi)
    assert abs(sigma - expected) <= 1e-15 * expected
    assert abs(expected - 180705.45087914678) < 1e-9


def test_radar_cross_section_does_not_depend_on_the_intensity():
    """ComputeRadarCrossSectionTest.m:13-23: excitation and current scale with the intensity."""
    rng = np.random.default_rng(3)
    b = rng.standard_normal(5) + 1j * rng.standard_normal(5)
    current = rng.standard_normal(5) + 1j * rng.standard_normal(5)
    weak = X.radar_cross_section(2 * np.pi, b, current, 1.0)
    strong = X.radar_cross_section(2 * np.pi, 3 * b, 3 * current, 3.0)
    assert abs(strong - weak) <= 1e-14 * abs(weak)
