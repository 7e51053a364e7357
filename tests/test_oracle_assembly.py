"""Pins the oracle's generators, spectra, products, dense matrix and spatial blocks (CPU)."""
import numpy as np
import pytest

from conftest import rel
from oracle import assembly as A
from oracle import excitation as X
from oracle import quadrature as Q
from oracle import spectra as S
from oracle.singular import evaluate_remainder

GRIDS = {'plate': (5, 4, 0.1, 0.15), 'oddPlate': (11, 7, 0.1, 0.1), 'twoFacesWide': (2, 5, 0.1, 0.1),
         'stripX': (9, 1, 0.1, 0.02), 'stripY': (1, 6, 0.03, 0.1)}


def _generate(cx, cy, a, b, orders, k=2 * np.pi):
    grid = A.Grid(cx, cy, a, b)
    return A.generate(grid, k, orders), grid


# --- reference.mat (tests/+reference) ---------------------------------------------------------------------------

def test_products_match_reference_mat(reference_data):
    """tests/+reference/+assembler/+greensMatrix/ApplicatorTest.m:5-21 at 1e-13."""
    for case in reference_data:
        grid = A.Grid(case['countX'], case['countY'], case['sizeX'], case['sizeY'])
        g = {'magneticXX': case['magneticXX'], 'magneticYY': case['magneticYY'], 'electric': case['electric']}
        v = case['vector']
        assert rel(S.fft_product(S.embedding_spectra(g, grid), grid, v), case['productFFT']) < 1e-13
        split = S.split_fft_product(S.subband_spectra(g, grid), grid, v)
        assert rel(split, case['productSplit']) < 1e-13
        assert rel(split, case['productLean']) < 1e-13
        assert rel(S.dense_matrix(g, grid) @ v, case['productFFT']) < 1e-13


def test_fft_product_matches_stored_dense_matrix(reference_data):
    """ApplicatorTest.m:23-34: the independent v2 dense matrix, accurate to about 4e-11."""
    case = reference_data[0]
    grid = A.Grid(case['countX'], case['countY'], case['sizeX'], case['sizeY'])
    g = {'magneticXX': case['magneticXX'], 'magneticYY': case['magneticYY'], 'electric': case['electric']}
    expected = case['impedanceMatrix'] @ case['vector']
    assert rel(S.fft_product(S.embedding_spectra(g, grid), grid, case['vector']), expected) < 1e-10


def test_plane_wave_matches_reference_mat(reference_data):
    """tests/+reference/+operator/GeneratePlaneWaveTest.m:5-13 at 1e-13."""
    for case in reference_data:
        grid = A.Grid(case['countX'], case['countY'], case['sizeX'], case['sizeY'])
        pw = X.plane_wave(grid, case['wavenumber'], case['direction'], case['polarization'], case['intensity'],
                          int(case['order']))
        assert rel(pw, case['planeWave']) < 1e-13


# --- products against the dense matrix of the generated generators (tests/+unit ApplicatorTest) ------------------

@pytest.mark.parametrize('name', sorted(GRIDS))
def test_products_match_dense_matrix(name):
    g, grid = _generate(*GRIDS[name], orders=(3, 3))
    rng = np.random.default_rng(1)
    v = rng.standard_normal(grid.count) + 1j * rng.standard_normal(grid.count)
    expected = S.dense_matrix(g, grid) @ v
    assert rel(S.fft_product(S.embedding_spectra(g, grid), grid, v), expected) < 1e-12
    assert rel(S.split_fft_product(S.subband_spectra(g, grid), grid, v), expected) < 1e-12


def test_modifier_selects_part_of_the_matrix():
    """OperatorTest.m:18-25: 'modifier' imag gives the imaginary part of the matrix."""
    g, grid = _generate(5, 4, 0.1, 0.15, (3, 3), k=4.0)
    gi = A.modify(g, 'imag')
    rng = np.random.default_rng(2)
    v = rng.standard_normal(grid.count) + 1j * rng.standard_normal(grid.count)
    expected = np.imag(S.dense_matrix(g, grid)) @ v
    assert rel(S.split_fft_product(S.subband_spectra(gi, grid), grid, v), expected) < 1e-12


# --- generators (MomentAssemblerTest) -----------------------------------------------------------------------------

def test_generator_sizes_and_content():
    g, grid = _generate(5, 4, 0.1, 0.1, (3, 3))
    assert g['magneticXX'].shape == (4, 4) and g['magneticYY'].shape == (5, 3) and g['electric'].shape == (5, 4)
    assert grid.content == 'BOTH'
    g, grid = _generate(6, 1, 0.1, 0.1, (3, 3))
    assert g['magneticXX'].shape == (5, 1) and g['magneticYY'] is None and grid.content == 'XONLY'


def test_entries_do_not_depend_on_grid_size():
    narrow, _ = _generate(2, 3, 0.1, 0.1, (3, 3))
    wide, _ = _generate(5, 3, 0.1, 0.1, (3, 3))
    np.testing.assert_allclose(narrow['magneticXX'], wide['magneticXX'][:1], rtol=1e-13)
    np.testing.assert_allclose(narrow['magneticYY'], wide['magneticYY'][:2], rtol=1e-13)
    np.testing.assert_allclose(narrow['electric'], wide['electric'][:2], rtol=1e-13)


def test_transposed_grid_transposes_generators():
    g, _ = _generate(5, 4, 0.1, 0.15, (4, 4))
    t, _ = _generate(4, 5, 0.15, 0.1, (4, 4))
    np.testing.assert_allclose(t['magneticYY'], g['magneticXX'].T, rtol=1e-13)
    np.testing.assert_allclose(t['electric'], g['electric'].T, rtol=1e-13)


def test_lattice_equals_per_offset_integration():
    """MomentAssemblerTest.m:55-77."""
    counts, cell, order, k = (9, 7), 0.1, 6, 2 * np.pi
    g, grid = _generate(counts[0], counts[1], cell, cell, (order, order))
    rule = Q.reduced_rule(order, (1, 1))
    dx, dy = np.meshgrid(np.arange(counts[0]), np.arange(counts[1]), indexing='ij')
    offsets = np.vstack([dx.ravel(order='F'), dy.ravel(order='F')])
    scales = (1j * A.Z0 / (4 * np.pi) * cell ** 4 / k, 1j * A.Z0 / (4 * np.pi) * cell ** 4 * k)
    parts = A.allocate_parts(grid)
    A.accumulate_moments(parts, A.integrate_moments(rule, offsets, (cell, cell), k), offsets, scales)
    from oracle.singular import extracted_moments
    A.accumulate_moments(parts, extracted_moments(cell, cell, k), np.array([[0, 1, 0, 1], [0, 0, 1, 1]]), scales)
    for key in ('electric', 'magneticXX', 'magneticYY'):
        np.testing.assert_allclose(g[key], parts[key], rtol=1e-14)


@pytest.mark.parametrize('case', [(12, 12, 0.1, 0.1), (6, 40, 0.1, 0.01), (30, 1, 0.025, 0.004)])
def test_default_orders_reach_rounding_level(case):
    """MomentAssemblerTest.m:79-91: orders [16, 8] against [32, 24]."""
    default, _ = _generate(*case, orders=(16, 8))
    precise, _ = _generate(*case, orders=(32, 24))
    for key in ('magneticXX', 'magneticYY', 'electric'):
        if default[key] is not None:
            np.testing.assert_allclose(default[key], precise[key], rtol=3e-14)


def test_integrate_moments_matches_definition_on_separated_cells():
    """IntegrateMomentsTest.m:5-28: brute-force tensor Gauss-Legendre on both cells."""
    sizes, k = (0.1, 0.15), 2 * np.pi
    offsets = np.array([[4, 0, 3], [1, 5, 4]])
    got = A.integrate_moments(Q.reduced_rule(8, (1, 1)), offsets, sizes, k)
    p, w = Q.gauss_legendre(10, 0, 1)
    xi, xo, e1, eo = np.meshgrid(p, p, p, p, indexing='ij')
    pw = w[:, None, None, None] * w[None, :, None, None] * w[None, None, :, None] * w[None, None, None, :]
    for n in range(3):
        d = np.sqrt((sizes[0] * (xi - xo - offsets[0, n])) ** 2 + (sizes[1] * (e1 - eo - offsets[1, n])) ** 2)
        v = pw * np.exp(-1j * k * d) / d
        expected = np.array([v.sum(), (xi * v).sum(), (xo * v).sum(), (xi * xo * v).sum(), (e1 * v).sum(),
                             (eo * v).sum(), (e1 * eo * v).sum()])
        np.testing.assert_allclose(got[:, n], expected, rtol=1e-13)


def test_touching_pairs_integrate_the_remainder():
    sizes, k = (0.1, 0.15), 2 * np.pi
    rule = Q.reduced_rule(6, (2, 1))
    nu, eta = np.meshgrid(rule['nodesX'], rule['nodesY'], indexing='ij')
    for off in ((0, 0), (1, 0), (0, 1), (1, 1)):
        d = np.sqrt((sizes[0] * (nu.ravel(order='F') - off[0])) ** 2 + (sizes[1] * (eta.ravel(order='F') - off[1])) ** 2)
        expected = rule['weights'] @ evaluate_remainder(d, k)
        np.testing.assert_allclose(A.integrate_moments(rule, np.array(off), sizes, k)[:, 0], expected, rtol=1e-14)


_W = (lambda nu: 1 - np.abs(nu),
      lambda nu: np.where(nu < 0, (1 + nu) ** 2 / 2, (1 - nu ** 2) / 2),
      lambda nu: np.where(nu < 0, (1 - nu ** 2) / 2, (1 - nu) ** 2 / 2),
      lambda nu: (1 - np.abs(nu)) ** 2 * (2 + np.abs(nu)) / 6)


def _duffy_moment(sizes, k, offset, kinds, order=48):
    """The moment of exp(-ikR)/R at a touching offset by Duffy's transform from the singular corner of every quadrant
    (an independent quadrature standing in for the reference's adaptive integral2)."""
    weight = _W[kinds[0] + 2 * kinds[1]]
    p, w = Q.gauss_legendre(order, 0, 1)
    s, t = np.meshgrid(p, p, indexing='ij')
    ws = np.outer(w, w)
    total = 0
    for n0, n1 in ((-1, 0), (0, 1)):
        for e0, e1 in ((-1, 0), (0, 1)):
            cn = n0 if abs(n0 - offset[0]) < abs(n1 - offset[0]) else n1
            ce = e0 if abs(e0 - offset[1]) < abs(e1 - offset[1]) else e1
            sn = 1 if cn == n0 else -1
            se = 1 if ce == e0 else -1
            for tri in range(2):
                u, v = (s, s * t) if tri == 0 else (s * t, s)
                nu = cn + sn * u
                eta = ce + se * v
                r = np.hypot(sizes[0] * (nu - offset[0]), sizes[1] * (eta - offset[1]))
                f = weight(nu) * (1 - np.abs(eta)) * np.exp(-1j * k * r) / r
                total += np.sum(ws * s * f)
    return total


def test_touching_entries_match_independent_quadrature():
    """MomentAssemblerTest.m:93-114 (reference tolerance 1e-10)."""
    sizes, k = (0.1, 0.15), 2 * np.pi
    g, _ = _generate(3, 3, sizes[0], sizes[1], (16, 8), k)
    scale = 1j * A.Z0 / (4 * np.pi) * (sizes[0] * sizes[1]) ** 2
    for off in ((0, 0), (1, 0), (0, 1), (1, 1)):
        np.testing.assert_allclose(g['electric'][off], scale / k * _duffy_moment(sizes, k, off, (0, 0)), rtol=1e-10)
    co = [_duffy_moment(sizes, k, (0, 0), kinds) for kinds in ((0, 0), (1, 0), (0, 1), (1, 1))]
    nb = _duffy_moment(sizes, k, (1, 0), (1, 0)) - _duffy_moment(sizes, k, (1, 0), (1, 1))
    expected = scale * k * (co[0] - co[1] - co[2] + 2 * co[3] + 2 * nb)
    np.testing.assert_allclose(g['magneticXX'][0, 0], expected, rtol=1e-10)


# --- spatial blocks (GenerateSpatialBlocksTest) --------------------------------------------------------------------

@pytest.mark.parametrize('counts', [(4, 3), (2, 3), (5, 1), (1, 4)])
def test_spatial_blocks_reproduce_dense_matrix(counts):
    g, grid = _generate(counts[0], counts[1], 0.1, 0.15, (3, 3))
    Z = S.dense_matrix(g, grid)
    blocks = S.spatial_blocks(g, grid)
    first = [0, grid.count_dir('x')]
    for m, dm in enumerate('xy'):
        for n, dn in enumerate('xy'):
            if grid.count_dir(dm) == 0 or grid.count_dir(dn) == 0:
                continue
            ro, co = grid.face_offsets(dm), grid.face_offsets(dn)
            block = blocks[2 * m + n][ro[0][:, None] - co[0][None, :] + counts[0] - 1,
                                      ro[1][:, None] - co[1][None, :] + counts[1] - 1]
            expected = Z[first[m]:first[m] + grid.count_dir(dm), first[n]:first[n] + grid.count_dir(dn)]
            np.testing.assert_allclose(block, expected, rtol=0, atol=1e-13 * np.max(np.abs(Z)))


# --- excitation (GeneratePlaneWaveTest, generateMiddlePort) ---------------------------------------------------------

def test_normal_incidence_gives_face_area():
    grid = A.Grid(4, 3, 0.1, 0.2)
    b = X.plane_wave(grid, 2 * np.pi, [0, 0, -1], [0.6, 0.8, 0], 2.0, order=3)
    np.testing.assert_allclose(b, np.r_[np.full(9, 2 * 0.6 * 0.02), np.full(8, 2 * 0.8 * 0.02)], rtol=1e-14)


def test_plane_wave_phase_steps():
    grid = A.Grid(4, 3, 0.1, 0.2)
    d = np.array([1, 2, -2]) / 3
    b = X.plane_wave(grid, 2 * np.pi, d, np.array([2, -1, 0]) / np.sqrt(5), 1.0, order=3)
    px = b[:9].reshape(3, 3, order='F')
    py = b[9:].reshape(4, 2, order='F')
    np.testing.assert_allclose(px[1:] / px[:-1], np.exp(-1j * 2 * np.pi * d[0] * 0.1), rtol=1e-13)
    np.testing.assert_allclose(py[:, 1] / py[:, 0], np.exp(-1j * 2 * np.pi * d[1] * 0.2), rtol=1e-13)


def test_middle_port_and_delta_gap():
    grid = A.Grid(64, 4, 0.01, 0.01)
    port = X.middle_port(grid, 'x')
    np.testing.assert_array_equal(port, 31 + 63 * np.arange(4))
    b = X.delta_gap(grid, port, 1.0)
    assert np.count_nonzero(b) == 4 and np.allclose(b[port], 0.01)
    assert abs(X.input_impedance(grid, np.where(b != 0, 2.0, 0), port, 1.0) - 1 / 0.08) < 1e-12
