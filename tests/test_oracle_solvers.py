"""Pins the oracle's COCG / GMRES / IterativeMethod and near-field preconditioner (CPU)."""
import numpy as np
import pytest

from conftest import rel
from oracle import assembly as A
from oracle import excitation as X
from oracle import nearfield as NF
from oracle import solvers as SV
from oracle import spectra as S


def symmetric_system(n, seed=1):
    """4 I + a symmetric random part (singular values within about 2..6), tests/+unit/+solver/Base.m:12-20."""
    rng = np.random.default_rng(seed)
    r = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    return 4 * np.eye(n) + (r + r.T) / 2, rng.standard_normal(n) + 1j * rng.standard_normal(n)


def nonsymmetric_system(n, seed=2):
    """tests/+unit/+solver/Base.m:22-29."""
    rng = np.random.default_rng(seed)
    m = 4 * np.eye(n) + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    return m, rng.standard_normal(n) + 1j * rng.standard_normal(n)


def test_cocg_solves_complex_symmetric_system():
    m, b = symmetric_system(40)
    out = SV.solve('cg', m, b, tol=1e-10, max_iter=100)
    assert out['converged'] and rel(out['solution'], np.linalg.solve(m, b)) < 1e-9
    assert out['residualNorms'].size == out['iterationCount'] + 1


def test_cocg_records_and_stops():
    m, b = symmetric_system(40)
    out = SV.solve('cg', m, b, tol=1e-8, max_iter=100)
    norms = out['residualNorms']
    assert norms[0] == 1 and norms[-1] < 1e-8 and np.all(norms[:-1] >= 1e-8)
    shorter = SV.solve('cg', m, b, tol=1e-8, max_iter=out['iterationCount'] - 1)
    assert not shorter['converged']


def test_cocg_diagonal_preconditioner_halves_iterations():
    m, b = symmetric_system(100)
    s = np.logspace(0, 2, 100)
    m = s[:, None] * m * s[None, :]
    diag = np.diag(m).copy()
    plain = SV.solve('cg', m, b, tol=1e-8, max_iter=1000)
    pre = SV.solve('cg', m, b, tol=1e-8, max_iter=1000, preconditioner=lambda r: r / diag)
    assert pre['converged'] and pre['residualNorm'] < 1e-8
    assert pre['iterationCount'] < plain['iterationCount'] / 2


def test_cocg_stops_on_breakdown():
    out = SV.solve('cg', np.diag([1.0, -1.0]), np.array([1.0, 1.0]), tol=1e-8, max_iter=100)
    assert not out['converged'] and out['iterationCount'] == 1


def test_gmres_solves_nonsymmetric_system():
    m, b = nonsymmetric_system(40)
    out = SV.solve('gmres', m, b, tol=1e-10, max_iter=100, restart=10)
    assert out['converged'] and out['iterationCount'] > 10
    assert rel(out['solution'], np.linalg.solve(m, b)) < 1e-9


def test_gmres_iteration_limit_and_restart_cap():
    m, b = nonsymmetric_system(60)
    out = SV.solve('gmres', m, b, tol=1e-14, max_iter=7, restart=5)
    assert not out['converged'] and out['iterationCount'] == 7
    m, b = nonsymmetric_system(20)
    a = SV.solve('gmres', m, b, tol=1e-10, max_iter=100, restart=20)
    c = SV.solve('gmres', m, b, tol=1e-10, max_iter=100, restart=50)
    assert a['converged'] and np.array_equal(a['solution'], c['solution'])


def test_gmres_residual_is_that_of_preconditioned_system():
    m, b = nonsymmetric_system(60)
    m = np.logspace(0, 3, 60)[:, None] * m
    d = np.diag(m).copy()
    P = lambda r: r / d
    plain = SV.solve('gmres', m, b, tol=1e-10, max_iter=200, restart=20)
    out = SV.solve('gmres', m, b, tol=1e-10, max_iter=200, restart=20, preconditioner=P)
    expected = np.linalg.norm(P(b - m @ out['solution'])) / np.linalg.norm(P(b))
    assert out['converged'] and abs(out['residualNorms'][0] - 1) < 1e-14
    assert abs(out['residualNorms'][-1] - expected) < 1e-12
    assert out['iterationCount'] < plain['iterationCount']


@pytest.mark.parametrize('method', ['cg', 'gmres'])
def test_iterative_method_common_behaviour(method):
    m, b = symmetric_system(30)
    out = SV.solve(method, m, b, tol=1e-6, max_iter=100, restart=np.inf)
    assert abs(out['residualNorm'] - np.linalg.norm(b - m @ out['solution']) / np.linalg.norm(b)) < 1e-12 * 1e-6 + 1e-18
    zero = SV.solve(method, np.eye(3), np.zeros(3), x0=np.ones(3), tol=1e-10, max_iter=100)
    assert zero['converged'] and zero['iterationCount'] == 0 and np.all(zero['solution'] == 0)
    exact = SV.solve(method, m, b, x0=np.linalg.solve(m, b), tol=1e-10, max_iter=100)
    assert exact['converged'] and exact['iterationCount'] == 0
    limited = SV.solve(method, m, b, tol=1e-14, max_iter=2, restart=np.inf)
    assert not limited['converged'] and limited['iterationCount'] == 2


@pytest.mark.parametrize('method', ['cg', 'gmres'])
def test_plate_solve_against_dense_matrix(method):
    """SolverTest.m:9-21: an 8x8 plate, plane wave, tolerance 1e-10, checked with the dense matrix."""
    grid = A.Grid(8, 8, 0.05, 0.05)
    g = A.generate(grid, 2 * np.pi, (3, 3))
    b = X.plane_wave(grid, 2 * np.pi, [0, 0, -1], [1, 0, 0], 1.0, order=3)
    spec = S.subband_spectra(g, grid)
    out = SV.solve(method, lambda x: S.split_fft_product(spec, grid, x), b, tol=1e-10, max_iter=1000, restart=50)
    Z = S.dense_matrix(g, grid)
    assert out['converged'] and rel(Z @ out['solution'], b) < 1e-10


def test_near_field_entries_and_symmetry():
    grid = A.Grid(7, 5, 0.1, 0.15)
    g = A.generate(grid, 2 * np.pi, (4, 4))
    Z = S.dense_matrix(g, grid)
    every = np.arange(grid.count)
    np.testing.assert_allclose(NF.entries(g, grid, every, every), Z, rtol=0, atol=1e-13 * np.abs(Z).max())
    P = NF.NearField(g, grid)
    assert P.matrix.getnnz(axis=1).max() == 18
    rng = np.random.default_rng(3)
    x = rng.standard_normal(grid.count) + 1j * rng.standard_normal(grid.count)
    y = rng.standard_normal(grid.count) + 1j * rng.standard_normal(grid.count)
    assert abs(x @ P.apply(y) - y @ P.apply(x)) < 1e-12 * abs(x @ P.apply(y))
    np.testing.assert_allclose(NF.jacobi_diagonal(g, grid), np.diag(Z), rtol=1e-14)


def test_near_field_rows_are_least_squares_fits():
    """Each SAI row is a least-squares solution: its fit residual is orthogonal to the columns (normal equations)."""
    grid = A.Grid(6, 6, 0.1, 0.1)
    g = A.generate(grid, 2 * np.pi, (4, 4))
    Z = S.dense_matrix(g, grid)
    M = NF.sai_matrix(g, grid).toarray()
    _, d, fx, fy = NF.rooftop_offsets(grid)
    for i in (0, 7, 20, grid.count - 1):
        near = np.nonzero((np.abs(fx - fx[i]) <= 1) & (np.abs(fy - fy[i]) <= 1))[0]
        fit = np.nonzero((np.abs(fx - fx[i]) <= 2) & (np.abs(fy - fy[i]) <= 2))[0]
        a = Z[np.ix_(fit, near)]
        res = a @ M[i, near] - (fit == i)
        assert np.linalg.norm(a.conj().T @ res) < 1e-12 * np.linalg.norm(a) ** 2
        assert np.count_nonzero(M[i]) == near.size


@pytest.mark.parametrize('method', ['cg', 'gmres'])
def test_near_field_preconditioner_cuts_iterations(method):
    grid = A.Grid(24, 24, 0.1, 0.1)
    g = A.generate(grid, 2 * np.pi, (6, 4))
    b = X.plane_wave(grid, 2 * np.pi, [0, 0, -1], [1, 0, 0], 1.0, order=4)
    spec = S.subband_spectra(g, grid)
    op = lambda x: S.split_fft_product(spec, grid, x)
    plain = SV.solve(method, op, b, tol=1e-6, max_iter=500, restart=50)
    pre = SV.solve(method, op, b, tol=1e-6, max_iter=500, restart=50, preconditioner=NF.NearField(g, grid).apply)
    assert pre['converged'] and pre['iterationCount'] < plain['iterationCount'] / 2


def test_near_field_by_class_equals_per_rooftop_fit():
    """Rooftops with the same existing neighbourhood share their SAI row (Z depends on offsets only)."""
    grid = A.Grid(11, 9, 0.1, 0.12)
    g = A.generate(grid, 2 * np.pi, (4, 4))
    mask = np.ones((11, 9), bool)
    mask[4:7, 3:5] = False
    mx, my = mask[:-1] & mask[1:], mask[:, :-1] & mask[:, 1:]
    for args in ((), (mx, my)):
        a = NF.sai_matrix(g, grid, *args).toarray()
        b = NF.sai_matrix_by_class(g, grid, *args).toarray()
        assert np.abs(a - b).max() < 1e-11 * np.abs(a).max()


def test_block_cocg_of_one_column_is_cocg():
    m, b = symmetric_system(40)
    X, conv, its, norms = SV.block_cocg(lambda v: m @ v, b[:, None], 1e-10, 100)
    ref = SV.solve('cg', m, b, tol=1e-10, max_iter=100)
    assert conv[0] and its[0] == ref['iterationCount']
    np.testing.assert_allclose(norms[0], ref['residualNorms'], rtol=1e-8)
    assert rel(X[:, 0], ref['solution']) < 1e-10


def test_block_cocg_terminates_within_n_over_s_iterations():
    """A block Krylov space of s columns gains s dimensions per iteration: on the 40 x 40 well-conditioned complex
    symmetric system, 8 right-hand sides converge to 1e-12 within 40 / 8 = 5 iterations (plus rounding), against
    backslash; COCG alone needs far more."""
    m, _ = symmetric_system(40)
    rng = np.random.default_rng(4)
    B = rng.standard_normal((40, 8)) + 1j * rng.standard_normal((40, 8))
    X, conv, its, norms = SV.block_cocg(lambda v: m @ v, B, 1e-12, 50)
    assert conv.all() and its.max() <= 6
    assert rel(X, np.linalg.solve(m, B)) < 1e-11
    single = SV.solve('cg', m, B[:, 0], tol=1e-12, max_iter=100)
    assert single['iterationCount'] > 2 * its.max()


def test_block_cocg_with_preconditioner_matches_backslash():
    m, _ = symmetric_system(60)
    s = np.logspace(0, 1, 60)
    m = s[:, None] * m * s[None, :]
    rng = np.random.default_rng(5)
    B = rng.standard_normal((60, 5)) + 1j * rng.standard_normal((60, 5))
    d = np.diag(m).copy()
    X, conv, its, norms = SV.block_cocg(lambda v: m @ v, B, 1e-10, 200, precondition=lambda v: v / d)
    assert conv.all() and rel(X, np.linalg.solve(m, B)) < 1e-8


def test_block_orth_basis_and_drop():
    """block_orth: orthonormal kept columns spanning W; a zero column and a column that is a combination of earlier
    kept ones (the last, in the order columns are taken) are dropped."""
    rng = np.random.default_rng(6)
    W = rng.standard_normal((50, 6)) + 1j * rng.standard_normal((50, 6))
    P, C, T, kept = SV.block_orth(W, 1e-6)
    assert kept.all()
    np.testing.assert_allclose(P.conj().T @ P, np.eye(6), atol=1e-13)
    assert rel(P @ (P.conj().T @ W), W) < 1e-13
    assert rel(P @ T, W) < 1e-13 and np.all(np.tril(T, -1) == 0)
    r = np.linalg.qr(W)[1]   # Householder QR, rows scaled to a positive real diagonal: the same factor
    d = np.diag(r)
    np.testing.assert_allclose(T, (d.conj() / abs(d))[:, None] * r, atol=1e-12)
    W[:, 3] = 2 * W[:, 0] - 1j * W[:, 5]
    W[:, 4] = 0
    P, C, T, kept = SV.block_orth(W, 1e-6)
    assert kept.tolist() == [True, True, True, True, False, False]
    assert rel(P @ T, W) < 1e-12
    k = np.flatnonzero(kept)
    np.testing.assert_allclose(P[:, k].conj().T @ P[:, k], np.eye(4), atol=1e-12)
    assert np.all(P[:, ~kept] == 0) and rel(P @ (P.conj().T @ W), W) < 1e-12


def test_block_cocg_deflates_a_dependent_block():
    """Right-hand sides with a repeated column and a zero column: the block drops the dependent directions and every
    column still matches backslash."""
    m, _ = symmetric_system(40)
    rng = np.random.default_rng(8)
    B = rng.standard_normal((40, 5)) + 1j * rng.standard_normal((40, 5))
    B[:, 3] = B[:, 1] * (1 - 2j)
    B[:, 4] = 0
    X, conv, its, norms = SV.block_cocg(lambda v: m @ v, B, 1e-12, 100)
    assert conv.all() and its[4] == 0 and np.all(X[:, 4] == 0)
    assert rel(X, np.linalg.solve(m, B)) < 1e-11
