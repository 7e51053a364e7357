"""GPU solves with the operator forms the reference's solver accepts besides its own products (convertToHandle.m:10-16,
Solver.m:146-154): a matrix, a function handle and an object with apply, as operator or preconditioner. The synthetic
systems of tests/+unit/+solver/Base.m:12-29 through rt_solve (dense operators on the device, callbacks on the host),
against backslash and the oracle's loops: ConjugateGradientTest.m, GeneralizedMinimumResidualTest.m and
IterativeMethodTest.m."""
import numpy as np
import pytest

from conftest import rel
from oracle import solvers as SV
from test_oracle_solvers import nonsymmetric_system, symmetric_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def rt():
    import __graft_entry__ as ge
    return ge.load()


def test_cocg_on_a_matrix_matches_backslash_and_the_oracle(rt):
    """ConjugateGradientTest.m:5-32: the complex symmetric system to 1e-10 (solution within 1e-9 of backslash), the
    record of every iteration, and the iteration limit one short of it."""
    m, b = symmetric_system(40)
    res = rt.Solver('cg', tolerance=1e-10, maxIterationCount=100).solve(m, b)
    assert res.converged() and rel(res.solution(), np.linalg.solve(m, b)) < 1e-9
    ref = SV.solve('cg', m, b, tol=1e-10, max_iter=100)
    assert res.iterationCount() == ref['iterationCount']
    keep = ref['residualNorms'] > 1e-8   # the tail is rounding-sized in both loops
    np.testing.assert_allclose(res.residualNorms()[keep], ref['residualNorms'][keep], rtol=1e-6)
    assert res.residualNorms()[0] == 1 and np.all(res.residualNorms()[:-1] >= 1e-10)
    shorter = rt.Solver('cg', tolerance=1e-10, maxIterationCount=res.iterationCount() - 1).solve(m, b)
    assert not shorter.converged() and shorter.iterationCount() == res.iterationCount() - 1


def test_cocg_breakdown_stops_after_one_iteration(rt):
    """ConjugateGradientTest.m:66-72: diag([1, -1]) with b = [1; 1] has b.' A b = 0."""
    res = rt.Solver('cg', tolerance=1e-8, maxIterationCount=100).solve(np.diag([1.0, -1.0]), np.array([1.0, 1.0]))
    assert not res.converged() and res.iterationCount() == 1


def test_function_handle_preconditioner_cuts_the_iterations(rt):
    """ConjugateGradientTest.m:55-64: a diagonal preconditioner given as a function handle (a host callback)."""
    m, b = symmetric_system(100)
    s = np.logspace(0, 2, 100)
    m = s[:, None] * m * s[None, :]
    diag = np.diag(m).copy()
    solver = rt.Solver('cg', tolerance=1e-8, maxIterationCount=1000)
    plain = solver.solve(m, b)
    pre = solver.solve(m, b, preconditioner=lambda r: r / diag)
    assert pre.converged() and pre.residualNorm() < 1e-8
    assert pre.iterationCount() < plain.iterationCount() / 2
    ref = SV.solve('cg', m, b, tol=1e-8, max_iter=1000, preconditioner=lambda r: r / diag)
    assert pre.iterationCount() == ref['iterationCount']


@pytest.mark.parametrize('restart', [10, 50, np.inf])
def test_gmres_on_a_nonsymmetric_matrix(rt, restart):
    """GeneralizedMinimumResidualTest.m: the non-symmetric system to 1e-10 against backslash, with restarts that cut
    the cycles short, the whole space, or none; the oracle's iteration record."""
    m, b = nonsymmetric_system(40)
    res = rt.Solver('gmres', tolerance=1e-10, maxIterationCount=200, restart=restart).solve(m, b)
    ref = SV.solve('gmres', m, b, tol=1e-10, max_iter=200, restart=restart)
    assert res.converged() and rel(res.solution(), np.linalg.solve(m, b)) < 1e-9
    assert res.iterationCount() == ref['iterationCount']
    keep = ref['residualNorms'] > 1e-8
    np.testing.assert_allclose(res.residualNorms()[keep], ref['residualNorms'][keep], rtol=1e-6)


def test_gmres_on_an_invariant_subspace_converges_at_the_breakdown(rt):
    """A matrix with two distinct eigenvalues: the Krylov space is invariant after two steps, the residual is zero
    there (a zero subdiagonal, iterate_.m:62-66), in the oracle's count of iterations."""
    n = 30
    m = np.diag(np.where(np.arange(n) % 2 == 0, 1.0, 2.0)).astype(complex)
    b = np.arange(1, n + 1) * (1 + 0.5j)
    for restart in (20, 100):
        res = rt.Solver('gmres', tolerance=1e-12, restart=restart).solve(m, b)
        ref = SV.solve('gmres', m, b, tol=1e-12, max_iter=1000, restart=restart)
        assert res.converged() and res.iterationCount() == ref['iterationCount'] == 2
        assert rel(res.solution(), np.linalg.solve(m, b)) < 1e-12


class _Applier:
    def __init__(self, matrix):
        self.matrix = matrix

    def apply(self, x):
        return self.matrix @ x


@pytest.mark.parametrize('method', ['cg', 'gmres'])
def test_every_operator_form_gives_the_same_solution(rt, method):
    """IterativeMethodTest.m:52-60: a matrix, a function handle and an object with apply solve the same system, and a
    matrix preconditioner is accepted too; a mixed solve (complex64 dense products) refines to the tolerance."""
    m, b = symmetric_system(30)
    solver = rt.Solver(method, tolerance=1e-10, maxIterationCount=200)
    forms = [m, lambda x: m @ x, _Applier(m), rt.DenseOperator(m)]
    sols = [solver.solve(form, b) for form in forms]
    for s in sols:
        assert s.converged() and rel(s.solution(), np.linalg.solve(m, b)) < 1e-9
    for s in sols[1:]:
        assert s.iterationCount() == sols[0].iterationCount()
    pre = solver.solve(m, b, preconditioner=np.diag(1 / np.diag(m)))
    assert pre.converged() and rel(m @ pre.solution(), b) < 1e-10
    mixed = rt.Solver(method, tolerance=1e-10, maxIterationCount=200, precision='c64').solve(
        rt.DenseOperator(m, precision='mixed'), b)
    assert mixed.converged() and mixed.outerCount >= 1 and rel(m @ mixed.solution(), b) < 1e-10


def test_operator_errors(rt):
    m, b = symmetric_system(10)
    with pytest.raises(rt.RooftopError) as e:
        rt.Solver().solve(m, np.ones(7))
    assert e.value.identifier == 'rooftopMoM:solver:sizeMismatch'
    with pytest.raises(rt.RooftopError) as e:
        rt.DenseOperator(np.ones((3, 4)))
    assert e.value.identifier == 'rooftopMoM:solver:invalidOperator'
    with pytest.raises(rt.RooftopError) as e:
        rt.Solver().solve(lambda x: x[:3], b)   # a handle that returns the wrong size
    assert e.value.identifier == 'rooftopMoM:solver:sizeMismatch'
    with pytest.raises(ZeroDivisionError):
        rt.Solver().solve(lambda x: 1 / 0, b)   # an error inside the handle reaches the caller
