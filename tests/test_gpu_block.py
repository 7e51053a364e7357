"""Block COCG on the GPU (rt_solve with RT_METHOD_BLOCK_CG): the right-hand sides of one complex symmetric system
share one block Krylov space. Against backslash, the oracle's block loop (oracle.solvers.block_cocg), the one-column
case that is COCG itself (ConjugateGradient.m:32-59), and the iteration counts of the lockstep COCG solves."""
import numpy as np
import pytest

from conftest import rel
from oracle import assembly as A
from oracle import excitation as X
from oracle import nearfield as NF
from oracle import solvers as SV
from oracle import spectra as S
from test_oracle_solvers import symmetric_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def rt():
    import __graft_entry__ as ge
    return ge.load()


def _columns(rng, n, s):
    return rng.standard_normal((n, s)) + 1j * rng.standard_normal((n, s))


def test_one_column_is_cocg(rt):
    m, b = symmetric_system(40)
    blk = rt.Solver('blockcg', tolerance=1e-10, maxIterationCount=100).solve(m, b)
    one = rt.Solver('cg', tolerance=1e-10, maxIterationCount=100).solve(m, b)
    assert blk.converged() and blk.iterationCount() == one.iterationCount()
    keep = one.residualNorms() > 1e-8   # the same recurrence with the direction normalized: equal up to rounding
    np.testing.assert_allclose(blk.residualNorms()[keep], one.residualNorms()[keep], rtol=1e-6)
    assert rel(blk.solution(), one.solution()) < 1e-9


def test_block_terminates_within_n_over_s_iterations(rt):
    """Eight right-hand sides of the 40 x 40 system: the block space gains eight dimensions per iteration, so every
    column converges to 1e-12 within about 40 / 8 iterations, as in the oracle's block loop, and against backslash."""
    m, _ = symmetric_system(40)
    B = _columns(np.random.default_rng(4), 40, 8)
    res = rt.Solver('blockcg', tolerance=1e-12, maxIterationCount=50).solve(m, B)
    _, conv, its, norms = SV.block_cocg(lambda v: m @ v, B, 1e-12, 50)
    ref = np.linalg.solve(m, B)
    for j, got in enumerate(res):
        assert got.converged() and conv[j]
        assert abs(got.iterationCount() - its[j]) <= 1
        assert rel(got.solution(), ref[:, j]) < 1e-10
        k = min(got.residualNorms().size, norms[j].size)
        keep = norms[j][:k] > 1e-6
        np.testing.assert_allclose(got.residualNorms()[:k][keep], norms[j][:k][keep], rtol=1e-6)
    assert max(r.iterationCount() for r in res) <= 6


def test_block_breakdown_stops_every_column(rt):
    """ConjugateGradientTest.m:66-72 as a block: diag([1, -1]) with b = [1; 1] has b.' A b = 0, a singular
    coefficient matrix in the first iteration."""
    res = rt.Solver('blockcg', tolerance=1e-8, maxIterationCount=100).solve(np.diag([1.0, -1.0]),
                                                                          np.array([1.0, 1.0]))
    assert not res.converged() and res.iterationCount() == 1


def test_dependent_and_zero_columns(rt):
    """A repeated column (a multiple of another) and a zero column: the block drops the dependent directions; every
    column matches backslash, the zero one with no iteration, as in the oracle's block loop."""
    m, _ = symmetric_system(40)
    B = _columns(np.random.default_rng(8), 40, 5)
    B[:, 3] = B[:, 1] * (1 - 2j)
    B[:, 4] = 0
    res = rt.Solver('blockcg', tolerance=1e-12, maxIterationCount=100).solve(m, B)
    _, conv, its, _ = SV.block_cocg(lambda v: m @ v, B, 1e-12, 100)
    ref = np.linalg.solve(m, B)
    for j, got in enumerate(res):
        assert got.converged() and abs(got.iterationCount() - its[j]) <= 1
        assert rel(got.solution(), ref[:, j]) < 1e-10 if j < 4 else np.all(got.solution() == 0)
    assert res[4].iterationCount() == 0


def _plate(rt, count, seed=7):
    cx, cy = 16, 14
    p = rt.Problem(rt.Mesh(cx, cy, 0.1, 0.1), 2 * np.pi, quadratureOrder=(6, 4))
    grid = A.Grid(cx, cy, 0.1, 0.1)
    gen = A.generate(grid, 2 * np.pi, (6, 4))
    rng = np.random.default_rng(seed)
    cols = []
    for _ in range(count):
        d = rng.standard_normal(3)
        d[2] = -abs(d[2]) - 0.5
        d /= np.linalg.norm(d)
        e = np.cross(d, [0.0, 1.0, 0.0])
        cols.append(X.plane_wave(grid, 2 * np.pi, d, e / np.linalg.norm(e), 1.0, order=4))
    return p, grid, gen, np.stack(cols, axis=1)


def test_plate_plane_waves_against_oracle_and_cocg(rt):
    """Six plane waves on the 16 x 14 plate with the near-field preconditioner to 1e-8: each column against the dense
    solve and the oracle's block loop; the block needs fewer iterations than the longest of the COCG solves."""
    p, grid, gen, B = _plate(rt, 6)
    Z = S.dense_matrix(gen, grid)
    P = NF.NearField(gen, grid)
    op, pc = rt.Operator(p), rt.Preconditioner(p, 'nearField')
    res = rt.Solver('blockcg', tolerance=1e-8, maxIterationCount=1000).solve(op, B, preconditioner=pc)
    _, conv, its, norms = SV.block_cocg(lambda v: Z @ v, B, 1e-8, 1000, precondition=P.apply)
    exact = np.linalg.solve(Z, B)
    for j, got in enumerate(res):
        print('column', j, got.iterationCount(), its[j], flush=True)
        assert got.converged() and conv[j]
        assert abs(got.iterationCount() - its[j]) <= 1
        assert got.residualNorm() < 1e-8 and rel(got.solution(), exact[:, j]) < 1e-6
        k = min(got.residualNorms().size, norms[j].size)
        keep = norms[j][:k] > 1e-4
        np.testing.assert_allclose(got.residualNorms()[:k][keep], norms[j][:k][keep], rtol=1e-6)
    lock = rt.Solver('cg', tolerance=1e-8, maxIterationCount=1000).solve(op, B, preconditioner=pc)
    assert max(r.iterationCount() for r in res) < max(r.iterationCount() for r in lock)


def test_refined_block_cycles_reach_the_tolerance(rt):
    """Block COCG as the inner solve of FP64 iterative refinement: each cycle reduces its residual by 1e-2 at most (or
    stops at a stall, handing back its best iterate) and the outer loop restarts it from the true residual; every
    column reaches 1e-8 against the dense matrix."""
    p, grid, gen, B = _plate(rt, 6)
    Z = S.dense_matrix(gen, grid)
    op, pc = rt.Operator(p), rt.Preconditioner(p, 'nearField')
    res = rt.Solver('blockcg', tolerance=1e-8, maxIterationCount=500, refine=True).solve(op, B, preconditioner=pc)
    for j, got in enumerate(res):
        assert got.converged() and got.outerCount >= 2
        assert rel(Z @ got.solution(), B[:, j]) < 1e-8


def test_mixed_block_refines_to_fp64_accuracy(rt):
    """Block COCG as the complex64 inner solve of FP64 iterative refinement: every column to 1e-10."""
    p, grid, gen, B = _plate(rt, 5, seed=9)
    Z = S.dense_matrix(gen, grid)
    op = rt.Operator(p, precision='mixed')
    pc = rt.Preconditioner(p, 'nearField', precision='mixed')
    res = rt.Solver('blockcg', tolerance=1e-10, precision='c64').solve(op, B, preconditioner=pc)
    for j, got in enumerate(res):
        assert got.converged() and got.outerCount >= 1
        assert rel(Z @ got.solution(), B[:, j]) < 1e-10


def test_block_limits(rt):
    m, _ = symmetric_system(20)
    with pytest.raises(rt.RooftopError) as e:
        rt.Solver('blockcg').solve(m, _columns(np.random.default_rng(1), 20, 65))
    assert e.value.identifier == 'rooftopMoM:solver:invalidRightHandSide'
    m, _ = symmetric_system(200)
    B = _columns(np.random.default_rng(2), 200, 64)
    res = rt.Solver('blockcg', tolerance=1e-10, maxIterationCount=100).solve(m, B)
    ref = np.linalg.solve(m, B)
    assert all(r.converged() for r in res)
    assert max(rel(r.solution(), ref[:, j]) for j, r in enumerate(res)) < 1e-8


def test_block_cycles_hand_stalled_systems_to_cocg(rt):
    """The 256 x 256 plate under 64 plane waves (seed 1), where the block recurrence of 64 right-hand sides loses its
    conjugacy (it diverges after about 60 iterations): the refined block COCG goes on with COCG, from their iterates,
    for the systems a block cycle does not halve, and every system reaches 1e-6 by the library's FP64 residual, in
    FP64 and mixed; two columns also by the oracle's product."""
    import bench
    n = 256
    p = rt.Problem(rt.Mesh(n, n, 0.1, 0.1), 2 * np.pi, quadratureOrder=(16, 8))
    grid = A.Grid(n, n, 0.1, 0.1)
    op = rt.Operator(p, precision='mixed')
    pc = rt.Preconditioner(p, 'nearField', precision='mixed')
    dirs, pols = bench.plane_wave_angles(64, 1)
    B = rt.plane_waves(p, dirs, pols)
    gen = p.generators()
    spectra = S.subband_spectra(gen, grid)
    for solver in (rt.Solver('blockcg', tolerance=1e-6, maxIterationCount=2000, refine=True),
                   rt.Solver('blockcg', tolerance=1e-6, maxIterationCount=2000, precision='c64')):
        print('plate 256 x 64 blockcg', solver.precision, flush=True)
        xb, results = solver.solveDevice(op, B, preconditioner=pc)
        its = [r.iterationCount() for r in results]
        print('converged %d of 64, iterations median %d max %d, outer max %d' % (
            sum(r.converged() for r in results), int(np.median(its)), max(its), max(r.outerCount for r in results)),
            flush=True)
        assert all(r.converged() and r.residualNorm() < 1e-6 for r in results)
        xh = xb.to_host()
        for q in (0, 63):
            b = X.plane_wave(grid, 2 * np.pi, dirs[q], pols[q], 1.0, order=8)
            assert rel(S.split_fft_product(spectra, grid, xh[:, q]), b) < 1.01e-6
        xb.free()
    B.free()
