"""GPU parity of the preconditioners and the Krylov loops (librooftop.so through the C ABI) against the oracle."""
import numpy as np
import pytest

from conftest import rel
from oracle import assembly as A
from oracle import excitation as X
from oracle import nearfield as NF
from oracle import solvers as SV
from oracle import spectra as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def rt():
    import __graft_entry__ as ge
    return ge.load()


def _setup(rt, cx, cy, a, b, orders, k=2 * np.pi, mask=None):
    p = rt.Problem(rt.Mesh(cx, cy, a, b), k, quadratureOrder=orders, mask=mask)
    grid = A.Grid(cx, cy, a, b)
    gen = A.generate(grid, k, orders)
    return p, grid, gen


def test_jacobi_matches_diagonal(rt):
    p, grid, gen = _setup(rt, 7, 5, 0.1, 0.15, (4, 4))
    Z = S.dense_matrix(gen, grid)
    x = np.random.default_rng(1).standard_normal(grid.count) + 1j
    assert rel(rt.Preconditioner(p, 'jacobi').apply(x), x / np.diag(Z)) < 1e-13


@pytest.mark.parametrize('size', [(7, 5, 0.1, 0.15), (12, 12, 0.1, 0.1), (2, 6, 0.1, 0.1), (9, 1, 0.1, 0.02)])
def test_near_field_matches_oracle(rt, size):
    print('near field', size, flush=True)
    p, grid, gen = _setup(rt, *size, (4, 4))
    P = NF.NearField(gen, grid)
    x = np.random.default_rng(2).standard_normal(grid.count) + 1j * np.random.default_rng(3).standard_normal(grid.count)
    pc = rt.Preconditioner(p, 'nearField', precision='mixed')
    assert rel(pc.apply(x), P.apply(x)) < 1e-11
    assert rel(pc.apply(x, precision='c64'), P.apply(x)) < 1e-5


def test_near_field_masked(rt):
    cx, cy = 10, 9
    mask = np.ones((cx, cy), bool)
    mask[4:7, 3:6] = False
    p, grid, gen = _setup(rt, cx, cy, 0.1, 0.1, (4, 4), mask=mask)
    mx = mask[:-1] & mask[1:]
    my = mask[:, :-1] & mask[:, 1:]
    P = NF.NearField(gen, grid, mx, my)
    x = np.random.default_rng(4).standard_normal(grid.count) + 0.5j
    assert rel(rt.Preconditioner(p, 'nearField').apply(x), P.apply(x)) < 1e-11


@pytest.mark.parametrize('cx,cy,seed,fill', [(40, 36, 5, 0.3), (33, 47, 6, 0.6), (64, 20, 7, 0.9), (1, 30, 8, 0.8)])
def test_near_field_on_blob_masks(rt, cx, cy, seed, fill):
    """Irregular masks (random 4-connected blobs, as the served requests carry): many stencil classes at the blob's
    boundary, formed on the device, against the oracle's per-rooftop fits."""
    print('near field blob', cx, cy, seed, fill, flush=True)
    mask = rt.blob_mask(cx, cy, seed, fill)
    p, grid, gen = _setup(rt, cx, cy, 0.1, 0.1, (4, 4), mask=mask)
    mx = mask[:-1] & mask[1:]
    my = mask[:, :-1] & mask[:, 1:]
    P = NF.NearField(gen, grid, mx, my)
    x = np.random.default_rng(seed).standard_normal(grid.count) + 1j * np.random.default_rng(seed + 1).standard_normal(
        grid.count)
    pc = rt.Preconditioner(p, 'nearField', precision='mixed')
    assert rel(pc.apply(x), P.apply(x)) < 1e-11
    assert rel(pc.apply(x, precision='c64'), P.apply(x)) < 1e-5


@pytest.mark.parametrize('method', ['cg', 'gmres'])
def test_plate_solve_matches_oracle(rt, method):
    """SolverTest.m:9-21 (8x8 plate, plane wave, 1e-10) plus the iteration record against the oracle's loop."""
    p, grid, gen = _setup(rt, 8, 8, 0.05, 0.05, (3, 3))
    exc = rt.Excitation.planeWave(p)
    b = X.plane_wave(grid, 2 * np.pi, [0, 0, -1], [1, 0, 0], 1.0, order=3)
    assert rel(exc.vector(), b) < 1e-13
    res = rt.Solver(method, tolerance=1e-10).solve(rt.Operator(p), exc)
    Z = S.dense_matrix(gen, grid)
    assert res.converged() and rel(Z @ res.solution(), b) < 1e-10
    ref = SV.solve(method, Z, b, tol=1e-10, max_iter=1000, restart=50)
    # The unpreconditioned EFIE of this plate is ill-conditioned (COCG most of all): product differences of 1e-15
    # grow along the iteration, so the monitored norms are compared while they are large, and the iteration counts
    # near 1e-10 may differ by a step or two. The preconditioned solve below matches the oracle's count exactly.
    k = min(res.residualNorms().size, ref['residualNorms'].size)
    keep = ref['residualNorms'][:k] > 1e-3
    np.testing.assert_allclose(res.residualNorms()[:k][keep], ref['residualNorms'][:k][keep], rtol=1e-6)
    assert abs(res.iterationCount() - ref['iterationCount']) <= max(2, ref['iterationCount'] // 10)
    assert res.residualNorm() < 1e-10


@pytest.mark.parametrize('method', ['cg', 'gmres'])
def test_preconditioned_solve_matches_oracle(rt, method):
    p, grid, gen = _setup(rt, 16, 14, 0.1, 0.1, (6, 4))
    Z = S.dense_matrix(gen, grid)
    P = NF.NearField(gen, grid)
    b = X.plane_wave(grid, 2 * np.pi, [0.3, 0.4, -np.sqrt(0.75)], [0.8, -0.6, 0], 1.0, order=4)
    res = rt.Solver(method, tolerance=1e-8, restart=20).solve(rt.Operator(p), b,
                                                            preconditioner=rt.Preconditioner(p, 'nearField'))
    ref = SV.solve(method, Z, b, tol=1e-8, max_iter=1000, restart=20, preconditioner=P.apply)
    assert res.converged() and res.iterationCount() == ref['iterationCount']
    assert rel(res.solution(), ref['solution']) < 1e-6
    plain = rt.Solver(method, tolerance=1e-8, restart=20).solve(rt.Operator(p), b)
    assert res.iterationCount() < plain.iterationCount()


def test_gmres_long_cycle_matches_oracle(rt):
    """GMRES(60) with the near field to 1e-10 on the 16 x 14 plate converges in one cycle of 49 steps: the steps past
    33, whose finalize reads the second half of the Hessenberg entries for H s, and every rotation of a long cycle,
    against the oracle's CGS2 loop (iterate_.m:38-71): the same iteration count, the monitored norms and the solution."""
    p, grid, gen = _setup(rt, 16, 14, 0.1, 0.1, (6, 4))
    Z = S.dense_matrix(gen, grid)
    P = NF.NearField(gen, grid)
    b = X.plane_wave(grid, 2 * np.pi, [0.3, 0.4, -np.sqrt(0.75)], [0.8, -0.6, 0], 1.0, order=4)
    res = rt.Solver('gmres', tolerance=1e-10, restart=60).solve(rt.Operator(p), b,
                                                              preconditioner=rt.Preconditioner(p, 'nearField'))
    ref = SV.solve('gmres', Z, b, tol=1e-10, max_iter=1000, restart=60, preconditioner=P.apply)
    assert ref['iterationCount'] > 40   # one cycle well past step 33
    assert res.converged() and res.iterationCount() == ref['iterationCount']
    k = ref['residualNorms'].size
    assert res.residualNorms().size == k
    np.testing.assert_allclose(res.residualNorms(), ref['residualNorms'], rtol=1e-6)
    assert rel(res.solution(), ref['solution']) < 1e-8


@pytest.mark.parametrize('method', ['cg', 'gmres'])
def test_mixed_precision_refinement_reaches_fp64_accuracy(rt, method):
    p, grid, gen = _setup(rt, 24, 20, 0.1, 0.1, (8, 4))
    Z = S.dense_matrix(gen, grid)
    b = X.plane_wave(grid, 2 * np.pi, [0, 0, -1], [1, 0, 0], 1.0, order=4)
    op = rt.Operator(p, precision='mixed')
    pc = rt.Preconditioner(p, 'nearField', precision='mixed')
    mixed = rt.Solver(method, tolerance=1e-10, precision='c64').solve(op, b, preconditioner=pc)
    assert mixed.converged() and mixed.outerCount >= 1
    assert mixed.residualNorm() < 1e-10
    assert rel(Z @ mixed.solution(), b) < 1e-10
    fp64 = rt.Solver(method, tolerance=1e-10, refine=True).solve(op, b, preconditioner=pc)
    assert rel(mixed.solution(), fp64.solution()) < 1e-8


def test_dipole_input_impedance(rt):
    """Config 1: 64 x 4 strip dipole, lambda/64 at 1 GHz, middle port, COCG + Jacobi to 1e-10 vs dense solve."""
    k = 2 * np.pi * 1e9 / 299792458.0
    a = 2 * np.pi / k / 64
    p, grid, gen = _setup(rt, 64, 4, a, a, (16, 8), k=k)
    exc = rt.Excitation.deltaGap(p)
    res = rt.Solver('cg', tolerance=1e-10, maxIterationCount=5000).solve(
        rt.Operator(p), exc, preconditioner=rt.Preconditioner(p, 'jacobi'))
    Z = S.dense_matrix(gen, grid)
    port = X.middle_port(grid, 'x')
    b = X.delta_gap(grid, port, 1.0)
    assert rel(exc.vector(), b) < 1e-15
    current = np.linalg.solve(Z, b)
    assert res.converged() and rel(res.solution(), current) < 1e-8
    zin = res.inputImpedance()
    ref = X.input_impedance(grid, current, port, 1.0)
    assert abs(zin - ref) < 1e-8 * abs(ref)


def test_solver_errors(rt):
    p, grid, _ = _setup(rt, 5, 4, 0.1, 0.15, (3, 3))
    op = rt.Operator(p)
    with pytest.raises(rt.RooftopError) as e:
        rt.Solver().solve(op, np.ones(4))
    assert e.value.identifier == 'rooftopMoM:solver:sizeMismatch'
    with pytest.raises(rt.RooftopError) as e:
        rt.Solver().solve(op, np.r_[np.ones(30), np.nan])
    assert e.value.identifier == 'rooftopMoM:solver:invalidRightHandSide'
    zero = rt.Solver().solve(op, np.zeros(grid.count))
    assert zero.converged() and zero.iterationCount() == 0 and np.all(zero.solution() == 0)
    limited = rt.Solver('gmres', tolerance=1e-14, maxIterationCount=7, restart=5).solve(op, np.ones(grid.count))
    assert not limited.converged() and limited.iterationCount() == 7


@pytest.mark.parametrize('method', ['cg', 'gmres'])
def test_initial_guess(rt, method):
    """IterativeMethodTest.m:44-50: an exact initial guess needs no iteration; a rough one still converges."""
    p, grid, gen = _setup(rt, 10, 9, 0.1, 0.1, (4, 4))
    Z = S.dense_matrix(gen, grid)
    b = X.plane_wave(grid, 2 * np.pi, [0, 0, -1], [0, 1, 0], 1.0, order=4)
    exact = np.linalg.solve(Z, b)
    op = rt.Operator(p)
    pc = rt.Preconditioner(p, 'nearField')
    res = rt.Solver(method, tolerance=1e-10).solve(op, b, preconditioner=pc, initialGuess=exact)
    assert res.converged() and res.iterationCount() == 0 and rel(res.solution(), exact) < 1e-12
    rough = exact * (1 + 0.01 * np.random.default_rng(5).standard_normal(grid.count))
    res = rt.Solver(method, tolerance=1e-10).solve(op, b, preconditioner=pc, initialGuess=rough)
    cold = rt.Solver(method, tolerance=1e-10).solve(op, b, preconditioner=pc)
    assert res.converged() and rel(res.solution(), exact) < 1e-8
    assert res.iterationCount() <= cold.iterationCount()


@pytest.mark.parametrize('method,precision', [('cg', 'c128'), ('gmres', 'c128'), ('cg', 'c64'), ('gmres', 'c64')])
def test_batched_solve_matches_single_solves(rt, method, precision):
    """R right-hand sides solved together give, column by column, the results of solving each alone (iteration
    counts, monitored norms, solutions), including a zero column and an exact initial guess."""
    p, grid, gen = _setup(rt, 16, 14, 0.1, 0.1, (6, 4))
    rng = np.random.default_rng(11)
    cols = []
    for q in range(5):
        d = rng.standard_normal(3)
        d[2] = -abs(d[2]) - 0.5
        d /= np.linalg.norm(d)
        e = np.cross(d, [0.0, 1.0, 0.0])
        cols.append(X.plane_wave(grid, 2 * np.pi, d, e / np.linalg.norm(e), 1.0, order=4))
    cols.insert(2, np.zeros(grid.count, complex))
    B = np.stack(cols, axis=1)
    op = rt.Operator(p, precision='mixed')
    pc = rt.Preconditioner(p, 'nearField', precision='mixed')
    solver = rt.Solver(method, tolerance=1e-9, restart=12, precision=precision)
    Z = S.dense_matrix(gen, grid)
    exact = np.linalg.solve(Z, B[:, 4])
    X0 = np.zeros_like(B)
    X0[:, 4] = exact
    batch = solver.solve(op, B, preconditioner=pc, initialGuess=X0)
    assert len(batch) == B.shape[1]
    for q in range(B.shape[1]):
        print('column', q, flush=True)
        alone = solver.solve(op, B[:, q], preconditioner=pc, initialGuess=X0[:, q])
        got = batch[q]
        assert got.converged() == alone.converged()
        assert got.iterationCount() == alone.iterationCount()
        np.testing.assert_allclose(got.residualNorms(), alone.residualNorms(), rtol=1e-9, atol=1e-300)
        assert rel(got.solution(), alone.solution()) < 1e-12
        if q == 2:
            assert got.iterationCount() == 0 and np.all(got.solution() == 0)
        elif q == 4:
            assert got.iterationCount() == 0
        else:
            assert got.converged() and rel(Z @ got.solution(), B[:, q]) < 1e-9



def test_row_block_gram_schmidt_matches_single_solves(rt):
    """Enough systems at once (a 128 x 128 plate, 17 right-hand sides) that the GMRES dot and update passes run on
    row blocks give, column by column, the results of the single solves, whose dot pass runs on strided 64-row tiles
    and whose update pass takes one tile per block: the same iteration counts, monitored norms and solutions up to
    the order of the sums (FP64 basis), and the same solutions to the mixed form's accuracy (complex64 basis)."""
    cx = cy = 128
    p = rt.Problem(rt.Mesh(cx, cy, 0.1, 0.1), 2 * np.pi, quadratureOrder=(6, 4))
    grid = A.Grid(cx, cy, 0.1, 0.1)
    rng = np.random.default_rng(17)
    cols = []
    for q in range(17):
        d = rng.standard_normal(3)
        d[2] = -abs(d[2]) - 0.5
        d /= np.linalg.norm(d)
        e = np.cross(d, [0.0, 1.0, 0.0])
        cols.append(X.plane_wave(grid, 2 * np.pi, d, e / np.linalg.norm(e), 1.0, order=4))
    B = np.stack(cols, axis=1)
    op = rt.Operator(p, precision='mixed')
    pc = rt.Preconditioner(p, 'nearField', precision='mixed')
    for precision in ('c128', 'c64'):
        print('precision', precision, flush=True)
        solver = rt.Solver('gmres', tolerance=1e-8, restart=30, precision=precision)
        batch = solver.solve(op, B, preconditioner=pc)
        for q in (0, 7, 16):
            alone = solver.solve(op, B[:, q], preconditioner=pc)
            got = batch[q]
            assert got.converged() and alone.converged()
            if precision == 'c128':
                assert got.iterationCount() == alone.iterationCount()
                np.testing.assert_allclose(got.residualNorms(), alone.residualNorms(), rtol=1e-9)
                assert rel(got.solution(), alone.solution()) < 1e-10
            else:
                assert abs(got.iterationCount() - alone.iterationCount()) <= 1
                assert rel(got.solution(), alone.solution()) < 1e-6


@pytest.mark.parametrize('precision', ['c128', 'c64'])
def test_batched_solve_with_chunked_products(rt, precision):
    """A batch whose matvec runs in right-hand-side chunks on two streams gives bitwise the results of the unchunked
    batch (each column's arithmetic is the same), including a zero column that the loop skips."""
    p, grid, gen = _setup(rt, 16, 14, 0.1, 0.1, (6, 4))
    rng = np.random.default_rng(12)
    B = rng.standard_normal((grid.count, 5)) + 1j * rng.standard_normal((grid.count, 5))
    B[:, 1] = 0
    op = rt.Operator(p, precision='mixed')
    pc = rt.Preconditioner(p, 'nearField', precision='mixed')
    solver = rt.Solver('gmres', tolerance=1e-9, restart=12, precision=precision)
    prec = 'c64' if precision == 'c64' else 'c128'
    base = {'bits': 3, 'group': 2, 'tile': 1, 'teams': 1, 'chunks': 1}
    op.setShape(base, nrhs=5, precision=prec)
    one = solver.solve(op, B, preconditioner=pc)
    for chunks in (2, 3, 5):
        print('chunks', chunks, flush=True)
        op.setShape(dict(base, chunks=chunks), nrhs=5, precision=prec)
        got = solver.solve(op, B, preconditioner=pc)
        for q in range(5):
            assert got[q].iterationCount() == one[q].iterationCount()
            assert np.array_equal(got[q].solution(), one[q].solution())
        x = rng.standard_normal((grid.count, 5)) + 0j
        op.setShape(base, nrhs=5, precision=prec)
        y1 = op.apply(x, precision=prec)
        op.setShape(dict(base, chunks=chunks), nrhs=5, precision=prec)
        assert np.array_equal(op.apply(x, precision=prec), y1)

def test_device_plane_waves_and_observables(rt):
    """Plane waves of many incidences built on the device equal the oracle's (generatePlaneWave.m), masked and not;
    RCS and input impedance computed on the device from a device-resident batch solve equal the host formulas."""
    rng = np.random.default_rng(21)
    cases = []
    for q in range(6):
        th, ph = rng.uniform(0, np.radians(85)), rng.uniform(0, 2 * np.pi)
        d = -np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
        te = np.array([-np.sin(ph), np.cos(ph), 0.0])
        cases.append((d, te * (1.5 if q % 2 else 1.0)))   # unnormalised polarization on odd columns
    D = np.array([c[0] for c in cases])
    Pz = np.array([c[1] for c in cases])
    E = np.linspace(0.5, 2.0, len(cases))
    for size, mask in (((13, 9, 0.1, 0.12), None), ((10, 8, 0.1, 0.1), 'blob')):
        print('plane waves', size, mask, flush=True)
        m = rt.blob_mask(size[0], size[1], 5, 0.6) if mask else None
        p, grid, gen = _setup(rt, *size, (6, 5), mask=m)
        batch = rt.plane_waves(p, D, Pz, E)
        got = batch.to_host()
        exist = rt.rooftop_existence(p)
        for r, (d, e) in enumerate(cases):
            ref = X.plane_wave(grid, 2 * np.pi, d, e / np.linalg.norm(e), E[r], order=5) * exist
            assert rel(got[:, r], ref) < 1e-13, (size, r, rel(got[:, r], ref))
            host = rt.Excitation.planeWave(p, d, e, E[r]).vector()
            assert rel(got[:, r], host) < 1e-13
        op = rt.Operator(p)
        pc = rt.Preconditioner(p, 'nearField')
        solver = rt.Solver('gmres', tolerance=1e-10, restart=30)
        xb, res = solver.solveDevice(op, batch, preconditioner=pc)
        assert all(q.converged() for q in res)
        sol = xb.to_host()
        sigma = rt.radar_cross_sections(p, batch, xb, E)
        for r in range(len(cases)):
            ref = X.radar_cross_section(2 * np.pi, got[:, r], sol[:, r], E[r])
            assert abs(sigma[r] - ref) < 1e-12 * abs(ref), (r, sigma[r], ref)
        if mask is None:
            port = X.middle_port(grid, 'x')
            V = np.array([1.0, 2.0, 0.5, 1.0, 1.0, 3.0])
            zin = rt.input_impedances(p, port, xb, V)
            for r in range(len(cases)):
                ref = X.input_impedance(grid, sol[:, r], port, V[r])
                assert abs(zin[r] - ref) < 1e-12 * abs(ref)
    p, _, _ = _setup(rt, 5, 4, 0.1, 0.1, (3, 3))
    with pytest.raises(rt.RooftopError) as e:
        rt.plane_waves(p, [[0, 0, -1]], [[0, 0, 1]])
    assert e.value.identifier == 'rooftopMoM:excitation:notPerpendicular'
    with pytest.raises(rt.RooftopError) as e:
        rt.plane_waves(p, [[0, 0, 0]], [[1, 0, 0]])
    assert e.value.identifier == 'rooftopMoM:excitation:zeroVector'
    xb = rt.DeviceBatch.from_host(p.context, np.ones((p.rooftopCount(), 2)))
    with pytest.raises(rt.RooftopError) as e:
        rt.input_impedances(p, [0, 0], xb)
    assert e.value.identifier == 'rooftopMoM:excitation:invalidPort'


@pytest.mark.parametrize('size', [(8, 8, 0.1, 0.1), (16, 4, 0.1, 0.12), (4, 16, 0.05, 0.1), (32, 32, 0.1, 0.1),
                                  (8, 1, 0.1, 0.02), (1, 8, 0.03, 0.1), (5, 4, 0.1, 0.15), (12, 9, 0.1, 0.1),
                                  (7, 1, 0.1, 0.02), (1, 6, 0.03, 0.1), (24, 16, 0.1, 0.1)])
def test_circulant_preconditioner_matches_oracle(rt, size):
    """rooftopMoM.Preconditioner(problem, 'kernel', k, 'builder', b) on the GPU (Preconditioner.m:7-70) against the
    oracle's circulants (pinned to reference.mat, the superoptimal definition and the dense transpose): every kernel,
    the three builders, both precisions, on face grids of power-of-two counts (the passes on the face grid) and of
    other counts (the inverse re-embedded as a Toeplitz product)."""
    print('circulant', size, flush=True)
    from oracle import circulant as CI
    p, grid, gen = _setup(rt, *size, (4, 4))
    g = p.generators()   # the oracle gets the GPU's generators: this test is about the preconditioner
    x = np.random.default_rng(7).standard_normal((grid.count, 2)) + 1j * np.random.default_rng(8).standard_normal(
        (grid.count, 2))
    for kernel, bw, builder in (('strang', None, 'partwise'), ('tChan', None, 'partwise'), ('rChan', None, 'partwise'),
                                ('huckle', (3, 5), 'partwise'), ('huckle', None, 'partwise'),
                                ('rChan', None, 'hybrid'), ('strang', None, 'hybrid'), ('huckle', (2, 4), 'hybrid'),
                                ('superoptimal', None, 'partwise'), ('superoptimal', None, 'hybrid'),
                                ('rChan', None, 'blockwise'), ('strang', None, 'blockwise'), ('tChan', None, 'blockwise'),
                                ('huckle', (2, 4), 'blockwise'), ('superoptimal', None, 'blockwise')):
        pc = rt.Preconditioner(p, kernel=kernel, bandwidth=bw, builder=builder, precision='mixed')
        inv = CI.circulant_preconditioner(g, grid, kernel, bw, builder)
        ref = np.stack([CI.apply(inv, grid, x[:, r]) for r in range(2)], axis=1)
        assert rel(pc.apply(x), ref) < 1e-12, (kernel, bw, builder)
        assert rel(pc.apply(x, precision='c64'), ref) < 2e-5, (kernel, bw, builder)
    imag = {k: None if v is None else v.imag + 0j for k, v in g.items()}
    pc = rt.Preconditioner(p, kernel='tChan', modifier='imag')
    assert rel(pc.apply(x[:, 0]), CI.apply(CI.circulant_preconditioner(imag, grid, 'tChan'), grid, x[:, 0])) < 1e-12


def test_circulant_preconditioner_errors_and_solve(rt):
    """A singular approximation raises rooftopMoM:preconditioner:singular (Preconditioner.m:57-65), with every
    builder; preconditioned COCG and GMRES converge in fewer
    iterations (PreconditionerTest.m:44-55, the partwise and the blockwise builders) and GMRES follows the oracle's
    iterations with the same preconditioner."""
    from oracle import circulant as CI
    p, grid, gen = _setup(rt, 16, 16, 0.1, 0.1, (4, 4))
    with pytest.raises(rt.RooftopError) as e:
        rt.Preconditioner(p, kernel='huckle', bandwidth=(2,))
    assert e.value.identifier == 'rooftopMoM:preconditioner:invalidBandwidth'
    z, zgrid, _ = _setup(rt, 8, 4, 0.1, 0.1, (3, 3))
    zeros = z.generators()
    z.setGenerators(*[None if zeros[k] is None else np.zeros_like(zeros[k])
                      for k in ('magneticXX', 'magneticYY', 'electric')])
    for builder in ('partwise', 'blockwise', 'hybrid'):
        with pytest.raises(rt.RooftopError) as e:
            rt.Preconditioner(z, kernel='rChan', builder=builder)
        assert e.value.identifier == 'rooftopMoM:preconditioner:singular', builder
    Z = S.dense_matrix(gen, grid)
    b = X.plane_wave(grid, 2 * np.pi, [0, 0, -1], [1, 0, 0], 1.0, order=4)
    op = rt.Operator(p)
    pc = rt.Preconditioner(p)   # without a kernel this library's default, the near field
    circ = rt.Preconditioner(p, kernel='rChan')
    inv = CI.circulant_preconditioner(p.generators(), grid, 'rChan')
    block = rt.Preconditioner(p, kernel='rChan', builder='blockwise')
    for method in ('cg', 'gmres'):
        res = rt.Solver(method, tolerance=1e-8, restart=30).solve(op, b, preconditioner=block)
        assert res.converged() and rel(Z @ res.solution(), b) < 1e-8
        plain = rt.Solver(method, tolerance=1e-8, restart=30).solve(op, b)
        res = rt.Solver(method, tolerance=1e-8, restart=30).solve(op, b, preconditioner=circ)
        assert res.converged() and res.iterationCount() < plain.iterationCount()
        assert rel(Z @ res.solution(), b) < 1e-8
        if method == 'gmres':
            ref = SV.solve(method, Z, b, tol=1e-8, max_iter=1000, restart=30,
                           preconditioner=lambda v: CI.apply(inv, grid, v))
            # The monitored norms agree; the last step may land on either side of 1e-8 (a flat tail).
            k = min(res.residualNorms().size, ref['residualNorms'].size)
            keep = ref['residualNorms'][:k] > 1e-3
            np.testing.assert_allclose(res.residualNorms()[:k][keep], ref['residualNorms'][:k][keep], rtol=1e-6)
            assert abs(res.iterationCount() - ref['iterationCount']) <= 1
    assert pc.kind == 'nearField' and circ.kind == 'circulant'


@pytest.mark.parametrize('restart', [100, np.inf])
def test_gmres_long_restarts_match_oracle(rt, restart):
    """Solver('gmres', 'restart', 100 or Inf) (Solver.m:31; Inf is min(N, maxIterationCount),
    GeneralizedMinimumResidual.m:15-23, iterate_.m:7): cycles longer than 62 steps run the CGS2 kernels. Without a
    preconditioner (208 and 134 iterations) the monitored norms follow the oracle's while they are large and the
    counts agree to rounding; with the near field the count is the oracle's exactly; a batch of three equals its
    single solves."""
    p, grid, gen = _setup(rt, 16, 14, 0.1, 0.1, (6, 4))
    Z = S.dense_matrix(gen, grid)
    P = NF.NearField(gen, grid)
    b = X.plane_wave(grid, 2 * np.pi, [0.3, 0.4, -np.sqrt(0.75)], [0.8, -0.6, 0], 1.0, order=4)
    op = rt.Operator(p)
    pc = rt.Preconditioner(p, 'nearField')
    solver = rt.Solver('gmres', tolerance=1e-8, restart=restart)
    print('plain, restart', restart, flush=True)
    res = solver.solve(op, b)
    ref = SV.solve('gmres', Z, b, tol=1e-8, max_iter=1000, restart=restart)
    assert res.converged() and rel(Z @ res.solution(), b) < 1e-8
    k = min(res.residualNorms().size, ref['residualNorms'].size)
    keep = ref['residualNorms'][:k] > 1e-3
    np.testing.assert_allclose(res.residualNorms()[:k][keep], ref['residualNorms'][:k][keep], rtol=1e-6)
    assert abs(res.iterationCount() - ref['iterationCount']) <= max(2, ref['iterationCount'] // 20)
    print('near field, restart', restart, flush=True)
    res = solver.solve(op, b, preconditioner=pc)
    ref = SV.solve('gmres', Z, b, tol=1e-8, max_iter=1000, restart=restart, preconditioner=P.apply)
    assert res.converged() and res.iterationCount() == ref['iterationCount']
    np.testing.assert_allclose(res.residualNorms(), ref['residualNorms'], rtol=1e-7, atol=1e-300)
    assert rel(res.solution(), ref['solution']) < 1e-6
    B = np.stack([b, 1j * b[::-1].copy(), np.zeros_like(b)], axis=1)
    B[:5, 2] = 1.0
    batch = solver.solve(op, B, preconditioner=pc)
    for q in range(3):
        alone = solver.solve(op, B[:, q], preconditioner=pc)
        assert batch[q].converged() and batch[q].iterationCount() == alone.iterationCount()
        assert rel(batch[q].solution(), alone.solution()) < 1e-12


def test_mixed_gmres_with_complex128_basis(rt):
    """precision='c64' with basis='c128' (RT_C64_FP64_BASIS): complex64 products, the Gram-Schmidt basis in FP64;
    it reaches the tolerance by the dense matrix and agrees with the complex64-basis and the FP64 solves."""
    p, grid, gen = _setup(rt, 24, 20, 0.1, 0.1, (8, 4))
    Z = S.dense_matrix(gen, grid)
    b = X.plane_wave(grid, 2 * np.pi, [0, 0, -1], [1, 0, 0], 1.0, order=4)
    op = rt.Operator(p, precision='mixed')
    pc = rt.Preconditioner(p, 'nearField', precision='mixed')
    wide = rt.Solver('gmres', tolerance=1e-10, precision='c64', basis='c128').solve(op, b, preconditioner=pc)
    narrow = rt.Solver('gmres', tolerance=1e-10, precision='c64').solve(op, b, preconditioner=pc)
    fp64 = rt.Solver('gmres', tolerance=1e-10, refine=True).solve(op, b, preconditioner=pc)
    assert wide.converged() and rel(Z @ wide.solution(), b) < 1e-10
    assert rel(wide.solution(), fp64.solution()) < 1e-8 and rel(narrow.solution(), fp64.solution()) < 1e-8
