"""GPU tests of the full-size workloads: the 256 x 256 plate solve (GMRES(50) + near field, mixed and FP64), the
512 x 512 plate under 64 plane waves solved together, the served stream at grid sides up to 1024 and the
4096 x 4096 plate (33.5 M unknowns). Results are checked against the oracle's split-FFT product on the GPU's own
generators (whose assembly is pinned separately), against brute-force integrals of the generator definition, and
between the batched, served and single forms. SolverTest.m:9-21 is the pattern: a solve checked by its residual
against an independent product."""
import numpy as np
import pytest

from conftest import rel
from oracle import assembly as A
from oracle import excitation as X
from oracle import nearfield as NF
from oracle import spectra as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def rt():
    import __graft_entry__ as ge
    return ge.load()


def _oracle_residual(gen, grid, x, b, exists=None):
    """norm(b - Z x) / norm(b) with the oracle's split-FFT product (FP64 on the host); with a mask, Z is restricted
    to the existing rooftops (exists: 1 or 0 per rooftop)."""
    ax = S.split_fft_product(S.subband_spectra(gen, grid), grid, x)
    if exists is not None:
        ax = exists * ax
    return rel(ax, b)


def test_plate_256_gmres_near_field(rt):
    """The 256 x 256 plate of lambda/10 under normal incidence, GMRES(50) + near field to 1e-6: the FP64 and the
    mixed (complex64 products in FP64 refinement) forms reach the tolerance by the oracle's product, and agree."""
    n = 256
    print('plate 256 solve', flush=True)
    p = rt.Problem(rt.Mesh(n, n, 0.1, 0.1), 2 * np.pi, quadratureOrder=(16, 8))
    grid = A.Grid(n, n, 0.1, 0.1)
    gen = A.generate(grid, 2 * np.pi, (16, 8))
    gpu = p.generators()
    for key in ('magneticXX', 'magneticYY', 'electric'):
        assert rel(gpu[key], gen[key]) < 1e-13
    op = rt.Operator(p, precision='mixed')
    pc = rt.Preconditioner(p, 'nearField', precision='mixed')
    exc = rt.Excitation.planeWave(p)
    b = X.plane_wave(grid, 2 * np.pi, [0, 0, -1], [1, 0, 0], 1.0, order=8)
    assert rel(exc.vector(), b) < 1e-13
    fp64 = rt.Solver('gmres', tolerance=1e-6, restart=50, refine=True).solve(op, exc, preconditioner=pc)
    mixed = rt.Solver('gmres', tolerance=1e-6, restart=50, precision='c64').solve(op, exc, preconditioner=pc)
    for res in (fp64, mixed):
        print('iterations %d, outer %d, residual %.3e' % (res.iterationCount(), res.outerCount, res.residualNorm()),
              flush=True)
        assert res.converged() and res.residualNorm() < 1e-6
        assert _oracle_residual(gen, grid, res.solution(), b) < 1.01e-6
    assert rel(mixed.solution(), fp64.solution()) < 1e-5


def test_near_field_batched_apply_on_face_tiles(rt):
    """The near field on the 256 x 256 plate with 8 right-hand sides (2048 blocks: a thread per face on 32 x 8 tiles)
    equals, column by column, its single applies (a thread per rooftop on 16 x 8 tiles), and the oracle's
    near field, in both precisions."""
    n = 256
    print('near field 256 x 8', flush=True)
    p = rt.Problem(rt.Mesh(n, n, 0.1, 0.1), 2 * np.pi, quadratureOrder=(16, 8))
    grid = A.Grid(n, n, 0.1, 0.1)
    pc = rt.Preconditioner(p, 'nearField', precision='mixed')
    rng = np.random.default_rng(31)
    x = rng.standard_normal((grid.count, 8)) + 1j * rng.standard_normal((grid.count, 8))
    P = NF.NearField(p.generators(), grid, by_class=True)
    for prec, tol in (('c128', 1e-11), ('c64', 1e-5)):
        batch = pc.apply(x, precision=prec)
        for r in range(8):
            single = pc.apply(x[:, r], precision=prec)
            assert rel(batch[:, r], single) < (1e-14 if prec == 'c128' else 1e-6), (prec, r)
        assert rel(batch[:, 3], P.apply(x[:, 3])) < tol, prec


def _plane_wave_angles(count, seed):
    import bench
    return bench.plane_wave_angles(count, seed)


@pytest.mark.parametrize('method', ['gmres', 'cg'])
def test_plate_512_with_64_plane_waves(rt, method):
    """The 512 x 512 plate under 64 plane waves (theta U[0, 80], phi U[0, 360), TE/TM, seed 1) built on the device and
    solved together, mixed: every system reaches 1e-6 by the library's FP64 residual, three of them by the oracle's
    product, and those three match their single solves."""
    n = 512
    print('plate 512 x 64', method, flush=True)
    p = rt.Problem(rt.Mesh(n, n, 0.1, 0.1), 2 * np.pi, quadratureOrder=(16, 8))
    grid = A.Grid(n, n, 0.1, 0.1)
    op = rt.Operator(p, precision='mixed')
    pc = rt.Preconditioner(p, 'nearField', precision='mixed')
    dirs, pols = _plane_wave_angles(64, 1)
    B = rt.plane_waves(p, dirs, pols)
    solver = rt.Solver(method, tolerance=1e-6, restart=50, maxIterationCount=2000, precision='c64')
    xb, results = solver.solveDevice(op, B, preconditioner=pc)
    its = [r.iterationCount() for r in results]
    print('converged %d of 64, iterations median %d max %d' % (sum(r.converged() for r in results),
                                                                int(np.median(its)), max(its)), flush=True)
    assert all(r.converged() and r.residualNorm() < 1e-6 for r in results)
    bh = B.to_host()
    xh = xb.to_host()
    gen = p.generators()
    for q in (0, 31, 63):
        b = X.plane_wave(grid, 2 * np.pi, dirs[q], pols[q], 1.0, order=8)
        assert rel(bh[:, q], b) < 1e-13
        assert _oracle_residual(gen, grid, xh[:, q], b) < 1.01e-6
        alone = solver.solve(op, b, preconditioner=pc)
        assert alone.converged()
        assert abs(alone.iterationCount() - results[q].iterationCount()) <= max(2, its[q] // 20)
        assert rel(xh[:, q], alone.solution()) < 1e-4
    B.free()
    xb.free()


def _served(rt, index, side):
    import bench
    return bench.served_request(rt, index, side=side)


def test_served_requests_up_to_side_1024(rt):
    """Served-stream requests (blob masks, random frequencies and feeds) of sides 100, 300, 700 and 1024, GMRES(50)
    + near field to 1e-6, mixed: solved together through the server they equal their solves alone; the masked product
    of the largest equals the oracle's restriction of the split-FFT product."""
    import bench
    qs = [_served(rt, i, side) for i, side in enumerate((100, 300, 700, 1024))]
    solver = rt.Solver('gmres', tolerance=1e-6, restart=50, maxIterationCount=3000, precision='c64')
    out, makespan, stats = bench.serve_stream(rt, qs, solver, slots=4, arrivals=[0.0] * len(qs))
    assert stats['completed'] == len(qs) and stats['peak_running'] >= 2
    for q, (x, res, times) in zip(qs, out):
        print('request L=%d fill %.2f: %d iterations, latency %.3f s' % (q['L'], q['fill'], res.iterationCount(),
                                                                          times['latency']), flush=True)
        assert res.converged() and res.residualNorm() < 1e-6
        p = rt.Problem(rt.Mesh(q['L'], q['L'], q['a'], q['a']), q['k'], mask=q['mask'])
        op = rt.Operator(p, precision='mixed')
        pc = rt.Preconditioner(p, 'nearField', precision='mixed')
        alone = solver.solve(op, q['rhs'], preconditioner=pc)
        assert res.iterationCount() == alone.iterationCount()
        assert rel(x, alone.solution()) < 1e-14
        if q['L'] == 1024:
            grid = A.Grid(q['L'], q['L'], q['a'], q['a'])
            m = q['mask']
            ex = np.concatenate([(m[:-1] & m[1:]).ravel(order='F'), (m[:, :-1] & m[:, 1:]).ravel(order='F')])
            v = np.random.default_rng(41).standard_normal(grid.count) * ex + 0j
            ref = ex * S.split_fft_product(S.subband_spectra(p.generators(), grid), grid, v)
            assert rel(op.apply(v), ref) < 1e-12
            assert rel(op.apply(v, precision='c64'), ref) < 5e-6
            assert _oracle_residual(p.generators(), grid, x, q['rhs'], ex) < 1.01e-6


@pytest.mark.parametrize('side', [32, 128, 512, 1024])
def test_request_bytes_cover_what_a_request_holds(rt, side):
    """The admission size of a request (rt_request_bytes) is at least the device bytes its solve holds at its peak,
    masked and unmasked, GMRES(50) and COCG, mixed and FP64."""
    for masked in (False, True):
        q = _served(rt, 5, side)
        mask = q['mask'] if masked else None
        rhs = q['rhs'] if masked else np.ones(2 * side * (side - 1), complex) * q['a']
        for solver in (rt.Solver('gmres', tolerance=1e-6, restart=50, maxIterationCount=50, precision='c64'),
                       rt.Solver('gmres', tolerance=1e-6, restart=50, maxIterationCount=50, refine=True),
                       rt.Solver('cg', tolerance=1e-6, maxIterationCount=50, precision='c64')):
            ctx = rt.Context()
            p = rt.Problem(rt.Mesh(side, side, q['a'], q['a']), q['k'], mask=mask, context=ctx)
            held = 'mixed' if solver.precision == 'c64' else 'c128'
            op = rt.Operator(p, precision=held)
            pc = rt.Preconditioner(p, 'nearField', precision=held)
            solver.solve(op, rhs, preconditioner=pc)
            peak = ctx.peakMemory()
            estimate = rt.Server.requestBytes(side, side, solver)
            print('side %d masked %s %s %s: peak %.1f MB, estimate %.1f MB' % (
                side, masked, solver.method, solver.precision, peak / 1e6, estimate / 1e6), flush=True)
            assert estimate >= peak
            del op, pc, p


def test_plate_4096_memory_ceiling(rt):
    """The 4096 x 4096 plate (33,546,240 rooftops, 8192-point transforms), mixed GMRES(30) + near field with FP64
    refinement to 1e-6: far generator entries against the brute-force definition, the solve's residual by the
    oracle's split-FFT product on the GPU's generators, and the bytes held against the device."""
    from test_oracle_pins import Z0, _a_by_definition, _phi_by_definition
    n, k = 4096, 2 * np.pi
    print('plate 4096: assembly', flush=True)
    ctx = rt.Context()
    p = rt.Problem(rt.Mesh(n, n, 0.1, 0.1), k, quadratureOrder=(16, 8), context=ctx)
    gen = p.generators()
    for d in ((4094, 4095), (100, 3000), (2, 4000), (1000, 7), (5, 5)):
        expected = 1j * Z0 * k / (4 * np.pi) * _a_by_definition('x', d, (0.1, 0.1), k)
        assert abs(gen['magneticXX'][d] - expected) < 1e-12 * abs(expected), d
        expected = 1j * Z0 * k / (4 * np.pi) * _a_by_definition('y', d[::-1], (0.1, 0.1), k)
        assert abs(gen['magneticYY'][d[::-1]] - expected) < 1e-12 * abs(expected), d
        expected = 1j * Z0 / (4 * np.pi * k) * _phi_by_definition(d, (0.1, 0.1), k)
        assert abs(gen['electric'][d] - expected) < 1e-12 * abs(expected), d
    op = rt.Operator(p, precision='mixed')
    pc = rt.Preconditioner(p, 'nearField', precision='mixed')
    exc = rt.Excitation.planeWave(p)
    print('plate 4096: solve', flush=True)
    res = rt.Solver('gmres', tolerance=1e-6, restart=30, precision='c64', maxIterationCount=3000).solve(
        op, exc, preconditioner=pc)
    held, device = ctx.memory()
    print('iterations %d, outer %d, residual %.3e, %.1f GB held, peak %.1f GB of %.1f GB' % (
        res.iterationCount(), res.outerCount, res.residualNorm(), held / 1e9, ctx.peakMemory() / 1e9, device / 1e9),
        flush=True)
    assert res.converged() and res.residualNorm() < 1e-6
    assert ctx.peakMemory() < 0.5 * device
    x = res.solution()
    b = exc.vector()
    del op, pc
    print('plate 4096: oracle residual', flush=True)
    grid = A.Grid(n, n, 0.1, 0.1)
    S.set_workers(16)
    try:
        assert _oracle_residual(gen, grid, x, b) < 1.01e-6
    finally:
        S.set_workers(1)
