"""GPU checks of the serving loop (rt_server_*): requests of different grids, frequencies and masks solved
concurrently give the results of solving each alone; admission follows the HBM budget."""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def rt():
    import __graft_entry__ as ge
    return ge.load()


def _requests(rt, count):
    import bench
    return [bench.served_request(rt, i, max_side=64) for i in range(count)]


def _alone(rt, q, solver):
    p = rt.Problem(rt.Mesh(q['L'], q['L'], q['a'], q['a']), q['k'], mask=q['mask'])
    precision = 'mixed' if solver.precision == 'c64' else 'c128'
    op = rt.Operator(p, precision=precision)
    pc = rt.Preconditioner(p, 'nearField', precision=precision)
    return solver.solve(op, q['rhs'], preconditioner=pc)


@pytest.mark.parametrize('precision', ['c128', 'c64'])
def test_requests_in_the_mix_equal_alone(rt, precision):
    import bench
    qs = _requests(rt, 8)
    solver = rt.Solver('cg', tolerance=1e-6, maxIterationCount=2000, precision=precision,
                       refine=precision == 'c64')
    out, makespan, stats = bench.serve_stream(rt, qs, solver, slots=4, arrivals=[0.0] * len(qs))
    assert stats['completed'] == len(qs) and stats['peak_running'] >= 2
    for q, (x, res, times) in zip(qs, out):
        print('request L=%d fill %.2f: %d iterations, latency %.3f s' % (q['L'], q['fill'], res.iterationCount(),
                                                                          times['latency']), flush=True)
        assert res.converged() and res.residualNorm() < 1e-6
        alone = _alone(rt, q, solver)
        assert rel(x, alone.solution()) < 1e-14
        assert res.iterationCount() == alone.iterationCount()


def test_budget_admission(rt):
    import bench
    qs = _requests(rt, 4)
    solver = rt.Solver('cg', tolerance=1e-6)
    # A budget that holds one request at a time: the slots never run two together.
    sizes = [rt.Server.requestBytes(q['L'], q['L'], solver) for q in qs]
    assert max(sizes) < 2 * min(sizes)
    small = rt.Server(slots=4, budget=max(sizes))
    tickets = [small.submit(q['L'], q['L'], q['a'], q['a'], q['k'], q['rhs'], mask=q['mask'], solver=solver)
               for q in qs]
    for q, t in zip(qs, tickets):
        x, res, _ = small.wait(t, q['n'])
        assert res.converged()
    assert small.stats()['peak_running'] == 1
    small.close()
    # A request that can never fit is refused at submission.
    tiny = rt.Server(slots=1, budget=1 << 20)
    with pytest.raises(rt.RooftopError) as e:
        tiny.submit(qs[0]['L'], qs[0]['L'], qs[0]['a'], qs[0]['a'], qs[0]['k'], qs[0]['rhs'], mask=qs[0]['mask'])
    assert e.value.identifier == 'rooftopMoM:resources:insufficientMemory'
    tiny.close()


def test_requests_of_one_embedding_size_are_solved_together(rt):
    """Batching by embedding size (rt_server_set_batching): six queued requests of sides 65 .. 128 (one FFT period,
    256) are solved as one group on the largest grid, each with its own mask, frequency, cell size and feed, and a
    request of another embedding alone; every result converges, equals its solve alone to the solve's accuracy, and
    has the oracle's residual on its own grid."""
    import bench
    from oracle import assembly as A
    from oracle import spectra as S
    sides = (70, 90, 128, 100, 120, 65, 40)
    qs = [bench.served_request(rt, 20 + i, side=s) for i, s in enumerate(sides)]
    solver = rt.Solver('gmres', tolerance=1e-6, restart=50, maxIterationCount=3000, precision='c64')
    server = rt.Server(slots=1, batch=8)
    # Submitted while the only slot is busy with a first request, so that the rest queue up together.
    first = server.submit(qs[-1]['L'], qs[-1]['L'], qs[-1]['a'], qs[-1]['a'], qs[-1]['k'], qs[-1]['rhs'],
                          mask=qs[-1]['mask'], solver=solver, preconditioner='nearField')
    tickets = [server.submit(q['L'], q['L'], q['a'], q['a'], q['k'], q['rhs'], mask=q['mask'], solver=solver,
                             preconditioner='nearField') for q in qs[:-1]]
    out = [server.wait(t, q['n']) for t, q in zip(tickets, qs[:-1])]
    server.wait(first, qs[-1]['n'])
    stats = server.stats()
    print('batched solves %d holding %d requests' % (stats['batched_solves'], stats['batched_requests']), flush=True)
    assert stats['batched_solves'] >= 1 and stats['batched_requests'] >= 4
    server.close()
    for q, (x, res, times) in zip(qs[:-1], out):
        p = rt.Problem(rt.Mesh(q['L'], q['L'], q['a'], q['a']), q['k'], mask=q['mask'])
        alone = solver.solve(rt.Operator(p, precision='mixed'), q['rhs'],
                             preconditioner=rt.Preconditioner(p, 'nearField', precision='mixed'))
        print('L=%d: %d iterations batched, %d alone; difference %.2e' % (
            q['L'], res.iterationCount(), alone.iterationCount(), rel(x, alone.solution())), flush=True)
        assert res.converged() and res.residualNorm() < 1e-6
        assert rel(x, alone.solution()) < 1e-6
        assert abs(res.iterationCount() - alone.iterationCount()) <= 2
        grid = A.Grid(q['L'], q['L'], q['a'], q['a'])
        m = q['mask']
        ex = np.concatenate([(m[:-1] & m[1:]).ravel(order='F'), (m[:, :-1] & m[:, 1:]).ravel(order='F')])
        ax = ex * S.split_fft_product(S.subband_spectra(p.generators(), grid), grid, x)
        assert rel(ax, q['rhs']) < 1.01e-6
        assert np.all(x[ex == 0] == 0)


def test_large_grids_are_solved_alone(rt):
    """Grids of FFT periods above 256 (sides above 128) keep the device busy alone: queued requests of one such period
    are not grouped (the batch statistics stay at zero), and each result converges."""
    import bench
    qs = [bench.served_request(rt, 40 + i, side=s) for i, s in enumerate((140, 150, 160, 40))]
    solver = rt.Solver('gmres', tolerance=1e-6, restart=50, maxIterationCount=3000, precision='c64')
    server = rt.Server(slots=1, batch=8)
    first = server.submit(qs[-1]['L'], qs[-1]['L'], qs[-1]['a'], qs[-1]['a'], qs[-1]['k'], qs[-1]['rhs'],
                          mask=qs[-1]['mask'], solver=solver, preconditioner='nearField')
    tickets = [server.submit(q['L'], q['L'], q['a'], q['a'], q['k'], q['rhs'], mask=q['mask'], solver=solver,
                             preconditioner='nearField') for q in qs[:-1]]
    out = [server.wait(t, q['n']) for t, q in zip(tickets, qs[:-1])]
    server.wait(first, qs[-1]['n'])
    stats = server.stats()
    server.close()
    assert stats['batched_solves'] == 0 and stats['batched_requests'] == 0
    for x, res, times in out:
        assert res.converged() and res.residualNorm() < 1e-6
