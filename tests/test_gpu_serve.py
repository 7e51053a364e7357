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
