"""Benchmark of the iterative-solve path on one MI355X.

Workload (one step = one solve): a square PEC plate of 256 x 256 cells of lambda/10 (k = 2 pi, a = b = 0.1 m,
130,560 rooftops, quadrature orders [16, 8]), a normally incident x-polarized plane wave, GMRES(50) with the
near-field preconditioner on 3x3-cell neighbourhoods, to a true relative residual of 1e-6. Two forms:
  fp64   complex128 products inside an FP64 refinement loop on the true residual;
  mixed  complex64 products (operator and preconditioner) inside the same FP64 refinement loop.
The right-hand side and the solution stay in HBM (rt_solve_device); one solve is timed from enqueue to its
convergence. value = mixed solves/s over the timed steps.

The JSON line also carries p50/p99 latency per form, the error of the mixed solution against the FP64 one, a live
roofline of the dominant kernel (by total GPU time in the mixed form), measured with HIP events on the library's
stream over the timed solves, and a CPU baseline: the oracle's FP64 split-FFT GMRES on a bounded
sample of the same solve, timed on this host.

Batched forms solve the same plate under 64 plane waves together (GMRES(50) and COCG, each mixed and FP64;
batch solves/s, p50/p99 batch latency), the matvec is profiled at 1 and 64 right-hand sides, and a memory-ceiling
run solves the 4096 x 4096 plate (33.5 M unknowns) once. A served stream of 32 requests (random grids, frequencies,
masks and feeds) runs through the library's request loop with Poisson arrivals (p50/p99 latency, solves/s).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--batch 64] [--no-cpu]
"""

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'solves/s + p50/p99 latency at rel.res 1e-6, mixed vs FP64; matvec HBM GB/s'
PEAK_HBM_GBS = 8000.0


# Timed kernel slots of the library (RT_TIMED_* of include/rooftop.h) and the kernel symbols they time.
TIMED = ['pass_a_k<double> (x FFT)', 'pass_b_k<double> (y FFT + multiply)', 'pass_c_k<double> (inverse x FFT)',
         'pass_a_k<float> (x FFT)', 'pass_b_k<float> (y FFT + multiply)', 'pass_c_k<float> (inverse x FFT)',
         'k_dcgs_dots ([s z] = V^H [p w])',
         'k_dcgs_update (q = (p - V s)/beta; p = (w - V z)/(beta nu); Hessenberg columns, rotations, stopping test)',
         'sai_apply<double> (near field)', 'sai_apply<float> (near field)',
         'k_dgmres_sums (block sums; beta, z_j, nu)']


class KernelClock:
    """The library's live kernel timing (HIP events around each timed launch on the library's stream)."""

    def __init__(self, rt, ctx):
        self.rt, self.ctx, self.lib = rt, ctx, rt.library()
        self.n = len(TIMED)
        self.launches = (C.c_int64 * self.n)()
        self.ms = (C.c_double * self.n)()
        self.by = (C.c_double * self.n)()

    def start(self):
        self.rt.check(self.lib.rt_ctx_set_timing(self.ctx.handle, 0))
        self.rt.check(self.lib.rt_ctx_kernel_timing(self.ctx.handle, self.launches, self.ms, self.by))
        self.rt.check(self.lib.rt_ctx_set_timing(self.ctx.handle, 1))

    def stop(self):
        """{slot name: {launches, ms_per_launch, bytes_per_launch, GBps, total_ms}} of the kernels that ran."""
        self.rt.check(self.lib.rt_ctx_set_timing(self.ctx.handle, 0))
        self.rt.check(self.lib.rt_ctx_kernel_timing(self.ctx.handle, self.launches, self.ms, self.by))
        out = {}
        for k, name in enumerate(TIMED):
            if self.launches[k]:
                out[name] = {'launches': int(self.launches[k]), 'ms_per_launch': round(self.ms[k], 6),
                             'bytes_per_launch': round(self.by[k]),
                             'GBps': round(self.by[k] / (self.ms[k] * 1e-3) / 1e9, 1) if self.ms[k] > 0 else 0.0,
                             'total_ms': round(self.launches[k] * self.ms[k], 3)}
        return out


def roofline_of(kernels, name, traffic=None):
    k = kernels[name]
    achieved = k['GBps']
    return {'bound': 'hbm', 'kernel': name, 'achieved': achieved, 'peak': PEAK_HBM_GBS, 'unit': 'GB/s',
            'frac': round(achieved / PEAK_HBM_GBS, 4), 'traffic': traffic, 'launches': k['launches'],
            'ms_per_launch': k['ms_per_launch'], 'bytes_per_launch': k['bytes_per_launch']}


def measured_traffic(kernel_symbol, workload):
    """HBM bytes per launch of a kernel from the committed rocprofv3 counter summary of this workload, if any:
    profiles/traffic_<kernel>_<workload>.txt holds a line 'traffic_bytes_per_launch: <bytes>'."""
    path = os.path.join(ROOT, 'profiles', 'traffic_%s_%s.txt' % (kernel_symbol, workload))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith('traffic_bytes_per_launch:'):
                return float(line.split(':', 1)[1])
    return None


def percentile(values, q):
    return float(np.percentile(np.asarray(values), q)) if values else None


def run_form(rt, ctx, op, pc, solver, d_b, d_x, n, steps, warmup):
    lib = rt.library()
    o = solver.options()
    out = rt._abi.rt_output()
    times, iters, outers, residuals = [], [], [], []
    for s in range(warmup + steps):
        ctx.synchronize()
        t0 = time.perf_counter()
        rt.check(lib.rt_solve_device(ctx.handle, op.handle, pc.handle, C.byref(o), C.c_void_p(d_b), n, 1, None,
                                     C.c_void_p(d_x), C.byref(out)))
        ctx.synchronize()
        dt = time.perf_counter() - t0
        if s >= warmup:
            times.append(dt)
            iters.append(out.iteration_count)
            outers.append(out.outer_count)
            residuals.append(out.residual_norm)
        if not out.converged:
            raise RuntimeError('solve did not converge: residual %.3e' % out.residual_norm)
    return times, iters, outers, residuals


def plane_wave_angles(count, seed=0):
    """(directions, polarizations), count x 3 each: (theta, phi) uniform in [0, 80] x [0, 360) degrees and TE or TM
    polarization drawn from the seed; incidence from above the plate (z > 0)."""
    rng = np.random.default_rng(seed)
    dirs, pols = [], []
    for _ in range(count):
        th, ph = np.radians(rng.uniform(0, 80)), np.radians(rng.uniform(0, 360))
        d = -np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
        te = np.array([-np.sin(ph), np.cos(ph), 0.0])
        tm = np.array([np.cos(th) * np.cos(ph), np.cos(th) * np.sin(ph), -np.sin(th)])
        dirs.append(d)
        pols.append(te if rng.random() < 0.5 else tm)
    return np.array(dirs), np.array(pols)


def plane_wave_batch(rt, problem, count, seed=0):
    """The plane waves of plane_wave_angles(count, seed) as a host N x count array (host formulas)."""
    dirs, pols = plane_wave_angles(count, seed)
    cols = [rt.Excitation.planeWave(problem, d, p).vector() for d, p in zip(dirs, pols)]
    return np.asfortranarray(np.stack(cols, axis=1), dtype=np.complex128)


def run_batch(rt, ctx, op, pc, solver, d_b, d_x, n, count, steps, warmup, allow_unconverged=False):
    """Solves the count right-hand sides in d_b together, steps + warmup times; returns per-batch seconds and the
    per-system outputs (iterations, outer steps, residual, converged) of the last batch. A system that does not
    converge raises, unless allow_unconverged."""
    lib = rt.library()
    o = solver.options()
    outs = (rt._abi.rt_output * count)()
    times = []
    for s in range(warmup + steps):
        ctx.synchronize()
        t0 = time.perf_counter()
        rt.check(lib.rt_solve_device(ctx.handle, op.handle, pc.handle, C.byref(o), C.c_void_p(d_b), n, count, None,
                                     C.c_void_p(d_x), outs))
        ctx.synchronize()
        if s >= warmup:
            times.append(time.perf_counter() - t0)
        bad = [r for r in range(count) if not outs[r].converged]
        if bad and not allow_unconverged:
            raise RuntimeError('batched solve: %d systems did not converge' % len(bad))
    return times, [(o_.iteration_count, o_.outer_count, o_.residual_norm, bool(o_.converged)) for o_ in outs]


C0 = 299792458.0


def served_request(rt, i, max_side=1024, side=None):
    """Request i of the served stream (seed 1000 + i): a square grid of side L = round(exp(U[ln 32, ln max_side])),
    frequency U[0.5, 3] GHz, cell lambda / U[10, 20], a 4-connected blob mask grown to a fill U[0.3, 0.9], and a 1 V
    delta gap at a uniformly drawn existing rooftop (b = V * edge length there). side, if given, replaces the drawn L
    (the other draws are unchanged)."""
    rng = np.random.default_rng(1000 + i)
    L = int(round(np.exp(rng.uniform(np.log(32), np.log(max_side)))))
    if side is not None:
        L = int(side)
    f = rng.uniform(0.5e9, 3e9)
    a = C0 / f / rng.uniform(10, 20)
    fill = rng.uniform(0.3, 0.9)
    mask = rt.blob_mask(L, L, 1000 + i, fill)
    ex = np.concatenate([(mask[:-1] & mask[1:]).ravel(order='F'), (mask[:, :-1] & mask[:, 1:]).ravel(order='F')])
    port = int(rng.choice(np.flatnonzero(ex)))
    rhs = np.zeros(ex.size, np.complex128)
    rhs[port] = 1.0 * a   # square cells: every edge has length a
    return {'L': L, 'a': a, 'k': 2 * np.pi * f / C0, 'mask': mask, 'rhs': rhs, 'n': ex.size, 'fill': fill,
            'port': port}


def serve_stream(rt, requests, solver, slots, arrivals, batch=1):
    """Submits requests[i] at arrivals[i] seconds after the start to a server with `slots` workers (batching up to
    `batch` queued requests of one embedding size per solve); returns the per-request (solution, Result, times) and
    the makespan (first arrival to last completion)."""
    import threading
    server = rt.Server(slots=slots, batch=batch)
    tickets = [None] * len(requests)
    t_start = time.perf_counter()

    def feeder():
        for i, q in enumerate(requests):
            delay = arrivals[i] - (time.perf_counter() - t_start)
            if delay > 0:
                time.sleep(delay)
            tickets[i] = server.submit(q['L'], q['L'], q['a'], q['a'], q['k'], q['rhs'], mask=q['mask'],
                                       solver=solver, preconditioner='nearField')

    th = threading.Thread(target=feeder)
    th.start()
    out = []
    for i, q in enumerate(requests):
        while tickets[i] is None:
            time.sleep(0.0005)
        out.append(server.wait(tickets[i], q['n']))
    th.join()
    makespan = time.perf_counter() - t_start - arrivals[0]
    stats = server.stats()
    server.close()
    return out, makespan, stats


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(n, budget_s=10.0):
    """The oracle's FP64 path (scipy FFTs on the host cores of this GPU's share, split-FFT product, sparse near-field
    apply, GMRES(50) with two-pass Gram-Schmidt and the same tolerance) on the same plate: whole solves repeated until
    about budget_s seconds of CPU work (at least one), setup excluded. The threads are capped at 16, the host CPU share
    of one GPU on an 8-GPU node; the host's own count and CPU model are reported beside them."""
    from oracle import assembly as A
    from oracle import excitation as X
    from oracle import nearfield as NF
    from oracle import solvers as SV
    from oracle import spectra as S
    cores = max(1, min(16, os.cpu_count() or 1))
    S.set_workers(cores)
    grid = A.Grid(n, n, 0.1, 0.1)
    gen = A.generate(grid, 2 * np.pi, (16, 8))
    spec = S.subband_spectra(gen, grid)
    pre = NF.NearField(gen, grid, by_class=True)
    b = X.plane_wave(grid, 2 * np.pi, [0, 0, -1], [1, 0, 0], 1.0, order=8)
    op = lambda x: S.split_fft_product(spec, grid, x)
    SV.solve('gmres', op, b, tol=1e-6, max_iter=2, restart=50, preconditioner=pre.apply)   # warm-up
    count, iters, res = 0, 0, 0.0
    t0 = time.perf_counter()
    while count == 0 or time.perf_counter() - t0 < budget_s:
        out = SV.solve('gmres', op, b, tol=1e-6, max_iter=1000, restart=50, preconditioner=pre.apply)
        count += 1
        iters = out['iterationCount']
        res = out['residualNorm']
    dt = time.perf_counter() - t0
    S.set_workers(1)
    cpu_baseline.solution = out['solution']   # the oracle's FP64 solution, for the GPU forms' error against it
    return {'value': count / dt, 'unit': 'solves/s', 'cores': cores, 'kind': 'port', 'cpu_model': cpu_model(),
            'host_cpus': os.cpu_count(),
            'sample': '%d whole solves of the same %dx%d plate in %.1f s: oracle FP64 split-FFT GMRES(50) + '
                      'near-field, %d iterations each, true residual %.2e (scipy FFT on %d threads)'
                      % (count, n, n, dt, iters, res, cores)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=30)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--n', type=int, default=256)
    ap.add_argument('--batch', type=int, default=64, help='right-hand sides of the batched form (0: skip it)')
    ap.add_argument('--batch-n', type=int, default=512, help='plate side (cells) of the batched form')
    ap.add_argument('--serve', type=int, default=32, help='requests of the served stream (0: skip it)')
    ap.add_argument('--ceiling', type=int, default=4096,
                    help='side of the memory-ceiling plate (one mixed GMRES(30) solve; 0: skip it)')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--copy-only', action='store_true',
                    help='only the plain 1 GiB device copy (calibrates the HBM byte counters of rocprofv3)')
    ap.add_argument('--matvec-only', nargs=3, metavar=('SIDE', 'RHS', 'PRECISION'),
                    help='only the matvec of the SIDE x SIDE plate (lambda/10, orders [16, 8]) on RHS right-hand '
                         'sides in c64 or c128, tuned, then applied --steps times (the counter profiles)')
    ap.add_argument('--shape', type=str, default=None,
                    help='with --matvec-only: this launch shape (JSON {"bits", "group", "tile", "teams", "chunks"}, '
                         'as a plain run reports it) instead of tuning, so that a profiled run launches the same '
                         'kernels as the plain one')
    ap.add_argument('--no-kernel-timing', action='store_true',
                    help='with --batched-gmres-only: no second, event-timed solve (the counter runs)')
    ap.add_argument('--batched-gmres-only', action='store_true',
                    help='only one batched mixed GMRES(50) + near-field solve of the --batch plane waves on the '
                         '--batch-n plate (the counter profiles of its Gram-Schmidt passes)')
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) != 1:
        print('bench.py measures one GPU: this build runs the path on a single MI355X and does not shard a solve '
              'across GPUs (independent solves would shard with no exchange); run it with --gpus 1.',
              file=sys.stderr)
        sys.exit(2)

    import __graft_entry__ as ge
    rt = ge.load()
    lib = rt.library()
    ctx = rt.Context.default()
    copy_bytes = 1 << 30
    copy_ms = C.c_double()
    rt.check(lib.rt_bandwidth_probe(ctx.handle, copy_bytes, 1 if args.copy_only else 20, C.byref(copy_ms)))
    copy_gbs = 2 * copy_bytes / (copy_ms.value * 1e-3) / 1e9
    if args.copy_only:
        print(json.dumps({'copy_bytes': copy_bytes, 'ms': copy_ms.value, 'GBps': round(copy_gbs, 1)}), flush=True)
        return
    if args.batched_gmres_only:
        R, side = args.batch, args.batch_n
        print('bench: batched mixed GMRES, %dx%d plate, %d plane waves' % (side, side, R), flush=True)
        bp = rt.Problem(rt.Mesh(side, side, 0.1, 0.1), 2 * np.pi, quadratureOrder=(16, 8))
        bo = rt.Operator(bp, precision='mixed')
        bq = rt.Preconditioner(bp, 'nearField', precision='mixed')
        bo.tune(nrhs=R, precision='c64')
        dirs, pols = plane_wave_angles(R, seed=1)
        bb = rt.plane_waves(bp, dirs, pols)
        bx = rt.DeviceBatch(ctx, bp.rooftopCount(), R)
        solver = rt.Solver('gmres', tolerance=1e-6, restart=50, precision='c64')
        bt, outs = run_batch(rt, ctx, bo, bq, solver, bb.ptr, bx.ptr, bp.rooftopCount(), R, 1, 0)
        line = {'workload': 'plate %dx%d, %d plane waves, mixed GMRES(50) + near field' % (side, side, R),
                'seconds': bt[0], 'iterations_max': max(q[0] for q in outs)}
        if not args.no_kernel_timing:   # a second solve with the live kernel timing
            clock = KernelClock(rt, ctx)
            clock.start()
            run_batch(rt, ctx, bo, bq, solver, bb.ptr, bx.ptr, bp.rooftopCount(), R, 1, 0)
            line['kernels'] = clock.stop()
        print(json.dumps(line), flush=True)
        return
    if args.matvec_only:
        side, R, prec = int(args.matvec_only[0]), int(args.matvec_only[1]), args.matvec_only[2]
        print('bench: matvec of the %dx%d plate, %d right-hand sides, %s' % (side, side, R, prec), flush=True)
        mp = rt.Problem(rt.Mesh(side, side, 0.1, 0.1), 2 * np.pi, quadratureOrder=(16, 8))
        mop = rt.Operator(mp, precision='mixed')
        if args.shape:
            shape = json.loads(args.shape)
            mop.setShape(shape, nrhs=R, precision=prec)
        else:
            shape, _ = mop.tune(nrhs=R, precision=prec)
        pms, pby, wall = mop.profile(nrhs=R, reps=max(1, args.steps), precision=prec)
        print(json.dumps({'workload': 'plate %dx%d, %d right-hand sides, %s' % (side, side, R, prec),
                          'shape': shape, 'pass_ms': pms, 'pass_bytes': pby, 'wall_ms': wall,
                          'GBps': sum(pby) / (wall * 1e-3) / 1e9,
                          'frac': sum(pby) / (wall * 1e-3) / 1e9 / PEAK_HBM_GBS}), flush=True)
        return
    n = args.n
    print('bench: plate %dx%d, assembling generators and spectra' % (n, n), flush=True)
    t0 = time.perf_counter()
    problem = rt.Problem(rt.Mesh(n, n, 0.1, 0.1), 2 * np.pi, quadratureOrder=(16, 8))
    op = rt.Operator(problem, precision='mixed')
    pc = rt.Preconditioner(problem, 'nearField', precision='mixed')
    setup_s = time.perf_counter() - t0
    # Launch shapes of the matvec passes for the batch sizes timed below (the product's autotuner; a one-off per
    # product and batch size, like an FFT plan, not part of a solve).
    t0 = time.perf_counter()
    tuned = {}
    for nrhs in sorted({1, args.batch} - {0}):
        for prec in ('c64', 'c128'):
            tuned['%s x%d' % (prec, nrhs)] = op.tune(nrhs=nrhs, precision=prec)[0]
    if args.batch > 0 and args.batch_n != n:
        # The batched form's own plate (the multi-right-hand-side configuration).
        bproblem = rt.Problem(rt.Mesh(args.batch_n, args.batch_n, 0.1, 0.1), 2 * np.pi, quadratureOrder=(16, 8))
        bop = rt.Operator(bproblem, precision='mixed')
        bpc = rt.Preconditioner(bproblem, 'nearField', precision='mixed')
        for prec in ('c64', 'c128'):
            tuned['%s x%d plate%d' % (prec, args.batch, args.batch_n)] = bop.tune(nrhs=args.batch, precision=prec)[0]
    elif args.batch > 0:
        bproblem, bop, bpc = problem, op, pc
    tune_s = time.perf_counter() - t0
    exc = rt.Excitation.planeWave(problem)
    N = problem.rooftopCount()
    b = np.ascontiguousarray(exc.vector(), dtype=np.complex128)
    d_b, d_x, d_x64 = C.c_void_p(), C.c_void_p(), C.c_void_p()
    for ptr in (d_b, d_x, d_x64):
        rt.check(lib.rt_device_alloc(ctx.handle, 16 * N, C.byref(ptr)))
    rt.check(lib.rt_copy_to_device(ctx.handle, d_b, b.ctypes.data_as(C.c_void_p), 16 * N))

    fp64 = rt.Solver('gmres', tolerance=1e-6, restart=50, refine=True)
    mixed = rt.Solver('gmres', tolerance=1e-6, restart=50, precision='c64')
    mixed_b64 = rt.Solver('gmres', tolerance=1e-6, restart=50, precision='c64', basis='c128')
    print('bench: fp64 form', flush=True)
    t64, i64, o64, r64 = run_form(rt, ctx, op, pc, fp64, d_b.value, d_x64.value, N, args.steps, args.warmup)
    print('bench: mixed form', flush=True)
    tmx, imx, omx, rmx = run_form(rt, ctx, op, pc, mixed, d_b.value, d_x.value, N, args.steps, args.warmup)
    x64 = np.empty(N, np.complex128)
    xmx = np.empty(N, np.complex128)
    rt.check(lib.rt_copy_to_host(ctx.handle, x64.ctypes.data_as(C.c_void_p), d_x64, 16 * N))
    rt.check(lib.rt_copy_to_host(ctx.handle, xmx.ctypes.data_as(C.c_void_p), d_x, 16 * N))
    err = float(np.linalg.norm(xmx - x64) / np.linalg.norm(x64))
    # The mixed form with its GMRES basis in complex128 (every Gram-Schmidt product and the basis in FP64).
    print('bench: mixed form, complex128 basis', flush=True)
    d_xb = C.c_void_p()
    rt.check(lib.rt_device_alloc(ctx.handle, 16 * N, C.byref(d_xb)))
    tb6, ib6, ob6, rb6 = run_form(rt, ctx, op, pc, mixed_b64, d_b.value, d_xb.value, N, max(5, args.steps // 3),
                                  args.warmup)
    xb6 = np.empty(N, np.complex128)
    rt.check(lib.rt_copy_to_host(ctx.handle, xb6.ctypes.data_as(C.c_void_p), d_xb, 16 * N))
    rt.check(lib.rt_device_free(ctx.handle, d_xb))
    # Kernel table: a separate run of the mixed form with the library's event timing on (the events themselves add
    # launch gaps, so the headline solves above run without them). The dominant kernel by total GPU time gives the
    # roofline.
    print('bench: mixed form, kernel timing', flush=True)
    clock = KernelClock(rt, ctx)
    clock.start()
    run_form(rt, ctx, op, pc, mixed, d_b.value, d_x.value, N, max(3, args.steps // 4), 0)
    kernels = clock.stop()
    dominant = max(kernels, key=lambda k: kernels[k]['total_ms'])
    workload = 'plate%d' % n
    roofline = roofline_of(kernels, dominant, measured_traffic(dominant.split(' ')[0].split('<')[0], workload))

    matvec = {}
    for R in sorted({1, args.batch} - {0}):
        for prec in ('c64', 'c128'):
            pms, pby, wall = op.profile(nrhs=R, reps=50 if R == 1 else 10, precision=prec)
            matvec['%s x%d' % (prec, R)] = {'ms': round(wall, 5), 'bytes': sum(pby),
                                            'GBps': round(sum(pby) / (wall * 1e-3) / 1e9, 1),
                                            'frac': round(sum(pby) / (wall * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                                            'pass_ms': [round(v, 5) for v in pms]}

    # Batched form: the same plate under `batch` plane waves, solved together (one product launch per pass for all).
    batch = {}
    if args.batch > 0:
        R = args.batch
        BN = bproblem.rooftopCount()
        for prec in ('c64', 'c128'):
            pms, pby, wall = bop.profile(nrhs=R, reps=10, precision=prec)
            batch['matvec %s x%d' % (prec, R)] = {'ms': round(wall, 5), 'bytes': sum(pby),
                                                 'GBps': round(sum(pby) / (wall * 1e-3) / 1e9, 1),
                                                 'frac': round(sum(pby) / (wall * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                                                 'pass_ms': [round(v, 5) for v in pms]}
        # The excitations are built on the device (rt_plane_waves_device), and the radar cross sections of the
        # solutions are computed there too: no N x R transfer.
        dirs, pols = plane_wave_angles(R, seed=1)
        t0 = time.perf_counter()
        b_batch = rt.plane_waves(bproblem, dirs, pols)
        batch['excitation_build_ms'] = round(1e3 * (time.perf_counter() - t0), 3)
        x_batch = rt.DeviceBatch(ctx, BN, R)
        d_B, d_X = C.c_void_p(b_batch.ptr), C.c_void_p(x_batch.ptr)
        rcs = {}
        bsteps, bwarm = max(1, args.steps // 10), 1
        # Block COCG (one block Krylov space for the batch) next to the lockstep COCG of independent recurrences; its
        # iteration limit is lower (a block iteration costs several lockstep ones).
        forms = (('gmres mixed', mixed), ('gmres fp64', fp64),
                 ('cocg mixed', rt.Solver('cg', tolerance=1e-6, maxIterationCount=2000, precision='c64')),
                 ('cocg fp64', rt.Solver('cg', tolerance=1e-6, maxIterationCount=2000, refine=True)),
                 ('blockcg mixed', rt.Solver('blockcg', tolerance=1e-6, maxIterationCount=1000, precision='c64')),
                 ('blockcg fp64', rt.Solver('blockcg', tolerance=1e-6, maxIterationCount=1000, refine=True)))
        for name, solver in forms:
            print('bench: batched %s form, %d right-hand sides' % (name, R), flush=True)
            slow = name.startswith('blockcg')   # one timed batch, no warmup: seconds per batch
            bt, outs = run_batch(rt, ctx, bop, bpc, solver, d_B.value, d_X.value, BN, R, 1 if slow else bsteps,
                                 0 if slow else bwarm, allow_unconverged=True)
            its = [q[0] for q in outs]
            conv = sum(q[3] for q in outs)
            rcs[name] = rt.radar_cross_sections(bproblem, b_batch, x_batch)
            # Only the systems that reach the tolerance count as solves.
            batch[name] = {'solves_per_s': round(conv * len(bt) / sum(bt), 3), 'converged': '%d of %d' % (conv, R),
                           'batch_p50_ms': round(1e3 * percentile(bt, 50), 3),
                           'batch_p99_ms': round(1e3 * percentile(bt, 99), 3),
                           'iterations_median': int(np.median(its)), 'iterations_max': int(max(its)),
                           'outer_max': int(max(q[1] for q in outs)),
                           'residual_max': max(q[2] for q in outs), 'steps': len(bt)}
        print('bench: batched gmres mixed form, kernel timing', flush=True)
        clock.start()
        run_batch(rt, ctx, bop, bpc, mixed, d_B.value, d_X.value, BN, R, 1, 0)
        batch['gmres mixed']['kernels'] = clock.stop()
        # The batched form's own dominant kernel (by total time) against the HBM peak: at 64 right-hand sides the
        # Gram-Schmidt passes move gigabytes per launch.
        bk = batch['gmres mixed']['kernels']
        bdom = max(bk, key=lambda k: bk[k]['total_ms'])
        batch['roofline'] = roofline_of(bk, bdom, measured_traffic(bdom.split(' ')[0].split('<')[0],
                                                                   'plate%dx%d' % (args.batch_n, R)))
        batch['rhs'] = R
        batch['plate'] = '%dx%d cells, lambda/10, %d rooftop unknowns' % (args.batch_n, args.batch_n, BN)
        batch['excitations'] = ('plane waves, theta U[0,80] deg, phi U[0,360) deg, TE/TM at random (seed 1), '
                                'solved together')
        for m in ('gmres', 'cocg', 'blockcg'):
            f64 = batch['%s fp64' % m]['solves_per_s']
            batch['%s mixed_vs_fp64_speedup' % m] = round(batch['%s mixed' % m]['solves_per_s'] / f64, 3) if f64 \
                else None   # no FP64 system converged
        best = lambda kind: max(batch['%s %s' % (m, kind)]['solves_per_s'] for m in ('gmres', 'cocg', 'blockcg'))
        batch['best mixed_vs_best_fp64'] = round(best('mixed') / max(best('fp64'), 1e-9), 3)
        # Radar cross sections of the mixed forms against the FP64 GMRES ones (every system converged there).
        for name in ('gmres mixed', 'cocg mixed', 'cocg fp64', 'blockcg mixed', 'blockcg fp64'):
            batch[name]['rcs_max_rel_diff_vs_gmres_fp64'] = float(np.max(np.abs(rcs[name] - rcs['gmres fp64']) /
                                                                         np.abs(rcs['gmres fp64'])))
        b_batch.free()
        x_batch.free()

    # Served stream: requests of random grids (32..1024), frequencies, blob masks and feeds; single-stream
    # throughput with one slot (requests back to back), all requests at once into 4 slots (saturation), then Poisson
    # arrivals at 0.7 x the mixed single-stream throughput into 4 slots, for the mixed and the FP64 GMRES(50) +
    # near-field solve to 1e-6; each
    # request's solution in the mix is compared with its solution alone (the single-stream run).
    serving = {}
    if args.serve > 0:
        Q = args.serve
        print('bench: served stream, %d requests' % Q, flush=True)
        reqs = [served_request(rt, i) for i in range(Q)]
        forms = (('mixed', rt.Solver('gmres', tolerance=1e-6, restart=50, maxIterationCount=3000, precision='c64')),
                 ('fp64', rt.Solver('gmres', tolerance=1e-6, restart=50, maxIterationCount=3000, refine=True)))
        alone, rate = {}, None
        for name, solver in forms:
            out1, span1, _ = serve_stream(rt, reqs, solver, 1, [0.0] * Q)
            alone[name] = out1
            serving[name] = {'single_stream_solves_per_s': round(Q / span1, 3),
                             'service_ms_mean': round(1e3 * np.mean([o[2]['service'] for o in out1]), 3)}
            if rate is None:
                rate = 0.7 * Q / span1
        arrivals = np.concatenate([[0.0], np.cumsum(np.random.default_rng(7).exponential(1.0 / rate, Q - 1))])
        for name, solver in forms:
            print('bench: served stream, %s form, saturation' % name, flush=True)
            sat, sat_span, sat_st = serve_stream(rt, reqs, solver, 4, [0.0] * Q)
            serving[name]['saturation_solves_per_s'] = round(Q / sat_span, 3)
            serving[name]['saturation_peak_concurrent'] = sat_st['peak_running']
            print('bench: served stream, %s form, Poisson arrivals' % name, flush=True)
            out, span, st = serve_stream(rt, reqs, solver, 4, arrivals)
            lat = [o[2]['latency'] for o in out]
            diff = max(float(np.linalg.norm(o[0] - a[0]) / np.linalg.norm(a[0])) for o, a in zip(out, alone[name]))
            serving[name].update({
                'solves_per_s': round(Q / span, 3), 'p50_ms': round(1e3 * percentile(lat, 50), 3),
                'p99_ms': round(1e3 * percentile(lat, 99), 3),
                'converged': int(sum(o[1].converged() for o in out)),
                'residual_max': max(o[1].residualNorm() for o in out),
                'iterations_median': int(np.median([o[1].iterationCount() for o in out])),
                'peak_concurrent': st['peak_running'], 'alone_vs_mix_max_rel_diff': diff})
        # Batching by embedding size: queued requests of one FFT period join one batched product and Krylov loop
        # (up to 8), on the same arrivals; their results against the ones alone (equal to the solve's accuracy).
        name, solver = forms[0]
        print('bench: served stream, mixed form batched by embedding size, saturation', flush=True)
        sat, sat_span, sat_st = serve_stream(rt, reqs, solver, 4, [0.0] * Q, batch=8)
        print('bench: served stream, mixed form batched by embedding size, Poisson arrivals', flush=True)
        out, span, st = serve_stream(rt, reqs, solver, 4, arrivals, batch=8)
        lat = [o[2]['latency'] for o in out]
        diff = max(float(np.linalg.norm(o[0] - a[0]) / np.linalg.norm(a[0])) for o, a in zip(out, alone[name]))
        serving['mixed batched'] = {
            'saturation_solves_per_s': round(Q / sat_span, 3), 'saturation_batched_requests':
                sat_st['batched_requests'], 'solves_per_s': round(Q / span, 3),
            'p50_ms': round(1e3 * percentile(lat, 50), 3), 'p99_ms': round(1e3 * percentile(lat, 99), 3),
            'converged': int(sum(o[1].converged() for o in out)),
            'residual_max': max(o[1].residualNorm() for o in out), 'batched_requests': st['batched_requests'],
            'batched_solves': st['batched_solves'], 'alone_vs_mix_max_rel_diff': diff}
        serving['requests'] = Q
        serving['arrival_rate_per_s'] = round(rate, 3)
        serving['workload'] = ('square grids L = round(exp(U[ln 32, ln 1024])), f ~ U[0.5, 3] GHz, cell lambda/U[10,'
                               ' 20], 4-connected blob masks filled U[0.3, 0.9], 1 V delta gap at a random existing '
                               'rooftop (seeds 1000+i); Poisson arrivals (seed 7) at 0.7 x the mixed single-stream '
                               'throughput; GMRES(50) + near field to 1e-6 (true residual); 4 slots; saturation = all '
                               'requests at once')
        serving['sides'] = [q['L'] for q in reqs]

    # Memory ceiling: the 4096 x 4096 plate (33.5 M unknowns), mixed GMRES(30) + near field, FP64 refinement to 1e-6,
    # one solve; the device bytes the library holds against the device's HBM.
    ceiling = {}
    if args.ceiling > 0:
        L = args.ceiling
        print('bench: memory ceiling, plate %dx%d' % (L, L), flush=True)
        del pc, op
        t0 = time.perf_counter()
        big = rt.Problem(rt.Mesh(L, L, 0.1, 0.1), 2 * np.pi, quadratureOrder=(16, 8))
        bop = rt.Operator(big, precision='mixed')
        bpc = rt.Preconditioner(big, 'nearField', precision='mixed')
        bsetup = time.perf_counter() - t0
        mv = {}
        for prec in ('c64', 'c128'):
            bop.tune(nrhs=1, precision=prec)
            pms, pby, wall = bop.profile(nrhs=1, reps=5, precision=prec)
            mv[prec] = {'ms': round(wall, 4), 'bytes': sum(pby), 'GBps': round(sum(pby) / (wall * 1e-3) / 1e9, 1),
                        'frac': round(sum(pby) / (wall * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                        'pass_ms': [round(v, 4) for v in pms]}
        bb = rt.Excitation.planeWave(big).vector()
        t0 = time.perf_counter()
        res = rt.Solver('gmres', tolerance=1e-6, restart=30, precision='c64', maxIterationCount=3000).solve(
            bop, bb, preconditioner=bpc)
        bt = time.perf_counter() - t0
        held, total = C.c_uint64(), C.c_uint64()
        rt.check(lib.rt_ctx_memory(ctx.handle, C.byref(held), C.byref(total)))
        ceiling = {'workload': 'plate %dx%d cells lambda/10, normal incidence, mixed GMRES(30) + near-field, FP64 '
                               'refinement to 1e-6' % (L, L),
                   'unknowns': big.rooftopCount(), 'solve_s': round(bt, 3), 'setup_s': round(bsetup, 3),
                   'converged': res.converged(), 'iterations': res.iterationCount(), 'outer': res.outerCount,
                   'residual': res.residualNorm(), 'held_GB': round(held.value / 1e9, 2),
                   'device_GB': round(total.value / 1e9, 1), 'matvec': mv}
        del bop, bpc, big

    steps = args.steps
    value = steps / sum(tmx)
    line = {
        'metric': METRIC, 'value': round(value, 3), 'unit': 'solves/s', 'n_gpus': 1, 'steps': steps,
        'warmup': args.warmup, 'ms_per_step': round(1e3 * sum(tmx) / steps, 3), 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'complex64 products, FP64 refinement',
        'data': 'synthetic',
        'config': {'workload': 'plate %dx%d cells lambda/10, normal-incidence x-polarized plane wave, GMRES(50) + '
                               'near-field 3x3, true rel.res 1e-6' % (n, n), 'unknowns': N, 'cells': [n, n],
                   'embedding': [2 * n, 2 * n], 'quadrature_orders': [16, 8]},
        'mixed': {'solves_per_s': round(value, 3), 'p50_ms': round(1e3 * percentile(tmx, 50), 3),
                  'p99_ms': round(1e3 * percentile(tmx, 99), 3), 'iterations': int(np.median(imx)),
                  'outer': int(np.median(omx)), 'residual': max(rmx)},
        'fp64': {'solves_per_s': round(steps / sum(t64), 3), 'p50_ms': round(1e3 * percentile(t64, 50), 3),
                 'p99_ms': round(1e3 * percentile(t64, 99), 3), 'iterations': int(np.median(i64)),
                 'outer': int(np.median(o64)), 'residual': max(r64)},
        'mixed_complex128_basis': {'solves_per_s': round(len(tb6) / sum(tb6), 3),
                                   'p50_ms': round(1e3 * percentile(tb6, 50), 3), 'iterations': int(np.median(ib6)),
                                   'residual': max(rb6),
                                   'error_vs_fp64': float(np.linalg.norm(xb6 - x64) / np.linalg.norm(x64))},
        'mixed_vs_fp64_speedup': round(sum(t64) / sum(tmx), 3),
        'error_mixed_vs_fp64': err,
        'roofline': roofline,
        'kernels': kernels,
        'matvec': matvec,
        'copy_ceiling_GBps': round(copy_gbs, 1),
        'batched': batch,
        'serving': serving,
        'memory_ceiling': ceiling,
        'setup_s': round(setup_s, 3),
        'tuning': {'seconds': round(tune_s, 3), 'shapes': tuned},
    }
    if not args.no_cpu:
        print('bench: cpu baseline', flush=True)
        line['cpu_baseline'] = cpu_baseline(n)
        # Each GPU form's solution against the oracle's FP64 solution of the same plate (all at residual 1e-6).
        xo = cpu_baseline.solution
        line['error_vs_fp64_oracle'] = {
            name: float(np.linalg.norm(v - xo) / np.linalg.norm(xo))
            for name, v in (('mixed', xmx), ('fp64', x64), ('mixed_complex128_basis', xb6))}
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
