"""GPU test of rt_product_apply_device_noalloc, the allocation-free application of a product: through the C ABI, two
calls in a row (the second with a smaller input) equal rt_product_apply_device bit for bit and the oracle's product
to the FP64 bar, hold no new device bytes in the context, and refuse what they cannot do without allocating."""
import ctypes as C

import numpy as np
import pytest

from conftest import rel
from oracle import assembly as A
from oracle import spectra as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def rt():
    import __graft_entry__ as ge
    return ge.load()


def _apply(rt, fn, op, x_dev, y_dev, n, nrhs, prec):
    code = {'c128': rt._abi.RT_C128, 'c64': rt._abi.RT_C64}[prec]
    rt.check(getattr(rt.library(), fn)(op.handle, C.c_void_p(x_dev.ptr), C.c_void_p(y_dev.ptr), n, nrhs, code))


def test_noalloc_apply_equals_apply_device(rt):
    ctx = rt.Context()
    cases = [((40, 33, 0.05, 0.05), 5), ((9, 7, 0.1, 0.12), 2)]   # the second call: a smaller input
    for (cx, cy, a, b), nrhs in cases:
        print('noalloc product', (cx, cy), nrhs, flush=True)
        p = rt.Problem(rt.Mesh(cx, cy, a, b), 2 * np.pi, quadratureOrder=(4, 4), context=ctx)
        op = rt.Operator(p, precision='mixed')
        n = p.rooftopCount()
        rng = np.random.default_rng(cx)
        x = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
        grid = A.Grid(cx, cy, a, b)
        spectra = S.subband_spectra(p.generators(), grid)
        ref = np.stack([S.split_fft_product(spectra, grid, x[:, r]) for r in range(nrhs)], axis=1)
        xd = rt.DeviceBatch.from_host(ctx, x)
        yn = rt.DeviceBatch(ctx, n, nrhs)
        yd = rt.DeviceBatch(ctx, n, nrhs)
        for prec, tol in (('c128', 1e-12), ('c64', 5e-6)):
            held = ctx.memory()[0]
            _apply(rt, 'rt_product_apply_device_noalloc', op, xd, yn, n, nrhs, prec)
            ctx.synchronize()
            assert ctx.memory()[0] == held
            got = yn.to_host()
            _apply(rt, 'rt_product_apply_device', op, xd, yd, n, nrhs, prec)
            assert np.array_equal(got, yd.to_host()), prec
            assert rel(got, ref) < tol, prec
        for buf in (xd, yn, yd):
            buf.free()


def test_noalloc_apply_refuses_what_needs_memory(rt):
    ctx = rt.Context()
    p = rt.Problem(rt.Mesh(512, 512, 0.1, 0.1), 2 * np.pi, quadratureOrder=(3, 3), context=ctx)
    op = rt.Operator(p, precision='mixed')
    n = p.rooftopCount()
    x = rt.DeviceBatch.from_host(ctx, np.ones((n, 8)))
    y = rt.DeviceBatch(ctx, n, 8)
    # 8 right-hand sides of the 512 x 512 plate in complex128: 8 x 16 MiB of working memory, more than 64 MiB.
    with pytest.raises(rt.RooftopError) as e:
        _apply(rt, 'rt_product_apply_device_noalloc', op, x, y, n, 8, 'c128')
    assert e.value.identifier == 'rooftop:product:noAlloc'
    # A launch shape whose pass-B spectra are not built yet (complex128 1024-point transforms default to radices up
    # to 16, so the radix-8 plan's) is refused until the product has been applied once with it.
    op.setShape({'bits': 3, 'group': 2, 'tile': 1}, nrhs=2, precision='c128')
    with pytest.raises(rt.RooftopError) as e:
        _apply(rt, 'rt_product_apply_device_noalloc', op, x, y, n, 2, 'c128')
    assert e.value.identifier == 'rooftop:product:noAlloc'
    _apply(rt, 'rt_product_apply_device', op, x, y, n, 2, 'c128')
    first = y.to_host()[:, :2]
    _apply(rt, 'rt_product_apply_device_noalloc', op, x, y, n, 2, 'c128')
    assert np.array_equal(y.to_host()[:, :2], first)
    pc = rt.Preconditioner(p, 'nearField')
    with pytest.raises(rt.RooftopError) as e:
        _apply(rt, 'rt_product_apply_device_noalloc', pc, x, y, n, 1, 'c128')
    assert e.value.identifier == 'rooftop:product:noAlloc'
    x.free()
    y.free()
