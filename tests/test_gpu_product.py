"""GPU parity of the generators and the Toeplitz product (librooftop.so through the C ABI) against the oracle."""
import numpy as np
import pytest

from conftest import rel
from oracle import assembly as A
from oracle import spectra as S

pytestmark = pytest.mark.gpu

GRIDS = {'plate': (5, 4, 0.1, 0.15), 'oddPlate': (11, 7, 0.1, 0.1), 'twoFacesWide': (2, 5, 0.1, 0.1),
         'stripX': (9, 1, 0.1, 0.02), 'stripY': (1, 6, 0.03, 0.1)}


@pytest.fixture(scope='module')
def rt():
    import __graft_entry__ as ge
    return ge.load()


def _problem(rt, cx, cy, a, b, orders, k=2 * np.pi, mask=None):
    return rt.Problem(rt.Mesh(cx, cy, a, b), k, quadratureOrder=orders, mask=mask)


def _random(n, seed=1, nrhs=None):
    rng = np.random.default_rng(seed)
    shape = (n,) if nrhs is None else (n, nrhs)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_products_match_reference_mat(rt, reference_data):
    """tests/+reference/+assembler/+greensMatrix/ApplicatorTest.m:5-21 (1e-13) with the stored v2 generators."""
    for case in reference_data:
        print('reference case', case['countX'], case['countY'], flush=True)
        p = _problem(rt, int(case['countX']), int(case['countY']), case['sizeX'], case['sizeY'],
                     (int(case['order']),) * 2, case['wavenumber'])
        p.setGenerators(case['magneticXX'], case['magneticYY'], case['electric'])
        for product in ('fft', 'splitFFT', 'leanSplitFFT'):
            y = rt.Operator(p, product=product).apply(case['vector'])
            assert rel(y, case['productFFT']) < 1e-13
            assert rel(y, case['productSplit']) < 1e-13


@pytest.mark.parametrize('name', sorted(GRIDS))
def test_generators_match_oracle(rt, name):
    cx, cy, a, b = GRIDS[name]
    for orders in ((3, 3), (16, 8)):
        print('generators', name, orders, flush=True)
        gpu = _problem(rt, cx, cy, a, b, orders).generators()
        ref = A.generate(A.Grid(cx, cy, a, b), 2 * np.pi, orders)
        for key in ('magneticXX', 'magneticYY', 'electric'):
            if ref[key] is None:
                assert gpu[key] is None
                continue
            np.testing.assert_allclose(gpu[key], ref[key], rtol=1e-13, atol=1e-14 * np.abs(ref[key]).max())


def test_generators_elongated_cells_and_larger_grid(rt):
    for case in ((6, 40, 0.1, 0.01), (30, 1, 0.025, 0.004), (40, 33, 0.05, 0.05)):
        print('generators', case, flush=True)
        gpu = _problem(rt, *case, (16, 8)).generators()
        ref = A.generate(A.Grid(*case), 2 * np.pi, (16, 8))
        for key in ('magneticXX', 'magneticYY', 'electric'):
            if ref[key] is not None:
                assert rel(gpu[key], ref[key]) < 1e-13


@pytest.mark.parametrize('name', sorted(GRIDS))
def test_products_match_dense_matrix(rt, name):
    """tests/+unit/+assembler/+greensMatrix/ApplicatorTest.m:7-29 (1e-12)."""
    cx, cy, a, b = GRIDS[name]
    print('product', name, flush=True)
    p = _problem(rt, cx, cy, a, b, (3, 3))
    grid = A.Grid(cx, cy, a, b)
    Z = S.dense_matrix(A.generate(grid, 2 * np.pi, (3, 3)), grid)
    x = _random(grid.count)
    op = rt.Operator(p, precision='mixed')
    assert rel(op.apply(x), Z @ x) < 1e-12
    assert rel(op.apply(x, precision='c64'), Z @ x) < 1e-5


def test_multiple_rhs_and_modifier(rt):
    p = _problem(rt, 9, 7, 0.1, 0.12, (4, 4), k=4.0)
    grid = A.Grid(9, 7, 0.1, 0.12)
    gen = A.generate(grid, 4.0, (4, 4))
    Z = S.dense_matrix(gen, grid)
    X = _random(grid.count, 3, nrhs=5)
    Y = rt.Operator(p).apply(X)
    for r in range(5):
        assert rel(Y[:, r], Z @ X[:, r]) < 1e-12
    Yi = rt.Operator(p, modifier='imag').apply(X[:, 0])
    assert rel(Yi, np.imag(Z) @ X[:, 0]) < 1e-12


def test_plate_256_matches_oracle_split_fft(rt):
    """The 256 x 256 plate (lambda/10) at production orders against the oracle's split-FFT product."""
    n = 256
    print('plate 256', flush=True)
    p = _problem(rt, n, n, 0.1, 0.1, (16, 8))
    grid = A.Grid(n, n, 0.1, 0.1)
    gen = A.generate(grid, 2 * np.pi, (16, 8))
    gpu = p.generators()
    for key in ('magneticXX', 'magneticYY', 'electric'):
        assert rel(gpu[key], gen[key]) < 1e-13
    x = _random(grid.count, 5)
    ref = S.split_fft_product(S.subband_spectra(gen, grid), grid, x)
    op = rt.Operator(p, precision='mixed')
    assert rel(op.apply(x), ref) < 1e-12
    assert rel(op.apply(x, precision='c64'), ref) < 2e-6


def test_masked_product_is_restriction(rt):
    cx, cy = 10, 8
    mask = np.ones((cx, cy), bool)
    mask[3:6, 2:5] = False
    mask[0, :] = False
    p = _problem(rt, cx, cy, 0.1, 0.1, (4, 4), mask=mask)
    grid = A.Grid(cx, cy, 0.1, 0.1)
    Z = S.dense_matrix(A.generate(grid, 2 * np.pi, (4, 4)), grid)
    ex = np.concatenate([(mask[:-1] & mask[1:]).ravel(order='F'), (mask[:, :-1] & mask[:, 1:]).ravel(order='F')])
    x = _random(grid.count, 7)
    expected = ex * (Z @ (ex * x))
    assert rel(rt.Operator(p).apply(x), expected) < 1e-12


def test_apply_checks_the_input(rt):
    op = rt.Operator(_problem(rt, 5, 4, 0.1, 0.15, (3, 3)))
    for bad in (np.ones(3), np.ones((1, 31)), np.array(['a'] * 31)):
        with pytest.raises(rt.RooftopError) as e:
            op.apply(bad)
        assert e.value.identifier == 'rooftopMoM:product:sizeMismatch'
    assert op.memory() > 0 and op.workingMemory() > 16 * 31


def test_handles_may_be_destroyed_in_any_order(rt):
    """A product keeps its problem alive: destroying the problem handle first leaves the product usable."""
    p = _problem(rt, 6, 5, 0.1, 0.1, (3, 3))
    op = rt.Operator(p)
    x = _random(p.rooftopCount(), 9)
    before = op.apply(x)
    rt.library().rt_problem_destroy(p.handle)
    p.handle = None
    assert rel(op.apply(x), before) == 0
    rt.library().rt_product_destroy(op.handle)
    op.handle = None


def _oracle_product(p, cx, cy, a, b, x):
    """The oracle's split-FFT product on the GPU's own generators (isolates the product from the assembly)."""
    grid = A.Grid(cx, cy, a, b)
    spectra = S.subband_spectra(p.generators(), grid)
    return np.stack([S.split_fft_product(spectra, grid, x[:, r]) for r in range(x.shape[1])], axis=1)


@pytest.mark.parametrize('size', [(17, 33), (64, 3), (3, 100), (130, 40), (1, 40), (40, 1), (7, 200), (20, 300)])
def test_product_shapes_across_transform_lengths(rt, size):
    """Every launch shape the autotuner may pick (both plans, all column groups, tiles and pass B forms that fit:
    both columns in LDS, a team per column, the columns in turn through one column of LDS) gives the same product, on
    grids whose embeddings cover transform lengths 1 .. 1024 (a team per column needs y transforms of at least 512
    points, one column of LDS at least two rounds)."""
    cx, cy = size
    print('shapes', size, flush=True)
    p = _problem(rt, cx, cy, 0.1, 0.1, (3, 3))
    x = _random(p.rooftopCount(), 9, nrhs=3)
    ref = _oracle_product(p, cx, cy, 0.1, 0.1, x)
    op = rt.Operator(p, precision='mixed')
    for prec, tol in (('c128', 1e-12), ('c64', 5e-6)):
        for bits in (3, 4):
            for group in (1, 2, 4, 8, 16, -2, -4):
                for tile, teams, chunks in ((1, 1, 1), (4, 1, 1), (16, 1, 1), (1, 2, 1), (2, 2, 1), (1, 1, 2),
                                            (1, 2, 3), (1, 3, 1), (4, 3, 1), (1, 3, 2)):
                    try:
                        op.setShape({'bits': bits, 'group': group, 'tile': tile, 'teams': teams, 'chunks': chunks},
                                    nrhs=3, precision=prec)
                    except rt.RooftopError:
                        continue   # does not fit this product
                    assert rel(op.apply(x, precision=prec), ref) < tol, (prec, bits, group, tile, teams, chunks)


def test_plate_512_product_and_tuning(rt):
    """The 512 x 512 plate of the batched configuration (transforms of 1024): both precisions against the oracle,
    before and after tuning for the batch."""
    n = 512
    print('plate 512', flush=True)
    p = _problem(rt, n, n, 0.1, 0.1, (4, 4))
    x = _random(p.rooftopCount(), 11, nrhs=2)
    ref = _oracle_product(p, n, n, 0.1, 0.1, x)
    op = rt.Operator(p, precision='mixed')
    assert rel(op.apply(x), ref) < 1e-12
    assert rel(op.apply(x, precision='c64'), ref) < 5e-6
    for prec in ('c128', 'c64'):
        shape, ms = op.tune(nrhs=2, precision=prec)
        assert shape['bits'] in (3, 4) and ms > 0
    assert rel(op.apply(x), ref) < 1e-12
    assert rel(op.apply(x, precision='c64'), ref) < 5e-6


def test_long_columns_three_launch_pass_b(rt):
    """8192-point complex128 y transforms (cy = 4096): the two columns of a position exceed one block's LDS, so pass B
    runs as forward transforms, multiply and inverse transforms in three launches; both precisions against the
    oracle."""
    cx, cy = 3, 4096
    print('long columns', flush=True)
    p = _problem(rt, cx, cy, 0.1, 0.1, (3, 3))
    x = _random(p.rooftopCount(), 13, nrhs=2)
    ref = _oracle_product(p, cx, cy, 0.1, 0.1, x)
    op = rt.Operator(p, precision='mixed')
    assert rel(op.apply(x), ref) < 1e-12
    assert rel(op.apply(x, precision='c64'), ref) < 5e-6
    # The same columns through one column of LDS per position (one launch), in both plans and precisions.
    for prec, tol in (('c128', 1e-12), ('c64', 5e-6)):
        for bits in (3, 4):
            op.setShape({'bits': bits, 'group': 1, 'tile': 1, 'teams': 3, 'chunks': 1}, nrhs=2, precision=prec)
            assert rel(op.apply(x, precision=prec), ref) < tol, (prec, bits)


def test_long_rows_columns_in_turn(rt):
    """8192-point x transforms (cx = 4096): passes A and C with two or four columns per block in turn through one
    column of LDS, against the side-by-side form and the oracle, in both plans and precisions, masked and not."""
    cx, cy = 4096, 3
    print('long rows', flush=True)
    for masked in (False, True):
        mask = None
        if masked:
            mask = np.ones((cx, cy), bool)
            mask[1000:1400, 1] = False
        p = _problem(rt, cx, cy, 0.1, 0.1, (3, 3), mask=mask)
        x = _random(p.rooftopCount(), 14, nrhs=3)
        ref = _oracle_product(p, cx, cy, 0.1, 0.1, x) if not masked else None
        op = rt.Operator(p, precision='mixed')
        for prec, tol in (('c128', 1e-12), ('c64', 5e-6)):
            base = op.apply(x, precision=prec)
            if ref is not None:
                assert rel(base, ref) < tol, prec
            for bits in (3, 4):
                for group in (2, -2, -4):
                    try:
                        op.setShape({'bits': bits, 'group': group, 'tile': 1, 'teams': 1, 'chunks': 1}, nrhs=3,
                                    precision=prec)
                    except rt.RooftopError:
                        assert group == 2   # two teams of 512 or 1024 threads may not fit a block side by side
                        continue
                    assert rel(op.apply(x, precision=prec), base) < tol, (masked, prec, bits, group)
