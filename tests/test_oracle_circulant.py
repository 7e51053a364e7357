"""Pins the oracle's circulant preconditioners (CPU): reference.mat, brute-force folds, optimality, the inverse."""
import numpy as np
import pytest

from conftest import rel
from oracle import assembly as A
from oracle import circulant as C
from oracle import spectra as S


def test_partwise_strang_and_tchan_match_reference_mat(reference_data):
    """tests/+reference/+assembler/+preconditioner/GenerateCirculantTest.m:5-25 at 1e-12."""
    for case in reference_data:
        grid = A.Grid(case['countX'], case['countY'], case['sizeX'], case['sizeY'])
        g = {'magneticXX': case['magneticXX'], 'magneticYY': case['magneticYY'], 'electric': case['electric']}
        for kernel, key in (('strang', 'preconditionedStrang'), ('tChan', 'preconditionedTChan')):
            inv = C.circulant_preconditioner(g, grid, kernel)
            assert rel(C.apply(inv, grid, case['vector']), case[key]) < 1e-12, (kernel, grid.content)


def _even_generator(cx, cy, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((cx, cy)) + 1j * rng.standard_normal((cx, cy))


def _brute_fold(two_sided, weight):
    """c(p, q) = sum over the offsets p' = p (mod cx), |p'| < cx, and q' likewise of w(p', q') t(p', q')."""
    cx, cy = (two_sided.shape[0] + 1) // 2, (two_sided.shape[1] + 1) // 2
    c = np.zeros((cx, cy), complex)
    for ox in range(-(cx - 1), cx):
        for oy in range(-(cy - 1), cy):
            c[ox % cx, oy % cy] += weight(ox, oy) * two_sided[ox + cx - 1, oy + cy - 1]
    return c


@pytest.mark.parametrize('shape', [(5, 4), (4, 4), (1, 6), (7, 1), (6, 3)])
def test_folds_match_brute_force(shape):
    """R. Chan wraps every offset, Strang keeps the offsets below half the period (half weight at half), T. Chan
    weights by how often the offset occurs; Huckle at full bandwidth is T. Chan."""
    cx, cy = shape
    g = _even_generator(cx, cy, 3)
    t = S.extend_evenly(g)
    rchan = _brute_fold(t, lambda ox, oy: 1.0)
    assert np.allclose(C.fold_generator(g, C.window_rchan, even=True), rchan, rtol=0, atol=1e-13)
    assert np.allclose(C.fold_generator(t, C.window_rchan), rchan, rtol=0, atol=1e-13)

    def strang_w(o, n):
        return 1.0 if 2 * abs(o) < n else (0.5 if 2 * abs(o) == n else 0.0)
    strang = _brute_fold(t, lambda ox, oy: strang_w(ox, cx) * strang_w(oy, cy))
    assert np.allclose(C.fold_generator(t, C.window_strang), strang, rtol=0, atol=1e-13)
    tchan = _brute_fold(t, lambda ox, oy: (cx - abs(ox)) / cx * (cy - abs(oy)) / cy)
    assert np.allclose(C.fold_generator(g, C.window_tchan, even=True), tchan, rtol=0, atol=1e-13)
    assert np.allclose(C.fold_generator(t, C.window_tchan), tchan, rtol=0, atol=1e-13)
    assert np.allclose(C.fold_generator(t, C.window_huckle((cx, cy))), tchan, rtol=0, atol=1e-13)
    # Huckle with bandwidth 1 keeps the offset 0 only.
    h1 = C.fold_generator(t, C.window_huckle((1, 1)))
    assert h1[0, 0] == g[0, 0] and np.count_nonzero(h1) == 1


def test_tchan_is_the_frobenius_optimal_circulant():
    """T. Chan's circulant is the average of the Toeplitz matrix over every wrapped diagonal, the circulant closest
    in the Frobenius norm: perturbing its generator only moves away."""
    cx, cy = 5, 3
    g = _even_generator(cx, cy, 4)
    T = C.dense_toeplitz(S.extend_evenly(g))
    c = C.fold_generator(g, C.window_tchan, even=True)
    best = np.linalg.norm(T - C.dense_circulant(c))
    rng = np.random.default_rng(5)
    for _ in range(5):
        d = 1e-3 * (rng.standard_normal((cx, cy)) + 1j * rng.standard_normal((cx, cy)))
        assert np.linalg.norm(T - C.dense_circulant(c + d)) > best
    # The average itself, class by class.
    i = np.arange(cx * cy)
    ix, iy = i % cx, i // cx
    cls_x = (ix[:, None] - ix[None, :]) % cx
    cls_y = (iy[:, None] - iy[None, :]) % cy
    for p in range(cx):
        for q in range(cy):
            sel = (cls_x == p) & (cls_y == q)
            assert abs(T[sel].mean() - c[p, q]) < 1e-13


def test_rchan_of_a_circulant_is_itself_and_inverse_blocks():
    """R. Chan of the Toeplitz generator of a circulant returns that circulant; the inverse spectra are the 2x2
    inverses frequency by frequency; a zero block is singular."""
    cx, cy = 4, 6
    rng = np.random.default_rng(6)
    c = rng.standard_normal((cx, cy)) + 1j * rng.standard_normal((cx, cy))
    # A Toeplitz generator holding c at the offsets 0 .. count-1 and zero at the negative ones: every wrapped class
    # is hit once, so R. Chan's fold gives c back.
    t = np.zeros((2 * cx - 1, 2 * cy - 1), complex)
    t[cx - 1:, cy - 1:] = c
    assert np.allclose(C.fold_generator(t, C.window_rchan), c, rtol=0, atol=1e-14)
    grid = A.Grid(7, 5, 0.1, 0.12)
    g = A.generate(grid, 2 * np.pi, (3, 3))
    sp = C.partwise_spectra(g, grid, 'rChan')
    inv = C.invert(sp, grid.content)
    one = sp['XX'] * inv['XX'] + sp['XY'] * inv['YX']
    zero = sp['XX'] * inv['XY'] + sp['XY'] * inv['YY']
    assert np.allclose(one, 1, atol=1e-11) and np.allclose(zero, 0, atol=1e-11)
    assert not C.is_singular(inv)
    zeros = {k: None if v is None else np.zeros_like(v) for k, v in g.items()}
    assert C.is_singular(C.invert(C.partwise_spectra(zeros, grid, 'rChan'), grid.content))


def test_hybrid_builder_from_pinned_parts():
    """generateHybridSpectra.m: the coupling blocks are the partwise ones (the same electric circulant), the
    diagonal blocks the kernel's fold of the spatial blocks (pinned against the dense matrix) by brute force."""
    grid = A.Grid(6, 5, 0.1, 0.12)
    g = A.generate(grid, 2 * np.pi, (3, 3))
    for kernel in ('strang', 'tChan', 'rChan', 'huckle'):
        hy = C.hybrid_spectra(g, grid, kernel, (2, 3) if kernel == 'huckle' else None)
        pw = C.partwise_spectra(g, grid, kernel, (2, 3) if kernel == 'huckle' else None)
        assert np.allclose(hy['XY'], pw['XY'], rtol=0, atol=1e-13 * np.abs(pw['XY']).max())
        assert np.allclose(hy['YX'], pw['YX'], rtol=0, atol=1e-13 * np.abs(pw['YX']).max())
    xx, _, _, yy = S.spatial_blocks(g, grid)
    for blk, key in ((xx, 'XX'), (yy, 'YY')):
        brute = _brute_fold(blk, lambda ox, oy: 1.0)
        assert np.allclose(C.hybrid_spectra(g, grid, 'rChan')[key], np.fft.fft2(brute), rtol=0,
                           atol=1e-12 * np.abs(np.fft.fft2(brute)).max())
    strip = A.Grid(8, 1, 0.1, 0.02)
    gs = A.generate(strip, 2 * np.pi, (3, 3))
    inv = C.circulant_preconditioner(gs, strip, 'tChan', builder='hybrid')
    assert inv['XX'] is not None and inv['YY'] is None and not C.is_singular(inv)


def _fourier(cx, cy):
    """The 2D DFT matrix on vectors indexed x fastest."""
    fx = np.exp(-2j * np.pi * np.outer(np.arange(cx), np.arange(cx)) / cx)
    fy = np.exp(-2j * np.pi * np.outer(np.arange(cy), np.arange(cy)) / cy)
    return np.kron(fy, fx)


def test_superoptimal_matches_its_definition():
    """KernelsTest.m:104-115: the eigenvalues of C are diag(F T T^H F^H) / conj(diag(F T F^H)), from the dense
    two-level Toeplitz matrix of a random complex generator in the Fourier basis."""
    rng = np.random.default_rng(1)
    t = rng.standard_normal((7, 5)) + 1j * rng.standard_normal((7, 5))
    T = C.dense_toeplitz(t)
    F = _fourier(4, 3)
    expected = np.diag(F @ (T @ T.conj().T) @ F.conj().T) / np.conj(np.diag(F @ T @ F.conj().T))
    got = np.fft.fft2(C.superoptimal(t)).reshape(-1, order='F')
    assert np.abs(got - expected).max() < 1e-12 * np.abs(expected).max()


def test_superoptimal_of_an_even_generator_is_even():
    """KernelsTest.m:117-123: C(p, q) = C(-p, -q) for an even generator."""
    rng = np.random.default_rng(1)
    t = S.extend_evenly(rng.standard_normal((5, 4)) + 1j * rng.standard_normal((5, 4)))
    c = C.superoptimal(t)
    mirrored = c[np.r_[0, 4:0:-1]][:, np.r_[0, 3:0:-1]]
    assert np.abs(c - mirrored).max() < 1e-13 * np.abs(c).max()


def test_blockwise_coupling_is_the_transposed_circulant():
    """generateBlockwiseSpectra.m: YX is the transpose of the XY circulant, its spectrum taken at the negative
    frequencies; checked against the dense transpose."""
    grid = A.Grid(5, 4, 0.1, 0.15)
    g = A.generate(grid, 4, (3, 3))
    _, xy, yx, _ = S.spatial_blocks(g, grid)
    for kernel in ('strang', 'rChan', 'superoptimal'):
        sp = C.blockwise_spectra(g, grid, kernel)
        cxy = np.fft.ifft2(sp['XY'])
        cyx = np.fft.ifft2(sp['YX'])
        assert np.abs(C.dense_circulant(cyx) - C.dense_circulant(cxy).T).max() < 1e-13 * np.abs(cxy).max()
    # R. Chan wraps every offset: the XY circulant is the brute-force fold of the spatial XY block.
    brute = _brute_fold(xy, lambda ox, oy: 1.0)
    assert np.allclose(np.fft.ifft2(C.blockwise_spectra(g, grid, 'rChan')['XY']), brute, rtol=0,
                       atol=1e-13 * np.abs(brute).max())


@pytest.mark.parametrize('builder', ['partwise', 'blockwise', 'hybrid'])
def test_every_kernel_and_builder_is_complex_symmetric(builder):
    """PreconditionerTest.m testIsSymmetric on its 5 x 4 problem (wavenumber 4, faces 0.1 x 0.15): x.' P y = y.' P x,
    which the conjugate gradients need."""
    grid = A.Grid(5, 4, 0.1, 0.15)
    g = A.generate(grid, 4, (3, 3))
    rng = np.random.default_rng(1)
    x = rng.standard_normal(grid.count) + 1j * rng.standard_normal(grid.count)
    y = rng.standard_normal(grid.count) + 1j * rng.standard_normal(grid.count)
    for kernel in C.KERNELS:
        inv = C.circulant_preconditioner(g, grid, kernel, (5, 4) if kernel == 'huckle' else None, builder)
        assert not C.is_singular(inv)
        a, b = x @ C.apply(inv, grid, y), y @ C.apply(inv, grid, x)
        assert abs(a - b) < 1e-12 * abs(a), (kernel, builder)


def test_blockwise_approximation_cuts_the_iterations():
    """PreconditionerTest.m testBlockwiseApproximationCutsTheIterations: 16 x 16 plate, R. Chan blockwise, GMRES to
    1e-8 needs fewer iterations than without."""
    from oracle import excitation as X
    from oracle import solvers as SV
    grid = A.Grid(16, 16, 0.05, 0.05)
    g = A.generate(grid, 2 * np.pi, (3, 3))
    Z = S.dense_matrix(g, grid)
    b = X.plane_wave(grid, 2 * np.pi, [0, 0, -1], [1, 0, 0], 1.0, order=3)
    inv = C.circulant_preconditioner(g, grid, 'rChan', builder='blockwise')
    plain = SV.solve('gmres', Z, b, tol=1e-8, max_iter=1000, restart=50)
    pre = SV.solve('gmres', Z, b, tol=1e-8, max_iter=1000, restart=50,
                   preconditioner=lambda v: C.apply(inv, grid, v))
    assert pre['converged'] and pre['iterationCount'] < plain['iterationCount']
