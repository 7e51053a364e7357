"""Circulant preconditioners: the kernels' folds, the superoptimal kernel, the partwise, blockwise and hybrid builders,
the inverse and its product (oracle; test infrastructure only).

Follows src/+rooftopMoM/+assembler/+preconditioner/: generateCirculant.m:16-24, generatePartwiseSpectra.m:16-30,
generateBlockwiseSpectra.m, generateHybridSpectra.m, +kernels/generate{Strang,TChan,RChan,Huckle,Superoptimal}.m,
+utils/foldGenerator.m:27-61, +utils/generatePartSpectra.m, +utils/generateOptimalOfProduct.m and
@Spectra/invert.m:12-20, applied by the FFTApplicator of the face grid (FFTApplicator.m:14-87).

Pinned: reference.mat preconditioned{1,2} = partwise Strang and T. Chan at 1e-12 (GenerateCirculantTest.m:5-25); the
folds against dense circulants built entry by entry, T. Chan's Frobenius optimality, Huckle at full bandwidth = T. Chan,
R. Chan of a circulant = that circulant, the inverse blocks frequency by frequency; the hybrid builder from those pinned
parts (its coupling is the partwise one, its diagonal blocks folds of the spatial blocks, pinned against the dense
matrix). The superoptimal kernel against its definition, the eigenvalues of c(T T^H) / c(T^H) from dense matrices in
the Fourier basis (KernelsTest.m:104-115), and its evenness for an even generator (KernelsTest.m:117-123); the
blockwise builder's transposed coupling against the dense transpose, and every kernel and builder's preconditioner
against complex symmetry (PreconditionerTest.m).
"""

import numpy as np

from . import spectra as S


# --- windows of the kernels (weights of the offsets -(count-1) .. count-1) ------------------------------------------

def window_strang(offsets, count, dimension):
    """+kernels/generateStrang.m:20-22: offsets below half the period; half weight at exactly half."""
    a = np.abs(offsets)
    return (a < count / 2).astype(float) + 0.5 * (a == count / 2)


def window_tchan(offsets, count, dimension):
    """+kernels/generateTChan.m:22-24: (count - |p|) / count, how often the offset occurs."""
    return (count - np.abs(offsets)) / count


def window_rchan(offsets, count, dimension):
    """+kernels/generateRChan.m: every offset with weight one."""
    return np.ones(np.shape(offsets))


def window_huckle(bandwidths):
    """+kernels/generateHuckle.m: max(0, 1 - |offset| / bandwidth) per direction."""
    def window(offsets, count, dimension):
        return np.maximum(0.0, 1.0 - np.abs(offsets) / bandwidths[dimension - 1])
    return window


def fold_generator(generator, window, even=False):
    """The generator of the circulant of period count per axis from a Toeplitz generator: every offset weighted by
    window(offsets, count, dimension) and added to its offset modulo the period (+utils/foldGenerator.m:27-61).
    generator: two-sided (2 count - 1 per axis, offset 0 in the middle), or with even=True the one-sided
    count_x x count_y generator of an even Toeplitz matrix."""
    g = np.asarray(generator, dtype=np.complex128)
    if even:
        cx, cy = g.shape
    else:
        cx, cy = (g.shape[0] + 1) // 2, (g.shape[1] + 1) // 2
    wx = np.asarray(window(np.arange(-(cx - 1), cx), cx, 1), dtype=float).reshape(-1, 1)
    wy = np.asarray(window(np.arange(-(cy - 1), cy), cy, 2), dtype=float).reshape(1, -1)
    if even:
        # The offset k keeps w(k) and gains its partner k - count, held by the even generator at count - k.
        px = np.r_[0.0, wx[:cx - 1, 0]].reshape(-1, 1)
        py = np.r_[0.0, wy[0, :cy - 1]].reshape(1, -1)
        rx = np.r_[0, np.arange(cx - 1, 0, -1)]
        ry = np.r_[0, np.arange(cy - 1, 0, -1)]
        c = wx[cx - 1:] * g + px * g[rx, :]
        return c * wy[:, cy - 1:] + c[:, ry] * py
    weighted = wx * g * wy

    def fold_rows(h):
        count = (h.shape[0] + 1) // 2
        out = h[count - 1:, :].copy()
        out[1:, :] += h[:count - 1, :]
        return out

    return fold_rows(fold_rows(weighted).T).T


# --- the superoptimal kernel of Tyrtyshnikov --------------------------------------------------------------------

def part_spectra(t):
    """+utils/generatePartSpectra.m: along each direction a Toeplitz matrix is a circulant plus a skew-circulant,
    c(k) = (t(k) + t(k - count)) / 2, s(k) = (t(k) - t(k - count)) / 2, k = 0 .. count - 1 (t(-count) = 0); the skew
    part is diagonal in the Fourier basis modulated by exp(i pi k / count). Returns {(m, n): eigenvalues}, m along x
    and n along y, 0 circulant and 1 skew."""
    t = np.asarray(t, dtype=np.complex128)
    cx, cy = (t.shape[0] + 1) // 2, (t.shape[1] + 1) // 2

    def split(h):   # along axis 0; row k of a part holds its generator at the offset k
        count = (h.shape[0] + 1) // 2
        positive = h[count - 1:]
        negative = np.vstack([np.zeros((1, h.shape[1]), dtype=np.complex128), h[:count - 1]])
        return (positive + negative) / 2, (positive - negative) / 2

    mod_x = np.exp(1j * np.pi * np.arange(cx) / cx)[:, None]
    mod_y = np.exp(1j * np.pi * np.arange(cy) / cy)[None, :]
    out = {}
    for m, along_x in enumerate(split(t)):
        for n, along_y in enumerate(split(along_x.T)):
            part = along_y.T
            if m == 1:
                part = mod_x * part
            if n == 1:
                part = part * mod_y
            out[m, n] = np.fft.fft2(part)
    return out


def _skew_kernel(count):
    """The spectrum of K(d) = 1 / (count^2 sin^2(pi (2 d + 1) / (2 count))) (generateOptimalOfProduct.m)."""
    d = np.arange(count)
    return np.fft.fft(1.0 / (count ** 2 * np.sin(np.pi * (2 * d + 1) / (2 * count)) ** 2))


def _convolve(values, kernel, axis):
    if axis == 0:
        return np.fft.ifft(kernel[:, None] * np.fft.fft(values, axis=0), axis=0)
    return np.fft.ifft(kernel[None, :] * np.fft.fft(values, axis=1), axis=1)


def optimal_of_product(parts_a, parts_b):
    """+utils/generateOptimalOfProduct.m: the eigenvalues of T. Chan's c(A B^H) from the part spectra, without the
    product: along a direction a circulant factor passes through c, a skew-circulant one is projected by the circular
    convolution of its eigenvalues with K; every pair of parts adds the A part's eigenvalues times the conjugate of the
    B part's, convolved first where one of the two is skew, the product convolved where both are."""
    cx, cy = parts_a[0, 0].shape
    kernels = (_skew_kernel(cx), _skew_kernel(cy))
    types = ((0, 0), (1, 0), (0, 1), (1, 1))
    total = np.zeros((cx, cy), dtype=np.complex128)
    for ta in types:
        for tb in types:
            va, vb = parts_a[ta], parts_b[tb]
            for d in range(2):
                if ta[d] == 1 and tb[d] == 0:
                    va = _convolve(va, kernels[d], d)
                if ta[d] == 0 and tb[d] == 1:
                    vb = _convolve(vb, kernels[d], d)
            product = va * np.conj(vb)
            for d in range(2):
                if ta[d] == 1 and tb[d] == 1:
                    product = _convolve(product, kernels[d], d)
            total = total + product
    return total


def superoptimal(t):
    """+kernels/generateSuperoptimal.m: the circulant C minimizing |I - C^-1 T|_F, C = c(T T^H) / c(T^H), from a
    two-sided generator; c(T^H) has the conjugate eigenvalues of c(T). Returns the count_x x count_y generator."""
    parts = part_spectra(t)
    squared = optimal_of_product(parts, parts)
    optimal = np.fft.fft2(fold_generator(t, window_tchan))
    return np.fft.ifft2(squared / np.conj(optimal))


# --- partwise builder, inverse, product -------------------------------------------------------------------------

KERNELS = ('strang', 'tChan', 'rChan', 'huckle', 'superoptimal')
BUILDERS = ('partwise', 'blockwise', 'hybrid')


def approximate(generator, kernel, count_x, count_y, bandwidth=None):
    """fft2 of the kernel's circulant of one generator zero-padded to the face grid (generatePartwiseSpectra.m:21-30:
    T. Chan and R. Chan fold the one-sided even generator, Strang and Huckle its even two-sided extension)."""
    if generator is None:
        return None
    g = S._pad(generator, count_x, count_y)
    if kernel == 'tChan':
        c = fold_generator(g, window_tchan, even=True)
    elif kernel == 'rChan':
        c = fold_generator(g, window_rchan, even=True)
    elif kernel == 'strang':
        c = fold_generator(S.extend_evenly(g), window_strang)
    elif kernel == 'huckle':
        bw = bandwidth if bandwidth is not None else (count_x, count_y)
        c = fold_generator(S.extend_evenly(g), window_huckle(bw))
    elif kernel == 'superoptimal':
        c = superoptimal(S.extend_evenly(g))
    else:
        raise ValueError(kernel)
    return np.fft.fft2(c)


def partwise_spectra(generators, grid, kernel, bandwidth=None):
    """The block spectra of the partwise circulant approximation on the face grid."""
    cx, cy = generators['electric'].shape
    sx, sy = S.shift_phases(cx, cy, 0, 0)
    return S.block_spectra(approximate(generators['magneticXX'], kernel, cx, cy, bandwidth),
                           approximate(generators['magneticYY'], kernel, cx, cy, bandwidth),
                           approximate(generators['electric'], kernel, cx, cy, bandwidth),
                           grid.content, grid.size_x, grid.size_y, sx, sy)


def _fold_two_sided(t, kernel, bandwidth, count_x, count_y):
    """A kernel applied to a two-sided generator (the blocks of the blockwise and hybrid builders)."""
    if t is None:
        return None
    if kernel == 'superoptimal':
        return superoptimal(t)
    windows = {'strang': window_strang, 'tChan': window_tchan, 'rChan': window_rchan,
               'huckle': window_huckle(bandwidth if bandwidth is not None else (count_x, count_y))}
    return fold_generator(t, windows[kernel])


def negate_frequencies(spectrum):
    """The spectrum at the negative frequencies (generateBlockwiseSpectra.m transpose_): the transposed circulant has
    the generator c(-k, -l)."""
    if spectrum is None:
        return None
    rows = np.r_[0, np.arange(spectrum.shape[0] - 1, 0, -1)]
    cols = np.r_[0, np.arange(spectrum.shape[1] - 1, 0, -1)]
    return spectrum[rows][:, cols]


def blockwise_spectra(generators, grid, kernel, bandwidth=None):
    """generateBlockwiseSpectra.m: each of the spatial blocks (generateSpatialBlocks.m) replaced by its kernel's
    circulant; the YX block is the transpose of the XY block, and so is its circulant."""
    cx, cy = generators['electric'].shape
    xx, xy, _, yy = S.spatial_blocks(generators, grid)

    def spectrum(t):
        c = _fold_two_sided(t, kernel, bandwidth, cx, cy)
        return None if c is None else np.fft.fft2(c)

    coupling = spectrum(xy)
    return {'XX': spectrum(xx), 'XY': coupling, 'YX': negate_frequencies(coupling), 'YY': spectrum(yy)}


def hybrid_spectra(generators, grid, kernel, bandwidth=None):
    """generateHybridSpectra.m: the kernel's circulants of the spatial diagonal blocks (generateSpatialBlocks.m) and
    the coupling of the electric generator's circulant through the block spectra."""
    e = generators['electric']
    cx, cy = e.shape
    xx, _, _, yy = S.spatial_blocks(generators, grid)
    sx, sy = S.shift_phases(cx, cy, 0, 0)
    electric = np.fft.fft2(_fold_two_sided(S.extend_evenly(e), kernel, bandwidth, cx, cy))
    coupling = S.block_spectra(None, None, electric, grid.content, grid.size_x, grid.size_y, sx, sy)
    fx = _fold_two_sided(xx, kernel, bandwidth, cx, cy)
    fy = _fold_two_sided(yy, kernel, bandwidth, cx, cy)
    return {'XX': None if fx is None else np.fft.fft2(fx), 'XY': coupling['XY'], 'YX': coupling['YX'],
            'YY': None if fy is None else np.fft.fft2(fy)}


def invert(spectra, content):
    """@Spectra/invert.m:12-20: the 2x2 inverse per frequency, or the reciprocal of the single block of a strip."""
    with np.errstate(divide='ignore', invalid='ignore'):
        if content == 'XONLY':
            return {'XX': 1 / spectra['XX'], 'XY': None, 'YX': None, 'YY': None}
        if content == 'YONLY':
            return {'XX': None, 'XY': None, 'YX': None, 'YY': 1 / spectra['YY']}
        det = spectra['XX'] * spectra['YY'] - spectra['XY'] * spectra['YX']
        return {'XX': spectra['YY'] / det, 'XY': -spectra['XY'] / det, 'YX': -spectra['YX'] / det,
                'YY': spectra['XX'] / det}


def is_singular(inverse):
    """Preconditioner.m:57-65: a zero determinant leaves Inf or NaN in the inverse."""
    return not all(np.all(np.isfinite(b)) for b in inverse.values() if b is not None)


def circulant_preconditioner(generators, grid, kernel='rChan', bandwidth=None, builder='partwise'):
    """The inverse spectra of generateCirculant.m:16-24 with the partwise, blockwise or hybrid builder."""
    build = {'partwise': partwise_spectra, 'blockwise': blockwise_spectra, 'hybrid': hybrid_spectra}[builder]
    return invert(build(generators, grid, kernel, bandwidth), grid.content)


def apply(inverse, grid, x):
    """The product with the preconditioner: the FFTApplicator of the face grid."""
    return S.fft_product(inverse, grid, x)


def dense_circulant(c):
    """The dense block of a 2D circulant with generator c (count_x x count_y): entry ((i, j), (k, l)) =
    c((i - k) mod count_x, (j - l) mod count_y), both indices x fastest."""
    cx, cy = c.shape
    i = np.arange(cx * cy)
    ix, iy = i % cx, i // cx
    return c[(ix[:, None] - ix[None, :]) % cx, (iy[:, None] - iy[None, :]) % cy]


def dense_toeplitz(two_sided):
    """The dense two-level Toeplitz block of a two-sided generator (offset 0 in the middle), x fastest."""
    cx, cy = (two_sided.shape[0] + 1) // 2, (two_sided.shape[1] + 1) // 2
    i = np.arange(cx * cy)
    ix, iy = i % cx, i // cx
    return two_sided[(ix[:, None] - ix[None, :]) + cx - 1, (iy[:, None] - iy[None, :]) + cy - 1]
