"""Circulant preconditioners: the kernels' folds, the partwise builder, the inverse and its product (oracle; test
infrastructure only).

Follows src/+rooftopMoM/+assembler/+preconditioner/: generateCirculant.m:16-24, generatePartwiseSpectra.m:16-30,
generateHybridSpectra.m, +kernels/generate{Strang,TChan,RChan,Huckle}.m, +utils/foldGenerator.m:27-61 and
@Spectra/invert.m:12-20, applied by the FFTApplicator of the face grid (FFTApplicator.m:14-87).

Pinned: reference.mat preconditioned{1,2} = partwise Strang and T. Chan at 1e-12 (GenerateCirculantTest.m:5-25); the
folds against dense circulants built entry by entry, T. Chan's Frobenius optimality, Huckle at full bandwidth = T. Chan,
R. Chan of a circulant = that circulant, the inverse blocks frequency by frequency; the hybrid builder from those pinned
parts (its coupling is the partwise one, its diagonal blocks folds of the spatial blocks, pinned against the dense
matrix).
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


# --- partwise builder, inverse, product -------------------------------------------------------------------------

KERNELS = ('strang', 'tChan', 'rChan', 'huckle')


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
    windows = {'strang': window_strang, 'tChan': window_tchan, 'rChan': window_rchan,
               'huckle': window_huckle(bandwidth if bandwidth is not None else (count_x, count_y))}
    return fold_generator(t, windows[kernel])


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
    """The inverse spectra of generateCirculant.m:16-24 with the partwise or hybrid builder."""
    build = partwise_spectra if builder == 'partwise' else hybrid_spectra
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
