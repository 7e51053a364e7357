"""Spectra, the FFT / split-FFT products, the dense matrix and the spatial blocks (oracle; test infrastructure only).

FFTs are scipy.fft (pocketfft); `workers` sets its thread count for the timed CPU baseline.
"""

import numpy as np
import scipy.fft as sfft

from .assembly import Grid

WORKERS = 1


def set_workers(count):
    """Thread count of the oracle's FFTs (1 by default)."""
    global WORKERS
    WORKERS = int(count)


def _fft(x, axis):
    return sfft.fft(x, axis=axis, workers=WORKERS)


def _ifft(x, axis):
    return sfft.ifft(x, axis=axis, workers=WORKERS)


def _fft2(x, shape=None):
    return sfft.fft2(x, s=shape, workers=WORKERS)


def _ifft2(x):
    return sfft.ifft2(x, workers=WORKERS)


def fast_fft_size(count):
    """Smallest size >= count with no prime factor above 7. +greensMatrix/+utils/generateFastFFTSize.m:11-15."""
    n = int(count)
    while True:
        m = n
        for p in (2, 3, 5, 7):
            while m % p == 0:
                m //= p
        if m == 1:
            return n
        n += 1


def shift_phases(size_x, size_y, parity_x, parity_y):
    """exp(1i*pi*(2m+parity)/size). +greensMatrix/+utils/generateShiftPhases.m:16-17."""
    sx = np.exp(1j * np.pi * (2 * np.arange(size_x) + parity_x) / size_x)[:, None]
    sy = np.exp(1j * np.pi * (2 * np.arange(size_y) + parity_y) / size_y)[None, :]
    return sx, sy


def subband_modulation(fft_size):
    """exp(-1i*pi*d/fftSize). +greensMatrix/+utils/generateSubbandModulation.m:12-15 (as a flat vector)."""
    return np.exp(-1j * np.pi * np.arange(fft_size) / fft_size)


def block_spectra(mxx, myy, e, content, face_x, face_y, sx, sy):
    """2x2 block spectra (XX, XY, YX, YY; None for a missing direction). +greensMatrix/+utils/generateBlockSpectra.m:19-50.

    Pinned through the products; the three-plane identity out_x = Axx X - Dx E (conj(Dx) X + conj(Dy) Y) used by the
    HIP kernels is this algebra with Dx = (1 - sx)/a.
    """
    mxx = 0 if mxx is None else mxx
    myy = 0 if myy is None else myy
    has_x = content != 'YONLY'
    has_y = content != 'XONLY'
    xx = mxx - (np.abs(1 - sx) / face_x) ** 2 * e if has_x else None
    yy = myy - (np.abs(1 - sy) / face_y) ** 2 * e if has_y else None
    xy = yx = None
    if has_x and has_y:
        xy = ((1 - sx) / (face_x * face_y)) * (np.conj(sy) - 1) * e
        yx = ((np.conj(sx) - 1) / (face_x * face_y)) * (1 - sy) * e
    return {'XX': xx, 'XY': xy, 'YX': yx, 'YY': yy}


def _pad(g, rows, cols):
    """+assembler/+utils/padWithZeros.m:12-20."""
    if g is None:
        return None
    out = np.zeros((rows, cols), dtype=np.complex128)
    r, c = min(g.shape[0], rows), min(g.shape[1], cols)
    out[:r, :c] = g[:r, :c]
    return out


def embedding_spectra(generators, grid):
    """Spectra of the period-(2n-2) circulant embedding (the 'fft' product).

    Follows +greensMatrix/+utils/generateEmbeddingSpectra.m:15-40.
    """
    e = generators['electric']
    cx, cy = e.shape
    mx = np.r_[np.arange(cx), np.arange(cx - 2, 0, -1)]
    my = np.r_[np.arange(cy), np.arange(cy - 2, 0, -1)]

    def embed(g):
        g = _pad(g, cx, cy)
        return None if g is None else _fft2(g[np.ix_(mx, my)])

    sx, sy = shift_phases(mx.size, my.size, 0, 0)
    return block_spectra(embed(generators['magneticXX']), embed(generators['magneticYY']), embed(e),
                         grid.content, grid.size_x, grid.size_y, sx, sy)


def fft_product(spectra, grid, x):
    """The FFTApplicator product. +greensMatrix/FFTApplicator.m:14-87.

    Pinned: reference.mat productFFT at 1e-13 (tests/+reference/+assembler/+greensMatrix/ApplicatorTest.m:5-21) and
    the dense matrix at 1e-12 (tests/+unit/+assembler/+greensMatrix/ApplicatorTest.m:7-29).
    """
    x = np.asarray(x, dtype=np.complex128)
    content = grid.content
    if content == 'BOTH':
        rx, cxx = grid.grid_size('x')
        ry, cyy = grid.grid_size('y')
        nx = rx * cxx
        shape = spectra['XX'].shape
        spec_x = _fft2(x[:nx].reshape(rx, cxx, order='F'), shape)
        ox = spectra['XX'] * spec_x
        oy = spectra['YX'] * spec_x
        spec_y = _fft2(x[nx:].reshape(ry, cyy, order='F'), shape)
        ox = ox + spectra['XY'] * spec_y
        oy = oy + spectra['YY'] * spec_y
        ox = _ifft2(ox)[:rx, :cxx]
        oy = _ifft2(oy)[:ry, :cyy]
        return np.concatenate([ox.ravel(order='F'), oy.ravel(order='F')])
    d = 'x' if content == 'XONLY' else 'y'
    block = spectra['XX'] if d == 'x' else spectra['YY']
    r, c = grid.grid_size(d)
    out = _ifft2(block * _fft2(x.reshape(r, c, order='F'), block.shape))
    return out[:r, :c].ravel(order='F')


def subband_spectra(generators, grid):
    """Subband block spectra of the split-FFT product (list of 4, or 2 for a strip).

    Follows +greensMatrix/+utils/generateSubbandSpectra.m:20-77. The undecimated array is the fft2 of the even
    embedding of period 2*fast(n) (offset fast(n) zero), divided by the number of subbands.
    """
    e = generators['electric']
    cx, cy = e.shape
    fx = fast_fft_size(cx) if cx > 1 else 1
    fy = fast_fft_size(cy) if cy > 1 else 1
    step_x = 2 if fx > 1 else 1
    step_y = 2 if fy > 1 else 1
    parities = [(px, py) for py in range(step_y) for px in range(step_x)]   # ndgrid order: x fastest
    kernels = [generators['magneticXX'], generators['magneticYY'], e]
    subbands = [[None] * len(parities) for _ in range(3)]
    for n, g in enumerate(kernels):
        if g is None or g.size == 0:
            continue
        rows, cols = g.shape
        emb = np.zeros((step_x * fx, step_y * fy), dtype=np.complex128)
        emb[:rows, :cols] = g / len(parities)
        if rows > 1:
            emb[emb.shape[0] - rows + 1:, :cols] = emb[rows - 1:0:-1, :cols]
        if cols > 1:
            emb[:, emb.shape[1] - cols + 1:] = emb[:, cols - 1:0:-1]
        emb = _fft2(emb)
        if np.all(np.imag(g) == 0):
            emb = np.real(emb).astype(np.complex128)
        for k, (px, py) in enumerate(parities):
            subbands[n][k] = emb[px::step_x, py::step_y]
    out = []
    for k, (px, py) in enumerate(parities):
        sx, sy = shift_phases(fx, fy, px, py)
        out.append(block_spectra(subbands[0][k], subbands[1][k], subbands[2][k], grid.content,
                                 grid.size_x, grid.size_y, sx, sy))
    return out


def split_fft_product(spectra, grid, x):
    """The SplitFFTApplicator product (24 1D FFT passes). @SplitFFTApplicator/apply.m:10-19,
    applyBothDirections_.m:14-85, applySingleDirection_.m:14-26.

    Pinned: reference.mat productSplit / productLean at 1e-13 and the dense matrix at 1e-12.
    """
    x = np.asarray(x, dtype=np.complex128)
    content = grid.content
    fx, fy = spectra[0]['XX' if content != 'YONLY' else 'YY'].shape
    if content != 'BOTH':
        d = 'x' if content == 'XONLY' else 'y'
        key = 'XX' if d == 'x' else 'YY'
        even = spectra[0][key].ravel()
        odd = spectra[1][key].ravel()
        mod = subband_modulation(even.size)
        count = x.size
        buf = np.zeros(even.size, dtype=np.complex128)
        buf[:count] = x
        out = _ifft(even * _fft(buf, 0), 0)
        out = out + _ifft(odd * _fft(buf * mod, 0), 0) * np.conj(mod)
        return out[:count]
    rx, cxx = grid.grid_size('x')
    ry, cyy = grid.grid_size('y')
    nx = rx * cxx
    inx = np.zeros((fx, fy), dtype=np.complex128)
    inx[:rx, :cxx] = x[:nx].reshape(rx, cxx, order='F')
    iny = np.zeros((fx, fy), dtype=np.complex128)
    iny[:ry, :cyy] = x[nx:].reshape(ry, cyy, order='F')
    mod_x = subband_modulation(fx)[:, None]
    mod_y = subband_modulation(fy)[None, :]
    out_x = out_y = None
    for parity_y in range(2):
        if parity_y:
            inx = inx * mod_y
            iny = iny * mod_y
        spec_x = _fft(inx, 1)
        spec_y = _fft(iny, 1)
        so_x = so_y = None
        for parity_x in range(2):
            if parity_x:
                spec_x = spec_x * mod_x
                spec_y = spec_y * mod_x
            sub_x = _fft(spec_x, 0)
            sub_y = _fft(spec_y, 0)
            b = spectra[parity_x + 2 * parity_y]
            o_x = _ifft(b['XX'] * sub_x + b['XY'] * sub_y, 0)
            o_y = _ifft(b['YX'] * sub_x + b['YY'] * sub_y, 0)
            if parity_x:
                so_x = so_x + o_x * np.conj(mod_x)
                so_y = so_y + o_y * np.conj(mod_x)
            else:
                so_x, so_y = o_x, o_y
        so_x = _ifft(so_x, 1)
        so_y = _ifft(so_y, 1)
        if parity_y:
            out_x = out_x + so_x * np.conj(mod_y)
            out_y = out_y + so_y * np.conj(mod_y)
        else:
            out_x, out_y = so_x, so_y
    return np.concatenate([out_x[:rx, :cxx].ravel(order='F'), out_y[:ry, :cyy].ravel(order='F')])


def dense_matrix(generators, grid):
    """The dense Green's matrix by definition: blkdiag(Mx, My) - D.' E D.

    Follows tests/+unit/Base.m:83-126 (generateDenseMatrix, toeplitz_), the principal oracle of the reference's product
    tests. Pinned: reference.mat impedanceMatrix (independent v2 assembly) at 1e-10, and the three reference products.
    """
    cx, cy = grid.count_x, grid.count_y
    fx, fy = np.meshgrid(np.arange(cx), np.arange(cy), indexing='ij')
    faces = np.vstack([fx.ravel(order='F'), fy.ravel(order='F')])

    def toeplitz(g, off):
        return g[np.abs(off[0][:, None] - off[0][None, :]), np.abs(off[1][:, None] - off[1][None, :])]

    face_matrix = toeplitz(generators['electric'], faces)
    blocks = []
    div_cols = []
    for d, key, size, second in (('x', 'magneticXX', grid.size_x, (1, 0)), ('y', 'magneticYY', grid.size_y, (0, 1))):
        count = grid.count_dir(d)
        if count == 0:
            continue
        first = grid.face_offsets(d)
        sec = first + np.array(second)[:, None]
        blocks.append(toeplitz(generators[key], first))
        div = np.zeros((cx * cy, count))
        div[first[0] + cx * first[1], np.arange(count)] = 1.0 / size
        div[sec[0] + cx * sec[1], np.arange(count)] = -1.0 / size
        div_cols.append(div)
    n = sum(b.shape[0] for b in blocks)
    mag = np.zeros((n, n), dtype=np.complex128)
    o = 0
    for b in blocks:
        mag[o:o + b.shape[0], o:o + b.shape[0]] = b
        o += b.shape[0]
    div = np.hstack(div_cols)
    return mag - div.T @ face_matrix @ div


def extend_evenly(g):
    """Two-sided generator (2c-1 per axis, offset 0 in the middle). +preconditioner/+utils/extendEvenly.m:12-13."""
    cx, cy = g.shape
    ix = np.r_[np.arange(cx - 1, 0, -1), np.arange(cx)]
    iy = np.r_[np.arange(cy - 1, 0, -1), np.arange(cy)]
    return g[np.ix_(ix, iy)]


def spatial_blocks(generators, grid):
    """Two-sided generators of the four blocks on the face grid: (XX, XY, YX, YY), None for a missing direction.

    Follows +preconditioner/+utils/generateSpatialBlocks.m:19-61. Entry Z(r, s) is the block generator at the offset
    between the first faces of r and s, index offset + count - 1.
    Pinned: the blocks reproduce the dense matrix at 1e-13 * max (GenerateSpatialBlocksTest.m:10-33).
    """
    e = generators['electric']
    cx, cy = e.shape
    a, b = grid.size_x, grid.size_y
    has_x = grid.content != 'YONLY'
    has_y = grid.content != 'XONLY'
    bordered = np.zeros((2 * cx + 1, 2 * cy + 1), dtype=np.complex128)
    bordered[1:-1, 1:-1] = extend_evenly(e)

    def shifted(da, db):
        return bordered[1 + da:2 * cx + da, 1 + db:2 * cy + db]

    p = np.arange(-(cx - 1), cx)[:, None]
    q = np.arange(-(cy - 1), cy)[None, :]
    xx = xy = yx = yy = None
    if has_x:
        xx = extend_evenly(_pad(generators['magneticXX'], cx, cy)) - (2 * shifted(0, 0) - shifted(-1, 0)
                                                                      - shifted(1, 0)) / a ** 2
        xx[(np.abs(p) > cx - 2).ravel(), :] = 0
    if has_y:
        yy = extend_evenly(_pad(generators['magneticYY'], cx, cy)) - (2 * shifted(0, 0) - shifted(0, -1)
                                                                      - shifted(0, 1)) / b ** 2
        yy[:, (np.abs(q) > cy - 2).ravel()] = 0
    if has_x and has_y:
        xy = -(shifted(0, 0) - shifted(0, -1) - shifted(1, 0) + shifted(1, -1)) / (a * b)
        xy[(p > cx - 2).ravel(), :] = 0
        xy[:, (q < -(cy - 2)).ravel()] = 0
        yx = np.rot90(xy, 2).copy()
    return xx, xy, yx, yy
