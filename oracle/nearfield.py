"""Near-field preconditioner on 3x3-cell neighbourhoods and the Jacobi diagonal (oracle; test infrastructure only).

The reference has no near-field preconditioner (its preconditioners are circulant ones, src/+rooftopMoM/
Preconditioner.m:41-51); this one is defined by this project. It is a symmetrized sparse approximate inverse (SAI):

  * the pattern S_i of rooftop i holds every rooftop (of either direction) whose first face lies within one cell of
    the first face of i in x and in y -- the rooftops of the 3x3-cell neighbourhood, at most 18;
  * the fit set F_i holds the rooftops whose first face lies within two cells (the 5x5 window, at most 50);
  * row i of M minimizes || Z[F_i, S_i] m - e_i ||_2 over F_i (complex least squares), so M Z ~ I on the near field;
  * P = (M + M^T) / 2, complex symmetric like Z, so COCG stays valid; it has the pattern of M.

Entries of Z come from the two-sided spatial blocks (generateSpatialBlocks.m:19-61: the entry of two rooftops is the
block generator at the offset between their first faces). Non-overlapping block Jacobi on the same 3x3 tiles
(block_jacobi below) is kept for comparison: on the EFIE it raises the iteration counts, because inverting a truncated
block amplifies the weak loop (divergence-free) modes of each tile.

Pinned by construction: entries against the dense matrix (tests/+unit/Base.m:83-114), each row against numpy's
LAPACK least squares, the symmetry x.' P y = y.' P x (as PreconditionerTest.m:33-41 checks the circulant ones), and
the residual property of the fit.
The Jacobi diagonal is diag(Z), the reference's `@(r) r ./ diagonal` (ConjugateGradientTest.m:55-59).
"""

import numpy as np
import scipy.sparse as sp

from .spectra import spatial_blocks


def rooftop_offsets(grid, mask_x=None, mask_y=None):
    """Per existing rooftop: global index, direction (0 x, 1 y) and first-face offset (fx, fy)."""
    idx, dirs, fx, fy = [], [], [], []
    base = 0
    for n, (d, mask) in enumerate((('x', mask_x), ('y', mask_y))):
        rows, cols = grid.grid_size(d)
        off = grid.face_offsets(d)
        keep = np.ones(rows * cols, dtype=bool) if mask is None else np.asarray(mask).ravel(order='F').astype(bool)
        idx.append(base + np.nonzero(keep)[0])
        dirs.append(np.full(int(keep.sum()), n))
        fx.append(off[0][keep])
        fy.append(off[1][keep])
        base += rows * cols
    return np.concatenate(idx), np.concatenate(dirs), np.concatenate(fx), np.concatenate(fy)


def entries(generators, grid, rows, cols):
    """Green's matrix entries Z[rows, cols] (global rooftop indices) from the spatial blocks."""
    blocks = spatial_blocks(generators, grid)
    table = ((blocks[0], blocks[1]), (blocks[2], blocks[3]))
    _, d, fx, fy = rooftop_offsets(grid)
    rows, cols = np.asarray(rows), np.asarray(cols)
    out = np.zeros((rows.size, cols.size), dtype=np.complex128)
    for a in range(2):
        for b in range(2):
            g = table[a][b]
            if g is None:
                continue
            ri = np.nonzero(d[rows] == a)[0]
            ci = np.nonzero(d[cols] == b)[0]
            if ri.size == 0 or ci.size == 0:
                continue
            px = fx[rows[ri]][:, None] - fx[cols[ci]][None, :] + grid.count_x - 1
            py = fy[rows[ri]][:, None] - fy[cols[ci]][None, :] + grid.count_y - 1
            out[np.ix_(ri, ci)] = g[px, py]
    return out


def sai_matrix(generators, grid, mask_x=None, mask_y=None, radius=1, fit_radius=2):
    """The unsymmetrized SAI M as a scipy CSR matrix over all grid rooftops (rows of missing rooftops are empty)."""
    idx, d, fx, fy = rooftop_offsets(grid, mask_x, mask_y)
    n = grid.count
    rows, cols, vals = [], [], []
    for k in range(idx.size):
        near = np.nonzero((np.abs(fx - fx[k]) <= radius) & (np.abs(fy - fy[k]) <= radius))[0]
        fit = np.nonzero((np.abs(fx - fx[k]) <= fit_radius) & (np.abs(fy - fy[k]) <= fit_radius))[0]
        a = entries(generators, grid, idx[fit], idx[near])
        e = (fit == k).astype(np.complex128)
        m = np.linalg.lstsq(a, e, rcond=None)[0]
        rows.append(np.full(near.size, idx[k]))
        cols.append(idx[near])
        vals.append(m)
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))


def sai_matrix_by_class(generators, grid, mask_x=None, mask_y=None):
    """The same M as sai_matrix, fitting one row per class of rooftops whose 5x5 windows hold the same existing
    rooftops (Z depends on offsets only, so such rooftops share their row). Vectorized for large grids."""
    idx, d, fx, fy = rooftop_offsets(grid, mask_x, mask_y)
    n = grid.count
    cx, cy = grid.count_x, grid.count_y
    present = np.zeros((2, cx + 4, cy + 4), dtype=bool)     # existence padded by 2 on every side
    present[d, fx + 2, fy + 2] = True
    sig = np.zeros(idx.size, dtype=np.uint64)
    for dd in range(2):
        for dy in range(-2, 3):
            for dx in range(-2, 3):
                bit = np.uint64(dd * 25 + (dx + 2) + 5 * (dy + 2))
                sig |= present[dd, fx + 2 + dx, fy + 2 + dy].astype(np.uint64) << bit
    key = sig | (d.astype(np.uint64) << np.uint64(60))
    _, first, cls = np.unique(key, return_index=True, return_inverse=True)
    slot_coef = np.zeros((first.size, 18), dtype=np.complex128)
    for c, k in enumerate(first):
        near = np.nonzero((np.abs(fx - fx[k]) <= 1) & (np.abs(fy - fy[k]) <= 1))[0]
        fit = np.nonzero((np.abs(fx - fx[k]) <= 2) & (np.abs(fy - fy[k]) <= 2))[0]
        a = entries(generators, grid, idx[fit], idx[near])
        m = np.linalg.lstsq(a, (fit == k).astype(np.complex128), rcond=None)[0]
        slot_coef[c, d[near] * 9 + (fx[near] - fx[k] + 1) + 3 * (fy[near] - fy[k] + 1)] = m
    # Global index of every (direction, first-face offset), -1 where no rooftop exists.
    where = np.full((2, cx + 2, cy + 2), -1, dtype=np.int64)
    where[d, fx + 1, fy + 1] = idx
    rows, cols, vals = [], [], []
    for dd in range(2):
        for dy in range(-1, 2):
            for dx in range(-1, 2):
                s = dd * 9 + (dx + 1) + 3 * (dy + 1)
                j = where[dd, fx + 1 + dx, fy + 1 + dy]
                ok = j >= 0
                rows.append(idx[ok])
                cols.append(j[ok])
                vals.append(slot_coef[cls[ok], s])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))


class NearField:
    """P = (M + M^T) / 2, the symmetrized near-field SAI (see the module docstring)."""

    def __init__(self, generators, grid, mask_x=None, mask_y=None, by_class=False):
        build = sai_matrix_by_class if by_class else sai_matrix
        m = build(generators, grid, mask_x, mask_y)
        self.matrix = ((m + m.T) * 0.5).tocsr()

    def apply(self, r):
        return self.matrix @ np.asarray(r, dtype=np.complex128)


def tile_members(grid, tile=3):
    """Global rooftop indices per tile x tile cell tile (a rooftop belongs to the tile of its first face)."""
    tiles_x = -(-grid.count_x // tile)
    idx, _, fx, fy = rooftop_offsets(grid)
    key = fx // tile + tiles_x * (fy // tile)
    return [idx[key == t] for t in np.unique(key)]


class BlockJacobi:
    """Non-overlapping 3x3-tile block Jacobi (comparison only): sum over tiles of E_t inv(Z_tt) E_t^T."""

    def __init__(self, generators, grid, tile=3):
        self.members = tile_members(grid, tile)
        self.inverses = [np.linalg.inv(entries(generators, grid, m, m)) for m in self.members]

    def apply(self, r):
        r = np.asarray(r, dtype=np.complex128)
        z = np.zeros_like(r)
        for m, inv in zip(self.members, self.inverses):
            z[m] = inv @ r[m]
        return z


def jacobi_diagonal(generators, grid):
    """diag(Z): XX(0,0) on the x-directed rooftops, YY(0,0) on the y-directed ones."""
    xx, _, _, yy = spatial_blocks(generators, grid)
    out = []
    if grid.count_dir('x'):
        out.append(np.full(grid.count_dir('x'), xx[grid.count_x - 1, grid.count_y - 1]))
    if grid.count_dir('y'):
        out.append(np.full(grid.count_dir('y'), yy[grid.count_x - 1, grid.count_y - 1]))
    return np.concatenate(out)
