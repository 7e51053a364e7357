"""The MomentAssembler: the three Toeplitz generators of the Green's matrix (oracle; test infrastructure only).

Grids are numpy arrays indexed [x, y] (MATLAB column-major storage: flatten with order='F', x fastest).
"""

import numpy as np

from .quadrature import matlab_round, reduced_rule
from .singular import evaluate_remainder, extracted_moments

Z0 = 376.730313668          # src/+rooftopMoM/+assembler/Constants.m:11 (CODATA 2018)
C0 = 299792458.0            # src/+rooftopMoM/+assembler/Constants.m:8
NEAR_GAP = 3                # @MomentAssembler/MomentAssembler.m:13


class Grid:
    """Uniform countX x countY grid of sizeX x sizeY faces with rooftops on its interior edges.

    Layout follows src/+rooftopMoM/+basisFunctions/UniformRooftops.m:31-101: x-directed rooftops first on a
    (countX-1) x countY grid, then y-directed ones on countX x (countY-1), x index fastest.
    """

    def __init__(self, count_x, count_y, size_x, size_y):
        self.count_x, self.count_y = int(count_x), int(count_y)
        self.size_x, self.size_y = float(size_x), float(size_y)

    def grid_size(self, direction):
        if direction == 'x':
            return self.count_x - 1, self.count_y
        return self.count_x, self.count_y - 1

    def count_dir(self, direction):
        r, c = self.grid_size(direction)
        return r * c

    @property
    def count(self):
        return self.count_dir('x') + self.count_dir('y')

    @property
    def content(self):
        has_x, has_y = self.count_dir('x') > 0, self.count_dir('y') > 0
        return 'BOTH' if has_x and has_y else 'XONLY' if has_x else 'YONLY' if has_y else 'NONE'

    def face_offsets(self, direction):
        """2 x count first-face offsets (m; n) of the rooftops of one direction (UniformRooftops.m:74-87)."""
        r, c = self.grid_size(direction)
        idx = np.arange(r * c)
        return np.vstack([idx % r, idx // r]) if r > 0 else np.zeros((2, 0), dtype=int)

    def edge_lengths(self, indices):
        """UniformRooftops.m:89-101 (indices 0-based)."""
        indices = np.asarray(indices)
        out = np.full(indices.shape, self.size_x)
        out[indices < self.count_dir('x')] = self.size_y
        return out


def count_near_offsets(counts, sizes, gap=NEAR_GAP):
    """Numbers of near offsets along x and y. Follows @MomentAssembler/private/countNearOffsets_.m:5-10."""
    out = []
    for n in range(2):
        gaps = np.maximum(np.arange(counts[n]) - 1, 0) * sizes[n]
        out.append(int(np.count_nonzero(gaps < gap * max(sizes))))
    return out


def integrate_moments(rule, offsets, sizes, wavenumber):
    """7 x C moments of exp(-ikR)/R over the cell pairs at the offsets; the remainder on touching offsets.

    Follows src/+rooftopMoM/+assembler/+greensMatrix/+utils/integrateMoments.m:17-38.
    Pinned: brute-force tensor Gauss-Legendre on separated cells at 1e-13 and the remainder on touching cells
    (IntegrateMomentsTest.m:5-54).
    """
    offsets = np.asarray(offsets, dtype=np.float64).reshape(2, -1)
    k = float(wavenumber)
    nx, ny = rule['nodesX'], rule['nodesY']
    dist = np.hypot(sizes[0] * (nx[:, None, None] - offsets[0][None, None, :]),
                    sizes[1] * (ny[None, :, None] - offsets[1][None, None, :]))
    kernel = np.exp(-1j * k * dist) / dist
    touching = np.all(offsets <= 1, axis=0)
    if np.any(touching):
        kernel[:, :, touching] = evaluate_remainder(dist[:, :, touching], k)
    c = offsets.shape[1]
    return rule['weights'] @ kernel.reshape(nx.size * ny.size, c, order='F')


def _quadrants(rule, order):
    """The four quadrants of a one-part rule: (7 x order^2 weights x fastest, shift). integrateLattice_.m:92-111."""
    w = rule['weights'].reshape(7, 2 * order, 2 * order, order='F')
    halves = (slice(0, order), slice(order, 2 * order))
    quads = []
    for hy in range(2):
        for hx in range(2):
            quads.append((w[:, halves[hx], halves[hy]].reshape(7, order * order, order='F'), (hx == 0, hy == 0)))
    return quads


def integrate_lattice(rule, counts, sizes, wavenumber):
    """Moments (7, countX, countY) of every offset from the kernel on the shared far lattice.

    Follows @MomentAssembler/private/integrateLattice_.m:17-90 (one tile covering the whole grid; the reference's
    tiling only bounds its memory, MomentAssemblerTest.m:43-53 shows the result does not depend on it).
    Pinned: lattice == per-offset integration with the same rule (MomentAssemblerTest.m:55-77, 1e-14).
    """
    order = rule['nodesX'].size // 2
    points = rule['nodesX'][order:]
    bx = np.arange(counts[0] + 1)
    by = np.arange(counts[1] + 1)
    x = (sizes[0] * (points[:, None] - bx[None, :])).reshape(order, 1, bx.size, 1)
    y = (sizes[1] * (points[:, None] - by[None, :])).reshape(1, order, 1, by.size)
    dist = np.hypot(x, y)
    kernel = (np.exp(-1j * wavenumber * dist) / dist).reshape(order * order, -1, order='F')
    moments = np.zeros((7, counts[0], counts[1]), dtype=np.complex128)
    for weights, shift in _quadrants(rule, order):
        sums = (weights @ kernel).reshape(7, bx.size, by.size, order='F')
        moments += sums[:, int(shift[0]):int(shift[0]) + counts[0], int(shift[1]):int(shift[1]) + counts[1]]
    return moments


_COMBINATION = np.array([
    [1, 0, 0, 0, 0, 0, 0],
    [1, -1, -1, 2, 0, 0, 0],
    [0, 1, 0, -1, 0, 0, 0],
    [0, 0, 1, -1, 0, 0, 0],
    [1, 0, 0, 0, -1, -1, 2],
    [0, 0, 0, 0, 1, 0, -1],
    [0, 0, 0, 0, 0, 1, -1]], dtype=np.float64)


def accumulate_moments(parts, moments, offsets, scales):
    """Adds the moments of cell pairs at the given distinct offsets into the generators.

    Follows src/+rooftopMoM/+assembler/+greensMatrix/+utils/accumulateMoments.m:27-133: electric += s1*M00;
    magneticXX at e += M00-M10-M01+2M11, at e-u += M10-M11 (the dX = 1 row counted twice for d = 0), at e+u += M01-M11;
    magneticYY the same with N and columns. The terms are laid out on the whole offset grid (zero where no offset is
    given), which is the reference's chunk with width = countY.
    Pinned: entry by entry against a brute-force 4D Gauss-Legendre integral of the rooftop definition of A^xx, A^yy
    and Phi (02_setting.tex:42-58) at far and separated near offsets, 4e-15 (a swapped M10/M01 row fails it by 0.26);
    against the note's formula on random moments added in chunks (AccumulateMomentsTest.m:5-60), 1e-14
    (tests/test_oracle_pins.py).
    """
    e = parts['electric']
    cx, cy = e.shape
    offsets = np.asarray(offsets, dtype=int).reshape(2, -1)
    weights = _COMBINATION * np.array([scales[0]] + [scales[1]] * 6)[:, None]
    terms = np.zeros((7, cx, cy), dtype=np.complex128)
    terms[:, offsets[0], offsets[1]] = weights @ moments
    e += terms[0]
    if parts['magneticXX'] is not None and parts['magneticXX'].size:
        rising = terms[2].copy()
        if cx > 1:
            rising[1, :] *= 2
        values = terms[1][:-1, :] + rising[1:, :]
        values[1:, :] += terms[3][:-2, :]
        parts['magneticXX'] += values
    if parts['magneticYY'] is not None and parts['magneticYY'].size:
        rising = terms[5].copy()
        if cy > 1:
            rising[:, 1] *= 2
        values = terms[4][:, :-1] + rising[:, 1:]
        values[:, 1:] += terms[6][:, :-2]
        parts['magneticYY'] += values
    return parts


def allocate_parts(grid):
    """Zero generators; the one of a missing direction is None. @MomentAssembler/private/allocateParts_.m:5-15."""
    cx, cy = grid.count_x, grid.count_y
    return {'electric': np.zeros((cx, cy), dtype=np.complex128),
            'magneticXX': np.zeros((cx - 1, cy), dtype=np.complex128) if grid.count_dir('x') else None,
            'magneticYY': np.zeros((cx, cy - 1), dtype=np.complex128) if grid.count_dir('y') else None}


def generate(grid, wavenumber, orders=(16, 8)):
    """The generators magneticXX, magneticYY, electric of the Green's matrix (dict; a missing direction is None).

    Follows src/+rooftopMoM/+assembler/+greensMatrix/@MomentAssembler/generate.m:10-58: near pass over the offsets
    below nearCounts with the near rule (split by the rounded aspect ratio), far pass on the lattice for all other
    offsets, then the closed-form touching moments. Chunking is left out (result-invariant).
    Pinned: reference.mat products through it; grid-size independence, transpose symmetry (1e-13), default orders vs
    [32, 24] (3e-14), the touching entries vs adaptive quadrature (1e-10), MomentAssemblerTest.m:22-114.
    """
    k = float(wavenumber)
    counts = (grid.count_x, grid.count_y)
    sizes = (grid.size_x, grid.size_y)
    orders = (int(orders[0]), int(orders[-1]))
    scales = (1j * Z0 / (4 * np.pi) * (sizes[0] * sizes[1]) ** 2 / k,
              1j * Z0 / (4 * np.pi) * (sizes[0] * sizes[1]) ** 2 * k)
    parts = allocate_parts(grid)
    near = count_near_offsets(counts, sizes)
    near_rule = reduced_rule(orders[0], matlab_round(np.array(sizes) / min(sizes)).astype(int))
    far_rule = reduced_rule(orders[1], (1, 1))

    dx, dy = np.meshgrid(np.arange(near[0]), np.arange(near[1]), indexing='ij')
    near_offsets = np.vstack([dx.ravel(order='F'), dy.ravel(order='F')])
    accumulate_moments(parts, integrate_moments(near_rule, near_offsets, sizes, k), near_offsets, scales)

    lattice = integrate_lattice(far_rule, counts, sizes, k)
    dx, dy = np.meshgrid(np.arange(counts[0]), np.arange(counts[1]), indexing='ij')
    far = (dx >= near[0]) | (dy >= near[1])
    far_offsets = np.vstack([dx[far], dy[far]])
    if far_offsets.size:
        accumulate_moments(parts, lattice[:, far], far_offsets, scales)

    touching = np.array([[0, 1, 0, 1], [0, 0, 1, 1]])
    inside = np.all(touching < np.array(counts)[:, None], axis=0)
    accumulate_moments(parts, extracted_moments(sizes[0], sizes[1], k)[:, inside], touching[:, inside], scales)
    return parts


def modify(generators, modifier):
    """Applies 'real', 'imag' or None to every generator (Generators.m:25-37, the Operator 'modifier')."""
    if modifier is None:
        return generators
    f = {'real': np.real, 'imag': np.imag}[modifier]
    return {key: (None if v is None else f(v).astype(np.complex128)) for key, v in generators.items()}
