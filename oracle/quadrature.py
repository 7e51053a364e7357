"""Gauss-Legendre rules and the reduced cell-pair moment rule (oracle; test infrastructure only)."""

import numpy as np


def matlab_round(x):
    """MATLAB round: half away from zero (Python's round is half to even)."""
    x = np.asarray(x, dtype=np.float64)
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def gauss_legendre(order, a=0.0, b=1.0):
    """Gauss-Legendre points (ascending) and weights on [a, b] by Golub-Welsch.

    Follows src/+rooftopMoM/+basisFunctions/+utils/generateGaussLegendre1D.m:16-25. Pinned against
    numpy.polynomial.legendre.leggauss (tests/test_oracle_quadrature.py).
    """
    if order == 1:
        values = np.zeros(1)
        vectors = np.ones((1, 1))
    else:
        n = np.arange(1, order, dtype=np.float64)
        off = n / np.sqrt(4.0 * n * n - 1.0)
        jacobi = np.diag(off, 1) + np.diag(off, -1)
        values, vectors = np.linalg.eigh(jacobi)
    permutation = np.argsort(values)
    values = values[permutation]
    weights = 2.0 * vectors[0, permutation] ** 2
    points = (a + b) / 2.0 + (b - a) / 2.0 * values
    weights = (b - a) / 2.0 * weights
    return points, weights


def tensor_rule(order):
    """Tensor Gauss-Legendre rule on the unit square, u fastest.

    Follows src/+rooftopMoM/+basisFunctions/GaussLegendre2RectQuadrature.m:13-22.
    Returns (pointsU, pointsV, weights2D), each of length order^2.
    """
    points, weights = gauss_legendre(order, 0.0, 1.0)
    pu, pv = np.meshgrid(points, points, indexing='ij')
    wu, wv = np.meshgrid(weights, weights, indexing='ij')
    return pu.ravel(order='F'), pv.ravel(order='F'), (wu * wv).ravel(order='F')


def _axis(order, part_count):
    """Nodes on [-1, 1] and Gauss-Legendre weights times the correlation weights W00, W10, W01, W11.

    Follows src/+rooftopMoM/+assembler/+greensMatrix/+utils/generateReducedRule.m:39-66 (generateAxis_).
    """
    points, point_weights = gauss_legendre(order, 0.0, 1.0)
    right = ((points[:, None] + np.arange(part_count)[None, :]) / part_count).ravel(order='F')
    nodes = np.concatenate([right - 1.0, right])
    pw = np.tile(point_weights / part_count, 2 * part_count)
    left = nodes < 0
    w10 = (1.0 - nodes ** 2) / 2.0
    w10[left] = (1.0 + nodes[left]) ** 2 / 2.0
    w01 = (1.0 - nodes) ** 2 / 2.0
    w01[left] = (1.0 - nodes[left] ** 2) / 2.0
    dist = np.abs(nodes)
    weights = pw[:, None] * np.stack([1.0 - dist, w10, w01, (1.0 - dist) ** 2 * (2.0 + dist) / 6.0], axis=1)
    return nodes, weights


def reduced_rule(order, part_counts):
    """The 2D rule of the seven cell-pair moments over nu = xi - xi', eta = eta - eta'.

    Follows src/+rooftopMoM/+assembler/+greensMatrix/+utils/generateReducedRule.m:21-37.
    Returns a dict with nodesX (Px,), nodesY (Py,), weights (7, Px*Py) with nu running fastest; the rows are the
    moments M00, M10, M01, M11, N10, N01, N11.
    Pinned: exact for polynomial kernels against the tensor rule on both cells (GenerateReducedRuleTest.m:19-35), and
    by the brute-force separated-cell check of integrate_moments.
    """
    nx, wx = _axis(order, int(part_counts[0]))
    ny, wy = _axis(order, int(part_counts[1]))
    px, py = nx.size, ny.size
    weights = np.empty((7, px * py))
    # kron(wy[:, 0], wx) -> rows M00..M11; kron(wy[:, 1:4], wx[:, 0]) -> rows N10..N11 (nu fastest).
    for r in range(4):
        weights[r] = (wy[:, 0][:, None] * wx[:, r][None, :]).ravel()
    for r in range(3):
        weights[4 + r] = (wy[:, 1 + r][:, None] * wx[:, 0][None, :]).ravel()
    return {'nodesX': nx, 'nodesY': ny, 'weights': weights}
