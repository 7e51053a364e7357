"""Closed-form touching moments and the extraction remainder (oracle; test infrastructure only)."""

import numpy as np

from .quadrature import gauss_legendre


def expm1_imag(theta):
    """expm1(-1i * theta) for real theta without cancellation: -2 sin^2(theta/2) - 1i sin(theta)."""
    s = np.sin(0.5 * theta)
    return -2.0 * s * s - 1j * np.sin(theta)


def evaluate_remainder(distance, wavenumber):
    """exp(-ikR)/R - (1/R - k^2 R/2 + k^4 R^3/24), accurate as R -> 0 (tends to -ik).

    Follows src/+rooftopMoM/+assembler/+singularIntegrals/evaluateRemainder.m:13-14 (expm1 of an imaginary argument).
    Pinned: against the direct difference at moderate R and the Taylor series at R = 1e-8..1e-3
    (EvaluateRemainderTest.m:5-21).
    """
    r = np.asarray(distance, dtype=np.float64)
    k = float(wavenumber)
    return expm1_imag(k * r) / r + k * k / 2.0 * r - k ** 4 / 24.0 * r ** 3


def integrate_monomial(power, m, n, a, b):
    """Mixed antiderivative G(a, b) = integral of u^m v^n (u^2+v^2)^(power/2) over [0,a]x[0,b], closed form.

    Follows src/+rooftopMoM/+assembler/+singularIntegrals/+utils/integrateMonomial.m:18-80 (the closed forms of the
    near-field note, appendix A). Pinned: through power_moments against exactMoments.csv and the thesis self terms, and
    directly against 2D Gauss-Legendre quadrature on rectangles away from the origin (tests).
    """
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    with np.errstate(divide='ignore', invalid='ignore'):
        R = np.sqrt(a ** 2 + b ** 2)
        lb = np.arcsinh(b / a)
        la = np.arcsinh(a / b)
        key = (power, m, n)
        if key == (-1, 0, 0):
            v = a * lb + b * la
        elif key == (-1, 0, 1):
            v = (-a ** 2 + a * R + b ** 2 * la) / 2
        elif key == (-1, 1, 0):
            v = (-b ** 2 + b * R + a ** 2 * lb) / 2
        elif key == (-1, 1, 1):
            v = (-a ** 3 - b ** 3 + (a ** 2 + b ** 2) * R) / 3
        elif key == (-1, 2, 0):
            v = (a * b * R + 2 * a ** 3 * lb - b ** 3 * la) / 6
        elif key == (-1, 2, 1):
            v = (-2 * a ** 4 + a * (2 * a ** 2 + b ** 2) * R - b ** 4 * la) / 8
        elif key == (-1, 3, 0):
            v = (2 * b ** 4 + b * (a ** 2 - 2 * b ** 2) * R + 3 * a ** 4 * lb) / 12
        elif key == (-1, 3, 1):
            v = (-3 * a ** 5 + 2 * b ** 5 + (3 * a ** 4 + a ** 2 * b ** 2 - 2 * b ** 4) * R) / 15
        elif key == (1, 0, 0):
            v = (2 * a * b * R + a ** 3 * lb + b ** 3 * la) / 6
        elif key == (1, 0, 1):
            v = (-2 * a ** 4 + a * (2 * a ** 2 + 5 * b ** 2) * R + 3 * b ** 4 * la) / 24
        elif key == (1, 1, 0):
            v = (-2 * b ** 4 + b * (5 * a ** 2 + 2 * b ** 2) * R + 3 * a ** 4 * lb) / 24
        elif key == (1, 1, 1):
            v = (-a ** 5 - b ** 5 + (a ** 4 + 2 * a ** 2 * b ** 2 + b ** 4) * R) / 15
        elif key == (1, 2, 0):
            v = (a * b * (6 * a ** 2 + b ** 2) * R + 4 * a ** 5 * lb - b ** 5 * la) / 40
        elif key == (1, 2, 1):
            v = (-8 * a ** 6 + a * (8 * a ** 4 + 14 * a ** 2 * b ** 2 + 3 * b ** 4) * R - 3 * b ** 6 * la) / 144
        elif key == (1, 3, 0):
            v = (4 * b ** 6 + b * (21 * a ** 4 + 2 * a ** 2 * b ** 2 - 4 * b ** 4) * R + 15 * a ** 6 * lb) / 180
        elif key == (1, 3, 1):
            v = (-5 * a ** 7 + 2 * b ** 7 + (5 * a ** 6 + 8 * a ** 4 * b ** 2 + a ** 2 * b ** 4 - 2 * b ** 6) * R) / 105
        elif key == (3, 0, 0):
            v = (7 * a * b * (a ** 2 + b ** 2) * R + 3 * a ** 5 * lb + 3 * b ** 5 * la) / 40
        elif key == (3, 0, 1):
            v = (-8 * a ** 6 + a * (8 * a ** 4 + 26 * a ** 2 * b ** 2 + 33 * b ** 4) * R + 15 * b ** 6 * la) / 240
        elif key == (3, 1, 0):
            v = (-8 * b ** 6 + b * (33 * a ** 4 + 26 * a ** 2 * b ** 2 + 8 * b ** 4) * R + 15 * a ** 6 * lb) / 240
        elif key == (3, 1, 1):
            v = (-a ** 7 - b ** 7 + (a ** 6 + 3 * a ** 4 * b ** 2 + 3 * a ** 2 * b ** 4 + b ** 6) * R) / 35
        elif key == (3, 2, 0):
            v = (a * b * (38 * a ** 4 + 26 * a ** 2 * b ** 2 + 3 * b ** 4) * R + 18 * a ** 7 * lb
                 - 3 * b ** 7 * la) / 336
        elif key == (3, 2, 1):
            v = (-48 * a ** 8 + a * (48 * a ** 6 + 136 * a ** 4 * b ** 2 + 118 * a ** 2 * b ** 4 + 15 * b ** 6) * R
                 - 15 * b ** 8 * la) / 1920
        elif key == (3, 3, 0):
            v = (16 * b ** 8 + b * (215 * a ** 6 + 134 * a ** 4 * b ** 2 + 8 * a ** 2 * b ** 4 - 16 * b ** 6) * R
                 + 105 * a ** 8 * lb) / 2240
        elif key == (3, 3, 1):
            v = (-7 * a ** 9 + 2 * b ** 9
                 + (7 * a ** 8 + 19 * a ** 6 * b ** 2 + 15 * a ** 4 * b ** 4 + a ** 2 * b ** 6 - 2 * b ** 8) * R) / 315
        else:
            raise ValueError('no closed form for power %d, m %d, n %d' % key)
    v = np.array(v, dtype=np.float64)
    v[(a == 0) | (b == 0)] = 0.0
    return v


# Correlation weights W00, W10, W01, W11 as ascending polynomial coefficients in nu on the parts [-1, 0] and [0, 1]
# (generatePowerMoments.m:66-71, the near-field note's "Weights of the rooftop moments").
_WEIGHTS = (
    ((1.0, 1.0), (1.0, -1.0)),
    ((0.5, 1.0, 0.5), (0.5, 0.0, -0.5)),
    ((0.5, 0.0, -0.5), (0.5, -1.0, 0.5)),
    ((1.0 / 3, 0.5, 0.0, -1.0 / 6), (1.0 / 3, -0.5, 0.0, 1.0 / 6)),
)
_TOUCHING = ((0, 0), (1, 0), (0, 1), (1, 1))


def _compose_part(weight, part, offset, cell_size, count):
    """One weight on one part of [-1, 1] as a polynomial in u = sigma*cellSize*(nu - offset) >= 0, and its interval.

    Follows generatePowerMoments.m:121-155 (composePart_). part is 0 for [-1, 0] and 1 for [0, 1]; the returned
    interval is 0 for [0, cellSize] and 1 for [cellSize, 2 cellSize].
    """
    ends = ((-1.0, 0.0), (0.0, 1.0))[part]
    shifted = (ends[0] - offset, ends[1] - offset)
    sigma = 1.0 - 2.0 * (shifted[1] <= 0)
    interval = int(min(sigma * shifted[0], sigma * shifted[1]))
    coefficients = np.zeros(count)
    term = np.array([1.0])
    for k, w in enumerate(weight):
        coefficients[:k + 1] += w * term
        term = np.convolve(term, np.array([offset, sigma / cell_size]))
    return coefficients, interval


def _integrate_rectangles(power, ratio):
    """Integrals of u^m v^n R^power over [0,1] or [1,2] times [0,ratio] or [ratio, 2 ratio], shape (4, 2, 2, 2).

    Follows generatePowerMoments.m:102-153 (integrateRectangles_): corner differences of the closed form, replaced by
    20-point Gauss-Legendre on the rectangles at least their longer side away from the origin.
    """
    as_, bs = np.meshgrid(np.array([0.0, 1.0, 2.0]), ratio * np.array([0.0, 1.0, 2.0]), indexing='ij')
    rect = np.zeros((4, 2, 2, 2))
    for m in range(4):
        for n in range(2):
            c = integrate_monomial(power, m, n, as_, bs)
            rect[m, n] = np.diff(np.diff(c, axis=0), axis=1)
    points, pw = gauss_legendre(20, 0.0, 1.0)
    for k in range(2):
        for l in range(2):
            if np.hypot(k, ratio * l) < max(1.0, ratio):
                continue
            u = (k + points)[:, None]
            v = ratio * (l + points)[None, :]
            kernel = np.sqrt(u ** 2 + v ** 2) ** power
            for m in range(4):
                for n in range(2):
                    rect[m, n, k, l] = ratio * pw @ (u ** m * v ** n * kernel) @ pw
    return rect


def _along_x(a, b, power):
    """M00, M10, M01, M11 at the four touching offsets (4x4: weight x offset). generatePowerMoments.m:57-100."""
    ratio = b / a
    rect = _integrate_rectangles(power, ratio)
    moments = np.zeros((4, 4))
    for k in range(4):
        for l, (ox, oy) in enumerate(_TOUCHING):
            for s in range(2):
                ca, ia = _compose_part(_WEIGHTS[k][s], s, ox, 1.0, 4)
                for t in range(2):
                    cb, ib = _compose_part(_WEIGHTS[0][t], t, oy, ratio, 2)
                    moments[k, l] += ca @ rect[:, :, ia, ib] @ cb
    return a ** power * moments / ratio


def power_moments(size_x, size_y, power):
    """Seven moments (rows M00 M10 M01 M11 N10 N01 N11) of R^power at the touching offsets (0,0),(1,0),(0,1),(1,1).

    Divided by (ab)^2. Follows src/+rooftopMoM/+assembler/+singularIntegrals/generatePowerMoments.m:17-55.
    Pinned: tests/data/exactMoments.csv (30 digits) at 1e-13 (1e-12 for aspect > 4), the bachelor-thesis self terms at
    1e-14, transpose and homogeneity (GeneratePowerMomentsTest.m:9-62).
    """
    along_x = _along_x(size_x, size_y, power)
    along_y = _along_x(size_y, size_x, power)
    return np.vstack([along_x, along_y[1:4][:, [0, 2, 1, 3]]])


def extracted_moments(size_x, size_y, wavenumber):
    """Moments of S(R) = 1/R - k^2 R/2 + k^4 R^3/24 at the touching offsets (7x4 complex).

    Follows src/+rooftopMoM/+assembler/+singularIntegrals/generateExtractedMoments.m:15-18.
    """
    k = float(wavenumber)
    return (power_moments(size_x, size_y, -1) - k * k / 2 * power_moments(size_x, size_y, 1)
            + k ** 4 / 24 * power_moments(size_x, size_y, 3)).astype(np.complex128)
