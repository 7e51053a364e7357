"""Right-hand sides and outputs (oracle; test infrastructure only)."""

import numpy as np

from .assembly import Z0
from .quadrature import tensor_rule


def plane_wave(grid, wavenumber, direction, polarization, intensity=1.0, order=8):
    """Plane-wave excitation, O(N) by translation phases; x-directed rooftops first.

    Follows src/+rooftopMoM/+operator/generatePlaneWave.m:18-42 (the rooftop quadrature is the problem's far order,
    Problem.m:62-64). Pinned: reference.mat planeWave at 1e-13, normal incidence = face area, phase steps
    (GeneratePlaneWaveTest.m:5-48).
    """
    direction = np.asarray(direction, dtype=np.float64)
    polarization = np.asarray(polarization, dtype=np.float64)
    a, b = grid.size_x, grid.size_y
    area = a * b
    pu, pv, w = tensor_rule(order)
    steps = -1j * wavenumber * direction[:2] * np.array([a, b])
    point_phases = np.exp(steps[0] * pu + steps[1] * pv)
    parts = []
    for n, (d, coord, second) in enumerate((('x', pu, (1, 0)), ('y', pv, (0, 1)))):
        rows, cols = grid.grid_size(d)
        if rows * cols == 0:
            continue
        second_phase = np.exp(steps[0] * second[0] + steps[1] * second[1])
        amplitudes = coord + second_phase * (1 - coord)
        reference = intensity * polarization[n] * area * (point_phases @ (w * amplitudes))
        trans = np.exp(steps[0] * np.arange(rows))[:, None] * np.exp(steps[1] * np.arange(cols))[None, :]
        parts.append(reference * trans.ravel(order='F'))
    return np.concatenate(parts)


def middle_port(grid, direction='x'):
    """0-based rooftop indices of the middle line of edges. +operator/+utils/generateMiddlePort.m:14-28."""
    rows, cols = grid.grid_size(direction)
    if direction == 'x':
        r = np.full(cols, grid.count_x // 2)          # 1-based row floor(countX/2)
        c = np.arange(1, cols + 1)
        first = 0
    else:
        r = np.arange(1, rows + 1)
        c = np.full(rows, grid.count_y // 2)
        first = grid.count_dir('x')
    return first + (r - 1) + rows * (c - 1)


def delta_gap(grid, port, voltage=1.0):
    """Voltage times edge length on the port rooftops. +operator/generateDeltaGap.m:13-14."""
    out = np.zeros(grid.count, dtype=np.complex128)
    out[port] = voltage * grid.edge_lengths(port)
    return out


def input_impedance(grid, current, port, voltage=1.0):
    """V / (edge lengths . I(port)). +operator/computeInputImpedance.m:17."""
    return voltage / (grid.edge_lengths(port) @ np.asarray(current)[port])


def radar_cross_section(wavenumber, excitation, current, intensity=1.0):
    """(k Z0)^2 |b.' I|^2 / (4 pi E^4). +operator/computeRadarCrossSection.m:18-19.

    Pinned: the closed form of ComputeRadarCrossSectionTest.m:5-11 at 1e-15 and intensity independence (:13-23)."""
    return (wavenumber * Z0) ** 2 * abs(np.asarray(excitation) @ np.asarray(current)) ** 2 / (4 * np.pi * intensity ** 4)
