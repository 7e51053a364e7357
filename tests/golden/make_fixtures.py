"""Converts the reference's own data fixtures into small committed numpy files.

Run once in a container where the reference tree is mounted:

    python tests/golden/make_fixtures.py [/root/reference]

Inputs (data files only, no reference code is read or run):
  * tests/data/reference.mat  -- three grids with stored generators, a random input vector, the three products of the
    reference package ('fft', 'splitFFT', 'leanSplitFFT'), two circulant-preconditioned vectors, an oblique plane wave
    and, for the first grid, a 31x31 dense impedance matrix (tests/data/generateReferenceData.m:19-96 made them).
    It is read with scipy.io.loadmat, a plain MATLAB-v5 data reader that executes nothing from the file.
  * tests/data/exactMoments.csv -- touching-cell moments of R^p for five cells, 30 digits (mpmath, 40-digit arithmetic;
    tests/data/generateExactMoments.py made it).

Outputs: tests/golden/reference_v2.npz and tests/golden/exact_moments.csv.
"""

import os
import shutil
import sys

import numpy as np
import scipy.io as sio

HERE = os.path.dirname(os.path.abspath(__file__))


def _array(value, dtype):
    value = np.asarray(value)
    if value.size == 0:
        return np.zeros((0, 0), dtype=dtype)
    return value.astype(dtype)


def main(reference_root):
    data = sio.loadmat(os.path.join(reference_root, 'tests', 'data', 'reference.mat'), squeeze_me=False)
    cases = data['reference']
    out = {'case_count': np.int64(cases.shape[1])}
    for n in range(cases.shape[1]):
        case = cases[0, n]
        prefix = 'case%d_' % n
        for name in ('countX', 'countY', 'order'):
            out[prefix + name] = np.int64(case[name].ravel()[0])
        for name in ('sizeX', 'sizeY', 'wavenumber', 'intensity'):
            out[prefix + name] = np.float64(case[name].ravel()[0])
        for name in ('magneticXX', 'magneticYY', 'electric', 'impedanceMatrix'):
            out[prefix + name] = _array(case[name], np.complex128)
        for name in ('vector', 'productFFT', 'productSplit', 'productLean', 'planeWave'):
            out[prefix + name] = _array(case[name], np.complex128).ravel()
        for name in ('direction', 'polarization'):
            out[prefix + name] = _array(case[name], np.float64).ravel()
        cells = case['preconditioned']
        out[prefix + 'preconditionedStrang'] = _array(cells[0, 0], np.complex128).ravel()
        out[prefix + 'preconditionedTChan'] = _array(cells[1, 0], np.complex128).ravel()
    np.savez_compressed(os.path.join(HERE, 'reference_v2.npz'), **out)
    shutil.copyfile(os.path.join(reference_root, 'tests', 'data', 'exactMoments.csv'),
                    os.path.join(HERE, 'exact_moments.csv'))


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
