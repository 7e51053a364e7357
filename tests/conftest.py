import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (gfx950) GPU and the built librooftop.so')


GOLDEN = os.path.join(ROOT, 'tests', 'golden')


@pytest.fixture(scope='session')
def reference_data():
    """tests/golden/reference_v2.npz: the reference's reference.mat cases (see tests/golden/make_fixtures.py)."""
    data = np.load(os.path.join(GOLDEN, 'reference_v2.npz'))
    cases = []
    for n in range(int(data['case_count'])):
        p = 'case%d_' % n
        case = {k[len(p):]: data[k] for k in data.files if k.startswith(p)}
        for key in ('magneticXX', 'magneticYY'):
            if case[key].size == 0:
                case[key] = None
        cases.append(case)
    return cases


def rel(a, b):
    """norm(a - b) / norm(b); 0 when both vanish, inf when only b does."""
    diff = float(np.linalg.norm(np.asarray(a) - np.asarray(b)))
    ref = float(np.linalg.norm(np.asarray(b)))
    if ref == 0:
        return 0.0 if diff == 0 else float('inf')
    return diff / ref
