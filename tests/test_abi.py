"""CPU checks of the C-ABI boundary: librooftop.so loads and exports every function include/rooftop.h declares, and
the ctypes signatures cover exactly that list. No compute call is made (no GPU here)."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, 'include', 'rooftop.h')


def _declared():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(rt_[a-z0-9_]+)\s*\(', text)))


@pytest.fixture(scope='module')
def rt():
    import __graft_entry__ as ge
    lib = os.path.join(ge.PACKAGE_DIR, 'librooftop.so')
    if not os.path.exists(lib):
        ge.build()
    return ge.load()


def test_header_declares_the_abi():
    names = _declared()
    for name in ('rt_ctx_create', 'rt_problem_create', 'rt_product_create', 'rt_precond_create', 'rt_product_apply',
                 'rt_solve', 'rt_last_error_id'):
        assert name in names


def test_library_exports_every_declared_symbol(rt):
    out = subprocess.run(['nm', '-D', '--defined-only', rt._abi.LIBRARY], capture_output=True, text=True, check=True)
    exported = set(re.findall(r'\bT (rt_[a-z0-9_]+)', out.stdout))
    missing = [n for n in _declared() if n not in exported]
    assert not missing, missing


def test_ctypes_signatures_match_header(rt):
    assert sorted(name for name, _, _ in rt._abi.SIGNATURES) == _declared()
    lib = rt.library()
    for name, _, _ in rt._abi.SIGNATURES:
        assert hasattr(lib, name)
    assert lib.rt_version().decode().startswith('rooftop-hip')


def test_host_validation_mirrors_reference_errors(rt):
    with pytest.raises(rt.RooftopError) as e:
        rt.Mesh(1, 1, 0.1, 0.1)
    assert e.value.identifier == 'rooftopMoM:mesh:noRooftops'
    with pytest.raises(rt.RooftopError) as e:
        rt.Solver(method='bicg')
    assert e.value.identifier == 'MATLAB:validators:mustBeMember'
    with pytest.raises(rt.RooftopError) as e:
        rt.Solver(restart=2.5)
    assert e.value.identifier == 'rooftopMoM:solver:invalidRestart'
    with pytest.raises(rt.RooftopError) as e:
        rt.Solver(tolerance=2)
    assert e.value.identifier == 'MATLAB:validators:mustBeLessThan'


def test_missing_library_fails_loudly(rt, monkeypatch):
    monkeypatch.setattr(rt._abi, '_lib', None)
    monkeypatch.setattr(rt._abi, 'LIBRARY', '/nonexistent/librooftop.so')
    with pytest.raises(RuntimeError, match='no CPU fallback'):
        rt.library()


def test_blob_mask_is_a_seeded_connected_blob(rt):
    """rt_blob_mask (host code): the requested fill, 4-connected, reproducible per seed."""
    import collections
    for seed, (cx, cy), fill in ((1000, (40, 30), 0.3), (1001, (17, 64), 0.9), (7, (5, 5), 1.0)):
        m = rt.blob_mask(cx, cy, seed, fill)
        assert m.shape == (cx, cy) and m.sum() == max(1, int(fill * cx * cy + 0.5))
        start = tuple(np.argwhere(m)[0])
        seen, todo = {start}, collections.deque([start])
        while todo:
            i, j = todo.popleft()
            for a, b in ((i + 1, j), (i - 1, j), (i, j + 1), (i, j - 1)):
                if 0 <= a < cx and 0 <= b < cy and m[a, b] and (a, b) not in seen:
                    seen.add((a, b))
                    todo.append((a, b))
        assert len(seen) == m.sum()
        assert np.array_equal(m, rt.blob_mask(cx, cy, seed, fill))
    assert not np.array_equal(rt.blob_mask(40, 30, 1, 0.5), rt.blob_mask(40, 30, 2, 0.5))
