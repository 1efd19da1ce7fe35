"""The C-ABI library loads and exports every function include/*.h declares;
host-side pattern construction (no GPU needed)."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp

from engine_path import REPO, ngmi, workloads


def _declared_functions():
    names = []
    for h in glob.glob(os.path.join(REPO, 'include', '*.h')):
        src = open(h).read()
        src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(ngmi_\w+|ngSpice_\w+|ngGet_\w+)\s*\(", src, flags=re.M):
            names.append(m.group(1))
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = ngmi.lib()
    names = _declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


@pytest.mark.parametrize('seed', [0, 1, 2])
def test_csc_from_elements_matches_scipy(seed):
    g = np.random.default_rng(seed)
    n = 500
    ne = 6000
    r = g.integers(0, n, ne).astype(np.int32)
    c = g.integers(0, n, ne).astype(np.int32)
    Ap, Ai, e2z = ngmi.csc_from_elements(n, r, c)
    ref = sp.coo_matrix((np.ones(ne), (r, c)), shape=(n, n)).tocsc()
    ref.sum_duplicates()
    ref.sort_indices()
    assert np.array_equal(Ap, ref.indptr)
    assert np.array_equal(Ai, ref.indices)
    # every element binds to the entry holding its own (row, col)
    col_of = np.repeat(np.arange(n), np.diff(Ap))
    assert np.array_equal(Ai[e2z], r)
    assert np.array_equal(col_of[e2z], c)


def test_csc_rejects_out_of_range():
    with pytest.raises(ngmi.EngineError):
        ngmi.csc_from_elements(4, np.array([0, 4]), np.array([0, 1]))


def test_csc_empty():
    Ap, Ai, e2z = ngmi.csc_from_elements(3, np.zeros(0), np.zeros(0))
    assert list(Ap) == [0, 0, 0, 0] and len(Ai) == 0 and len(e2z) == 0


def test_eval_workload_is_deterministic_and_well_formed():
    a = workloads.device_eval_workload(n=4096, seed=1)
    b = workloads.device_eval_workload(n=4096, seed=1)
    for k in ('vterm', 'rows_e', 'cols_e'):
        assert np.array_equal(a[k], b[k])
    t = a['terms']
    for x in range(4):
        for y in range(x):
            assert not np.any(t[x] == t[y])
    assert a['params']['l'].min() >= 45e-9 and a['params']['l'].max() <= 1e-6
    assert a['params']['w'].min() >= 100e-9 and a['params']['w'].max() <= 10e-6
