"""The static LU's pivot sequence against the reference KLU (CPU only).

ngspice's KLU path fixes its pivot sequence in klu_factor (SMPreorder,
klusmp.c:752-806) and every later klu_refactor reuses it.  The engine can
take that sequence (ngmi_circuit_set_pivot_order); these checks show that
its symbolic analysis then reproduces KLU's own factor structure:

* the strictly-lower L pattern has exactly KLU's nnz(L) - n entries
  (Numeric->lnz counts the unit diagonal), i.e. the same fill inside KLU's
  diagonal blocks;
* when the off-diagonal blocks take no fill (the ring oscillators), the
  engine's L+U slot count equals (lnz - n) + unz + nzoff, KLU's total.

The reference KLU is compiled from the reference's own sources into
oracle/_ref (oracle/Makefile `ref`); without it these tests skip.
"""
import os
import sys

import numpy as np
import pytest

from engine_path import ngmi, workloads

sys.path.insert(0, os.path.join(os.path.dirname(__file__), '..', 'oracle'))
import pyoracle as po  # noqa: E402


def _klu_order(desc, tstop):
    T = po.Transient(desc)
    if not T.use_reference_klu():
        pytest.skip('reference KLU (oracle/_ref) not built')
    rc, _, _ = T.run(desc['tstep'], tstop)
    assert rc == 0
    got = T.klu_order()
    assert got is not None
    return got


CASES = [
    ('ring11', lambda: workloads.ring_oscillator(stages=11, tstop=2e-9), True),
    ('ring101', lambda: workloads.ring_oscillator(), True),
    ('adder', lambda: workloads.mc_adder(0, tstop=0.2e-9), False),
]


@pytest.mark.parametrize('name,make,no_offblock_fill', CASES, ids=[c[0] for c in CASES])
def test_klu_pivot_sequence_reproduces_klu_structure(name, make, no_offblock_fill):
    desc = make()
    rows, cols, cnt = _klu_order(desc, 0.05e-9)
    c = ngmi.Circuit(desc)
    own = c.analyze()
    n = own['unknowns']
    assert sorted(rows) == list(range(1, n + 1)) and sorted(cols) == list(range(1, n + 1))
    c.set_pivot_order(rows, cols)
    a = c.analyze()
    assert (a['rows'] == rows).all() and (a['cols'] == cols).all()
    print('%s: n %d, KLU lnz %d unz %d nzoff %d blocks %d; engine with KLU order: L %d, L+U %d; '
          'own order: L+U %d' % (name, n, cnt['lnz'], cnt['unz'], cnt['nzoff'], cnt['nblocks'],
                                 a['l_nonzeros'], a['lu_nonzeros'], own['lu_nonzeros']))
    assert a['l_nonzeros'] == cnt['lnz'] - n
    total = cnt['lnz'] - n + cnt['unz'] + cnt['nzoff']
    if no_offblock_fill:
        assert a['lu_nonzeros'] == total
    else:
        assert a['lu_nonzeros'] >= total
    c.set_pivot_order(None, None)
    assert c.analyze()['lu_nonzeros'] == own['lu_nonzeros']


def test_pivot_order_rejects_non_permutations():
    desc = workloads.ring_oscillator(stages=11, tstop=2e-9)
    c = ngmi.Circuit(desc)
    n = c.analyze()['unknowns']
    rows = np.arange(1, n + 1)
    with pytest.raises(ngmi.EngineError):
        c.set_pivot_order(rows[:-1], rows[:-1])
    bad = rows.copy()
    bad[0] = bad[1]
    with pytest.raises(ngmi.EngineError):
        c.set_pivot_order(bad, rows)
    with pytest.raises(ngmi.EngineError):
        c.set_pivot_order(rows, rows + 1)
