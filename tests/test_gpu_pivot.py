"""KLU's numeric pivoting in the batched engine, on the GPU, against the
oracle (the reference KLU from oracle/_ref where a test says so).

Where NIiter factors afresh (INITJCT, the first INITTRAN iteration, the
iteration after either, and after a refactor met a zero pivot,
niiter.c:107-195, 335-347), the step kernel tests whether klu_factor would
keep the static pivots (klu_kernel.c:370-530 on the row-scaled matrix); if
not, or a static pivot is zero, that circuit's iteration continues next
pass with KLU's pivot search on a dense workspace, and refactors on those
rows until the next fresh factorization.

Bars: the waveform bar of the transient tests (same accepted timepoints,
values within 1e-6 of full scale and of themselves above 1e-2 of full
scale); on KLU's ordering, the pivot rows in use at the end equal the
reference KLU's Numeric->Pnum and the counts of fresh factorizations and
refactors equal the reference's; a circuit whose operating point is
singular ends with the oracle's return code.
"""
import os
import sys

import numpy as np
import pytest

from engine_path import ngmi, workloads

sys.path.insert(0, os.path.join(os.path.dirname(__file__), '..', 'oracle'))
import pyoracle as po  # noqa: E402

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu

WAVE_TOL = 1e-6


def _oracle(desc, klu=False):
    T = po.Transient(desc)
    if klu:
        assert T.use_reference_klu()
    rc, t, x = T.run(desc['tstep'], desc['tstop'])
    return T, rc, t, x


def _engine(descs, klu_order=False, order=None):
    circ = []
    for d in descs:
        c = ngmi.Circuit(d)
        if klu_order:
            c.use_klu_order()
        if order is not None:
            c.set_pivot_order(*order)
        circ.append(c)
    b = ngmi.Batch(circ, max_points=100000)
    rc, passes = b.tran(descs[0]['tstep'], descs[0]['tstop'])
    return b, rc, passes


def _compare(tag, b, c, t, x, nn):
    st = b.stats(c)
    assert st['status'] == 0, st
    te, xe = b.waveform(c, b.info()['unknowns'])
    print('%s: engine %d pts, %d iterations; oracle %d pts' % (tag, len(te), st['iterations'], len(t)),
          flush=True)
    assert len(te) == len(t)
    assert np.allclose(te, t, rtol=1e-9, atol=1e-21)
    xo = x[:, 1:]
    scale = np.ones(xo.shape[1])
    scale[nn:] = np.maximum(np.abs(xo[:, nn:]).max(axis=0), 1e-30)
    d = np.abs(xe - xo)
    err = (d / scale).max()
    big = np.abs(xo) >= 1e-2 * scale
    rel = float((d[big] / np.abs(xo[big])).max()) if big.any() else 0.0
    print('%s: max scaled error %.3e, relative %.3e' % (tag, err, rel), flush=True)
    assert err <= WAVE_TOL and rel <= WAVE_TOL
    return st


def test_klu_order_pivots_as_the_reference_klu():
    if po.klu_ref() is None:
        pytest.skip('reference KLU (oracle/_ref) not built')
    desc = workloads.ring_with_divider()
    print('case: ring11 with a floating source, KLU ordering, 3 circuits', flush=True)
    T, rc, t, x = _oracle(desc, klu=True)
    assert rc == 0
    rows, cols, cnt = T.klu_order()
    b, rc, passes = _engine([desc] * 3, klu_order=True)
    assert rc == 0, ngmi.last_error()
    for c in range(3):
        st = _compare('divider[%d]' % c, b, c, t, x, desc['nnodes'])
        pv = b.pivot_stats(c)
        print('  engine pivots: pv %d, %d fresh factorizations, %d refactors; reference %s'
              % (pv['pv'], pv['factors'], pv['refactors'], cnt), flush=True)
        assert pv['pv'] == 1
        assert (pv['rows'] == rows).all() and (pv['cols'] == cols).all()
        assert pv['factors'] == cnt['factors'] == cnt['offdiag_factors']
        assert pv['refactors'] == cnt['refactors']


def test_own_order_with_pivoting():
    desc = workloads.ring_with_divider()
    print('case: ring11 with a floating source, engine ordering', flush=True)
    _, rc, t, x = _oracle(desc)
    assert rc == 0
    b, rc, _ = _engine([desc])
    assert rc == 0, ngmi.last_error()
    _compare('divider', b, 0, t, x, desc['nnodes'])
    pv = b.pivot_stats(0)
    print('  engine pivots: pv %d, %d fresh factorizations, %d refactors'
          % (pv['pv'], pv['factors'], pv['refactors']), flush=True)


def test_zero_static_pivot_is_factored_afresh():
    """A pivot sequence whose first pivot is the source branch's empty
    diagonal: every static factorization meets an exact zero, and the
    iteration goes on with KLU's pivot search (niiter.c:160-195) instead of
    a rejected step."""
    desc = workloads.ring_oscillator(stages=11, tstop=0.3e-9)
    n = desc['nnodes'] + len(desc['vsrc'])
    order = [n] + list(range(1, n))
    print('case: ring11, pivot order starting at the branch diagonal', flush=True)
    _, rc, t, x = _oracle(desc)
    assert rc == 0
    b, rc, _ = _engine([desc, desc], order=(order, order))
    assert rc == 0, ngmi.last_error()
    for c in range(2):
        st = _compare('zero-pivot[%d]' % c, b, c, t, x, desc['nnodes'])
        assert st['rejected'] == 0
        pv = b.pivot_stats(c)
        print('  engine pivots: pv %d, %d fresh factorizations, %d refactors'
              % (pv['pv'], pv['factors'], pv['refactors']), flush=True)
        assert pv['pv'] == 1 and pv['factors'] >= 4 and pv['refactors'] > 0
        assert pv['rows'][0] != n


def test_floating_node():
    """A capacitor-only node: the operating point's matrix is singular, the
    Newton run and gmin and source stepping fail, and CKTop ends with
    optran (cktop.c:93-105, optran.c:285-851); with optran deselected the
    run ends with the stepping's error, as the oracle's."""
    chain = workloads.inverter_chain()
    print('case: inverter chain with a capacitor-only node, optran deselected', flush=True)
    desc = workloads.with_floating_node(chain, node=3)
    desc['options'] = dict(desc['options'], optran=None)
    T, rco, _, _ = _oracle(desc)
    b, rc, _ = _engine([desc])
    st = b.stats(0)
    print('  oracle rc %d %s; engine %s' % (rco, T.stats(), st), flush=True)
    assert rco != 0 and st['status'] == rco and st['iterations'] == T.stats()['iterations']
    print('case: the same through optran (the default), 2 circuits', flush=True)
    desc = workloads.with_floating_node(chain, node=3)
    T, rco, t, x = _oracle(desc)
    assert rco == 0
    b, rc, _ = _engine([desc, desc])
    assert rc == 0, ngmi.last_error()
    for c in range(2):
        st = _compare('optran[%d]' % c, b, c, t, x, desc['nnodes'])
        assert st['iterations'] == T.stats()['iterations']
    print('case: ring11 with a capacitor-only node and gshunt = 1e-12', flush=True)
    desc = workloads.with_floating_node(workloads.ring_oscillator(stages=11, tstop=0.3e-9), node=1,
                                        gshunt=1e-12)
    _, rco, t, x = _oracle(desc)
    assert rco == 0
    b, rc, _ = _engine([desc])
    assert rc == 0, ngmi.last_error()
    _compare('floating+gshunt', b, 0, t, x, desc['nnodes'])


def test_workloads_pivoting():
    """The benchmark circuits on KLU's ordering keep the static pivots at
    every fresh factorization, as the reference KLU does (no off-diagonal
    pivot, the reference's count where oracle/_ref is built); on the
    engine's own ordering KLU's test may reject a diagonal at a few fresh
    factorizations (at most the four of a short run), and the run goes on
    from the dense factorization."""
    descs = [('ring101', workloads.ring_oscillator(tstop=0.2e-9))] + \
            [('adder%d' % s, workloads.mc_adder(s, tstop=0.1e-9)) for s in range(2)]
    for name, d in descs:
        ref = None
        if po.klu_ref() is not None:
            T, rc, _, _ = _oracle(d, klu=True)
            assert rc == 0
            ref = T.klu_order()[2]['offdiag_factors']
        b, rc, _ = _engine([d], klu_order=True)
        assert rc == 0, ngmi.last_error()
        pk = b.pivot_stats(0)
        b, rc, _ = _engine([d])
        assert rc == 0, ngmi.last_error()
        po_ = b.pivot_stats(0)
        print('  %s: KLU ordering %d pivoted factorizations (reference %s); own ordering %d, pv %d'
              % (name, pk['factors'], ref, po_['factors'], po_['pv']), flush=True)
        assert pk['factors'] == 0 and pk['pv'] == 0
        if ref is not None:
            assert ref == 0
        assert po_['factors'] <= 4
