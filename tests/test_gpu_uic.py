""".tran ... uic in the batched engine, on the GPU, against the oracle.

Under UIC (MODEUIC) the operating point is a single load on the .ic vector
(CKTic, cktic.c:19-80; niiter.c:41-46), BSIM4 instances without their own
icvds/icvgs/icvbs take them from their terminals' .ic values
(b4getic.c:32-55), capacitors their initial voltage (capgetic.c:19-35);
the time-0 point is not output (dctran.c:416-419) and the first step sets
a breakpoint at tstep (dctran.c:509-512).

Bars: the waveform bar of the transient tests (same accepted timepoints,
values within 1e-6 of full scale and of themselves above 1e-2 of full
scale) and the oracle's Newton iteration count, on RC charging from
.ic v(2) = 0.5 and from no .ic (two topology groups in one batch), ring11
with every node's .ic given, ring11 with the supply node's .ic left out
(its PMOS start from 0 V on VDD), ring11 with the instance ICs given
explicitly, and ring11 under Gear integration.
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


def _oracle(desc):
    T = po.Transient(desc)
    rc, t, x = T.run(desc['tstep'], desc['tstop'])
    assert rc == 0, rc
    return T, t, x


def _compare(tag, b, c, T, t, x, nn):
    st = b.stats(c)
    assert st['status'] == 0, st
    te, xe = b.waveform(c, b.circuit_info(c)['unknowns'])
    print('%s: engine %d pts from t = %.3g, %d iterations; oracle %d pts, %d iterations'
          % (tag, len(te), te[0], st['iterations'], len(t), T.stats()['iterations']), flush=True)
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
    assert st['iterations'] == T.stats()['iterations']


def _run_engine(descs):
    b = ngmi.Batch([ngmi.Circuit(d) for d in descs], max_points=100000)
    rc, _ = b.tran(descs[0]['tstep'], descs[0]['tstop'])
    assert rc == 0, ngmi.last_error()
    return b


def test_rc_uic():
    descs = [workloads.rc_charge(ic={2: 0.5}, uic=True), workloads.rc_charge(uic=True)]
    print('case: RC uic from v(2) = 0.5 and from 0, one batch', flush=True)
    b = _run_engine(descs)
    for c, d in enumerate(descs):
        T, t, x = _oracle(d)
        assert t[0] > 0
        _compare('rc[%d]' % c, b, c, T, t, x, d['nnodes'])


def _ring_ic():
    d = workloads.ring_oscillator(stages=11, tstop=0.3e-9)
    vdd = d['nnodes']
    ic = {k: (0.0 if k % 2 else 1.0) for k in range(1, vdd)}
    ic[1] = 0.51
    ic[vdd] = 1.0
    return d, ic


def test_ring_uic():
    d, ic = _ring_ic()
    vdd = d['nnodes']
    cases = [('all nodes', workloads.with_uic(d, ic), 3),
             ('supply left out', workloads.with_uic(d, {k: v for k, v in ic.items() if k != vdd}), 1)]
    explicit = workloads.with_uic(d, ic)
    mos = []
    for key, dn, g, s, bb, p in explicit['mos']:
        v = lambda u: ic.get(u, 0.0)  # noqa: E731
        mos.append((key, dn, g, s, bb, dict(p, icvds=v(dn) - v(s), icvgs=v(g) - v(s), icvbs=v(bb) - v(s))))
    cases.append(('instance ICs given', dict(explicit, mos=mos), 1))
    gear = workloads.with_uic(d, ic)
    gear['options'] = dict(gear['options'], method=2)
    cases.append(('Gear', gear, 2))
    for name, desc, copies in cases:
        print('case: ring11 uic, %s, %d circuit(s)' % (name, copies), flush=True)
        T, t, x = _oracle(desc)
        b = _run_engine([desc] * copies)
        for c in range(copies):
            _compare('%s[%d]' % (name, c), b, c, T, t, x, desc['nnodes'])
