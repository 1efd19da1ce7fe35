""".tran ... uic in the oracle's DCtran, on the CPU.

Under UIC the operating point is one load on the initial conditions
(niiter.c:41-46): CKTic puts the .ic values into CKTrhs/CKTrhsOld, 0 at the
other unknowns (cktic.c:19-80), BSIM4getic gives each instance the terminal
differences of that vector as icvds/icvgs/icvbs where they are not given
(b4getic.c:32-55), CAPgetic each capacitor its initial voltage
(capgetic.c:19-35); the time-0 point is not output (dctran.c:416-419) and
the first step sets a breakpoint at tstep (dctran.c:509-512).

Bars:
* RC charging through 1 kOhm into 1 pF from .ic v(2) = 0.5 (and from no
  .ic, i.e. 0; and from 0.5 under Gear integration): v(2) within 5e-5 V
  of v - (v - v0) exp(-t/RC), the closed
  form, at every output point; the first output point after t = 0; tstep
  among the output times; without uic the same .ic gives the t = 0 point
  at v0 (the .ic row of the operating point, cktload.c:146-173);
* ring11 with every node's .ic given: the instance ICs taken from the
  nodes equal, bit for bit in the waveform, the instance ICs given
  explicitly as v(d) - v(s), v(g) - v(s), v(b) - v(s); the same run with
  the drain and source of that difference swapped differs (so the
  comparison is sensitive to what the instance ICs are).
"""
import os
import sys

import numpy as np

from engine_path import workloads

sys.path.insert(0, os.path.join(os.path.dirname(__file__), '..', 'oracle'))
import pyoracle as po  # noqa: E402

R, C = 1e3, 1e-12


def _run(desc):
    T = po.Transient(desc)
    rc, t, x = T.run(desc['tstep'], desc['tstop'])
    assert rc == 0, rc
    return T, t, x


def test_rc_charge_closed_form():
    for v0, ic, method in ((0.5, {2: 0.5}, 1), (0.0, {}, 1), (0.5, {2: 0.5}, 2)):
        d = workloads.rc_charge(ic=ic, uic=True)
        d['options'] = dict(d['options'], method=method)
        T, t, x = _run(d)
        exact = 1.0 - (1.0 - v0) * np.exp(-t / (R * C))
        err = np.abs(x[:, 2] - exact).max()
        print('uic, %s, v(2) from %.2f: %d points from t = %.3g, max error %.2e V, %s'
              % ('gear' if method == 2 else 'trap', v0, len(t), t[0], err, T.stats()))
        assert t[0] > 0
        assert np.isclose(t, d['tstep'], rtol=1e-9, atol=0).any()
        assert np.abs(x[:, 1] - 1.0).max() < 1e-12
        assert err <= 5e-5
    d = workloads.rc_charge(ic={2: 0.5}, uic=False)
    T, t, x = _run(d)
    assert t[0] == 0 and abs(x[0, 2] - 0.5) < 1e-9 and abs(x[0, 1] - 1.0) < 1e-12


def _ring_all_ic():
    d = workloads.ring_oscillator(stages=11, tstop=0.3e-9)
    vdd = d['nnodes']
    ic = {k: (0.0 if k % 2 else 1.0) for k in range(1, vdd)}
    ic[1] = 0.51
    ic[vdd] = 1.0
    return workloads.with_uic(d, ic), ic


def _with_instance_ic(d, ic, swap=False):
    mos = []
    for m in d['mos']:
        key, dn, g, s, b, p = m
        v = lambda u: ic.get(u, 0.0)  # noqa: E731
        if swap:
            dn, s = s, dn
        p = dict(p, icvds=v(dn) - v(s), icvgs=v(g) - v(s), icvbs=v(b) - v(s))
        mos.append((key, m[1], g, m[3], b, p))
    return dict(d, mos=mos)


def test_bsim4_initial_conditions_from_nodes():
    d, ic = _ring_all_ic()
    _, t, x = _run(d)
    _, t2, x2 = _run(_with_instance_ic(d, ic))
    assert len(t) == len(t2) and (t == t2).all() and (x == x2).all()
    _, t3, x3 = _run(_with_instance_ic(d, ic, swap=True))
    n = min(len(t), len(t3))
    diff = np.abs(x3[:n] - x[:n]).max()
    print('ring11 uic: %d points; instance ICs from the nodes equal the explicit ones; '
          'with drain and source swapped the waveform moves by %.3g V' % (len(t), diff))
    assert diff > 1e-3
