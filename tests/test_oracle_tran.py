"""Pin the oracle's DCtran / NIiter / CKTterr restatement (oracle/tran_ref.c).

* RC step response: the closed form, and the reference's own golden
  (tests/general/rc.out, 4 significant digits on the 0.1 s print grid).
  That golden was produced by an older ngspice whose PULSE defaults were
  PW = PER = TSTOP; the same waveform is given here explicitly.
* The reference's KLU (oracle/_ref, when built) and the oracle's dense LU
  give the same accepted timepoints and waveforms to solver rounding.
* The 16-bit mirror adder settles to the arithmetic sum of its inputs before
  and after the input switch; the 11-stage ring oscillates rail to rail.
"""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, '..', 'oracle'))
from engine_path import workloads  # noqa: E402
import pyoracle as po  # noqa: E402

RC = {'nnodes': 2, 'models': {}, 'mos': [], 'caps': [(2, 0, 1.0)], 'res': [(1, 2, 1.0)],
      'vsrc': [(1, 0, 0.0, (0, 1, 0, 0.1, 0.1, 7, 7))], 'ic': {}, 'tstep': 0.1, 'tstop': 7.0}


def rc_closed_form(t):
    a = (0.1 - (1 - np.exp(-0.1))) / 0.1
    return np.where(t <= 0.1, (t - (1 - np.exp(-t))) / 0.1, 1 - (1 - a) * np.exp(-(t - 0.1)))


def test_rc_closed_form():
    T = po.Transient(RC)
    rc, t, x = T.run(0.1, 7.0)
    assert rc == 0
    err = np.abs(x[:, 2] - rc_closed_form(t)).max()
    # trapezoidal integration under LTE control at reltol 1e-3
    assert err < 1e-3, err
    assert abs(t[-1] - 7.0) < 1e-12


def test_rc_reference_golden():
    gold = json.load(open(os.path.join(HERE, 'golden', 'rc_golden.json')))
    tg = np.array([r[0] for r in gold['t_v2']])
    vg = np.array([r[1] for r in gold['t_v2']])
    T = po.Transient(RC)
    rc, t, x = T.run(0.1, 7.0)
    assert rc == 0
    v = np.interp(tg, t, x[:, 2])
    # 4 printed digits (half a unit of the last digit is 5e-5) plus the
    # difference of the interpolated print grid between versions
    assert np.abs(v - vg).max() < 2e-4, np.abs(v - vg).max()
    # accepted points: the older version reports 87 data rows
    assert abs(len(t) - gold['data_rows']) <= 2, (len(t), gold['data_rows'])


def test_reference_klu_matches_dense_lu():
    d = workloads.mc_adder(0)
    T1 = po.Transient(d)
    T2 = po.Transient(d)
    if not T2.use_reference_klu():
        pytest.skip('oracle/_ref not built (needs the reference sources)')
    _, t1, x1 = T1.run(d['tstep'], 2e-9)
    _, t2, x2 = T2.run(d['tstep'], 2e-9)
    assert len(t1) == len(t2)
    assert np.array_equal(t1, t2)
    assert np.abs(x1 - x2).max() < 1e-8


@pytest.mark.parametrize('seed', [0, 5])
def test_adder_settles_to_the_sum(seed):
    d = workloads.mc_adder(seed, tstop=3e-9)
    T = po.Transient(d)
    T.use_reference_klu()   # faster when oracle/_ref is built; dense LU otherwise
    rc, t, x = T.run(d['tstep'], 3e-9)
    assert rc == 0
    names = d['names']
    bits = workloads.ADDER_BITS
    s_nodes = [names['s%d' % i] for i in range(bits)]
    cout = names['co%d' % (bits - 1)]
    for vec, when in ((d['bits0'], 0.95e-9), (d['bits1'], 3e-9)):
        a = sum(int(vec[i]) << i for i in range(bits))
        b = sum(int(vec[bits + i]) << i for i in range(bits))
        s = a + b + int(vec[2 * bits])
        k = min(np.searchsorted(t, when), len(t) - 1)
        got = sum((1 << i) for i, n in enumerate(s_nodes) if x[k, n] > 0.5)
        got += (1 << bits) if x[k, cout] > 0.5 else 0
        assert got == s, (when, got, s)


def test_ring_oscillates():
    d = workloads.ring_oscillator(stages=11, tstop=2e-9)
    T = po.Transient(d)
    rc, t, x = T.run(d['tstep'], 2e-9)
    assert rc == 0
    v1 = x[:, 1]
    late = t > 0.5e-9
    assert v1[late].max() > 0.9 and v1[late].min() < 0.1


@pytest.mark.parametrize('order', [1, 2])
def test_gear_coefficients_differentiate_polynomials_exactly(order):
    """NIcomCof (nicomcof.c:60-115) for Gear: sum_k ag[k] q(t_k) is the
    derivative at the new point t_0 of every polynomial of degree <= order
    through the points t_0, t_0 - h, t_0 - h - h1 (variable-step BDF), and
    for order 2 equals the closed-form BDF2 weights."""
    rng = np.random.default_rng(order)
    for _ in range(50):
        h, h1, h2 = rng.uniform(1e-13, 1e-9, size=3)
        ag = po.comcof(2, order, h, [h, h1, h2])
        tk = np.array([0.0, -h, -h - h1])[:order + 1]
        w = ag[:order + 1]
        assert np.all(ag[order + 1:] == 0)
        for deg in range(order + 1):
            got = np.sum(w * tk ** deg)
            want = 1.0 if deg == 1 else 0.0
            scale = np.sum(np.abs(w * tk ** deg)) + 1.0
            assert abs(got - want) <= 1e-12 * scale, (deg, got, want)
        if order == 2:
            bdf2 = [(2 * h + h1) / (h * (h + h1)), -(h + h1) / (h * h1), h / (h1 * (h + h1))]
            assert np.allclose(w, bdf2, rtol=1e-12, atol=0)


def test_trapezoidal_coefficients():
    """NIcomCof trapezoidal (nicomcof.c:33-46): order 1 is backward Euler;
    order 2 with xmu = 0.5 makes NIintegrate's i0 = -i1 ag1 + ag0 (q0 - q1)
    the trapezoidal rule, exact for a quadratic charge."""
    h = 3.7e-11
    ag = po.comcof(1, 1, h, [h])
    assert ag[0] == 1 / h and ag[1] == -1 / h
    ag = po.comcof(1, 2, h, [h, 2 * h])
    t1, t0 = 1.3e-9, 1.3e-9 + h
    q = lambda t: 4.0 * t * t + 2.0 * t
    i = lambda t: 8.0 * t + 2.0
    i0 = -i(t1) * ag[1] + ag[0] * (q(t0) - q(t1))
    assert abs(i0 - i(t0)) <= 1e-9 * abs(i(t0))
    with pytest.raises(ValueError):
        po.comcof(1, 3, h, [h, h, h])


def test_gear_rc_transient():
    """Gear integration (method=gear) on the RC step against its closed form,
    and against trapezoidal integration of the same circuit."""
    desc = {'nnodes': 2, 'models': {}, 'mos': [], 'caps': [(2, 0, 1.0)], 'res': [(1, 2, 1.0)],
            'vsrc': [(1, 0, 0.0, (0, 1, 0, 0.1, 0.1, 7, 7))], 'ic': {}, 'tstep': 0.1,
            'options': {'method': 2, 'maxord': 6}}
    T = po.Transient(desc)
    rc, t, x = T.run(0.1, 7.0)
    assert rc == 0
    a = (0.1 - (1 - np.exp(-0.1))) / 0.1
    ref = np.where(t <= 0.1, (t - (1 - np.exp(-t))) / 0.1, 1 - (1 - a) * np.exp(-(t - 0.1)))
    assert np.abs(x[:, 2] - ref).max() < 2e-3


OP_STEPPING = {
    'gmin': {'noopiter': 1},
    'src': {'noopiter': 1, 'gminsteps': 0},
    's3gmin': {'noopiter': 1, 'gminsteps': 5},
    's3src': {'noopiter': 1, 'gminsteps': 0, 'srcsteps': 5},
}


@pytest.mark.parametrize('how', list(OP_STEPPING))
def test_operating_point_stepping_reaches_the_direct_solution(how):
    """CKTop's stepping restatement (dynamic_gmin + new_gmin, gillespie_src,
    spice3_gmin in 5 steps, spice3_src in 5 steps) forced with noopiter
    converges to the operating point the direct Newton finds, within the
    Newton tolerance, and the transient that follows accepts the same
    timepoints."""
    desc = workloads.ring_oscillator(stages=11)
    ref = po.Transient(dict(desc))
    rc, t0, x0 = ref.run(desc['tstep'], 0.5e-9)
    assert rc == 0
    d = dict(desc)
    d['options'] = dict(OP_STEPPING[how])
    T = po.Transient(d)
    rc, t, x = T.run(desc['tstep'], 0.5e-9)
    assert rc == 0
    assert T.stats()['iterations'] > ref.stats()['iterations']
    assert len(t) == len(t0) and np.allclose(t, t0, rtol=1e-9)
    assert np.abs(x[0] - x0[0]).max() < 1e-6 * 10
    assert np.abs(x - x0).max() < 1e-6


def test_sparse_mode_matches_dense_mode():
    """The oracle's sparse mode (the matrix only in CSC, for large meshes) adds
    every stamp in the same order as its dense mode, so with the reference
    KLU both give the same run bit for bit: the ring and the adder forced
    into sparse mode against their dense runs."""
    import pyoracle as po
    from engine_path import workloads
    for desc, tstop in ((workloads.ring_oscillator(stages=11, tstop=1e-9), 1e-9),
                        (workloads.mc_adder(1), 1e-9)):
        Td = po.Transient(desc)
        if not Td.use_reference_klu():
            pytest.skip('reference KLU not built')
        old = po.set_sparse_above(0)
        try:
            Ts = po.Transient(desc)
            assert Ts.use_reference_klu()
        finally:
            po.set_sparse_above(old)
        rd, td, xd = Td.run(desc['tstep'], tstop)
        rs, ts, xs = Ts.run(desc['tstep'], tstop)
        assert rd == 0 and rs == 0
        assert np.array_equal(td, ts) and np.array_equal(xd, xs)
        assert Td.stats() == Ts.stats()
