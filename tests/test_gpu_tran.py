"""GPU parity of the batched transient engine against the oracle's DCtran
restatement.

Bar: the engine accepts the same timepoints as the oracle (times equal to
1e-9 relative) and every recorded node voltage / branch current agrees
within 1e-6 of the waveform's full scale (1 V for node voltages; the
largest magnitude for branch currents), and within 1e-6 of its own value
wherever that value is at least 1e-2 of full scale (10 mV for a node).
"""
import numpy as np
import pytest

from engine_path import ngmi, workloads
import sys
import os

sys.path.insert(0, os.path.join(os.path.dirname(__file__), '..', 'oracle'))
import pyoracle as po  # noqa: E402

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu

WAVE_TOL = 1e-6


def _oracle(desc, tstop):
    T = po.Transient(desc)
    rc, t, x = T.run(desc['tstep'], tstop)
    assert rc == 0, rc
    return t, x, T.stats()


def _engine(descs, tstop):
    circ = [ngmi.Circuit(d) for d in descs]
    b = ngmi.Batch(circ, max_points=200000)
    rc, passes = b.tran(descs[0]['tstep'], tstop)
    assert rc == 0, (rc, ngmi.last_error())
    n = b.info()['unknowns']
    out = []
    for c in range(len(descs)):
        st = b.stats(c)
        assert st['status'] == 0, st
        t, x = b.waveform(c, n)
        out.append((t, x, st))
    return out, b.info(), passes


def _compare(tag, eng, ora, nn):
    te, xe, se = eng
    to, xo, so = ora
    print('%s: engine %d pts (%d rejected, %d iters), oracle %d pts (%d rejected, %d iters)'
          % (tag, len(te), se['rejected'], se['iterations'], len(to), so['rejected'],
             so['iterations']), flush=True)
    assert len(te) == len(to), (len(te), len(to))
    assert np.allclose(te, to, rtol=1e-9, atol=1e-21)
    # oracle x has a ground column 0; engine records unknowns 1..n
    xo = xo[:, 1:]
    scale = np.ones(xo.shape[1])
    scale[nn:] = np.maximum(np.abs(xo[:, nn:]).max(axis=0), 1e-30)
    d = np.abs(xe - xo)
    err = (d / scale).max()
    # relative to each value itself at the reference's timepoints, for the
    # values above 10 mV (node voltages) or 1e-2 of full scale (branch
    # currents); below that the full-scale bar above holds
    big = np.abs(xo) >= 1e-2 * scale
    rel = float((d[big] / np.abs(xo[big])).max()) if big.any() else 0.0
    print('%s: max scaled waveform error %.3e, max relative error %.3e (|x| >= 1e-2 of full scale)'
          % (tag, err, rel), flush=True)
    assert err <= WAVE_TOL
    assert rel <= WAVE_TOL


def test_rc_step_against_oracle_and_closed_form():
    desc = {'nnodes': 2, 'models': {}, 'mos': [], 'caps': [(2, 0, 1.0)], 'res': [(1, 2, 1.0)],
            'vsrc': [(1, 0, 0.0, (0, 1, 0, 0.1, 0.1, 7, 7))], 'ic': {}, 'tstep': 0.1,
            'tstop': 7.0}
    print('case: RC step, .tran 0.1 7', flush=True)
    (eng,), info, _ = _engine([desc], 7.0)
    ora = _oracle(desc, 7.0)
    _compare('rc', eng, ora, 2)
    t, x, _ = eng
    a = (0.1 - (1 - np.exp(-0.1))) / 0.1
    ref = np.where(t <= 0.1, (t - (1 - np.exp(-t))) / 0.1, 1 - (1 - a) * np.exp(-(t - 0.1)))
    assert np.abs(x[:, 1] - ref).max() < 1e-3


def test_small_ring_oscillator():
    desc = workloads.ring_oscillator(stages=11, tstop=2e-9)
    print('case: 11-stage ring, .tran 10p 2n', flush=True)
    (eng,), info, _ = _engine([desc], 2e-9)
    print('lu', info, flush=True)
    ora = _oracle(desc, 2e-9)
    _compare('ring11', eng, ora, desc['nnodes'])


def test_ring101_full_length():
    """The reference's own CPU-sized case at full length: 101-stage ring,
    .ic v(1)=0.51, .tran 10p 200n (about 45 oscillation periods, 2e4
    accepted points)."""
    desc = workloads.ring_oscillator()
    tstop = desc['tstop']
    print('case: 101-stage ring, .tran 10p 200n', flush=True)
    (eng,), info, passes = _engine([desc], tstop)
    print('lu', info, 'passes', passes, flush=True)
    ora = _oracle(desc, tstop)
    _compare('ring101', eng, ora, desc['nnodes'])


def test_adder_batch():
    descs = [workloads.mc_adder(s) for s in (0, 1, 2)]
    tstop = 2.5e-9
    print('case: 16-bit mirror adder, seeds 0..2, .tran 5p 2.5n', flush=True)
    engs, info, passes = _engine(descs, tstop)
    print('lu', info, 'passes', passes, flush=True)
    for s, d in enumerate(descs):
        ora = _oracle(d, tstop)
        _compare('adder%d' % s, engs[s], ora, d['nnodes'])


def test_klu_pivot_order():
    """The engine on KLU's own pivot sequence (Pnum/Q of the reference KLU,
    compiled into oracle/_ref) against the oracle solving with that KLU."""
    print('case: KLU pivot order, 11-stage ring .tran 10p 2n and adder seed 0 .tran 5p 2.5n',
          flush=True)
    for tag, desc, tstop in [('ring11-klu', workloads.ring_oscillator(stages=11, tstop=2e-9), 2e-9),
                             ('adder0-klu', workloads.mc_adder(0), 2.5e-9)]:
        T = po.Transient(desc)
        if not T.use_reference_klu():
            pytest.skip('reference KLU (oracle/_ref) not built')
        rc, t, x = T.run(desc['tstep'], tstop)
        assert rc == 0, rc
        rows, cols, cnt = T.klu_order()
        circ = ngmi.Circuit(desc)
        circ.set_pivot_order(rows, cols)
        b = ngmi.Batch([circ], max_points=200000)
        rc, _ = b.tran(desc['tstep'], tstop)
        assert rc == 0, (rc, ngmi.last_error())
        info = b.info()
        print(tag, 'lu', info, 'klu', cnt, flush=True)
        st = b.stats(0)
        assert st['status'] == 0, st
        te, xe = b.waveform(0, info['unknowns'])
        _compare(tag, (te, xe, st), (t, x, T.stats()), desc['nnodes'])


def test_klu_order_block_solve():
    """The engine on KLU's ordering computed by itself (use_klu_order: the
    restated klu_analyze), factoring only the diagonal blocks of KLU's block
    triangular form and solving block by block from the last, as klu_factor
    and klu_solve do, against the oracle solving with the reference KLU:
    the 101-stage ring (3 blocks) and the adder (69 blocks, 1042 entries
    off the blocks)."""
    print('case: KLU order + block solve, 101-stage ring .tran 10p 2n and adder seed 0 .tran 5p 2.5n',
          flush=True)
    for tag, desc, tstop in [('ring101-kluorder', workloads.ring_oscillator(tstop=2e-9), 2e-9),
                             ('adder0-kluorder', workloads.mc_adder(0), 2.5e-9)]:
        T = po.Transient(desc)
        if not T.use_reference_klu():
            pytest.skip('reference KLU (oracle/_ref) not built')
        rc, t, x = T.run(desc['tstep'], tstop)
        assert rc == 0, rc
        circ = ngmi.Circuit(desc)
        circ.use_klu_order()
        b = ngmi.Batch([circ, ngmi.Circuit(desc)], max_points=200000)
        rc, _ = b.tran(desc['tstep'], tstop)
        assert rc == 0, (rc, ngmi.last_error())
        info = b.info()
        print(tag, 'lu', info, flush=True)
        for c in range(2):
            st = b.stats(c)
            assert st['status'] == 0, st
            te, xe = b.waveform(c, info['unknowns'])
            _compare(tag, (te, xe, st), (t, x, T.stats()), desc['nnodes'])


def test_gear_method():
    """method=gear (NIcomCof's Gear branch, NIintegrate's Gear sum and
    CKTterr's Gear factor) on the engine against the oracle."""
    print('case: Gear integration, RC .tran 0.1 7, 11-stage ring .tran 10p 2n, adder seed 0 '
          '.tran 5p 2.5n', flush=True)
    rc_desc = {'nnodes': 2, 'models': {}, 'mos': [], 'caps': [(2, 0, 1.0)], 'res': [(1, 2, 1.0)],
               'vsrc': [(1, 0, 0.0, (0, 1, 0, 0.1, 0.1, 7, 7))], 'ic': {}, 'tstep': 0.1}
    cases = [('rc-gear', rc_desc, 7.0), ('ring11-gear', workloads.ring_oscillator(stages=11), 2e-9),
             ('adder0-gear', workloads.mc_adder(0), 2.5e-9)]
    for tag, desc, tstop in cases:
        desc = dict(desc)
        desc['options'] = dict(desc.get('options', {}), method=2, maxord=6)
        (eng,), info, _ = _engine([desc], tstop)
        ora = _oracle(desc, tstop)
        _compare(tag, eng, ora, desc['nnodes'])


def test_uneven_batch_rings():
    """Rings of 11, 21, 13 and 31 stages, two of each with different loads,
    interleaved in one batch: four topology groups advanced by one launch
    per pass; every circuit against the oracle."""
    print('case: uneven batch, rings of 11/21/13/31 stages x 2, .tran 10p 1n', flush=True)
    stages = [11, 21, 13, 31, 11, 13, 21, 31]
    loads = [1e-15, 1e-15, 1e-15, 1e-15, 1.5e-15, 1.5e-15, 1.5e-15, 1.5e-15]
    tstop = 1e-9
    descs = [workloads.ring_oscillator(stages=k, cload=cl, tstop=tstop)
             for k, cl in zip(stages, loads)]
    b = ngmi.Batch([ngmi.Circuit(d) for d in descs], max_points=200000)
    assert b.groups() == 4
    rc, passes = b.tran(descs[0]['tstep'], tstop)
    assert rc == 0, (rc, ngmi.last_error())
    for c, d in enumerate(descs):
        ci = b.circuit_info(c)
        assert ci['unknowns'] == d['nnodes'] + 1
        st = b.stats(c)
        assert st['status'] == 0, st
        te, xe = b.waveform(c)
        _compare('ring%d-c%d' % (stages[c], c), (te, xe, st), _oracle(d, tstop), d['nnodes'])


def _oracle_klu(desc, tstop):
    T = po.Transient(desc)
    T.use_reference_klu()
    rc, t, x = T.run(desc['tstep'], tstop)
    assert rc == 0, rc
    return t, x, T.stats()


def test_uneven_batch_long_rings():
    """The uneven batch's long end: rings of 151, 231 and 301 stages (the
    301-stage ring is the longest of the 2048-ring workload and its per-circuit
    LDS program the largest) beside an 11-stage ring, two loads each for the
    long ones, in one batch of four topology groups; .tran 10p 3n, every
    circuit against the oracle (which solves with the reference KLU)."""
    stages = [11, 151, 231, 301, 151, 301]
    loads = [1e-15, 1e-15, 1e-15, 1e-15, 1.5e-15, 1.5e-15]
    tstop = 3e-9
    print('case: uneven batch, rings of %s stages, .tran 10p 3n' % stages, flush=True)
    descs = [workloads.ring_oscillator(stages=k, cload=cl, tstop=tstop)
             for k, cl in zip(stages, loads)]
    b = ngmi.Batch([ngmi.Circuit(d) for d in descs], max_points=200000)
    assert b.groups() == 4
    rc, passes = b.tran(descs[0]['tstep'], tstop)
    assert rc == 0, (rc, ngmi.last_error())
    for c, d in enumerate(descs):
        ci = b.circuit_info(c)
        assert ci['unknowns'] == d['nnodes'] + 1
        st = b.stats(c)
        assert st['status'] == 0, st
        te, xe = b.waveform(c)
        _compare('ring%d-c%d' % (stages[c], c), (te, xe, st), _oracle_klu(d, tstop), d['nnodes'])


def test_operating_point_stepping():
    """CKTop's fallbacks (cktop.c:161-720) forced with noopiter: dynamic and
    true gmin stepping (gminsteps 1), Gillespie source stepping (gminsteps
    0, srcsteps 1), and the spice3 variants in 5 steps each, on the engine
    against the oracle; the Newton iteration counts, which include every
    stepping solve, must be equal."""
    print('case: operating point by gmin and source stepping (default and spice3), 11-stage ring '
          '.tran 10p 1n, adder seed 0 .tran 5p 1.5n', flush=True)
    for tag, desc, tstop in [('ring11', workloads.ring_oscillator(stages=11), 1e-9),
                             ('adder0', workloads.mc_adder(0), 1.5e-9)]:
        for how, opt in [('gmin', {'noopiter': 1}), ('src', {'noopiter': 1, 'gminsteps': 0}),
                         ('s3gmin', {'noopiter': 1, 'gminsteps': 5}),
                         ('s3src', {'noopiter': 1, 'gminsteps': 0, 'srcsteps': 5})]:
            d = dict(desc)
            d['options'] = dict(opt)
            (eng,), info, _ = _engine([d], tstop)
            ora = _oracle(d, tstop)
            assert eng[2]['iterations'] == ora[2]['iterations'], (eng[2], ora[2])
            _compare('%s-%s' % (tag, how), eng, ora, d['nnodes'])
