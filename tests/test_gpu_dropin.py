"""The engine behind the reference's own interfaces, one circuit at a time:
the oracle's DCtran / NIiter / CKTload (the restated reference loop) runs
with its MOSFET loads replaced by the engine's DEVload core
(ngmi_devload_run: BSIM4load for the circuit's instances on the GPU,
gathering CKTrhsOld and the GENstate slices, adding the stamps through the
bound matrix positions, voting into CKTnoncon) and its linear solve
replaced by the engine's SMP core (ngmi_smp_factor / ngmi_smp_solve:
LoadGmin_CSC + SMPluFac + SMPsolve on the GPU).  This is how ngspice would
call the engine through the shims a maintainer adds (DEVload devdefs.h:57,
SMP* smpdefs.h:61-86).

Bar: the same accepted timepoints as the pure oracle run and every unknown
within the waveform bar (1e-6 of full scale, and of its own value above
1e-2 of full scale), the same Newton iteration count."""
import numpy as np
import pytest

from engine_path import ngmi, workloads
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(__file__), '..', 'oracle'))
import pyoracle as po  # noqa: E402

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu

WAVE_TOL = 1e-6


def _dropin_objects(T, desc):
    """the engine's cores for the oracle circuit T: instance set in the order
    the oracle's CKTload adds instances (newest first), their nodes, state
    offsets and dense-matrix positions; the SMP object on T's pattern"""
    n = T.n
    order = list(range(len(desc['mos'])))[::-1]
    keys = list(desc['models'].keys())
    models = [ngmi.Model(desc['models'][k], temp=desc.get('temp', 300.15)) for k in keys]
    pnames = sorted(desc['mos'][0][5].keys())
    for m in desc['mos']:
        assert sorted(m[5].keys()) == pnames
    s = ngmi.InstanceSet()
    s.add(models, [keys.index(desc['mos'][k][0]) for k in order],
          {p: np.array([desc['mos'][k][5][p] for k in order]) for p in pnames})
    s.finalize()
    nodes, soff, pos = [], [], []
    for k in order:
        st, nd = T.mos_info(k)
        nodes.append(nd)
        soff.append(st)
        pos.append([r * (n + 1) + c if (r and c) else -1 for r in nd for c in nd])
    # the state vectors hold every device's states: 29 per MOSFET, then 2 per
    # capacitor (the oracle numbers them in that order)
    nstates = 29 * len(desc['mos']) + 2 * len(desc['caps'])
    dl = ngmi.DevLoad(s, n, nstates, np.array(nodes), np.array(soff), np.array(pos))
    Ap, Ai = T.csc()
    smp = ngmi.SMP(n, Ap, Ai)
    return s, models, dl, smp


def _run(desc, tstop):
    ora = po.Transient(desc)
    rc, to, xo = ora.run(desc['tstep'], tstop)
    assert rc == 0
    so = ora.stats()
    T = po.Transient(desc)
    keep = _dropin_objects(T, desc)
    dl, smp = keep[2], keep[3]
    T.use_dropin(ngmi.DevLoad.run_fn(), dl.h, ngmi.SMP.factor_fn(), ngmi.SMP.solve_fn(), smp.h)
    rc, te, xe = T.run(desc['tstep'], tstop)
    assert rc == 0, rc
    se = T.stats()
    return (te, xe, se), (to, xo, so)


def _compare(tag, eng, ora, nn):
    te, xe, se = eng
    to, xo, so = ora
    print('%s: drop-in %d pts (%d rejected, %d iters), oracle %d pts (%d rejected, %d iters)'
          % (tag, len(te), se['rejected'], se['iterations'], len(to), so['rejected'], so['iterations']),
          flush=True)
    assert len(te) == len(to)
    assert np.allclose(te, to, rtol=1e-9, atol=1e-21)
    xe, xo = xe[:, 1:], xo[:, 1:]
    scale = np.ones(xo.shape[1])
    scale[nn:] = np.maximum(np.abs(xo[:, nn:]).max(axis=0), 1e-30)
    d = np.abs(xe - xo)
    err = (d / scale).max()
    big = np.abs(xo) >= 1e-2 * scale
    rel = float((d[big] / np.abs(xo[big])).max()) if big.any() else 0.0
    print('%s: max scaled error %.3e, max relative error %.3e' % (tag, err, rel), flush=True)
    assert err <= WAVE_TOL and rel <= WAVE_TOL
    assert se['iterations'] == so['iterations']


def test_dropin_ring11():
    desc = workloads.ring_oscillator(stages=11, tstop=2e-9)
    print('case: drop-in DEVload + SMP, 11-stage ring .tran 10p 2n', flush=True)
    eng, ora = _run(desc, 2e-9)
    _compare('ring11-dropin', eng, ora, desc['nnodes'])


def test_dropin_ring101():
    """the reference's CPU-sized case (configuration 0's circuit), first 5 ns"""
    desc = workloads.ring_oscillator()
    print('case: drop-in DEVload + SMP, 101-stage ring .tran 10p 5n', flush=True)
    eng, ora = _run(desc, 5e-9)
    _compare('ring101-dropin', eng, ora, desc['nnodes'])


def test_dropin_adder():
    desc = workloads.mc_adder(0)
    print('case: drop-in DEVload + SMP, 16-bit adder seed 0 .tran 5p 1.5n', flush=True)
    eng, ora = _run(desc, 1.5e-9)
    _compare('adder0-dropin', eng, ora, desc['nnodes'])
