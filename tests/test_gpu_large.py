"""One large circuit on the whole GPU: the power-grid mesh's transient
(configuration 3's mesh: an RC mesh of 0.1 ohm segments and 10 fF per node,
1 V pads every 64 nodes, CMOS inverter loads driven by phased clocks)
through ngmi_large_*, against the oracle's DCtran restatement solving with
the reference's KLU on the same circuit.

Bar: the same accepted timepoints (1e-9 relative) and Newton iteration
count; node voltages within 1e-6 of full scale (1 V) and within 1e-6 of
their own value above 10 mV; source branch currents within 1e-6 of their
full scale, floored at 1 uA (a source that never carries more is held to
abstol = 1e-12 A, the current tolerance NIconvTest applies).  A pad's
current is the sum of G (V_pad - V_k) over its 0.1 ohm segments, a small
difference of nearly equal voltages, so its relative accuracy is the node
voltages' absolute agreement divided by the segments' IR drop, not a fixed
fraction: each pad current is held instead to the bound that agreement
implies, |dI| <= 4 G_pad max|dV| (G_pad the pad node's segment
conductance, the 4 a margin of 2 over both ends of every segment), and
its relative error is printed.  (The oracle run with the reference KLU and
the same oracle run with a dense LU differ by up to 5e-7 relative in these
currents on a 24 x 24 mesh, while their node voltages agree to 4e-14 V.)  At the full 1024 x 1024 size, where
the oracle would take hours, the final point must satisfy KCL: the circuit
equations assembled at the accepted solution balance to within the Newton
tolerance (scaled residual below 1e-4, the reltol-level change NIconvTest
allows over the row's largest current term)."""
import os
import sys
import time

import numpy as np
import pytest

from engine_path import ngmi, workloads

sys.path.insert(0, os.path.join(os.path.dirname(__file__), '..', 'oracle'))
import pyoracle as po  # noqa: E402

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu

WAVE_TOL = 1e-6


def _pad_conductance(d):
    """per voltage source, the summed segment conductance at its + node"""
    g = {}
    for a, b, r in d['res']:
        g[a] = g.get(a, 0.0) + 1.0 / r
        g[b] = g.get(b, 0.0) + 1.0 / r
    return np.array([g.get(v[0], 0.0) for v in d['vsrc']])


def _compare(tag, te, xe, ora, nn, gpad=None):
    to, xo, so = ora
    print('%s: engine %d pts, oracle %d pts (%d rejected, %d iters)'
          % (tag, len(te), len(to), so['rejected'], so['iterations']), flush=True)
    assert len(te) == len(to)
    assert np.allclose(te, to, rtol=1e-9, atol=1e-21)
    xo = xo[:, 1:]
    scale = np.ones(xo.shape[1])
    scale[nn:] = np.maximum(np.abs(xo[:, nn:]).max(axis=0), 1e-6)
    d = np.abs(xe - xo)
    sd = d / scale
    # relative bar where a value is at least 1e-2 of its full scale and, for
    # a branch current, at least 1 uA (below that the scaled bar's 1e-12 A
    # floor, abstol, is the meaningful one)
    big = (np.abs(xo) >= 1e-2 * scale) & (np.abs(xo) >= np.where(np.arange(xo.shape[1]) < nn, 0.0, 1e-6))
    rel = np.where(big, d / np.where(big, np.abs(xo), 1.0), 0.0)
    err_n, rel_n = sd[:, :nn].max(), rel[:, :nn].max()
    err_b = sd[:, nn:].max() if xo.shape[1] > nn else 0.0
    rel_b = rel[:, nn:].max() if xo.shape[1] > nn else 0.0
    print('%s: nodes: max |dv| %.3e V, max relative %.3e; branches: max scaled %.3e, max relative %.3e'
          % (tag, err_n, rel_n, err_b, rel_b), flush=True)
    assert err_n <= WAVE_TOL and rel_n <= WAVE_TOL
    assert err_b <= WAVE_TOL
    if gpad is not None and xo.shape[1] > nn:
        dv = d[:, :nn].max()
        db = d[:, nn:].max(axis=0)
        pads = gpad > 0
        bound = 4 * gpad * dv + 1e-12
        print('%s: pad currents: max |dI| %.3e A against the node-agreement bound %.3e A (max |dV| %.3e V)'
              % (tag, db[pads].max(), bound[pads].min(), dv), flush=True)
        assert (db[pads] <= bound[pads]).all()


@pytest.mark.parametrize('side,pad_every,tstop', [(48, 16, 1e-9), (128, 64, 0.05e-9)])
def test_mesh_against_oracle(side, pad_every, tstop):
    d = workloads.power_grid(side=side, pad_every=pad_every, tstop=tstop)
    loads = len(d['mos']) // 2
    print('case: power-grid mesh %dx%d, %d inverter loads, .tran 10p %gn' % (side, side, loads, tstop * 1e9),
          flush=True)
    t0 = time.time()
    lg = ngmi.Large(ngmi.Circuit(d), max_points=100000)
    info = lg.info()
    print('setup %.1fs: %s' % (time.time() - t0, info), flush=True)
    rc = lg.tran(d['tstep'], tstop)
    assert rc == 0, (rc, ngmi.last_error())
    st = lg.stats()
    te, xe = lg.waveform()
    print('engine: %s, times %s' % (st, lg.times()), flush=True)
    T = po.Transient(d)
    assert T.use_reference_klu()
    rc, to, xo = T.run(d['tstep'], tstop)
    assert rc == 0, rc
    _compare('mesh%d' % side, te, xe, (to, xo, T.stats()), d['nnodes'], _pad_conductance(d))
    assert st['iterations'] == T.stats()['iterations']
    # the static pivot order's largest multiplier: KLU (tol 1e-3) would
    # have kept every one of these pivots
    assert 0 < st['max_multiplier'] < 1e3, st
    r, per = lg.kcl()
    print('KCL at the final point: scaled %.2e, max |F| nodes %.2e A, branches %.2e V' % (r, per[0], per[1]),
          flush=True)
    assert r < 1e-4


def test_full_mesh_kcl():
    """the full configuration-3 mesh: 1024 x 1024, 16384 inverter loads,
    1.1M unknowns; the first 3 ps of its .tran 10p 10n (TMAX 10 ps; the
    16384 clock phases put a PULSE corner every ~15 fs)"""
    tstop = 3e-12
    print('case: power-grid mesh 1024x1024, 16384 loads, first 3 ps of .tran 10p 10n, KCL at the end', flush=True)
    t0 = time.time()
    d = workloads.power_grid(side=1024, tstop=tstop)
    probe = [1, 1 + 512 * 1024 + 512, d['nnodes']]
    lg = ngmi.Large(ngmi.Circuit(d), probe=probe, max_points=10000)
    print('setup %.1fs: %s' % (time.time() - t0, lg.info()), flush=True)
    del d
    t1 = time.time()
    rc = lg.tran(10e-12, tstop, 10e-12)
    assert rc == 0, (rc, ngmi.last_error())
    st = lg.stats()
    t, x = lg.waveform()
    print('run %.1fs: %s, times %s' % (time.time() - t1, st, lg.times()), flush=True)
    assert st['accepted'] >= 10 and abs(t[-1] - tstop) < 1e-18
    # pads hold the mesh near 1 V; every recorded node within the rails
    assert np.all(x[:, :2] > 0.9) and np.all(x[:, :2] <= 1.0 + 1e-6)
    assert np.all(x[:, 2] > -0.1) and np.all(x[:, 2] < 1.1)
    r, per = lg.kcl()
    print('KCL at the final point: scaled %.2e, max |F| nodes %.2e A, branches %.2e V' % (r, per[0], per[1]),
          flush=True)
    assert r < 1e-4
