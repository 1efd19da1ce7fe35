"""HIP-graph replay of the batched transient's pass chunks, and the
autotuner.

The graph path (knob "graph" = 1, the default) must reproduce the eager
launches bit for bit: same timepoints, same waveforms, same Newton counts,
for a single-topology batch on two streams, an uneven batch of topology
groups, and a batch with the learned guess (whose per-pass row counter is
reset by a kernel inside the graph).  Each run replays the captured graph
many times with host reads and eager work between replays."""
import os
import tempfile

import numpy as np
import pytest

from engine_path import ngmi, workloads

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu


def _run(descs, tstop, graph, net=None):
    ngmi.tune('graph', graph)
    try:
        b = ngmi.Batch([ngmi.Circuit(d) for d in descs], max_points=200000)
        if net is not None:
            b.set_guess(net)
        rc, passes = b.tran(descs[0]['tstep'], tstop)
        assert rc == 0, ngmi.last_error()
        out = [(b.stats(c), *b.waveform(c)) for c in range(len(descs))]
        return out, passes
    finally:
        ngmi.tune('graph', 1)


def _same(a, b, tag):
    (ra, pa), (rb, pb) = a, b
    assert pa == pb, (tag, pa, pb)
    for c, ((sa, ta, xa), (sb, tb, xb)) in enumerate(zip(ra, rb)):
        assert sa == sb, (tag, c, sa, sb)
        assert np.array_equal(ta, tb) and np.array_equal(xa, xb), (tag, c)
    print('%s: %d circuits, %d passes, bit-identical' % (tag, len(ra), pa), flush=True)


def test_graph_matches_eager_single_topology():
    print('case: graph vs eager, adders seeds 0..7 .tran 5p 2n (two streams)', flush=True)
    descs = [workloads.mc_adder(s) for s in range(8)]
    _same(_run(descs, 2e-9, 0), _run(descs, 2e-9, 1), 'adders')


def test_stream_parts_are_bit_identical():
    """knob "streams" 1..4: the single-topology batch in that many parts on
    as many streams, under graphs and eagerly; the circuits are independent,
    so every split gives the one-stream results bit for bit"""
    print('case: streams 1..4, adders seeds 0..6 .tran 5p 1n, graph and eager', flush=True)
    descs = [workloads.mc_adder(s) for s in range(7)]
    keep = ngmi.get_tune('streams')
    try:
        ngmi.tune('streams', 1)
        ref = _run(descs, 1e-9, 1)
        for n in (2, 3, 4):
            ngmi.tune('streams', n)
            _same(ref, _run(descs, 1e-9, 1), 'streams %d, graph' % n)
            _same(ref, _run(descs, 1e-9, 0), 'streams %d, eager' % n)
    finally:
        ngmi.tune('streams', keep)


def test_graph_matches_eager_uneven_batch():
    print('case: graph vs eager, rings of 11/21/13/31 stages .tran 10p 2n (topology groups)', flush=True)
    descs = [workloads.ring_oscillator(stages=k, tstop=2e-9) for k in (11, 21, 13, 31, 11)]
    _same(_run(descs, 2e-9, 0), _run(descs, 2e-9, 1), 'rings')


def test_graph_matches_eager_with_guess():
    print('case: graph vs eager with the guess network in the loop, adders 0..3 .tran 5p 1.5n', flush=True)
    descs = [workloads.mc_adder(s) for s in range(4)]
    net = ngmi.GuessNet(seed=3)
    _same(_run(descs, 1.5e-9, 0, net), _run(descs, 1.5e-9, 1, net), 'adders+guess')


def test_autotune_and_cache():
    print('case: autotuner on 256 adders, 96 trial passes; cache round trip', flush=True)
    descs = [workloads.mc_adder(s) for s in range(256)]
    circ = [ngmi.Circuit(d) for d in descs]
    before = {k: ngmi.get_tune(k) for k in ('order_slack', 'mos_waves', 'streams', 'graph')}
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, 'tune.txt')
        try:
            rep = ngmi.autotune(circ, workloads.ADDER_TSTEP, 20e-9, trial_passes=96, cache_path=path)
            print(rep, flush=True)
            chosen = {k: ngmi.get_tune(k) for k in before}
            assert 'chosen' in rep and os.path.exists(path)
            for k, v in before.items():
                ngmi.tune(k, v)
            assert ngmi.autotune_load(circ, path)
            assert {k: ngmi.get_tune(k) for k in before} == chosen
            assert not ngmi.autotune_load(circ[:3], path)
        finally:
            for k, v in before.items():
                ngmi.tune(k, v)
