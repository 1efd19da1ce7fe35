"""A deck through the front end runs like the same circuit built from its
description, and the raw file holds the engine's waveform.

Bar: the deck's circuit accepts the same timepoints as the described one
and as the oracle (1e-9 relative) and every unknown, matched by name,
agrees within 1e-6 of full scale (the waveform bar of the transient
tests); the ASCII raw file reproduces each value to its 15 printed digits
and the binary raw file bit for bit.
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


def _run(circuit, tstep, tstop):
    b = ngmi.Batch([circuit], max_points=200000)
    rc, _ = b.tran(tstep, tstop)
    assert rc == 0, (rc, ngmi.last_error())
    t, x = b.waveform(0)
    return b, t, x


def _desc_index(name, nnodes):
    """unknown of a deck name in the description's numbering"""
    if name.endswith('#branch'):
        return nnodes + 1 + int(name[1:-len('#branch')])
    return int(name)


def test_deck_runs_like_the_description_and_the_oracle():
    d = workloads.ring_oscillator(stages=11, tstop=2e-9)
    tstop = d['tstop']
    print('case: 11-stage ring as a deck, .tran 10p 2n', flush=True)
    nl = ngmi.Netlist(workloads.deck(d, 'ring11'))
    b1, t1, x1 = _run(nl.circuit(), d['tstep'], tstop)
    _, t2, x2 = _run(ngmi.Circuit(d), d['tstep'], tstop)
    T = po.Transient(d)
    rc, to, xo = T.run(d['tstep'], tstop)
    assert rc == 0
    names = nl.names()[1:]
    perm = [_desc_index(n, d['nnodes']) for n in names]
    print('deck %d pts, description %d pts, oracle %d pts' % (len(t1), len(t2), len(to)), flush=True)
    assert len(t1) == len(t2) == len(to)
    assert np.allclose(t1, t2, rtol=1e-9, atol=1e-21)
    assert np.allclose(t1, to, rtol=1e-9, atol=1e-21)
    nn = d['nnodes']
    for ref, tag in ((x2, 'description'), (xo[:, 1:], 'oracle')):
        err = 0.0
        for k, u in enumerate(perm):
            scale = 1.0 if u <= nn else max(np.abs(ref[:, u - 1]).max(), 1e-30)
            err = max(err, np.abs(x1[:, k] - ref[:, u - 1]).max() / scale)
        print('deck vs %s: max scaled error %.3e' % (tag, err), flush=True)
        assert err <= 1e-6


def test_raw_file_holds_the_waveform(tmp_path):
    d = workloads.ring_oscillator(stages=11, tstop=0.5e-9)
    print('case: raw files of an 11-stage ring, .tran 10p 0.5n', flush=True)
    nl = ngmi.Netlist(workloads.deck(d, 'ring11 raw'))
    b, t, x = _run(nl.circuit(), d['tstep'], d['tstop'])
    names = nl.names()[1:]
    path = tmp_path / 'ring.raw'
    b.write_raw(0, path, title=nl.title, names=names)
    text = path.read_text().split('\n')
    head = {ln.split(':', 1)[0]: ln.split(':', 1)[1].strip() for ln in text[:7]}
    assert head['Title'] == 'ring11 raw'
    assert head['Plotname'] == 'Transient Analysis'
    assert head['Flags'] == 'real'
    assert int(head['No. Variables']) == len(names) + 1
    assert int(head['No. Points']) == len(t)
    assert text[7] == 'Variables:'
    assert text[8] == '\t0\ttime\ttime'
    var = text[9:9 + len(names)]
    for k, (ln, nm) in enumerate(zip(var, names)):
        kind = 'current' if nm.endswith('#branch') else 'voltage'
        shown = 'i(%s)' % nm[:-7] if kind == 'current' else 'v(%s)' % nm
        assert ln == '\t%d\t%s\t%s' % (k + 1, shown, kind)
    assert text[9 + len(names)] == 'Values:'
    rows = text[10 + len(names):]
    vals = []
    p = 0
    while p < len(rows) and rows[p]:
        idx, first = rows[p].split('\t\t')
        assert int(idx) == len(vals)
        row = [float(first)] + [float(rows[p + 1 + k].strip()) for k in range(len(names))]
        vals.append(row)
        p += 1 + len(names)
    vals = np.array(vals)
    assert vals.shape == (len(t), len(names) + 1)
    assert np.allclose(vals[:, 0], t, rtol=1e-15, atol=0)
    assert np.allclose(vals[:, 1:], x, rtol=1e-15, atol=1e-300)
    bpath = tmp_path / 'ring_bin.raw'
    b.write_raw(0, bpath, title=nl.title, names=names, binary=True)
    blob = bpath.read_bytes()
    at = blob.index(b'Binary:\n') + len(b'Binary:\n')
    data = np.frombuffer(blob[at:], dtype=np.float64).reshape(len(t), len(names) + 1)
    assert np.array_equal(data[:, 0], t) and np.array_equal(data[:, 1:], x)
    print('raw files: %d points x %d variables, ASCII and binary' % data.shape, flush=True)


def test_live_data_stream_matches_the_waveform():
    """ngmi_batch_set_data_callback (the SendData / OUTpData stream): the
    points handed out during the run, concatenated, are the recorded
    waveform bit for bit, in order, for each watched circuit."""
    descs = [workloads.ring_oscillator(stages=11, tstop=2e-9),
             workloads.ring_oscillator(stages=11, cload=1.5e-15, tstop=2e-9)]
    print('case: live data stream, two 11-stage rings, .tran 10p 2n, watching circuit 1', flush=True)
    b = ngmi.Batch([ngmi.Circuit(d) for d in descs], max_points=200000)
    got = []
    b.set_data_callback(lambda c, first, t, x: got.append((c, first, t, x)), circuits=[1])
    rc, passes = b.tran(10e-12, 2e-9)
    assert rc == 0, (rc, ngmi.last_error())
    assert len(got) > 3 and all(g[0] == 1 for g in got)
    firsts = [g[1] for g in got]
    assert firsts[0] == 0
    assert all(firsts[i + 1] == firsts[i] + len(got[i][2]) for i in range(len(got) - 1))
    t = np.concatenate([g[2] for g in got])
    x = np.concatenate([g[3] for g in got])
    tw, xw = b.waveform(1)
    print('%d calls, %d points streamed, %d recorded' % (len(got), len(t), len(tw)), flush=True)
    assert np.array_equal(t, tw) and np.array_equal(x, xw)
    b.set_data_callback(None)
