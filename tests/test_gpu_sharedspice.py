"""The sharedspice embedding interface (ngSpice_Init / ngSpice_Command /
ngSpice_Circ / ngGet_Vec_Info, sharedspice.h:384-486) exported by the engine
library, driven through ctypes the way a host program drives the
reference's shared library: a deck sent line by line with circbyline, run
in the foreground and with bg_run on a background thread.

Bar: SendData delivers every accepted point once, in order, with the same
values as the vectors ngGet_Vec_Info serves afterwards (bit for bit); those
vectors equal the engine's batch run of the same deck (same timepoints,
values bit for bit) and the oracle within the waveform bar; BGThreadRunning
reports the thread's start (false) and end (true)."""
import ctypes
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


class VecValues(ctypes.Structure):
    _fields_ = [('name', ctypes.c_char_p), ('creal', ctypes.c_double), ('cimag', ctypes.c_double),
                ('is_scale', ctypes.c_bool), ('is_complex', ctypes.c_bool)]


class VecValuesAll(ctypes.Structure):
    _fields_ = [('veccount', ctypes.c_int), ('vecindex', ctypes.c_int),
                ('vecsa', ctypes.POINTER(ctypes.POINTER(VecValues)))]


class VectorInfo(ctypes.Structure):
    _fields_ = [('v_name', ctypes.c_char_p), ('v_type', ctypes.c_int), ('v_flags', ctypes.c_short),
                ('v_realdata', ctypes.POINTER(ctypes.c_double)), ('v_compdata', ctypes.c_void_p),
                ('v_length', ctypes.c_int)]


SEND_CHAR = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p)
SEND_STAT = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p)
CONTROLLED_EXIT = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_int, ctypes.c_bool, ctypes.c_bool, ctypes.c_int,
                                   ctypes.c_void_p)
SEND_DATA = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.POINTER(VecValuesAll), ctypes.c_int, ctypes.c_int,
                             ctypes.c_void_p)
SEND_INIT = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p)
BG_RUNNING = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_bool, ctypes.c_int, ctypes.c_void_p)


class Host:
    """a host program's side: callbacks recording what the library sends"""

    def __init__(self):
        self.L = ngmi.lib()
        L = self.L
        L.ngSpice_Init.argtypes = [SEND_CHAR, SEND_STAT, CONTROLLED_EXIT, SEND_DATA, SEND_INIT, BG_RUNNING,
                                   ctypes.c_void_p]
        L.ngSpice_Command.argtypes = [ctypes.c_char_p]
        L.ngGet_Vec_Info.argtypes = [ctypes.c_char_p]
        L.ngGet_Vec_Info.restype = ctypes.POINTER(VectorInfo)
        L.ngSpice_running.restype = ctypes.c_bool
        L.ngSpice_CurPlot.restype = ctypes.c_char_p
        self.out, self.points, self.bg, self.exits = [], [], [], []
        self.cbs = [SEND_CHAR(self._char), SEND_STAT(lambda s, i, u: 0), CONTROLLED_EXIT(self._exit),
                    SEND_DATA(self._data), SEND_INIT(lambda p, i, u: 0), BG_RUNNING(self._bg)]
        assert L.ngSpice_Init(*self.cbs, None) == 0

    def _char(self, s, ident, user):
        self.out.append(s.decode())
        return 0

    def _exit(self, status, immediate, quit_, ident, user):
        self.exits.append((status, quit_))
        return 0

    def _data(self, pall, count, ident, user):
        a = pall.contents
        self.points.append((a.vecindex, [a.vecsa[k].contents.creal for k in range(a.veccount)],
                            [a.vecsa[k].contents.name.decode() for k in range(a.veccount)]))
        return 0

    def _bg(self, exited, ident, user):
        self.bg.append(bool(exited))
        return 0

    def cmd(self, text):
        return self.L.ngSpice_Command(text.encode())

    def vec(self, name):
        p = self.L.ngGet_Vec_Info(name.encode())
        if not p:
            return None
        v = p.contents
        return np.ctypeslib.as_array(v.v_realdata, shape=(v.v_length,)).copy()


def _send_deck(h, text):
    for line in text.splitlines():
        assert h.cmd('circbyline ' + line) == 0, h.out[-3:]


def test_foreground_run_matches_the_batch_and_the_oracle():
    d = workloads.ring_oscillator(stages=11, tstop=2e-9)
    deck = workloads.deck(d, 'ring11 through sharedspice')
    print('case: sharedspice circbyline + run, 11-stage ring .tran 10p 2n', flush=True)
    h = Host()
    _send_deck(h, deck)
    assert any('Circuit:' in s for s in h.out)
    assert h.cmd('run') == 0, h.out[-3:]
    assert h.L.ngSpice_CurPlot() == b'tran1'
    t = h.vec('time')
    assert t is not None and len(t) > 10
    # SendData: every point once, in order, the same values as the vectors
    idx = [p[0] for p in h.points]
    assert idx == list(range(len(t)))
    names = h.points[0][2]
    assert names[0] == 'time'
    sent = np.array([p[1] for p in h.points])
    for k, nm in enumerate(names):
        v = h.vec(nm)
        assert v is not None, nm
        assert np.array_equal(v, sent[:, k]), nm
    # the same deck through the engine's batch API
    nl = ngmi.Netlist(deck)
    b = ngmi.Batch([nl.circuit()], max_points=200000)
    rc, _ = b.tran(d['tstep'], d['tstop'])
    assert rc == 0
    tb, xb = b.waveform(0)
    assert np.array_equal(tb, t)
    nlnames = nl.names()[1:]
    for k, nm in enumerate(nlnames):
        assert np.array_equal(h.vec(nm), xb[:, k]), nm
    # and against the oracle: node 1 by the v() form and the supply current by i()
    T = po.Transient(d)
    rc, to, xo = T.run(d['tstep'], d['tstop'])
    assert rc == 0 and len(to) == len(t) and np.allclose(to, t, rtol=1e-9, atol=1e-21)
    err = np.abs(h.vec('v(1)') - xo[:, 1]).max()
    print('%d points; |v(1) - oracle| %.2e V; SendData calls %d' % (len(t), err, len(h.points)), flush=True)
    assert err < 1e-6
    assert h.vec('nonexistent') is None


def test_background_run_and_quit():
    d = workloads.ring_oscillator(stages=11, tstop=1e-9)
    deck = workloads.deck(d, 'ring11 in the background')
    print('case: sharedspice bg_run, 11-stage ring .tran 10p 1n', flush=True)
    h = Host()
    _send_deck(h, deck)
    assert h.cmd('bg_run') == 0
    t0 = time.time()
    while (h.L.ngSpice_running() or not h.bg) and time.time() - t0 < 60:
        time.sleep(0.01)
    while len(h.bg) < 2 and time.time() - t0 < 60:
        time.sleep(0.01)
    assert h.bg == [False, True]
    t = h.vec('time')
    assert t is not None and len(h.points) == len(t)
    # an explicit tran command overrides the deck's analysis
    h.points.clear()
    assert h.cmd('tran 10p 0.5n') == 0
    t2 = h.vec('time')
    assert abs(t2[-1] - 0.5e-9) < 1e-15 and len(h.points) == len(t2)
    assert h.cmd('frobnicate') == 1 and any('frobnicate' in s for s in h.out)
    assert h.cmd('quit') == 0 and h.exits == [(0, True)]
