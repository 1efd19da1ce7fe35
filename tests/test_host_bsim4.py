"""The engine's BSIM4 code on the CPU against the oracle, bit for bit.

tests/host/b4_hostcheck.cpp compiles the engine's host setup
(bsim4_setup.cpp: defaults, binning, temperature, geometry, b4check clamps)
and its evaluation template (bsim4_eval.h) with g++.  For QA model cards of
the reference (with the selectors the engine builds) under 18 selector
variants, 3 temperatures and random biases, a sequence of loads in every
Newton mode must give exactly the oracle's Jacobian, RHS, charge and
terminal-voltage states and convergence votes: the restatement and the
engine follow the same operation order with contraction off.
"""
import ctypes
import json
import math
import os
import random
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(REPO, 'oracle'))
import pyoracle as po  # noqa: E402

SRC = os.path.join(HERE, 'host', 'b4_hostcheck.cpp')
CSRC = os.path.join(REPO, 'ngspice-sf-mirror_amd', 'csrc')
SO = os.path.join(HERE, 'host', '_build', 'libb4host.so')
D = ctypes.c_double
NS = 29


def _lib():
    deps = [SRC, os.path.join(CSRC, 'bsim4_setup.cpp'), os.path.join(CSRC, 'bsim4_setup.h'),
            os.path.join(CSRC, 'bsim4_eval.h'), os.path.join(CSRC, 'bsim4_fields.h'),
            os.path.join(CSRC, 'bsim4_tables.h')]
    if not os.path.exists(SO) or os.path.getmtime(SO) < max(os.path.getmtime(p) for p in deps):
        os.makedirs(os.path.dirname(SO), exist_ok=True)
        subprocess.check_call(['g++', '-std=c++17', '-O2', '-ffp-contract=off', '-fPIC', '-shared',
                               '-Wno-unknown-pragmas', '-o', SO, SRC,
                               os.path.join(CSRC, 'bsim4_setup.cpp')])
    L = ctypes.CDLL(SO)
    vp = ctypes.c_void_p
    dp = ctypes.POINTER(D)
    L.hc_model_new.restype = vp
    L.hc_model_free.argtypes = [vp]
    L.hc_model_param.argtypes = [vp, ctypes.c_char_p, D]
    L.hc_model_finish.argtypes = [vp, D, D]
    L.hc_inst_new.argtypes = [vp]
    L.hc_inst_new.restype = vp
    L.hc_inst_free.argtypes = [vp]
    L.hc_inst_param.argtypes = [vp, ctypes.c_char_p, D]
    L.hc_inst_finish.argtypes = [vp]
    L.hc_load.argtypes = [vp, ctypes.c_int, D, D, D, dp, ctypes.c_int, ctypes.c_int, D, dp, dp, dp,
                          dp, dp, dp, dp, dp]
    return L


DATA = json.load(open(os.path.join(HERE, 'golden', 'bsim4_qa.json')))['devices']
VARIANTS = [{}, {'igcmod': 1, 'igbmod': 1}, {'igcmod': 2}, {'gidlmod': 1}, {'mobmod': 1},
            {'mobmod': 2}, {'mobmod': 3}, {'mobmod': 4}, {'mobmod': 5}, {'mobmod': 6},
            {'diomod': 0}, {'diomod': 2}, {'tempmod': 1}, {'tempmod': 2}, {'tempmod': 3},
            {'cvchargemod': 1}, {'xpart': 0.0}, {'xpart': 1.0}]
MODES = [(0x10 | 0x200, None), (0x10 | 0x100, None), (0x40 | 0x100, None),
         (0x1 | 0x1000, 'inittran'), (0x1 | 0x2000, 'pred'), (0x1 | 0x100, 'float'),
         (0x1 | 0x100, 'float'), (0x1 | 0x2000, 'pred'), (0x1 | 0x100, 'float')]
ROLES = None


def _same(a, b):
    """bitwise-equal values; NaN equals NaN (a card forced into a selector it
    was not fitted for, e.g. tempMod 2 with the tempMod-0 `at`, gives a
    negative vsat and NaN currents in both, as it would in the reference)"""
    return a == b or (math.isnan(a) and math.isnan(b))


def _case(L, mp, ip, T, rng):
    om = po.Model(mp, nom_temp=300.15, temp=T)
    hm = L.hc_model_new()
    try:
        for k, v in mp.items():
            L.hc_model_param(hm, k.encode(), float(v))
        assert L.hc_model_finish(hm, 300.15, T) == 0
        oi = po.Instance(om, ip)
        hi = L.hc_inst_new(hm)
        for k, v in ip.items():
            L.hc_inst_param(hi, k.encode(), float(v))
        assert L.hc_inst_finish(hi) == 0
        assert oi.bits == 0
        oi.finish([1, 2, 3, 4, 1, 2, 2, 3, 4, 4, 4, 0], 0, T)
        ost = [(D * NS)() for _ in range(4)]
        hst = [(D * NS)() for _ in range(4)]
        cache = (D * 64)()
        sgn = -1 if mp.get('pmos') else 1
        roles = po.slot_roles()
        h = 1e-11
        for mode, tag in MODES:
            v = [0, sgn * rng.uniform(-0.2, 1.2), sgn * rng.uniform(-0.3, 1.3), 0.0,
                 sgn * rng.uniform(-0.5, 0.2)] + [0] * 11
            rhs = (D * 16)(*v)
            if tag == 'inittran':
                ag, order = (1 / h, -1 / h) + (0,) * 5, 1
            else:
                ag, order = (2 / h, 1.0) + (0,) * 5, 2
            dold = (h, 0.8 * h, 0.7 * h, 0, 0, 0, 0)
            ostamp, orhs, onc = oi.load(mode, T, rhs, ost, ag=ag, order=order, method=1, delta=h,
                                        delta_old=dold)
            J, R = (D * 16)(), (D * 4)()
            hnc = L.hc_load(hi, mode, 1e-12, 1e-3, 1e-12, (D * 7)(*ag), order, 1, h / dold[1],
                            (D * 4)(*v[1:5]), hst[0], hst[1], hst[2], hst[3], cache, J, R)
            M = [[0.0] * 5 for _ in range(5)]
            for s, (r, c) in enumerate(roles):
                nr, nc = oi.node(r), oi.node(c)
                if nr and nc:
                    M[nr][nc] += ostamp[s]
            Rn = [0.0] * 5
            for r in range(12):
                if oi.node(r):
                    Rn[oi.node(r)] += orhs[r]
            for a in range(4):
                for b in range(4):
                    assert _same(M[a + 1][b + 1], J[a * 4 + b]), (mode, a, b, M[a + 1][b + 1],
                                                                  J[a * 4 + b], v)
                assert _same(Rn[a + 1], R[a]), (mode, a)
            for lv in (0, 1):
                for k in (0, 1, 2, 3, 11, 12, 13, 14, 15, 16, 19, 21, 28):
                    assert _same(ost[lv][k], hst[lv][k]), (mode, lv, k)
            assert onc == hnc
        L.hc_inst_free(hi)
    finally:
        L.hc_model_free(hm)


@pytest.mark.parametrize('pol', ['nmos', 'pmos'])
def test_engine_bsim4_bitwise_equals_oracle(pol):
    L = _lib()
    rng = random.Random(5 if pol == 'nmos' else 6)
    n = 0
    for t in DATA[pol]['tests'][:4]:
        for var in VARIANTS:
            mp = dict(t['model'])
            mp[pol] = 1
            mp.update({'rgatemod': 0, 'rbodymod': 0, 'rdsmod': 0, 'rsh': 0})
            mp.update(var)
            for T in (300.15, 223.15, 423.15):
                _case(L, mp, dict(t['instance']), T, rng)
                n += 1
    assert n == 4 * len(VARIANTS) * 3
