"""Helpers that drive the oracle's BSIM4 restatement the way the reference's
CMC QA harness (tests/bin/ngspice.pm, tests/bin/compareSimulationResults.pl)
drives ngspice, and compare with its acceptance rule."""
import ctypes
import math
import os
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, 'oracle'))
import pyoracle as po  # noqa: E402

MODETRAN = 0x1
MODEDCTRANCURVE = 0x40
MODEINITFLOAT = 0x100
# integration coefficient used to read the capacitances out of the stamps
AG0_C = 1e15


def qa_match(ref, sim, clip, ndigit=6, reltol=1e-6):
    """compareSimulationResults.pl:178-236: equal after rounding to ndigit+1
    significant digits within 10 units, or relative error below reltol."""
    if ref == sim:
        return True
    if abs(ref) < clip and abs(sim) < clip:
        return True
    if ref * sim <= 0.0:
        return False
    lo, hi = sorted((abs(ref), abs(sim)))
    mag = int(math.log(lo) / math.log(10)) + 1
    if lo < 1:
        mag -= 1
    lo_i = int(0.5 + lo * 10 ** (ndigit + 1 - mag))
    hi_i = int(0.5 + hi * 10 ** (ndigit + 1 - mag))
    if abs(lo_i - hi_i) <= 10:
        return True
    a = abs(ref - sim)
    return a / (0.5 * (abs(ref) + abs(sim) + a)) < reltol


def model_for(test, pol, temp_c):
    mp = dict(test['model'])
    mp[pol] = 1
    return po.Model(mp, nom_temp=300.15, temp=temp_c + 273.15)


def ac_admittance(model, inst_params, temp_c, bias, omega):
    """4x4 terminal admittance Y (d,g,s,b) at one bias, internal nodes
    eliminated, from the DC operating point and the load's G and C."""
    T = temp_c + 273.15
    inst = po.Instance(model, inst_params)
    _, v = inst.qa_dc(T, bias['d'], bias['g'], bias['s'], bias['b'], want_nodes=True)
    nslot, nrole, nstate = po.counts()
    nn = max(inst.node(r) for r in range(nrole)) + 1
    D = ctypes.c_double
    rhs_old = (D * 16)(*v)
    states = [(D * nstate)() for _ in range(4)]
    inst.load(MODEDCTRANCURVE | MODEINITFLOAT, T, rhs_old, states, reltol=1e-5, abstol=1e-15)
    stamp0, _, _ = inst.load(MODETRAN | MODEINITFLOAT, T, rhs_old, states,
                             ag=(0.0,) * 7, reltol=1e-5, abstol=1e-15)
    # The stamps are G + ag0 C.  ag0 = 1 would leave C (~1e-15 F) at the
    # rounding level of G (~1e-4 S) in every entry holding both, so C is
    # taken with ag0 large enough that ag0 C dominates, and divided back out.
    stamp1, _, _ = inst.load(MODETRAN | MODEINITFLOAT, T, rhs_old, states,
                             ag=(AG0_C,) + (0.0,) * 6, reltol=1e-5, abstol=1e-15)
    G = np.zeros((nn, nn))
    C = np.zeros((nn, nn))
    for s, (r, c) in enumerate(po.slot_roles()):
        nr, nc = inst.node(r), inst.node(c)
        if nr == 0 or nc == 0:
            continue
        G[nr, nc] += stamp0[s]
        C[nr, nc] += (stamp1[s] - stamp0[s]) / AG0_C
    Y = G + 1j * omega * C
    ext = [1, 2, 3, 4]
    intn = list(range(5, nn))
    Yee = Y[np.ix_(ext, ext)]
    if intn:
        Yei = Y[np.ix_(ext, intn)]
        Yie = Y[np.ix_(intn, ext)]
        Yii = Y[np.ix_(intn, intn)]
        Yee = Yee - Yei @ np.linalg.solve(Yii, Yie)
    return Yee


PIN = {'d': 0, 'g': 1, 's': 2, 'b': 3}


def ac_quantity(Y, name, omega):
    """g(m,f) = Re Y[m,f]; c(m,f) = Im Y/omega (self) or -Im Y/omega."""
    kind, rest = name.split('(')
    m, f = rest.rstrip(')').split(',')
    y = Y[PIN[m], PIN[f]]
    if kind.lower() == 'g':
        return y.real
    return y.imag / omega if m == f else -y.imag / omega
