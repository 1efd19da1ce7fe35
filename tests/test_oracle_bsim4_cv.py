"""Pin the oracle's BSIM4 charges and capacitances (C-V) on the selector set
the engine builds (capMod 2, no rgate/rbody/NQS), where no reference golden
holds C-V at the 1e-6 bar:

  * charge conservation: the four terminal charges the load stores
    (qg, qd, qs, qb; b4ld.c:4581-4590) sum to zero to rounding;
  * every capacitive Jacobian entry the load stamps (b4ld.c:4184-4567,
    the ag0-proportional part of each matrix entry) equals the derivative
    of the stored terminal charge, dQ_i/dV_j, taken by a central finite
    difference with Richardson extrapolation (h = 1e-4 and 2e-4 V).

A dropped term or a wrong sign in any of the capMod-2 charge derivatives,
the overlap capacitances or the junction capacitances breaks the second
check (the derivatives are computed separately from the charges).

The capacitances are read out of the stamps as (stamp(ag0 = s) -
stamp(ag0 = 0)) / s with s = 1e15: with ag0 = 1 the capacitive part of an
entry (~1e-15) sits at the rounding level of its conductance (~1e-4).
"""
import ctypes

import numpy as np
import pytest

from engine_path import workloads
from qa_helpers import po

D = ctypes.c_double
NODES = [1, 2, 3, 4, 1, 2, 2, 3, 4, 4, 4, 0]   # d g s b, internal nodes collapsed
S_QB, S_QG, S_QD, S_QS = 11, 13, 15, 28
MODETRAN, MODEDCTRANCURVE, MODEINITFLOAT = 0x1, 0x40, 0x100
AG0 = 1e15
TEMP = 300.15

# the selector variants the engine's GPU parity test runs
VARIANTS = [{}, {'igcmod': 1, 'igbmod': 1}, {'igcmod': 2, 'gidlmod': 1}, {'mobmod': 2},
            {'mobmod': 3, 'diomod': 2}, {'mobmod': 6, 'tempmod': 3}, {'cvchargemod': 1},
            {'diomod': 0, 'tempmod': 1, 'xpart': 0.0}, {'lambda': 0.1, 'vtl': 2e5},
            {'xpart': 0.5}, {'xpart': 1.0}]
GEOMS = [{'w': 200e-9, 'l': 45e-9, 'ad': 2e-14, 'as': 2e-14, 'pd': 6e-7, 'ps': 6e-7},
         {'w': 1e-6, 'l': 100e-9, 'delvto': 0.03},
         {'w': 10e-6, 'l': 1e-6, 'ad': 5e-12, 'as': 5e-12, 'pd': 2e-5, 'ps': 2e-5, 'delvto': -0.03}]
# (vd, vg, vs, vb) in NMOS polarity; |vds| kept away from 0 where the
# drain/source roles swap (the load's mode flag), and both orientations used
BIASES = [(vd, vg, vs, vb) for vd, vs in ((0.1, 0.0), (0.5, 0.0), (1.0, 0.0), (0.0, 0.6))
          for vg in (-0.2, 0.2, 0.45, 0.7, 1.0) for vb in (0.0, -0.3)]

_NSLOT, _NROLE, _NSTATE = po.counts()
_ROLES = po.slot_roles()


def charges_and_caps(inst, v, typ):
    """terminal charges Q[d g s b] (circuit polarity) and the capacitance
    matrix C[i][j] = d(charge at i)/d(voltage at j) from the stamps"""
    rv = (D * 16)(0.0, *v)
    st = [(D * _NSTATE)() for _ in range(4)]
    inst.load(MODEDCTRANCURVE | MODEINITFLOAT, TEMP, rv, st)
    s0, _, _ = inst.load(MODETRAN | MODEINITFLOAT, TEMP, rv, st, ag=(0.0,) * 7)
    Q = typ * np.array([st[0][S_QD], st[0][S_QG], st[0][S_QS], st[0][S_QB]])
    s1, _, _ = inst.load(MODETRAN | MODEINITFLOAT, TEMP, rv, st, ag=(AG0,) + (0.0,) * 6)
    C = np.zeros((4, 4))
    for s, (r, c) in enumerate(_ROLES):
        nr, nc = NODES[r], NODES[c]
        if nr and nc:
            C[nr - 1, nc - 1] += (s1[s] - s0[s]) / AG0
    return Q, C


def fd_caps(inst, v, typ, h=2e-4):
    def central(hh):
        out = np.zeros((4, 4))
        for j in range(4):
            vp, vm = list(v), list(v)
            vp[j] += hh
            vm[j] -= hh
            out[:, j] = (charges_and_caps(inst, vp, typ)[0] - charges_and_caps(inst, vm, typ)[0]) / (2 * hh)
        return out
    return (4 * central(h / 2) - central(h)) / 3


@pytest.mark.parametrize('var', range(len(VARIANTS)))
@pytest.mark.parametrize('pol', ['nmos', 'pmos'])
def test_charge_conservation_and_cv_derivatives(var, pol):
    base = workloads.BASE_NMOS if pol == 'nmos' else workloads.BASE_PMOS
    typ = 1 if pol == 'nmos' else -1
    m = po.Model(dict(base, **VARIANTS[var]))
    worst_fd, worst_sum, n = 0.0, 0.0, 0
    for gi, geo in enumerate(GEOMS):
        inst = po.Instance(m, geo)
        assert inst.bits == 0
        inst.finish(NODES, 0, TEMP)
        for b in BIASES:
            v = [typ * x for x in b]
            Q, C = charges_and_caps(inst, v, typ)
            scale = np.abs(C).max()
            assert scale > 0
            qs = np.abs(Q).max()
            worst_sum = max(worst_sum, abs(Q.sum()) / qs if qs > 0 else 0.0)
            err = np.abs(fd_caps(inst, v, typ) - C).max() / scale
            worst_fd = max(worst_fd, err)
            n += 1
            assert err < 1e-8, (pol, VARIANTS[var], geo, v, err)
    print('%s %r: %d points, dQ/dV vs stamps worst %.1e of max|C|, charge sum worst %.1e of max|Q|'
          % (pol, VARIANTS[var], n, worst_fd, worst_sum))
    assert worst_sum < 1e-13


def test_fd_check_catches_a_planted_slip():
    """The check is sensitive at the level it claims: a capacitance matrix
    with one entry off by 1e-6 of max|C| (a slip the QA goldens' 1e-6 rule
    would barely see) fails it."""
    m = po.Model(dict(workloads.BASE_NMOS))
    inst = po.Instance(m, GEOMS[1])
    inst.finish(NODES, 0, TEMP)
    v = [0.5, 0.7, 0.0, -0.3]
    Q, C = charges_and_caps(inst, v, 1)
    R = fd_caps(inst, v, 1)
    assert np.abs(R - C).max() / np.abs(C).max() < 1e-8
    C[0, 2] += 1e-6 * np.abs(C).max()
    assert np.abs(R - C).max() / np.abs(C).max() > 1e-7
