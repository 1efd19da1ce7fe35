"""Synthetic workloads with seeded, counter-based random streams (numpy's
Philox), so the engine and the oracle see identical values.

device_eval_workload: 2**20 isolated BSIM4 instances of the 45 nm base card
(NMOS vth0 = 0.40 V, PMOS vth0 = -0.40 V, toxe = 1.8 nm, other parameters at
their BSIM4 4.8.3 defaults), Vgs, Vds ~ U[0, 1] V, Vbs ~ U[-0.3, 0] V,
W log-uniform on [100 nm, 10 um], L log-uniform on [45 nm, 1 um], a vth0
offset (instance parameter delvto) ~ N(0, 30 mV); the first half NMOS, the
second half PMOS.  Each instance's 16 Jacobian entries land in a random
2**21-row sparse matrix: its four terminals map to four distinct random rows.
"""
import numpy as np

BASE_NMOS = {'nmos': 1, 'vth0': 0.40, 'toxe': 1.8e-9}
BASE_PMOS = {'pmos': 1, 'vth0': -0.40, 'toxe': 1.8e-9}

EVAL_N = 1 << 20
EVAL_ROWS = 1 << 21
H_STEP = 1e-11          # 10 ps backward-Euler step used by the evaluation workload


def rng(config, seed):
    return np.random.Generator(np.random.Philox(key=[config, seed]))


def device_eval_workload(n=EVAL_N, seed=1, rows=EVAL_ROWS, with_matrix=True):
    """Instance parameters, terminal voltages and the element map.

    Returns a dict with
      model_idx[n] (0 NMOS, 1 PMOS), params {l, w, delvto},
      vterm[4, n] (node voltages d g s b in circuit polarity, source at 0),
      vds, vgs, vbs (device polarity), and, with_matrix, elements
      rows_e[16 n], cols_e[16 n] where element e = slot * n + i is stamp slot
      `slot` (row-major over d g s b) of instance i, and rhs_rows[4 n].
    """
    g = rng(1, seed)
    vgs = g.uniform(0.0, 1.0, n)
    vds = g.uniform(0.0, 1.0, n)
    vbs = g.uniform(-0.3, 0.0, n)
    w = np.exp(g.uniform(np.log(100e-9), np.log(10e-6), n))
    l = np.exp(g.uniform(np.log(45e-9), np.log(1e-6), n))
    delvto = g.normal(0.0, 0.030, n)
    model_idx = np.zeros(n, dtype=np.int32)
    model_idx[n // 2:] = 1
    typ = np.where(model_idx == 0, 1.0, -1.0)
    vterm = np.stack([typ * vds, typ * vgs, np.zeros(n), typ * vbs])
    out = dict(n=n, model_idx=model_idx, params={'l': l, 'w': w, 'delvto': delvto},
               vterm=np.ascontiguousarray(vterm), vds=vds, vgs=vgs, vbs=vbs, type=typ)
    if with_matrix:
        # four distinct rows per instance: draw, then resolve collisions by offset
        term = g.integers(0, rows, size=(4, n), dtype=np.int64)
        for a in range(1, 4):
            for _ in range(8):
                clash = np.zeros(n, dtype=bool)
                for b in range(a):
                    clash |= term[a] == term[b]
                if not clash.any():
                    break
                term[a, clash] = (term[a, clash] + 1 + a) % rows
        slot_r = np.repeat(np.arange(4), 4)
        slot_c = np.tile(np.arange(4), 4)
        out['rows_e'] = term[slot_r].reshape(-1).astype(np.int32)
        out['cols_e'] = term[slot_c].reshape(-1).astype(np.int32)
        out['rhs_rows'] = term.reshape(-1).astype(np.int64)
        out['nrows'] = rows
        out['terms'] = term
    return out


def eval_states(wl):
    """State levels 0..2 (29 x n each) for the evaluation workload: level 0
    holds the previous iterate's terminal voltages equal to this one's (so
    Newton limiting leaves them unchanged), charge history at zero."""
    n = wl['n']
    s0 = np.zeros((29, n))
    s0[3] = wl['vds']          # vds
    s0[2] = wl['vgs']          # vgs
    s0[1] = wl['vbs']          # vbs
    s0[0] = wl['vbs'] - wl['vds']   # vbd
    return [s0, np.zeros((29, n)), np.zeros((29, n))]
