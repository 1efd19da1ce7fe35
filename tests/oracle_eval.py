"""Per-instance BSIM4 evaluation through the oracle (test infrastructure):
one oracle instance per engine instance, nodes d g s b = 1 2 3 4, node
matrix and RHS summed from the oracle's pointer slots exactly as the
reference's CKTload accumulates them (b4ld.c:5435-5623 stamp targets)."""
import ctypes
import os
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, 'oracle'))
import pyoracle as po  # noqa: E402

D = ctypes.c_double
NODES = [1, 2, 3, 4, 1, 2, 2, 3, 4, 4, 4, 0]   # roles d g s b dp gp gm sp db bp sb q

# values the load keeps per instance for the next call (b4ld.c:444-507), in
# the engine's cache row order; the first 25 rows of its cache
CACHE_FIELDS = ['cd', 'csub', 'cbd', 'cbs', 'gbd', 'gbs', 'gm', 'gds', 'gmbs', 'gbbs', 'gbgs',
                'gbds', 'Igidl', 'Igisl', 'ggidld', 'ggidlg', 'ggidlb', 'ggisls', 'ggislg',
                'ggislb', 'Igs', 'Igd', 'Igcs', 'Igcd', 'Igb']
CURRENT_ROWS = [0, 1, 2, 3, 12, 13, 20, 21, 22, 23, 24]
CONDUCTANCE_ROWS = [4, 5, 6, 7, 8, 9, 10, 11, 14, 15, 16, 17, 18, 19]


class OracleBatch:
    """n oracle instances; states kept as [level][29][n] numpy arrays."""

    def __init__(self, model_params, model_idx, inst_params, temp=300.15, tnom=300.15):
        self.models = [po.Model(mp, nom_temp=tnom, temp=temp) for mp in model_params]
        self.n = len(model_idx)
        self.temp = temp
        self.inst = []
        for k in range(self.n):
            ip = {name: float(vals[k]) for name, vals in inst_params.items()}
            h = po.Instance(self.models[model_idx[k]], ip)
            if h.bits != 0:
                raise ValueError('instance %d needs internal nodes' % k)
            h.finish(NODES, 0, temp)
            self.inst.append(h)
        self.roles = po.slot_roles()
        # slot -> (row node-1, col node-1) for the default topology
        self.slot_rc = [(NODES[r] - 1, NODES[c] - 1) for r, c in self.roles]

    def load(self, mode, vterm, states, ag=(0.0,) * 7, order=1, method=1, delta=0.0,
             delta_old=(0.0,) * 7, gmin=1e-12, reltol=1e-3, abstol=1e-12):
        """states: list of 3 arrays [29, n] (modified in place).  Returns
        stamp[16, n], rhs[4, n], noncon[n]."""
        n = self.n
        stamp = np.zeros((16, n))
        rhs = np.zeros((4, n))
        nc = np.zeros(n, dtype=np.int32)
        self.cache = np.zeros((len(CACHE_FIELDS), n))
        for k in range(n):
            sv = [(D * 29)(*states[lv][:, k]) for lv in range(3)] + [(D * 29)()]
            rv = (D * 16)(0.0, *vterm[:, k])
            st, rh, c = self.inst[k].load(mode, self.temp, rv, sv, ag=ag, order=order,
                                          method=method, delta=delta, delta_old=delta_old,
                                          gmin=gmin, reltol=reltol, abstol=abstol)
            for s, (r, cc) in enumerate(self.slot_rc):
                if r >= 0 and cc >= 0:
                    stamp[r * 4 + cc, k] += st[s]
            for role in range(12):
                node = NODES[role]
                if node:
                    rhs[node - 1, k] += rh[role]
            for lv in range(3):
                states[lv][:, k] = np.frombuffer(sv[lv], dtype=np.float64)
            nc[k] = c
            for f, name in enumerate(CACHE_FIELDS):
                self.cache[f, k] = self.inst[k].get(name)
        return stamp, rhs, nc
