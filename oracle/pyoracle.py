"""ORACLE (test infrastructure only): ctypes access to liboracle.so (the C
restatement of the reference's BSIM4 / Newton / transient path) and to
_ref/libklu_ref.so (the reference's own KLU, compiled from its sources).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_klu = None

N_ROLES = ['d', 'g', 's', 'b', 'dp', 'gp', 'gm', 'sp', 'db', 'bp', 'sb', 'q']


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, 'liboracle.so')
        if not os.path.exists(path):
            raise RuntimeError('oracle not built: run make -C oracle')
        L = ctypes.CDLL(path)
        vp, d, i, cp = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_char_p
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.or_model_new.restype = vp
        L.or_model_free.argtypes = [vp]
        L.or_model_param.argtypes = [vp, cp, d]
        L.or_model_param.restype = i
        L.or_model_finish.argtypes = [vp, d, d]
        L.or_model_finish.restype = i
        L.or_model_get.argtypes = [vp, cp]
        L.or_model_get.restype = d
        L.or_inst_new.argtypes = [vp]
        L.or_inst_new.restype = vp
        L.or_inst_free.argtypes = [vp]
        L.or_inst_param.argtypes = [vp, cp, d]
        L.or_inst_param.restype = i
        L.or_inst_defaults.argtypes = [vp]
        L.or_inst_defaults.restype = i
        L.or_inst_finish.argtypes = [vp, ip, i, d]
        L.or_inst_finish.restype = i
        L.or_qa_dc.argtypes = [vp, d, d, d, d, d, d, dp, dp]
        L.or_qa_dc.restype = i
        L.or_inst_load.argtypes = [vp, i, d, d, d, d, d, dp, i, i, d, dp, dp, dp, dp, dp, dp, dp, dp]
        L.or_inst_load.restype = i
        L.or_inst_get.argtypes = [vp, cp]
        L.or_inst_get.restype = d
        L.or_inst_node.argtypes = [vp, i]
        L.or_inst_node.restype = i
        L.or_counts.argtypes = [ip, ip, ip]
        L.or_slot_rc.argtypes = [i, ip, ip]
        L.or_slot_rc.restype = i
        _lib = L
    return _lib


def klu_ref():
    """The reference's KLU, or None when oracle/_ref was not built."""
    global _klu
    if _klu is None:
        path = os.path.join(HERE, '_ref', 'libklu_ref.so')
        if not os.path.exists(path):
            return None
        _klu = ctypes.CDLL(path)
    return _klu


def counts():
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib().or_counts(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    return a.value, b.value, c.value


def slot_roles():
    n, _, _ = counts()
    out = []
    for s in range(n):
        r, c = ctypes.c_int(), ctypes.c_int()
        lib().or_slot_rc(s, ctypes.byref(r), ctypes.byref(c))
        out.append((r.value, c.value))
    return out


class Model:
    def __init__(self, params, nom_temp=300.15, temp=300.15):
        L = lib()
        self.h = L.or_model_new()
        for k, v in params.items():
            if L.or_model_param(self.h, k.encode(), float(v)) != 0:
                raise KeyError('unknown BSIM4 model parameter %s' % k)
        rc = L.or_model_finish(self.h, nom_temp, temp)
        if rc != 0:
            raise ValueError('model setup failed (%d)' % rc)

    def get(self, what):
        return lib().or_model_get(self.h, what.encode())

    def __del__(self):
        try:
            lib().or_model_free(self.h)
        except Exception:
            pass


class Instance:
    def __init__(self, model, params):
        L = lib()
        self.model = model
        self.h = L.or_inst_new(model.h)
        for k, v in params.items():
            if L.or_inst_param(self.h, k.encode(), float(v)) != 0:
                raise KeyError('unknown BSIM4 instance parameter %s' % k)
        self.bits = L.or_inst_defaults(self.h)
        if self.bits < 0:
            raise ValueError('instance setup failed (%d)' % self.bits)

    def finish(self, nodes, state, temp):
        arr = (ctypes.c_int * 12)(*nodes)
        rc = lib().or_inst_finish(self.h, arr, state, temp)
        if rc != 0:
            raise ValueError('instance temperature setup failed (%d)' % rc)

    def qa_dc(self, temp, vd, vg, vs, vb, gmin=1e-12, want_nodes=False):
        cur = (ctypes.c_double * 4)()
        vout = (ctypes.c_double * 16)()
        rc = lib().or_qa_dc(self.h, temp, vd, vg, vs, vb, gmin, cur, vout)
        if rc < 0:
            raise RuntimeError('QA DC solve failed (%d)' % rc)
        if want_nodes:
            return list(cur), list(vout)
        return list(cur)

    def node(self, role):
        return lib().or_inst_node(self.h, role)

    def load(self, mode, temp, rhs_old, states, ag=(0.0,) * 7, order=1, method=1,
             delta=0.0, delta_old=(0.0,) * 7, gmin=1e-12, reltol=1e-3, abstol=1e-12,
             vntol=1e-6):
        """One load; states is a list of 4 ctypes double arrays (state0..3)."""
        nslot, nrole, _ = counts()
        D = ctypes.c_double
        st = (D * nslot)()
        rh = (D * nrole)()
        agv = (D * 7)(*ag)
        dov = (D * 7)(*delta_old)
        nc = lib().or_inst_load(self.h, mode, temp, gmin, reltol, abstol, vntol, agv, order,
                                method, delta, dov, rhs_old, states[0], states[1], states[2],
                                states[3], st, rh)
        return list(st), list(rh), nc

    def get(self, what):
        return lib().or_inst_get(self.h, what.encode())

    def __del__(self):
        try:
            lib().or_inst_free(self.h)
        except Exception:
            pass


def _tran_protos():
    L = lib()
    if getattr(L, '_tran_ready', False):
        return L
    vp, d, i = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
    dp = ctypes.POINTER(ctypes.c_double)
    L.or_tran_new.argtypes = [i, i, i, i, i]
    L.or_tran_new.restype = vp
    L.or_tran_set_options.argtypes = [vp, d, d, d, d, d, d, i, i, i]
    L.or_tran_add_mos.argtypes = [vp, vp, i, i, i, i]
    L.or_tran_add_cap.argtypes = [vp, i, i, d]
    L.or_tran_add_res.argtypes = [vp, i, i, d]
    L.or_tran_add_vsrc.argtypes = [vp, i, i, d, dp, i]
    L.or_tran_set_ic.argtypes = [vp, i, d]
    L.or_tran_setup.argtypes = [vp]
    L.or_tran_run.argtypes = [vp, d, d, d]
    L.or_tran_npts.argtypes = [vp]
    L.or_tran_npts.restype = ctypes.c_int64
    L.or_tran_n.argtypes = [vp]
    L.or_tran_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                ctypes.POINTER(ctypes.c_int64)]
    L.or_tran_output.argtypes = [vp, dp, dp]
    L.or_tran_free.argtypes = [vp]
    L.or_tran_use_klu.argtypes = [vp, ctypes.c_char_p]
    L.or_comcof.argtypes = [i, i, d, dp, dp]
    L.or_tran_csc.argtypes = [vp, vp, vp]
    L.or_tran_csc.restype = ctypes.c_int64
    L.or_tran_mos_info.argtypes = [vp, i, ctypes.POINTER(i), vp]
    L.or_tran_use_dropin.argtypes = [vp, vp, vp, vp, vp, vp]
    L.or_set_sparse_above.argtypes = [i]
    L.or_tran_set_op_options.argtypes = [vp, i, i, i, d, d, i]
    ip = ctypes.POINTER(ctypes.c_int)
    L.or_tran_klu_order.argtypes = [vp, ip, ip, ip]
    L._tran_ready = True
    return L


def comcof(method, order, delta, delta_old):
    """NIcomCof restatement: ag[0..6] for method (1 trapezoidal, 2 Gear)."""
    import numpy as np
    L = _tran_protos()
    d = np.zeros(7)
    d[:len(delta_old)] = delta_old
    ag = np.zeros(7)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = L.or_comcof(method, order, delta, d.ctypes.data_as(dp), ag.ctypes.data_as(dp))
    if rc != 0:
        raise ValueError('order %d not available for method %d' % (order, method))
    return ag


def set_sparse_above(n):
    """circuits set up afterwards with more than n unknowns keep their matrix
    sparse (solved by the reference KLU); returns the previous threshold"""
    return _tran_protos().or_set_sparse_above(int(n))


class Transient:
    """DCtran restatement on one circuit (see tran_ref.c).

    circuit: dict with nnodes, models {key: params}, mos [(key, d, g, s, b,
    {inst params})], caps [(p, n, C)], res [(p, n, R)], vsrc [(p, n, dc,
    pulse tuple or None)], ic {node: v}, temp (K), options {reltol, abstol,
    vntol, chgtol, trtol, gmin, method, maxord, uic, noopiter, gminsteps,
    srcsteps, gminfactor, gshunt, itl2}."""

    def __init__(self, circuit):
        import numpy as np
        L = _tran_protos()
        c = circuit
        self.np = np
        temp = c.get('temp', 300.15)
        self.models = {k: Model(v, nom_temp=300.15, temp=temp) for k, v in c['models'].items()}
        self.h = L.or_tran_new(c['nnodes'], len(c['mos']), len(c['caps']), len(c.get('res', [])),
                               len(c['vsrc']))
        o = dict(reltol=1e-3, abstol=1e-12, vntol=1e-6, chgtol=1e-14, trtol=7.0, gmin=1e-12,
                 method=1, maxord=2, uic=0)
        o.update(c.get('options', {}))
        L.or_tran_set_options(self.h, o['reltol'], o['abstol'], o['vntol'], o['chgtol'],
                              o['trtol'], o['gmin'], o['method'], o['maxord'], o['uic'])
        op = dict(noopiter=0, gminsteps=1, srcsteps=1, gminfactor=10.0, gshunt=0.0, itl2=50)
        op.update({k: v for k, v in o.items() if k in op})
        if L.or_tran_set_op_options(self.h, op['noopiter'], op['gminsteps'], op['srcsteps'],
                                    op['gminfactor'], op['gshunt'], op['itl2']) != 0:
            raise ValueError('operating-point stepping options not restated')
        L.or_tran_set_pivrel.argtypes = [ctypes.c_void_p, ctypes.c_double]
        if L.or_tran_set_pivrel(self.h, float(o.get('pivrel', 1e-3))) != 0:
            raise ValueError('pivrel out of range')
        # optran (step, final time); None or step 0 deselects it
        opt = o.get('optran', (1e-7, 1e-5))
        L.or_tran_set_optran.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        if L.or_tran_set_optran(self.h, float(opt[0]) if opt else 0.0, float(opt[1]) if opt else 1.0) != 0:
            raise ValueError('optran parameters out of range')
        self.inst = []
        for key, dn, gn, sn, bn, ip in c['mos']:
            inst = Instance(self.models[key], ip)
            if L.or_tran_add_mos(self.h, inst.h, dn, gn, sn, bn) != 0:
                raise ValueError('MOSFET needs internal nodes')
            self.inst.append(inst)
        for p, n, cv in c['caps']:
            L.or_tran_add_cap(self.h, p, n, cv)
        for p, n, r in c.get('res', []):
            L.or_tran_add_res(self.h, p, n, r)
        for p, n, dc, pulse in c['vsrc']:
            if pulse is None:
                L.or_tran_add_vsrc(self.h, p, n, dc, None, 0)
            else:
                arr = (ctypes.c_double * len(pulse))(*pulse)
                L.or_tran_add_vsrc(self.h, p, n, dc, arr, len(pulse))
        for node, v in c.get('ic', {}).items():
            L.or_tran_set_ic(self.h, node, v)
        if L.or_tran_setup(self.h) != 0:
            raise RuntimeError('transient setup failed')
        self.solver = 'dense'

    def use_reference_klu(self):
        """Solve with the reference's KLU (oracle/_ref) when it was built;
        returns True when switched."""
        path = os.path.join(HERE, '_ref', 'libklu_ref.so')
        if not os.path.exists(path):
            return False
        if _tran_protos().or_tran_use_klu(self.h, path.encode()) != 0:
            return False
        self.solver = 'klu'
        return True

    def csc(self):
        """the circuit matrix's structural CSC pattern (Ap int64, Ai int32)"""
        L = _tran_protos()
        np = self.np
        nnz = L.or_tran_csc(self.h, None, None)
        Ap = np.zeros(L.or_tran_n(self.h) + 1, dtype=np.int64)
        Ai = np.zeros(max(nnz, 1), dtype=np.int32)
        L.or_tran_csc(self.h, Ap.ctypes.data, Ai.ctypes.data)
        return Ap, Ai[:nnz]

    def mos_info(self, k):
        """(state offset, [d, g, s, b]) of MOSFET k"""
        st = ctypes.c_int()
        nodes = (ctypes.c_int * 4)()
        if _tran_protos().or_tran_mos_info(self.h, k, ctypes.byref(st), nodes) != 0:
            raise IndexError(k)
        return st.value, list(nodes)

    @property
    def n(self):
        return _tran_protos().or_tran_n(self.h)

    def use_dropin(self, run_fn, dl_ctx, factor_fn, solve_fn, smp_ctx):
        """Replace the MOSFET loads and the linear solve by caller functions
        (addresses of C functions with the signatures or_tran_use_dropin
        documents) behind the same DCtran / NIiter."""
        if _tran_protos().or_tran_use_dropin(self.h, run_fn, dl_ctx, factor_fn, solve_fn, smp_ctx) != 0:
            raise RuntimeError('drop-in setup failed')
        self.solver = 'drop-in'

    def run(self, tstep, tstop, tmax=0.0):
        """Returns (rc, t[npts], x[npts, n+1]) with x[:, 0] = ground."""
        L = _tran_protos()
        np = self.np
        rc = L.or_tran_run(self.h, tstep, tstop, tmax)
        npts = L.or_tran_npts(self.h)
        n = L.or_tran_n(self.h)
        t = np.zeros(max(npts, 1))
        x = np.zeros((max(npts, 1), n + 1))
        if npts:
            L.or_tran_output(self.h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                             x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return rc, t[:npts], x[:npts]

    def klu_order(self):
        """The reference KLU's pivot sequence from its last klu_factor as
        1-based unknowns: (rows = Pnum + 1, cols = Q + 1, counts {lnz, unz,
        nzoff, nblocks}); None before a factor or without the reference KLU."""
        L = _tran_protos()
        np = self.np
        n = L.or_tran_n(self.h)
        rows = np.zeros(n, dtype=np.int32)
        cols = np.zeros(n, dtype=np.int32)
        cnt = np.zeros(8, dtype=np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        if L.or_tran_klu_order(self.h, rows.ctypes.data_as(ip), cols.ctypes.data_as(ip),
                               cnt.ctypes.data_as(ip)) != 0:
            return None
        return rows, cols, dict(lnz=int(cnt[0]), unz=int(cnt[1]), nzoff=int(cnt[2]),
                                nblocks=int(cnt[3]), factors=int(cnt[4]), offdiag_factors=int(cnt[5]),
                                refactors=int(cnt[6]), refactor_fails=int(cnt[7]))

    def stats(self):
        L = _tran_protos()
        a, r, it = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        L.or_tran_stats(self.h, ctypes.byref(a), ctypes.byref(r), ctypes.byref(it))
        return dict(accepted=a.value, rejected=r.value, iterations=it.value)

    def __del__(self):
        try:
            _tran_protos().or_tran_free(self.h)
        except Exception:
            pass


def klu_csc(n, Ap, Ai, Ax, b, pivrel=1e-3, refactors=0):
    """A x = b by the reference KLU: klu_analyze + klu_factor + klu_solve
    (through klu_glue.c), then `refactors` timed klu_refactor calls on the
    same values.  Returns (x, factor seconds, seconds per refactor), or
    None when oracle/_ref is not built."""
    import time
    import numpy as np
    K = klu_ref()
    if K is None:
        return None
    vp = ctypes.c_void_p
    K.kg_new.restype = vp
    K.kg_new.argtypes = [ctypes.c_int, vp, vp]
    K.kg_factor.argtypes = [vp, vp, ctypes.c_int, ctypes.c_double]
    K.kg_solve.argtypes = [vp, vp]
    K.kg_free.argtypes = [vp]
    Ap32 = np.ascontiguousarray(Ap, dtype=np.int32)
    Ai32 = np.ascontiguousarray(Ai, dtype=np.int32)
    Axc = np.ascontiguousarray(Ax, dtype=np.float64)
    h = K.kg_new(int(n), Ap32.ctypes.data, Ai32.ctypes.data)
    if not h:
        raise RuntimeError('klu_analyze failed')
    try:
        t = time.perf_counter()
        rc = K.kg_factor(h, Axc.ctypes.data, 1, pivrel)
        tf = time.perf_counter() - t
        if rc != 0:
            raise RuntimeError('klu_factor returned %d' % rc)
        tr = 0.0
        if refactors:
            t = time.perf_counter()
            for _ in range(refactors):
                if K.kg_factor(h, Axc.ctypes.data, 0, pivrel) != 0:
                    raise RuntimeError('klu_refactor failed')
            tr = (time.perf_counter() - t) / refactors
        x = np.array(b, dtype=np.float64, copy=True)
        if K.kg_solve(h, x.ctypes.data) != 0:
            raise RuntimeError('klu_solve failed')
        return x, tf, tr
    finally:
        K.kg_free(h)


def klu_pivot(A, P, Q, tol=1e-3, b=None):
    """The oracle's restatement of klu_factor's pivot choice on a dense
    matrix A (numpy, n x n) with the symbolic P, Q (0-based): returns
    (rc, Pnum, x) -- rc 0 or 102 (singular), x the solution of A x = b."""
    import numpy as np
    L = lib()
    vp = ctypes.c_void_p
    L.or_klu_pivot.argtypes = [ctypes.c_int, vp, vp, vp, ctypes.c_double, vp, vp, vp]
    n = A.shape[0]
    A = np.ascontiguousarray(A, dtype=np.float64)
    P32 = np.ascontiguousarray(P, dtype=np.int32)
    Q32 = np.ascontiguousarray(Q, dtype=np.int32)
    bb = np.ascontiguousarray(np.zeros(n) if b is None else b, dtype=np.float64)
    Pnum = np.zeros(max(n, 1), np.int32)
    x = np.zeros(max(n, 1))
    rc = L.or_klu_pivot(n, A.ctypes.data, P32.ctypes.data, Q32.ctypes.data, tol, bb.ctypes.data,
                        Pnum.ctypes.data, x.ctypes.data)
    return rc, Pnum[:n], x[:n]


def klu_factor_ref(n, Ap, Ai, Ax, tol=1e-3, b=None):
    """The reference's klu_analyze + klu_factor (+ klu_solve of b) on a CSC
    matrix: returns (rc, Pnum, Q, x) with rc 0 or 102, 0-based."""
    import numpy as np
    K = klu_ref()
    vp = ctypes.c_void_p
    K.kg_new.argtypes = [ctypes.c_int, vp, vp]
    K.kg_new.restype = vp
    K.kg_factor.argtypes = [vp, vp, ctypes.c_int, ctypes.c_double]
    K.kg_solve.argtypes = [vp, vp]
    K.kg_order.argtypes = [vp, vp, vp, vp]
    K.kg_free.argtypes = [vp]
    Ap32 = np.ascontiguousarray(Ap, dtype=np.int32)
    Ai32 = np.ascontiguousarray(Ai, dtype=np.int32)
    Ax64 = np.ascontiguousarray(Ax, dtype=np.float64).copy()
    h = K.kg_new(int(n), Ap32.ctypes.data, Ai32.ctypes.data)
    if not h:
        raise RuntimeError('klu_analyze failed')
    try:
        rc = K.kg_factor(h, Ax64.ctypes.data, 1, float(tol))
        rows = np.zeros(max(n, 1), np.int32)
        cols = np.zeros(max(n, 1), np.int32)
        cnt = np.zeros(8, np.int32)
        x = None
        if rc == 0:
            K.kg_order(h, rows.ctypes.data, cols.ctypes.data, cnt.ctypes.data)
            if b is not None:
                x = np.ascontiguousarray(b, dtype=np.float64).copy()
                K.kg_solve(h, x.ctypes.data)
        return rc, rows[:n], cols[:n], x
    finally:
        K.kg_free(h)


def klu_symbolic(n, Ap, Ai):
    """The reference klu_analyze (BTF + AMD, KLU's defaults) on a CSC
    pattern: returns (P, Q, R, nblocks, nzoff, structural_rank), 0-based."""
    import numpy as np
    K = klu_ref()
    vp = ctypes.c_void_p
    K.kg_symbolic.argtypes = [ctypes.c_int, vp, vp, vp, vp, vp, vp]
    Ap32 = np.ascontiguousarray(Ap, dtype=np.int32)
    Ai32 = np.ascontiguousarray(Ai, dtype=np.int32)
    P = np.zeros(max(n, 1), np.int32)
    Q = np.zeros(max(n, 1), np.int32)
    R = np.zeros(n + 1, np.int32)
    info = np.zeros(3, np.int32)
    if K.kg_symbolic(int(n), Ap32.ctypes.data, Ai32.ctypes.data, P.ctypes.data, Q.ctypes.data,
                     R.ctypes.data, info.ctypes.data) != 0:
        raise RuntimeError('klu_analyze failed')
    nb = int(info[0])
    return P[:n], Q[:n], R[:nb + 1], nb, int(info[1]), int(info[2])
