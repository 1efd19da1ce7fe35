"""Python host mirror of the engine's C ABI (include/ngspice_mi355x.h).

A thin ctypes layer: every call goes to libngspice_mi355x.so, which runs the
HIP kernels.  Device buffers are passed as integer addresses (for example
``torch.Tensor.data_ptr()`` of a tensor on ``cuda:0``).  There is no CPU
fallback: loading fails with an error when the library is missing.

Names mirror the reference's device API: ``Model.param`` is BSIM4mParam
(b4mpar.c:20), ``InstanceSet.add`` is BSIM4param (b4par.c:22) for many
instances, ``InstanceSet.finalize`` is the instance part of BSIM4temp
(b4temp.c:418-2400), ``InstanceSet.eval`` is BSIM4load (b4ld.c:69-5399).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libngspice_mi355x.so')

OK = 0
E_BADPARM = 7
E_UNSUPP = 10
E_SINGULAR = 102
E_ITERLIM = 103
E_TIMESTEP = 106
E_DEVICE = 120

MODETRAN = 0x1
MODEAC = 0x2
MODEDCOP = 0x10
MODETRANOP = 0x20
MODEDCTRANCURVE = 0x40
MODEINITFLOAT = 0x100
MODEINITJCT = 0x200
MODEINITFIX = 0x400
MODEINITSMSIG = 0x800
MODEINITTRAN = 0x1000
MODEINITPRED = 0x2000
MODEUIC = 0x10000

N_STATES = 29
N_CACHE = 38
N_STAMP = 16
N_RHS = 4

_lib = None


class EngineError(RuntimeError):
    def __init__(self, code, what):
        super().__init__('%s: error %d: %s' % (what, code, last_error()))
        self.code = code


class _Scalars(ctypes.Structure):
    _fields_ = [('mode', ctypes.c_int), ('order', ctypes.c_int), ('method', ctypes.c_int),
                ('ag', ctypes.c_double * 7), ('xfact', ctypes.c_double),
                ('gmin', ctypes.c_double), ('reltol', ctypes.c_double),
                ('abstol', ctypes.c_double)]


def lib():
    """The engine library; raises when it has not been built."""
    global _lib
    if _lib is None:
        # a process that uses torch must load torch's HIP runtime first, so
        # the engine binds to the same runtime instead of a second copy
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError('engine library %s not built (run make -C %s)' % (LIB_PATH, HERE))
        L = ctypes.CDLL(LIB_PATH)
        vp, d, i, cp = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_char_p
        i64, sz = ctypes.c_int64, ctypes.c_size_t
        L.ngmi_last_error.restype = cp
        L.ngmi_bsim4_model_new.restype = vp
        L.ngmi_bsim4_model_free.argtypes = [vp]
        L.ngmi_bsim4_model_param.argtypes = [vp, cp, d]
        L.ngmi_bsim4_model_setup.argtypes = [vp, d, d]
        L.ngmi_bsim4_model_get.argtypes = [vp, cp]
        L.ngmi_bsim4_model_get.restype = d
        L.ngmi_bsim4_set_new.restype = vp
        L.ngmi_bsim4_set_free.argtypes = [vp]
        L.ngmi_bsim4_set_add.argtypes = [vp, ctypes.POINTER(vp), i, ctypes.POINTER(i), i,
                                         ctypes.POINTER(cp), i, ctypes.POINTER(d)]
        L.ngmi_bsim4_set_finalize.argtypes = [vp]
        L.ngmi_bsim4_set_size.argtypes = [vp]
        L.ngmi_bsim4_set_field.argtypes = [vp, i, cp]
        L.ngmi_bsim4_set_field.restype = d
        L.ngmi_bsim4_set_layout.argtypes = [vp, ctypes.POINTER(i), ctypes.POINTER(i),
                                            ctypes.POINTER(sz)]
        L.ngmi_bsim4_eval.argtypes = [vp, ctypes.POINTER(_Scalars), vp, ctypes.POINTER(vp),
                                      vp, vp, vp, vp, vp]
        L.ngmi_csc_from_elements.argtypes = [i, i64, vp, vp, vp, vp, vp, ctypes.POINTER(i64)]
        L.ngmi_gather_new.argtypes = [i64, i64, vp]
        L.ngmi_gather_new.restype = vp
        L.ngmi_gather_free.argtypes = [vp]
        L.ngmi_gather_run.argtypes = [vp, vp, vp, vp]
        L.ngmi_device_info.argtypes = [cp, i, ctypes.POINTER(i), ctypes.POINTER(i)]
        L.ngmi_timer_new.restype = vp
        L.ngmi_timer_free.argtypes = [vp]
        L.ngmi_timer_start.argtypes = [vp, vp]
        L.ngmi_timer_stop.argtypes = [vp, vp]
        L.ngmi_timer_ms.argtypes = [vp]
        L.ngmi_timer_ms.restype = d
        L.ngmi_sync.argtypes = []
        _lib = L
    return _lib


def last_error():
    return (lib().ngmi_last_error() or b'').decode()


def check(rc, what):
    if rc != OK:
        raise EngineError(rc, what)


def device_info():
    buf = ctypes.create_string_buffer(256)
    cu, clk = ctypes.c_int(), ctypes.c_int()
    check(lib().ngmi_device_info(buf, 256, ctypes.byref(cu), ctypes.byref(clk)), 'device_info')
    return buf.value.decode(), cu.value, clk.value


def sync():
    check(lib().ngmi_sync(), 'sync')


def tune(knob, value):
    """launch-shape knob (ngmi_tune: 'eval_waves' and 'mos_waves' 0..4 with
    0 = automatic, 'streams', 'specialise', 'order_slack', 'graph',
    'guess_fused', 'guess_grid')"""
    L = lib()
    L.ngmi_tune.argtypes = [ctypes.c_char_p, ctypes.c_int]
    check(L.ngmi_tune(knob.encode(), int(value)), 'tune %s' % knob)


class Model:
    """One BSIM4 .model card at one temperature."""

    def __init__(self, params, tnom=300.15, temp=300.15):
        L = lib()
        self.h = L.ngmi_bsim4_model_new()
        for k, v in params.items():
            check(L.ngmi_bsim4_model_param(self.h, k.encode(), float(v)), 'model param %s' % k)
        check(L.ngmi_bsim4_model_setup(self.h, tnom, temp), 'model setup')

    def get(self, what):
        return lib().ngmi_bsim4_model_get(self.h, what.encode())

    def __del__(self):
        try:
            lib().ngmi_bsim4_model_free(self.h)
        except Exception:
            pass


class InstanceSet:
    """Device-resident BSIM4 instance tables."""

    def __init__(self):
        self.h = lib().ngmi_bsim4_set_new()
        self.models = []

    def add(self, models, model_idx, params):
        """params: dict name -> array of n values (instance parameters)."""
        L = lib()
        model_idx = np.ascontiguousarray(model_idx, dtype=np.int32)
        n = len(model_idx)
        names = list(params.keys())
        vals = np.ascontiguousarray(
            np.stack([np.broadcast_to(np.asarray(params[k], dtype=np.float64), (n,))
                      for k in names]) if names else np.zeros((0, n)))
        self.models.extend(models)
        mh = (ctypes.c_void_p * len(models))(*[m.h for m in models])
        nh = (ctypes.c_char_p * max(len(names), 1))(*[k.encode() for k in names])
        rc = L.ngmi_bsim4_set_add(self.h, mh, len(models),
                                  model_idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), n, nh,
                                  len(names), vals.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        if rc < 0:
            raise EngineError(-rc, 'set add')
        return rc

    def finalize(self):
        check(lib().ngmi_bsim4_set_finalize(self.h), 'set finalize')

    def __len__(self):
        return lib().ngmi_bsim4_set_size(self.h)

    def field(self, k, name):
        return lib().ngmi_bsim4_set_field(self.h, k, name.encode())

    def layout(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_size_t()
        check(lib().ngmi_bsim4_set_layout(self.h, ctypes.byref(a), ctypes.byref(b),
                                          ctypes.byref(c)), 'layout')
        return a.value, b.value, c.value

    def eval(self, mode, vterm, states, cache, stamp, rhs, noncon, ag=(0.0,) * 7, order=1,
             method=1, xfact=0.0, gmin=1e-12, reltol=1e-3, abstol=1e-12, stream=0):
        """BSIM4load over the set; all buffers are device addresses."""
        sc = _Scalars(mode, order, method, (ctypes.c_double * 7)(*ag), xfact, gmin, reltol,
                      abstol)
        st = (ctypes.c_void_p * 3)(*states)
        check(lib().ngmi_bsim4_eval(self.h, ctypes.byref(sc), vterm, st, cache, stamp, rhs,
                                    noncon, stream), 'bsim4 eval')

    def __del__(self):
        try:
            lib().ngmi_bsim4_set_free(self.h)
        except Exception:
            pass


def csc_from_elements(n, rows, cols):
    """SMPmakeElt + SMPconvertCOOtoCSC: returns (Ap, Ai, elt2nz)."""
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    ne = len(rows)
    Ap = np.zeros(n + 1, dtype=np.int64)
    Ai = np.zeros(max(ne, 1), dtype=np.int32)
    e2z = np.zeros(max(ne, 1), dtype=np.int64)
    nnz = ctypes.c_int64()
    check(lib().ngmi_csc_from_elements(n, ne, rows.ctypes.data, cols.ctypes.data, Ap.ctypes.data,
                                       Ai.ctypes.data, e2z.ctypes.data, ctypes.byref(nnz)),
          'csc_from_elements')
    return Ap, Ai[:nnz.value].copy(), e2z[:ne]


class Gather:
    """Deterministic assembly plan: Ax[z] = sum of src[e] over elements e -> z."""

    def __init__(self, nnz, elt2nz):
        e2z = np.ascontiguousarray(elt2nz, dtype=np.int64)
        self.nnz = nnz
        self.h = lib().ngmi_gather_new(nnz, len(e2z), e2z.ctypes.data)
        if not self.h:
            raise EngineError(1, 'gather plan')

    def run(self, src, Ax, stream=0):
        check(lib().ngmi_gather_run(self.h, src, Ax, stream), 'gather')

    def __del__(self):
        try:
            lib().ngmi_gather_free(self.h)
        except Exception:
            pass


class SparseLU:
    """One large sparse system factored by the whole GPU (ngmi_splu_*): the
    klu_analyze / klu_refactor / klu_solve trio for circuits beyond the
    per-circuit programs.  Ap (n+1) and Ai are a CSC pattern with ascending
    rows; analysis runs on the host at construction, factor() and solve()
    take device pointers (e.g. torch tensors' data_ptr())."""

    def __init__(self, n, Ap, Ai):
        L = lib()
        vp, i64 = ctypes.c_void_p, ctypes.c_int64
        L.ngmi_splu_analyze.argtypes = [ctypes.c_int, vp, vp, ctypes.POINTER(vp)]
        L.ngmi_splu_free.argtypes = [vp]
        L.ngmi_splu_info.argtypes = [vp] + [ctypes.POINTER(i64)] * 5 + [ctypes.POINTER(ctypes.c_double)]
        L.ngmi_splu_front.argtypes = [vp, ctypes.c_int] + [ctypes.POINTER(ctypes.c_int)] * 3 + [vp]
        L.ngmi_splu_rowsrc.argtypes = [vp, vp]
        L.ngmi_splu_factor.argtypes = [vp, vp, vp]
        L.ngmi_splu_solve.argtypes = [vp, vp, vp]
        L.ngmi_splu_multiplier.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
        self.n = int(n)
        Ap = np.ascontiguousarray(Ap, dtype=np.int64)
        Ai = np.ascontiguousarray(Ai, dtype=np.int32)
        h = vp()
        check(L.ngmi_splu_analyze(self.n, Ap.ctypes.data, Ai.ctypes.data, ctypes.byref(h)), 'splu_analyze')
        self.h = h.value

    def info(self):
        v = [ctypes.c_int64() for _ in range(5)]
        fl = ctypes.c_double()
        check(lib().ngmi_splu_info(self.h, *[ctypes.byref(x) for x in v], ctypes.byref(fl)), 'splu_info')
        keys = ('fronts', 'heights', 'max_front', 'front_doubles', 'lu_nonzeros')
        out = {k: x.value for k, x in zip(keys, v)}
        out['flops'] = fl.value
        return out

    def front(self, f):
        """(pivot count, boundary count, parent, unknowns) of front f"""
        ns, nb, par = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        L = lib()
        check(L.ngmi_splu_front(self.h, f, ctypes.byref(ns), ctypes.byref(nb), ctypes.byref(par), None),
              'splu_front')
        idx = np.zeros(ns.value + nb.value, dtype=np.int32)
        check(L.ngmi_splu_front(self.h, f, ctypes.byref(ns), ctypes.byref(nb), ctypes.byref(par),
                                idx.ctypes.data), 'splu_front')
        return ns.value, nb.value, par.value, idx

    def rowsrc(self):
        out = np.zeros(self.n, dtype=np.int32)
        check(lib().ngmi_splu_rowsrc(self.h, out.ctypes.data), 'splu_rowsrc')
        return out

    def factor(self, Ax_ptr, stream=0):
        check(lib().ngmi_splu_factor(self.h, Ax_ptr, stream), 'splu_factor')

    def solve(self, b_ptr, stream=0):
        check(lib().ngmi_splu_solve(self.h, b_ptr, stream), 'splu_solve')

    def multiplier(self):
        """the largest |l| multiplier of the last successful factor"""
        v = ctypes.c_double()
        check(lib().ngmi_splu_multiplier(self.h, ctypes.byref(v)), 'splu_multiplier')
        return v.value

    def __del__(self):
        try:
            if self.h:
                lib().ngmi_splu_free(self.h)
        except Exception:
            pass


class Timer:
    """HIP events recorded on the stream the timed kernels run on."""

    def __init__(self):
        self.h = lib().ngmi_timer_new()

    def start(self, stream=0):
        check(lib().ngmi_timer_start(self.h, stream), 'timer start')

    def stop(self, stream=0):
        check(lib().ngmi_timer_stop(self.h, stream), 'timer stop')

    def ms(self):
        return lib().ngmi_timer_ms(self.h)

    def __del__(self):
        try:
            lib().ngmi_timer_free(self.h)
        except Exception:
            pass


def _tran_protos(L):
    if getattr(L, '_tran_ready', False):
        return L
    vp, d, i, cp = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_char_p
    i64 = ctypes.c_int64
    L.ngmi_circuit_new.argtypes = [i]
    L.ngmi_circuit_new.restype = vp
    L.ngmi_circuit_free.argtypes = [vp]
    L.ngmi_circuit_add_bsim4.argtypes = [vp, vp, i, i, i, i, i, ctypes.POINTER(cp),
                                         ctypes.POINTER(d)]
    L.ngmi_circuit_add_cap.argtypes = [vp, i, i, d]
    L.ngmi_circuit_add_res.argtypes = [vp, i, i, d]
    L.ngmi_circuit_add_vsrc.argtypes = [vp, i, i, d, ctypes.POINTER(d), i]
    L.ngmi_circuit_set_ic.argtypes = [vp, i, d]
    L.ngmi_circuit_set_options.argtypes = [vp, d, d, d, d, d, d, i, i]
    L.ngmi_circuit_set_op_options.argtypes = [vp, i, i, i, d, d, i]
    L.ngmi_circuit_set_pivrel.argtypes = [vp, d]
    L.ngmi_circuit_set_optran.argtypes = [vp, d, d]
    L.ngmi_circuit_set_uic.argtypes = [vp, ctypes.c_int]
    ip = ctypes.POINTER(i)
    L.ngmi_circuit_set_pivot_order.argtypes = [vp, i, ip, ip]
    L.ngmi_circuit_analyze.argtypes = [vp, ip, ip, ip, ip, ip, ip, ip]
    L.ngmi_batch_new.argtypes = [ctypes.POINTER(vp), i, ctypes.POINTER(i), i, i64]
    L.ngmi_batch_new.restype = vp
    L.ngmi_batch_free.argtypes = [vp]
    L.ngmi_batch_info.argtypes = [vp, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i),
                                  ctypes.POINTER(i)]
    L.ngmi_batch_tran.argtypes = [vp, d, d, d, i64, vp, ctypes.POINTER(i64)]
    L.ngmi_batch_stats.argtypes = [vp, i, ctypes.POINTER(i), ctypes.POINTER(i64),
                                   ctypes.POINTER(i64), ctypes.POINTER(i64), ctypes.POINTER(i64)]
    L.ngmi_batch_waveform.argtypes = [vp, i, ctypes.POINTER(d), ctypes.POINTER(d), i64,
                                      ctypes.POINTER(i64)]
    L.ngmi_circuit_add_bsim4_many.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(i),
                                              ctypes.POINTER(i), i, i, ctypes.POINTER(cp),
                                              ctypes.POINTER(d)]
    L.ngmi_batch_groups.argtypes = [vp, ctypes.POINTER(i)]
    L.ngmi_batch_write_raw.argtypes = [vp, i, cp, cp, ctypes.POINTER(cp), i]
    L.ngmi_netlist_parse.argtypes = [cp, ctypes.c_uint64]
    L.ngmi_netlist_parse.restype = vp
    L.ngmi_netlist_free.argtypes = [vp]
    L.ngmi_netlist_title.argtypes = [vp]
    L.ngmi_netlist_title.restype = cp
    L.ngmi_netlist_size.argtypes = [vp, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i)]
    L.ngmi_netlist_name.argtypes = [vp, i]
    L.ngmi_netlist_name.restype = cp
    L.ngmi_netlist_find.argtypes = [vp, cp]
    L.ngmi_netlist_tran.argtypes = [vp] + [ctypes.POINTER(d)] * 4 + [ctypes.POINTER(i)]
    L.ngmi_netlist_element.argtypes = [vp, i, cp, ctypes.POINTER(i), ctypes.POINTER(d)]
    L.ngmi_netlist_element_name.argtypes = [vp, i]
    L.ngmi_netlist_element_name.restype = cp
    L.ngmi_netlist_circuit.argtypes = [vp]
    L.ngmi_netlist_circuit.restype = vp
    L.ngmi_batch_circuit_info.argtypes = [vp, i, ctypes.POINTER(i), ctypes.POINTER(i),
                                          ctypes.POINTER(i), ctypes.POINTER(i)]
    L.ngmi_batch_set_profile.argtypes = [vp, i]
    L.ngmi_batch_table_cols.argtypes = [vp, ctypes.POINTER(i), ctypes.POINTER(i)]
    u64 = ctypes.c_ulonglong
    L.ngmi_batch_kernel_times.argtypes = [vp, ctypes.POINTER(d), ctypes.POINTER(d),
                                          ctypes.POINTER(u64), ctypes.POINTER(u64)]
    L._tran_ready = True
    return L


class GuessNet:
    """The Newton initial-guess MLP (16 -> 256 x 3 SiLU -> 1) on the GPU."""

    IN, HID, IN_PAD = 16, 256, 32

    def __init__(self, seed=0):
        L = lib()
        vp, d, i = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
        fp = ctypes.POINTER(ctypes.c_float)
        L.ngmi_mlp_new.argtypes = [ctypes.c_ulonglong]
        L.ngmi_mlp_new.restype = vp
        L.ngmi_mlp_free.argtypes = [vp]
        L.ngmi_mlp_num_params.restype = ctypes.c_int64
        L.ngmi_mlp_get_params.argtypes = [vp, fp, i]
        L.ngmi_mlp_set_params.argtypes = [vp, fp]
        L.ngmi_mlp_forward.argtypes = [vp, vp, vp, i, vp]
        L.ngmi_mlp_train_step.argtypes = [vp, vp, vp, i, ctypes.c_float, fp, vp]
        self.h = L.ngmi_mlp_new(seed)
        if not self.h:
            raise EngineError(1, 'mlp')
        self.n = L.ngmi_mlp_num_params()

    def params(self, grads=False):
        out = np.zeros(self.n, dtype=np.float32)
        check(lib().ngmi_mlp_get_params(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                        1 if grads else 0), 'get params')
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(lib().ngmi_mlp_set_params(self.h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
              'set params')

    def unpack(self, p):
        """split a parameter vector into (W1[256, 32], b1, W2, b2, W3, b3, w4[256], b4)"""
        H, I = self.HID, self.IN_PAD
        sizes = [H * I, H, H * H, H, H * H, H, H, 1]
        out, o = [], 0
        for s in sizes:
            out.append(p[o:o + s])
            o += s
        W1, b1, W2, b2, W3, b3, w4, b4 = out
        return W1.reshape(H, I), b1, W2.reshape(H, H), b2, W3.reshape(H, H), b3, w4, b4

    def forward(self, X_dev, y_dev, M, stream=0):
        check(lib().ngmi_mlp_forward(self.h, X_dev, y_dev, M, stream), 'mlp forward')

    def train_step(self, X_dev, t_dev, M, lr, stream=0, want_loss=True):
        loss = ctypes.c_float()
        check(lib().ngmi_mlp_train_step(self.h, X_dev, t_dev, M, lr,
                                        ctypes.byref(loss) if want_loss else None, stream),
              'mlp train')
        return loss.value

    def __del__(self):
        try:
            lib().ngmi_mlp_free(self.h)
        except Exception:
            pass


class DevLoad:
    """BSIM4load for one circuit whose vectors live in host memory
    (ngmi_devload_*): the core a reference-shaped DEVload binding calls.
    set: a finalized InstanceSet; nodes [n,4] (d g s b, 0 = ground);
    state_off [n]; ax_pos [n,16] positions of each instance's Jacobian
    entries in the caller's matrix array (-1 none)."""

    def __init__(self, set_, n_unknowns, num_states, nodes, state_off, ax_pos):
        L = lib()
        vp = ctypes.c_void_p
        L.ngmi_devload_new.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp, vp]
        L.ngmi_devload_new.restype = vp
        L.ngmi_devload_free.argtypes = [vp]
        L.ngmi_devload_run.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        self.set = set_
        self._nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        self._soff = np.ascontiguousarray(state_off, dtype=np.int32)
        self._pos = np.ascontiguousarray(ax_pos, dtype=np.int64)
        self.h = L.ngmi_devload_new(set_.h, n_unknowns, num_states, self._nodes.ctypes.data,
                                    self._soff.ctypes.data, self._pos.ctypes.data)
        if not self.h:
            raise EngineError(1, 'devload_new')

    @staticmethod
    def run_fn():
        """address of ngmi_devload_run, for callers that take a C function"""
        return ctypes.cast(lib().ngmi_devload_run, ctypes.c_void_p).value

    def __del__(self):
        try:
            lib().ngmi_devload_free(self.h)
        except Exception:
            pass


class SMP:
    """SMPpreOrder / SMPluFac / SMPsolve for one circuit's host-memory CSC
    matrix on the GPU (ngmi_smp_*)."""

    def __init__(self, n, Ap, Ai):
        L = lib()
        vp = ctypes.c_void_p
        L.ngmi_smp_new.argtypes = [ctypes.c_int, vp, vp, ctypes.POINTER(vp)]
        L.ngmi_smp_free.argtypes = [vp]
        L.ngmi_smp_factor.argtypes = [vp, vp, ctypes.c_double]
        L.ngmi_smp_solve.argtypes = [vp, vp]
        Ap = np.ascontiguousarray(Ap, dtype=np.int64)
        Ai = np.ascontiguousarray(Ai, dtype=np.int32)
        h = vp()
        check(L.ngmi_smp_new(int(n), Ap.ctypes.data, Ai.ctypes.data, ctypes.byref(h)), 'smp_new')
        self.h = h.value

    @staticmethod
    def factor_fn():
        return ctypes.cast(lib().ngmi_smp_factor, ctypes.c_void_p).value

    @staticmethod
    def solve_fn():
        return ctypes.cast(lib().ngmi_smp_solve, ctypes.c_void_p).value

    def factor(self, Ax, gmin=0.0):
        Ax = np.ascontiguousarray(Ax, dtype=np.float64)
        check(lib().ngmi_smp_factor(self.h, Ax.ctypes.data, gmin), 'smp_factor')

    def solve(self, rhs):
        """rhs: float64 array [n+1], 1-based, solved in place"""
        check(lib().ngmi_smp_solve(self.h, rhs.ctypes.data), 'smp_solve')

    def __del__(self):
        try:
            lib().ngmi_smp_free(self.h)
        except Exception:
            pass


class Large:
    """One circuit too large for the batched programs (ngmi_large_*): the
    power-grid mesh's transient with DCtran / NIiter on the host and every
    vector operation, the BSIM4 loads and the whole-GPU LU on the device.
    probe: 1-based unknowns recorded at each accepted point (None: all)."""

    def __init__(self, circuit, probe=None, max_points=100000):
        L = lib()
        vp, i64 = ctypes.c_void_p, ctypes.c_int64
        L.ngmi_large_new.argtypes = [vp, vp, ctypes.c_int, i64]
        L.ngmi_large_new.restype = vp
        L.ngmi_large_free.argtypes = [vp]
        L.ngmi_large_info.argtypes = [vp] + [ctypes.POINTER(i64)] * 3 + [ctypes.POINTER(ctypes.c_double)]
        L.ngmi_large_tran.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, vp]
        L.ngmi_large_stats.argtypes = [vp] + [ctypes.POINTER(i64)] * 4
        L.ngmi_large_multiplier.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
        L.ngmi_large_waveform.argtypes = [vp, vp, vp, i64, ctypes.POINTER(i64)]
        L.ngmi_large_times.argtypes = [vp] + [ctypes.POINTER(ctypes.c_double)] * 4
        L.ngmi_large_kcl.argtypes = [vp, ctypes.POINTER(ctypes.c_double), vp]
        self.circuit = circuit
        pr = None if probe is None else np.ascontiguousarray(probe, dtype=np.int32)
        self.nprobe = None if pr is None else len(pr)
        self.h = L.ngmi_large_new(circuit.h, pr.ctypes.data if pr is not None else None,
                                  len(pr) if pr is not None else 0, max_points)
        if not self.h:
            raise EngineError(1, 'large_new')
        if self.nprobe is None:
            self.nprobe = self.info()['unknowns']

    def info(self):
        v = [ctypes.c_int64() for _ in range(3)]
        fl = ctypes.c_double()
        check(lib().ngmi_large_info(self.h, *[ctypes.byref(x) for x in v], ctypes.byref(fl)), 'large_info')
        return dict(unknowns=v[0].value, nonzeros=v[1].value, lu_nonzeros=v[2].value, flops=fl.value)

    def tran(self, tstep, tstop, tmax=0.0, stream=0):
        return lib().ngmi_large_tran(self.h, tstep, tstop, tmax, stream)

    def stats(self):
        v = [ctypes.c_int64() for _ in range(4)]
        check(lib().ngmi_large_stats(self.h, *[ctypes.byref(x) for x in v]), 'large_stats')
        lm = ctypes.c_double()
        check(lib().ngmi_large_multiplier(self.h, ctypes.byref(lm)), 'large_multiplier')
        return dict(points=v[0].value, accepted=v[1].value, rejected=v[2].value, iterations=v[3].value,
                    max_multiplier=lm.value)

    def waveform(self):
        n = self.stats()['points']
        t = np.zeros(max(n, 1))
        x = np.zeros((max(n, 1), self.nprobe))
        got = ctypes.c_int64()
        check(lib().ngmi_large_waveform(self.h, t.ctypes.data, x.ctypes.data, n, ctypes.byref(got)),
              'large_waveform')
        return t[:got.value], x[:got.value]

    def times(self):
        v = [ctypes.c_double() for _ in range(4)]
        check(lib().ngmi_large_times(self.h, *[ctypes.byref(x) for x in v]), 'large_times')
        return dict(load_ms=v[0].value, factor_ms=v[1].value, solve_ms=v[2].value, conv_ms=v[3].value)

    def kcl(self):
        """KCL residual of the final point: max |F(x)| over the unknowns, with
        F the circuit equations at the last accepted solution, scaled by the
        largest current term; returns (scaled residual, max |F| per row kind
        [nodes, branches])"""
        r = ctypes.c_double()
        per = np.zeros(2)
        check(lib().ngmi_large_kcl(self.h, ctypes.byref(r), per.ctypes.data), 'large_kcl')
        return r.value, per

    def __del__(self):
        try:
            lib().ngmi_large_free(self.h)
        except Exception:
            pass


def autotune(circuits, tstep, tstop, trial_passes=200, cache_path=None):
    """ngmi_autotune_batch over the Circuit objects; returns the timing
    report (the chosen knobs stay set process-wide)"""
    L = lib()
    vp = ctypes.c_void_p
    L.ngmi_autotune_batch.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int64,
                                      ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
    arr = (vp * len(circuits))(*[c.h for c in circuits])
    buf = ctypes.create_string_buffer(4096)
    check(L.ngmi_autotune_batch(arr, len(circuits), tstep, tstop, trial_passes,
                                cache_path.encode() if cache_path else None, buf, 4096), 'autotune')
    return buf.value.decode()


def autotune_load(circuits, cache_path):
    """apply a cached autotuner choice; False when the cache has none"""
    L = lib()
    vp = ctypes.c_void_p
    L.ngmi_autotune_load.argtypes = [ctypes.c_char_p, vp, ctypes.c_int]
    arr = (vp * len(circuits))(*[c.h for c in circuits])
    rc = L.ngmi_autotune_load(cache_path.encode(), arr, len(circuits))
    if rc == 14:
        return False
    check(rc, 'autotune_load')
    return True


def get_tune(knob):
    """current value of a launch knob"""
    L = lib()
    L.ngmi_tune_get.argtypes = [ctypes.c_char_p]
    v = L.ngmi_tune_get(knob.encode())
    if v < 0:
        raise EngineError(-v, 'tune get %s' % knob)
    return v


def klu_analyze(n, Ap, Ai):
    """KLU's symbolic analysis restated in the engine (ngmi_klu_analyze, host
    code): (P, Q, R, nblocks, nzoff, structural_rank), 0-based, for the n x n
    CSC pattern."""
    L = lib()
    vp = ctypes.c_void_p
    L.ngmi_klu_analyze.argtypes = [ctypes.c_int] + [vp] * 5 + [ctypes.POINTER(ctypes.c_int)] * 3
    Ap = np.ascontiguousarray(Ap, dtype=np.int32)
    Ai = np.ascontiguousarray(Ai, dtype=np.int32)
    P = np.zeros(max(n, 1), np.int32)
    Q = np.zeros(max(n, 1), np.int32)
    R = np.zeros(n + 1, np.int32)
    nb, nz, rk = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    check(L.ngmi_klu_analyze(int(n), Ap.ctypes.data, Ai.ctypes.data, P.ctypes.data, Q.ctypes.data, R.ctypes.data,
                             ctypes.byref(nb), ctypes.byref(nz), ctypes.byref(rk)), 'klu_analyze')
    return P[:n], Q[:n], R[:nb.value + 1], nb.value, nz.value, rk.value


def guess_features(t4, h, v4, nbv, nbmask, reltol=1e-3, vntol=1e-6, stream=0):
    """The engine's guess feature rows (the device function the in-loop
    features kernel uses) for caller-supplied windows, through
    ngmi_guess_features: t4[M,4] and v4[M,4] the last four accepted times
    and values (newest first), h[M] the step, nbv[M,8] neighbour values,
    nbmask[M] their presence bits.  Device pointers (torch tensors' data_ptr)
    in, (F [M,16] float32, act [M] float64) device tensors out."""
    import torch
    M = int(h.numel())
    F = torch.empty((M, 16), dtype=torch.float32, device=h.device)
    act = torch.empty(M, dtype=torch.float64, device=h.device)
    L = lib()
    vp = ctypes.c_void_p
    L.ngmi_guess_features.argtypes = [ctypes.c_int] + [vp] * 5 + [ctypes.c_double] * 2 + [vp] * 3
    check(L.ngmi_guess_features(M, t4.data_ptr(), h.data_ptr(), v4.data_ptr(), nbv.data_ptr(),
                                nbmask.data_ptr(), reltol, vntol, F.data_ptr(), act.data_ptr(), stream),
          'guess features')
    return F, act


class Circuit:
    """A netlist for the engine from a circuit description dict (nnodes,
    models, mos, caps, res, vsrc, ic, options); see workloads.py."""

    def __init__(self, desc, temp=300.15):
        L = _tran_protos(lib())
        self.models = {k: Model(v, temp=temp) for k, v in desc['models'].items()}
        self.h = L.ngmi_circuit_new(desc['nnodes'])
        if not self.h:
            raise EngineError(1, 'circuit')
        o = dict(reltol=1e-3, abstol=1e-12, vntol=1e-6, chgtol=1e-14, trtol=7.0, gmin=1e-12,
                 method=1, maxord=2)
        o.update(desc.get('options', {}))
        check(L.ngmi_circuit_set_options(self.h, o['reltol'], o['abstol'], o['vntol'],
                                         o['chgtol'], o['trtol'], o['gmin'], o['method'],
                                         o['maxord']), 'options')
        op = dict(noopiter=0, gminsteps=1, srcsteps=1, gminfactor=10.0, gshunt=0.0, itl2=50)
        op.update({k: v for k, v in o.items() if k in op})
        check(L.ngmi_circuit_set_op_options(self.h, op['noopiter'], op['gminsteps'],
                                            op['srcsteps'], op['gminfactor'], op['gshunt'],
                                            op['itl2']), 'operating-point options')
        check(L.ngmi_circuit_set_pivrel(self.h, o.get('pivrel', 1e-3)), 'pivrel')
        opt = o.get('optran', (1e-7, 1e-5))
        check(L.ngmi_circuit_set_optran(self.h, float(opt[0]) if opt else 0.0, float(opt[1]) if opt else 1.0),
              'optran')
        check(L.ngmi_circuit_set_uic(self.h, int(o.get('uic', 0))), 'uic')
        mos = desc['mos']
        if mos:
            keys = list(self.models)
            pnames = sorted({k for m in mos for k in m[5]})
            n = len(mos)
            midx = np.array([keys.index(m[0]) for m in mos], dtype=np.int32)
            nodes = np.array([m[1:5] for m in mos], dtype=np.int32).reshape(-1)
            vals = np.zeros((len(pnames), n))
            for k, m in enumerate(mos):
                for j, p in enumerate(pnames):
                    if p not in m[5]:
                        raise ValueError('every MOSFET of a circuit must give the same '
                                         'instance parameters (%s missing)' % p)
                    vals[j, k] = m[5][p]
            mh = (ctypes.c_void_p * len(keys))(*[self.models[k].h for k in keys])
            names = (ctypes.c_char_p * max(len(pnames), 1))(*[p.encode() for p in pnames])
            check(L.ngmi_circuit_add_bsim4_many(
                self.h, mh, midx.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                nodes.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), n, len(pnames), names,
                vals.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), 'add bsim4')
        for p, n, cv in desc['caps']:
            check(L.ngmi_circuit_add_cap(self.h, p, n, cv), 'add cap')
        for p, n, r in desc.get('res', []):
            check(L.ngmi_circuit_add_res(self.h, p, n, r), 'add res')
        for p, n, dc, pulse in desc['vsrc']:
            if pulse is None:
                check(L.ngmi_circuit_add_vsrc(self.h, p, n, dc, None, 0), 'add vsrc')
            else:
                arr = (ctypes.c_double * len(pulse))(*pulse)
                check(L.ngmi_circuit_add_vsrc(self.h, p, n, dc, arr, len(pulse)), 'add vsrc')
        for node, v in desc.get('ic', {}).items():
            check(L.ngmi_circuit_set_ic(self.h, node, v), 'set ic')

    def use_klu_order(self, on=True):
        """Static LU on KLU's own ordering, computed by the engine (the
        restated klu_analyze), instead of the engine's own."""
        L = lib()
        L.ngmi_circuit_use_klu_order.argtypes = [ctypes.c_void_p, ctypes.c_int]
        check(L.ngmi_circuit_use_klu_order(self.h, 1 if on else 0), 'klu order')

    def set_pivot_order(self, rows, cols):
        """Pivot sequence of the static LU (1-based unknowns), e.g. KLU's
        Pnum + 1 / Q + 1; None restores the engine's own ordering."""
        L = _tran_protos(lib())
        if rows is None:
            check(L.ngmi_circuit_set_pivot_order(self.h, 0, None, None), 'pivot order')
            return
        r = np.ascontiguousarray(rows, dtype=np.int32)
        c = np.ascontiguousarray(cols, dtype=np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        check(L.ngmi_circuit_set_pivot_order(self.h, len(r), r.ctypes.data_as(ip),
                                             c.ctypes.data_as(ip)), 'pivot order')

    def analyze(self):
        """Host-only symbolic analysis: dict with unknowns, lu_nonzeros,
        l_nonzeros (strictly lower), factor_terms, levels and the pivot
        sequence rows/cols (1-based)."""
        L = _tran_protos(lib())
        i = ctypes.c_int
        n, ns, nt, nl = i(), i(), i(), i()
        lev = (i * 3)()
        check(L.ngmi_circuit_analyze(self.h, ctypes.byref(n), ctypes.byref(ns), ctypes.byref(nt),
                                     ctypes.byref(nl), lev, None, None), 'analyze')
        rows = np.zeros(n.value, dtype=np.int32)
        cols = np.zeros(n.value, dtype=np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        check(L.ngmi_circuit_analyze(self.h, None, None, None, None, None, rows.ctypes.data_as(ip),
                                     cols.ctypes.data_as(ip)), 'analyze')
        return dict(unknowns=n.value, lu_nonzeros=ns.value, l_nonzeros=nl.value,
                    factor_terms=nt.value, levels=tuple(lev), rows=rows, cols=cols)

    def __del__(self):
        try:
            lib().ngmi_circuit_free(self.h)
        except Exception:
            pass


class Batch:
    """C circuits advanced together on the GPU; circuits of equal topology
    share one symbolic analysis (a group), one launch per pass covers all."""

    def __init__(self, circuits, probe=None, max_points=20000):
        L = _tran_protos(lib())
        self.circuits = circuits
        arr = (ctypes.c_void_p * len(circuits))(*[c.h for c in circuits])
        if probe is None:
            pp, npb = None, 0
        else:
            pp = (ctypes.c_int * len(probe))(*probe)
            npb = len(probe)
        self.h = L.ngmi_batch_new(arr, len(circuits), pp, npb, max_points)
        if not self.h:
            raise EngineError(1, 'batch')
        self.C = len(circuits)

    def info(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lv = (ctypes.c_int * 3)()
        check(lib().ngmi_batch_info(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                    lv), 'info')
        return dict(unknowns=a.value, lu_nonzeros=b.value, factor_terms=c.value,
                    levels=tuple(lv))

    def groups(self):
        g = ctypes.c_int()
        check(lib().ngmi_batch_groups(self.h, ctypes.byref(g)), 'groups')
        return g.value

    def circuit_info(self, c):
        g, n, p, s = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().ngmi_batch_circuit_info(self.h, c, ctypes.byref(g), ctypes.byref(n),
                                            ctypes.byref(p), ctypes.byref(s)), 'circuit info')
        return dict(group=g.value, unknowns=n.value, nprobe=p.value, lu_nonzeros=s.value)

    def tran(self, tstep, tstop, tmax=0.0, max_passes=10 ** 7, stream=0):
        passes = ctypes.c_int64()
        rc = lib().ngmi_batch_tran(self.h, tstep, tstop, tmax, max_passes, stream,
                                   ctypes.byref(passes))
        return rc, passes.value

    def table_cols(self):
        a, b2 = ctypes.c_int(), ctypes.c_int()
        check(lib().ngmi_batch_table_cols(self.h, ctypes.byref(a), ctypes.byref(b2)), 'cols')
        return a.value, b2.value

    def set_guess(self, net):
        """attach a GuessNet (None detaches) for the learned initial guess"""
        L = lib()
        L.ngmi_batch_set_guess.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.net = net
        check(L.ngmi_batch_set_guess(self.h, net.h if net is not None else None), 'set guess')

    def neighbours(self, nnodes):
        L = lib()
        L.ngmi_batch_neighbours.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        out = np.zeros((nnodes + 1) * 8, dtype=np.int32)
        check(L.ngmi_batch_neighbours(self.h, out.ctypes.data), 'neighbours')
        return out

    def guess_stats(self, c):
        L = lib()
        i64 = ctypes.c_int64
        L.ngmi_batch_guess_stats.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(i64),
                                             ctypes.POINTER(i64), ctypes.POINTER(i64),
                                             ctypes.POINTER(ctypes.c_double)]
        g, f, pz, ms = i64(), i64(), i64(), ctypes.c_double()
        check(L.ngmi_batch_guess_stats(self.h, c, ctypes.byref(g), ctypes.byref(f), ctypes.byref(pz),
                                       ctypes.byref(ms)), 'guess stats')
        return dict(guessed=g.value, fallbacks=f.value, paused=pz.value, guess_ms=ms.value)

    def pivot_stats(self, c):
        """KLU's pivoting in circuit c: pv (1 while it refactors on pivot
        rows of its own), the counts of pivoted factorizations and of
        refactors on their rows, and the pivot sequence in use (1-based
        unknowns of the k-th pivot's row and column)."""
        L = lib()
        i64 = ctypes.c_int64
        ip = ctypes.POINTER(ctypes.c_int)
        L.ngmi_batch_pivot_stats.argtypes = [ctypes.c_void_p, ctypes.c_int, ip, ctypes.POINTER(i64),
                                             ctypes.POINTER(i64), ip, ip]
        n = self.circuit_info(c)['unknowns']
        rows = np.zeros(n, dtype=np.int32)
        cols = np.zeros(n, dtype=np.int32)
        pv, f, r = ctypes.c_int(), i64(), i64()
        check(L.ngmi_batch_pivot_stats(self.h, c, ctypes.byref(pv), ctypes.byref(f), ctypes.byref(r),
                                       rows.ctypes.data_as(ip), cols.ctypes.data_as(ip)), 'pivot stats')
        return dict(pv=pv.value, factors=f.value, refactors=r.value, rows=rows, cols=cols)

    def guess_rows(self):
        """network rows of the last run, summed over its passes"""
        L = lib()
        L.ngmi_batch_guess_rows.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]
        r = ctypes.c_int64()
        check(L.ngmi_batch_guess_rows(self.h, ctypes.byref(r)), 'guess rows')
        return r.value

    def set_profile(self, on=True):
        check(lib().ngmi_batch_set_profile(self.h, 1 if on else 0), 'profile')

    def kernel_times(self):
        """(BSIM4 kernel ms, step kernel ms, BSIM4 evaluations, circuit Newton
        passes) summed over the last tran() run with profiling on."""
        a, b2 = ctypes.c_double(), ctypes.c_double()
        e, p = ctypes.c_ulonglong(), ctypes.c_ulonglong()
        check(lib().ngmi_batch_kernel_times(self.h, ctypes.byref(a), ctypes.byref(b2),
                                            ctypes.byref(e), ctypes.byref(p)), 'kernel times')
        return a.value, b2.value, e.value, p.value

    def stats(self, c):
        st = ctypes.c_int()
        a = [ctypes.c_int64() for _ in range(4)]
        check(lib().ngmi_batch_stats(self.h, c, ctypes.byref(st), *[ctypes.byref(x) for x in a]),
              'stats')
        return dict(status=st.value, npts=a[0].value, accepted=a[1].value, rejected=a[2].value,
                    iterations=a[3].value)

    _DATA_CB = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                ctypes.c_int, ctypes.c_void_p)

    def set_data_callback(self, fn, circuits=None):
        """Live data stream (ngmi_batch_set_data_callback): fn(circuit,
        first, t, x) receives each watched circuit's newly recorded points
        (numpy copies, x as [count, nprobe]) during tran(); None turns it
        off."""
        L = lib()
        L.ngmi_batch_set_data_callback.argtypes = [ctypes.c_void_p, self._DATA_CB, ctypes.c_void_p,
                                                   ctypes.c_int, ctypes.c_void_p]
        if fn is None:
            self._data_cb = None
            check(L.ngmi_batch_set_data_callback(self.h, self._DATA_CB(), None, 0, None), 'data callback')
            return

        def tramp(c, first, count, t, x, nprobe, user):
            fn(c, first, np.ctypeslib.as_array(t, (count,)).copy(),
               np.ctypeslib.as_array(x, (count, nprobe)).copy())
            return 0

        self._data_cb = self._DATA_CB(tramp)
        ids = None if circuits is None else np.ascontiguousarray(circuits, dtype=np.int32)
        check(L.ngmi_batch_set_data_callback(self.h, self._data_cb, None if ids is None else ids.ctypes.data,
                                             0 if ids is None else len(ids), None), 'data callback')

    def waveform(self, c, nprobe=None):
        if nprobe is None:
            nprobe = self.circuit_info(c)['nprobe']
        s = self.stats(c)
        n = max(s['npts'], 1)
        t = np.zeros(n)
        x = np.zeros((n, nprobe))
        got = ctypes.c_int64()
        check(lib().ngmi_batch_waveform(self.h, c, t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                        x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n,
                                        ctypes.byref(got)), 'waveform')
        return t[:got.value], x[:got.value]

    def write_raw(self, c, path, title='', names=None, binary=False):
        """Circuit c's waveform as an ngspice batch-mode raw file; names[k]
        names recorded unknown k (node name or '<source>#branch')."""
        arr = None
        if names is not None:
            arr = (ctypes.c_char_p * len(names))(*[n.encode() for n in names])
        check(lib().ngmi_batch_write_raw(self.h, c, str(path).encode(), title.encode(), arr,
                                         1 if binary else 0), 'write_raw')

    def __del__(self):
        try:
            lib().ngmi_batch_free(self.h)
        except Exception:
            pass


class Netlist:
    """A SPICE deck parsed by the engine's front end (ngmi_netlist_parse):
    MNA numbering as ngspice gives the same deck."""

    def __init__(self, text, seed=0):
        L = _tran_protos(lib())
        self.h = L.ngmi_netlist_parse(text.encode(), seed)
        if not self.h:
            raise EngineError(1, 'netlist: ' + last_error())

    @property
    def title(self):
        return lib().ngmi_netlist_title(self.h).decode()

    def size(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().ngmi_netlist_size(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
              'netlist size')
        return dict(nodes=a.value, unknowns=b.value, elements=c.value)

    def names(self):
        """unknown names, index 0 = ground"""
        n = self.size()['unknowns']
        return [lib().ngmi_netlist_name(self.h, k).decode() for k in range(n + 1)]

    def find(self, name):
        return lib().ngmi_netlist_find(self.h, name.encode())

    def tran(self):
        v = [ctypes.c_double() for _ in range(4)]
        u = ctypes.c_int()
        check(lib().ngmi_netlist_tran(self.h, *[ctypes.byref(x) for x in v], ctypes.byref(u)),
              'netlist tran')
        return dict(tstep=v[0].value, tstop=v[1].value, tstart=v[2].value, tmax=v[3].value,
                    uic=u.value)

    def elements(self):
        """flattened elements: (name, kind, terminal unknowns, value)"""
        out = []
        for k in range(self.size()['elements']):
            kind = ctypes.create_string_buffer(1)
            nodes = (ctypes.c_int * 4)()
            val = ctypes.c_double()
            n = lib().ngmi_netlist_element(self.h, k, kind, nodes, ctypes.byref(val))
            out.append((lib().ngmi_netlist_element_name(self.h, k).decode(), kind.raw.decode(),
                        list(nodes[:n]), val.value))
        return out

    def circuit(self):
        return NetlistCircuit(self)

    def __del__(self):
        try:
            lib().ngmi_netlist_free(self.h)
        except Exception:
            pass


class NetlistCircuit:
    """A circuit built from a Netlist (usable in Batch like Circuit)."""

    def __init__(self, netlist):
        self.netlist = netlist
        self.h = lib().ngmi_netlist_circuit(netlist.h)
        if not self.h:
            raise EngineError(1, 'netlist circuit: ' + last_error())

    def analyze(self):
        return Circuit.analyze(self)

    def __del__(self):
        try:
            lib().ngmi_circuit_free(self.h)
        except Exception:
            pass
