"""GPU parity of the BSIM4 evaluation kernel against the oracle.

Bar (per instance, per group of like quantities): the largest difference
between engine and oracle is at most 1e-12 of that instance's largest
magnitude in the group (or of 1e-6 of the group's largest magnitude in the
batch, for instances whose whole group is near zero).  Groups: terminal currents, conductances (the
per-instance values the load keeps for the next call), charges, charge
currents, terminal-voltage states, the 16 Jacobian entries and the 4 RHS
entries.  Per-group scaling is used because several outputs are sums of
terms of both signs (the equivalent current sources of the RHS, the
source-side Jacobian entries); an element that cancels to 1e-6 of its group
carries FP64 rounding of the group's size, not of its own.  Device
convergence votes must be identical.  The printed "rel" column is the
element-wise relative difference over entries above 1e-3 of the group scale,
for information.
"""
import numpy as np
import pytest

from engine_path import ngmi, workloads
from oracle_eval import OracleBatch, CACHE_FIELDS, CURRENT_ROWS, CONDUCTANCE_ROWS

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu

TOL_SCALED = 1e-12
CHARGE_STATES = [11, 13, 15, 28, 19, 21]   # qb qg qd qs qbs qbd
CHARGE_CURRENTS = [12, 14, 16]             # cqb cqg cqd
VOLT_STATES = [0, 1, 2, 3]


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _compare(name, gpu, ref):
    """gpu, ref: [rows, n]."""
    # an instance whose whole group is (near) zero -- e.g. charge currents
    # right after INITTRAN, where the oracle's ag0 q0 + ag1 q1 cancels to
    # exactly 0 and a fused multiply-add leaves a rounding residual -- is
    # scaled by 1e-6 of the group's largest magnitude in the batch
    scale = np.maximum(np.abs(ref).max(axis=0), 1e-6 * np.abs(ref).max())
    scale = np.where(scale == 0, 1e-300, scale)
    d = np.abs(gpu - ref)
    scaled = (d / scale).max()
    big = np.abs(ref) > 1e-3 * scale
    rel = (d[big] / np.abs(ref[big])).max() if big.any() else 0.0
    print('%-10s scaled %.2e rel %.2e' % (name, scaled, rel), flush=True)
    assert scaled <= TOL_SCALED, (name, scaled)
    return scaled


class GpuEval:
    def __init__(self, models, model_idx, params):
        self.m = [ngmi.Model(mp) for mp in models]
        self.s = ngmi.InstanceSet()
        self.s.add(self.m, model_idx, params)
        self.s.finalize()
        n = len(model_idx)
        self.n = n
        self.cache = torch.zeros(38 * n, dtype=torch.float64, device='cuda')
        self.stamp = torch.zeros(16 * n, dtype=torch.float64, device='cuda')
        self.rhs = torch.zeros(4 * n, dtype=torch.float64, device='cuda')
        self.nc = torch.zeros(n, dtype=torch.int32, device='cuda')

    def load(self, mode, vterm, states, **kw):
        st = [_dev(s) for s in states]
        vt = _dev(vterm)
        self.s.eval(mode, vt.data_ptr(), [x.data_ptr() for x in st], self.cache.data_ptr(),
                    self.stamp.data_ptr(), self.rhs.data_ptr(), self.nc.data_ptr(), **kw)
        torch.cuda.synchronize()
        for lv in range(3):
            states[lv][:] = st[lv].cpu().numpy().reshape(29, self.n)
        return (self.stamp.cpu().numpy().reshape(16, self.n),
                self.rhs.cpu().numpy().reshape(4, self.n), self.nc.cpu().numpy(),
                self.cache.cpu().numpy().reshape(38, self.n)[:len(CACHE_FIELDS)])


def _check_all(g, gst, o, ocache, os_, or_, onc, ost):
    gs, gr, gnc, gcache = g
    _compare('currents', gcache[CURRENT_ROWS], ocache[CURRENT_ROWS])
    _compare('conduct', gcache[CONDUCTANCE_ROWS], ocache[CONDUCTANCE_ROWS])
    _compare('stamp', gs, os_)
    _compare('rhs', gr, or_)
    for lv in range(3):
        _compare('charge%d' % lv, gst[lv][CHARGE_STATES], ost[lv][CHARGE_STATES])
        _compare('cq%d' % lv, gst[lv][CHARGE_CURRENTS], ost[lv][CHARGE_CURRENTS])
        _compare('volt%d' % lv, gst[lv][VOLT_STATES], ost[lv][VOLT_STATES])
    assert np.array_equal(gnc, onc)


def test_device_eval_workload_slice():
    """The device-evaluation workload's distribution (first 1024 NMOS and
    1024 PMOS draws of seed 1), backward-Euler transient load."""
    wl = workloads.device_eval_workload(n=2048, seed=1, with_matrix=False)
    models = [workloads.BASE_NMOS, workloads.BASE_PMOS]
    print('case: device_eval_workload n=2048 seed=1', flush=True)
    g = GpuEval(models, wl['model_idx'], wl['params'])
    o = OracleBatch(models, wl['model_idx'], wl['params'])
    h = workloads.H_STEP
    kw = dict(ag=(1 / h, -1 / h, 0, 0, 0, 0, 0), order=1)
    gst = workloads.eval_states(wl)
    ost = [s.copy() for s in gst]
    mode = ngmi.MODETRAN | ngmi.MODEINITFLOAT
    gout = g.load(mode, wl['vterm'], gst, **kw)
    os_, or_, onc = o.load(mode, wl['vterm'], ost, **kw)
    _check_all(gout, gst, o, o.cache, os_, or_, onc, ost)
    assert np.abs(gout[0]).max() > 0 and np.abs(gst[0][CHARGE_STATES]).max() > 0


VARIANTS = [{}, {'igcmod': 1, 'igbmod': 1}, {'igcmod': 2, 'gidlmod': 1}, {'mobmod': 2},
            {'mobmod': 3, 'diomod': 2}, {'mobmod': 6, 'tempmod': 3}, {'cvchargemod': 1},
            {'diomod': 0, 'tempmod': 1, 'xpart': 0.0}, {'lambda': 0.1, 'vtl': 2e5}]


@pytest.mark.parametrize('var', range(len(VARIANTS)))
def test_newton_mode_sequence(var):
    """INITJCT -> DC INITFLOAT -> INITTRAN -> INITPRED -> transient INITFLOAT
    iterations (trapezoidal order 2), each compared after the call, with the
    cache of the previous call feeding limiting and the convergence test."""
    rng = np.random.default_rng(100 + var)
    n = 256
    nm = dict(workloads.BASE_NMOS, **VARIANTS[var])
    pm = dict(workloads.BASE_PMOS, **VARIANTS[var])
    model_idx = (np.arange(n) >= n // 2).astype(np.int32)
    params = {'l': np.exp(rng.uniform(np.log(45e-9), np.log(1e-6), n)),
              'w': np.exp(rng.uniform(np.log(100e-9), np.log(10e-6), n)),
              'delvto': rng.normal(0, 0.03, n),
              'ad': rng.uniform(0, 1e-12, n), 'as': rng.uniform(0, 1e-12, n),
              'pd': rng.uniform(0, 2e-6, n), 'ps': rng.uniform(0, 2e-6, n),
              'nf': rng.integers(1, 4, n).astype(float)}
    print('case: mode sequence variant %r' % (VARIANTS[var],), flush=True)
    g = GpuEval([nm, pm], model_idx, params)
    o = OracleBatch([nm, pm], model_idx, params)
    typ = np.where(model_idx == 0, 1.0, -1.0)
    gst = [np.zeros((29, n)) for _ in range(3)]
    ost = [np.zeros((29, n)) for _ in range(3)]
    h = 1e-11
    steps = [
        (ngmi.MODEDCOP | ngmi.MODEINITJCT, dict(ag=(0,) * 7, order=1)),
        (ngmi.MODEDCOP | ngmi.MODEINITFLOAT, dict(ag=(0,) * 7, order=1)),
        (ngmi.MODETRAN | ngmi.MODEINITTRAN, dict(ag=(1 / h, -1 / h, 0, 0, 0, 0, 0), order=1)),
        (ngmi.MODETRAN | ngmi.MODEINITPRED, dict(ag=(2 / h, 1.0, 0, 0, 0, 0, 0), order=2)),
        (ngmi.MODETRAN | ngmi.MODEINITFLOAT, dict(ag=(2 / h, 1.0, 0, 0, 0, 0, 0), order=2)),
        (ngmi.MODETRAN | ngmi.MODEINITFLOAT, dict(ag=(2 / h, 1.0, 0, 0, 0, 0, 0), order=2)),
    ]
    for k, (mode, kw) in enumerate(steps):
        v = np.stack([typ * rng.uniform(-0.2, 1.1, n), typ * rng.uniform(-0.2, 1.2, n),
                      typ * rng.uniform(-0.05, 0.05, n), typ * rng.uniform(-0.4, 0.1, n)])
        okw = dict(kw)
        gkw = dict(kw)
        if mode & ngmi.MODEINITPRED:
            okw.update(delta=h, delta_old=(h, 0.8 * h, 0.7 * h, 0, 0, 0, 0))
            gkw.update(xfact=h / (0.8 * h))
        print('  step %d mode %#x' % (k, mode), flush=True)
        gout = g.load(mode, v, gst, **gkw)
        os_, or_, onc = o.load(mode, v, ost, **okw)
        _check_all(gout, gst, o, o.cache, os_, or_, onc, ost)


def test_gather_assembly_matches_ordered_sum():
    """Deterministic assembly into a random 2**14-row CSC equals numpy's
    ordered accumulation (np.add.at visits elements in ascending order)."""
    wl = workloads.device_eval_workload(n=4096, seed=3, rows=1 << 14)
    n = wl['n']
    Ap, Ai, e2z = ngmi.csc_from_elements(wl['nrows'], wl['rows_e'], wl['cols_e'])
    nnz = len(Ai)
    gat = ngmi.Gather(nnz, e2z)
    src = np.random.default_rng(0).normal(size=16 * n)
    d_src = _dev(src)
    d_ax = torch.zeros(nnz, dtype=torch.float64, device='cuda')
    gat.run(d_src.data_ptr(), d_ax.data_ptr())
    torch.cuda.synchronize()
    ref = np.zeros(nnz)
    np.add.at(ref, e2z, src)
    assert np.array_equal(d_ax.cpu().numpy(), ref)
