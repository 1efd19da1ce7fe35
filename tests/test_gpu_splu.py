"""The whole-GPU multifrontal LU (ngmi_splu_factor / ngmi_splu_solve)
against the reference KLU compiled into oracle/_ref (SuperLU from scipy
where that is not built), on the power-grid matrix.

Bar: the two direct solvers pivot differently (KLU numerically, this one on
a static nested-dissection sequence), so solutions agree to rounding, held
to 1e-9 of the solution's largest entry; at the full 1024 x 1024 size the
scaled residual |A x - b| / (|A| |x| + |b|) stays below 1e-14, and a second
refactor of the same values reproduces the solution bit for bit.
"""
import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from engine_path import ngmi, workloads

sys.path.insert(0, os.path.join(os.path.dirname(__file__), '..', 'oracle'))
import pyoracle as po  # noqa: E402

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu


def _gpu(s, Ax, b):
    dA = torch.from_numpy(np.ascontiguousarray(Ax)).cuda()
    db = torch.from_numpy(np.ascontiguousarray(b)).cuda()
    s.factor(dA.data_ptr())
    s.solve(db.data_ptr())
    torch.cuda.synchronize()
    return db.cpu().numpy()


@pytest.mark.parametrize('side,loads,pad', [(48, 64, 16), (256, 1024, 64)])
def test_power_grid_against_reference_klu(side, loads, pad):
    print('case: power grid %dx%d, %d inverters, pads every %d' % (side, side, loads, pad), flush=True)
    n, Ap, Ai, Ax = workloads.power_grid_jacobian(side, loads, pad, seed=side)
    b = workloads.rng(5, side).uniform(-1, 1, n)
    s = ngmi.SparseLU(n, Ap, Ai)
    x = _gpu(s, Ax, b)
    got = po.klu_csc(n, Ap, Ai, Ax, b)
    if got is not None:
        ref, tag = got[0], 'reference KLU'
    else:
        ref, tag = spla.spsolve(sp.csc_matrix((Ax, Ai, Ap), shape=(n, n)), b), 'SuperLU'
    err = np.abs(x - ref).max() / np.abs(ref).max()
    print('n %d, %s: max difference to %s %.2e of max |x|' % (n, s.info(), tag, err), flush=True)
    assert err < 1e-9
    # the same pattern with doubled values: every rounding doubles too
    x2 = _gpu(s, 2.0 * Ax, b)
    assert np.array_equal(x2, 0.5 * x)


def test_full_power_grid_residual_and_repeatability():
    print('case: power grid 1024x1024, 16384 inverters, pads every 64', flush=True)
    n, Ap, Ai, Ax = workloads.power_grid_jacobian()
    b = workloads.rng(5, 1024).uniform(-1, 1, n)
    s = ngmi.SparseLU(n, Ap, Ai)
    print('n %d, %s' % (n, s.info()), flush=True)
    x = _gpu(s, Ax, b)
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    res = np.abs(A @ x - b).max()
    scale = abs(A).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max()
    print('scaled residual %.2e' % (res / scale), flush=True)
    assert res / scale < 1e-14
    x_again = _gpu(s, Ax, b)
    assert np.array_equal(x, x_again)


def test_zero_pivot_is_reported():
    print('case: 2x2 with a zero first pivot', flush=True)
    Ap = np.array([0, 2, 4])
    Ai = np.array([0, 1, 0, 1])
    Ax = np.array([0.0, 1.0, 1.0, 1.0])
    s = ngmi.SparseLU(2, Ap, Ai)
    dA = torch.from_numpy(Ax).cuda()
    db = torch.ones(2, dtype=torch.float64, device='cuda')
    # no factorization yet: the solve refuses
    with pytest.raises(ngmi.EngineError) as e:
        s.solve(db.data_ptr())
    assert e.value.code == ngmi.E_SINGULAR
    with pytest.raises(ngmi.EngineError) as e:
        s.factor(dA.data_ptr())
    assert e.value.code == ngmi.E_SINGULAR
    # after a failed factorization the solve refuses too, instead of running on inf/NaN fronts
    with pytest.raises(ngmi.EngineError) as e:
        s.solve(db.data_ptr())
    assert e.value.code == ngmi.E_SINGULAR
    # a good matrix afterwards factors and solves
    dA.copy_(torch.tensor([2.0, 1.0, 1.0, 1.0], dtype=torch.float64))
    s.factor(dA.data_ptr())
    s.solve(db.data_ptr())
    torch.cuda.synchronize()
    assert np.allclose(db.cpu().numpy(), np.linalg.solve(np.array([[2.0, 1.0], [1.0, 1.0]]), np.ones(2)))


def test_largest_multiplier():
    """ngmi_splu_multiplier: the largest |l| of the factorization equals the
    one numpy's dense-front factorization over the library's own fronts
    gives (test_splu.frontal_solve); a 2x2 whose pivots are both 1e-6 of
    their columns reports 1e6, the multiplier KLU's threshold test
    (klu_kernel.c:370-530) would have refused"""
    from test_splu import frontal_solve
    print('case: largest multiplier, power grid 48x48 and a 2x2 with small pivots', flush=True)
    n, Ap, Ai, Ax = workloads.power_grid_jacobian(48, 64, 16, seed=48)
    b = workloads.rng(5, 48).uniform(-1, 1, n)
    s = ngmi.SparseLU(n, Ap, Ai)
    _gpu(s, Ax, b)
    st = {}
    frontal_solve(s, n, Ap, Ai, Ax, b, st)
    got = s.multiplier()
    print('power grid: library %.15g, numpy fronts %.15g' % (got, st['lmax']), flush=True)
    assert abs(got - st['lmax']) <= 1e-12 * st['lmax']
    eps = 1e-6
    s2 = ngmi.SparseLU(2, np.array([0, 2, 4]), np.array([0, 1, 0, 1]))
    _gpu(s2, np.array([eps, 1.0, 1.0, eps]), np.ones(2))
    print('2x2: %.17g' % s2.multiplier(), flush=True)
    assert s2.multiplier() == 1.0 / eps


def test_concurrent_with_a_transient_batch():
    """The large-system LU refactoring and solving on its own stream and host
    thread while a ring batch runs its transient on another (how the
    configuration-3 bench runs them): both give bit for bit what they give
    alone."""
    import threading
    print('case: 128x128 power grid LU x6 beside four rings .tran 10p 1n', flush=True)
    n, Ap, Ai, Ax = workloads.power_grid_jacobian(128, 256, 32, seed=7)
    b = workloads.rng(5, 7).uniform(-1, 1, n)
    s = ngmi.SparseLU(n, Ap, Ai)
    x_alone = _gpu(s, Ax, b)
    descs = [workloads.ring_oscillator(stages=k, tstop=1e-9) for k in (11, 13, 21, 31)]

    def rings():
        bt = ngmi.Batch([ngmi.Circuit(d) for d in descs], max_points=200000)
        return bt

    b1 = rings()
    rc, _ = b1.tran(10e-12, 1e-9)
    assert rc == 0
    alone = [b1.waveform(c) for c in range(len(descs))]
    b2 = rings()
    dA = torch.from_numpy(Ax).cuda()
    db0 = torch.from_numpy(b).cuda()
    ms = torch.cuda.Stream()
    outs = []

    def mesh_loop():
        with torch.cuda.stream(ms):
            for _ in range(6):
                db = db0.clone()
                s.factor(dA.data_ptr(), ms.cuda_stream)
                s.solve(db.data_ptr(), ms.cuda_stream)
                ms.synchronize()
                outs.append(db.cpu().numpy())

    rs = torch.cuda.Stream()
    th = threading.Thread(target=mesh_loop)
    th.start()
    rc, _ = b2.tran(10e-12, 1e-9, stream=rs.cuda_stream)
    th.join()
    torch.cuda.synchronize()
    assert rc == 0, ngmi.last_error()
    assert len(outs) == 6 and all(np.array_equal(o, x_alone) for o in outs)
    for c in range(len(descs)):
        t, x = b2.waveform(c)
        assert np.array_equal(t, alone[c][0]) and np.array_equal(x, alone[c][1])
    print('6 concurrent refactor+solves and 4 ring waveforms identical to the runs alone', flush=True)
