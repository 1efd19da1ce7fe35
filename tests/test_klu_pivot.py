"""KLU's numeric pivoting, on the CPU: the oracle's restatement of
klu_factor's pivot choice (or_klu_pivot: threshold partial pivoting with
diagonal preference on the row-scaled matrix, klu_kernel.c:370-530,
klu_scale.c:19-159) against the reference's own klu_factor, compiled from
its sources into oracle/_ref (without it these tests skip); and NIiter's
handling of a singular matrix and of fresh factorizations in the oracle's
transient, against the same transient solved by the reference KLU.

Bars:
* pivot rows (Numeric->Pnum) equal KLU's, except where the first
  difference falls on an exact tie between the two rows' scaled candidates
  (KLU takes the first in its column pattern, the restatement the lowest
  row), at most 5 % of the matrices; singularity (KLU_SINGULAR) agrees on
  every matrix; the restatement's solution has a scaled residual within
  1e-12 of the working precision bound;
* the ring with a floating 0.3 V source (workloads.ring_with_divider): the
  reference KLU picks an off-diagonal pivot at every fresh factorization
  (four: INITJCT and the iteration after it, the first INITTRAN iteration
  and the one after it, niiter.c:107-146, 335-347), and the oracle's dense
  solve gives the same waveform to 1e-9 of full scale;
* a floating node on an inverter chain (workloads.with_floating_node):
  with optran deselected both solvers end the operating point the same way
  (the same return code after the same Newton iterations); with optran
  (the default) both find it through the transient operating point, with
  the same Newton iterations and waveforms within 1e-12, the nodes with a
  DC path within 1 mV of the operating point that gshunt = 1e-12 gives;
  with gshunt the waveforms agree to 1e-12.
"""
import os
import sys

import numpy as np
import pytest

from engine_path import workloads

sys.path.insert(0, os.path.join(os.path.dirname(__file__), '..', 'oracle'))
import pyoracle as po  # noqa: E402

sp = pytest.importorskip('scipy.sparse')


def _ref_or_skip():
    if po.klu_ref() is None:
        pytest.skip('reference KLU (oracle/_ref) not built')


def _matrix(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 60))
    A = sp.random(n, n, density=float(rng.uniform(0.05, 0.4)), format='csc', random_state=seed)
    A = A + sp.diags(rng.uniform(0.5, 2, n) * (rng.random(n) < 0.8))
    A = A.tocsc()
    A.sum_duplicates()
    A.sort_indices()
    # magnitudes over ten orders, either sign; every third matrix with its
    # diagonal shrunk a millionfold
    A.data = A.data * 10.0 ** rng.uniform(-5, 5, A.data.size) * np.where(rng.random(A.data.size) < .5, -1, 1)
    if seed % 3 == 0:
        d = A.diagonal()
        A = (A + sp.diags(d * 1e-6 - d)).tocsc()
        A.sort_indices()
    return A, rng.standard_normal(n)


def _scaled_candidate(A, P, Q, Pnum, k0, row):
    """|Schur complement (row, Q[k0])| / Rs[row] after KLU's first k0 pivots"""
    Ad = A.toarray()
    Rs = np.abs(Ad).max(axis=1)
    Rs[Rs == 0] = 1
    S = Ad / Rs[:, None]
    done = set()
    for k in range(k0):
        col, p = Q[k], Pnum[k]
        for i in range(Ad.shape[0]):
            if i in done or i == p or S[i, col] == 0:
                continue
            S[i, :] -= S[i, col] / S[p, col] * S[p, :]
        done.add(p)
    return abs(S[row, Q[k0]])


def test_pivot_rule_matches_reference_klu():
    _ref_or_skip()
    total = agree = ties = offdiag = singular = 0
    for seed in range(400):
        A, b = _matrix(seed)
        n = A.shape[0]
        P, Q, R, nb, nzoff, rank = po.klu_symbolic(n, A.indptr, A.indices)
        rc, Pk, Qk, xk = po.klu_factor_ref(n, A.indptr, A.indices, A.data, 1e-3, b)
        rc2, Pn, x = po.klu_pivot(A.toarray(), P, Q, 1e-3, b)
        total += 1
        assert (rc != 0) == (rc2 != 0), (seed, rc, rc2)
        if rc:
            singular += 1
            continue
        assert (Qk == Q).all()
        offdiag += bool((Pk != P).any())
        Ad = A.toarray()
        res = np.abs(Ad @ x - b).max() / (np.abs(Ad).max() * np.abs(x).max() + np.abs(b).max())
        assert res < 1e-12 * n, (seed, res)
        if (Pk == Pn).all():
            agree += 1
            continue
        k0 = int(np.nonzero(Pk != Pn)[0][0])
        a = _scaled_candidate(A, P, Q, Pk, k0, Pk[k0])
        c = _scaled_candidate(A, P, Q, Pk, k0, Pn[k0])
        assert abs(a - c) <= 1e-12 * max(a, c), (seed, k0, a, c)
        ties += 1
    print('%d matrices: %d singular, %d with off-diagonal pivots; pivot rows equal KLU on %d, '
          '%d differ first at an exact tie' % (total, singular, offdiag, agree, ties))
    assert offdiag > 200 and singular > 20
    assert ties <= 0.05 * (total - singular)


def _run(desc, klu):
    T = po.Transient(desc)
    if klu:
        assert T.use_reference_klu()
    rc, t, x = T.run(desc['tstep'], desc['tstop'])
    return T, rc, t, x


def test_divider_pivots_off_the_diagonal():
    _ref_or_skip()
    desc = workloads.ring_with_divider()
    T, rc, t, x = _run(desc, True)
    assert rc == 0
    rows, cols, cnt = T.klu_order()
    Ap, Ai = T.csc()
    P = po.klu_symbolic(T.n, Ap, Ai)[0]
    print('reference KLU: %s; last Pnum differs from P at %d steps'
          % (cnt, int((rows - 1 != P).sum())))
    assert cnt['factors'] == 4 and cnt['offdiag_factors'] == 4 and cnt['refactor_fails'] == 0
    T2, rc2, t2, x2 = _run(desc, False)
    assert rc2 == 0 and len(t2) == len(t)
    scale = np.maximum(np.abs(x).max(axis=0), 1.0)
    err = (np.abs(x2 - x) / scale).max()
    print('dense solve against the reference KLU: %.2e of full scale' % err)
    assert err <= 1e-9


def test_floating_node_singular_operating_point():
    """A capacitor-only node makes the operating point's matrix singular:
    the Newton run and gmin and source stepping fail (E_SINGULAR, then
    no convergence), and CKTop ends with optran (cktop.c:93-105)."""
    _ref_or_skip()
    chain = workloads.inverter_chain()
    desc = workloads.with_floating_node(chain, node=3)
    desc['options'] = dict(desc['options'], optran=None)
    Tk, rck, _, _ = _run(desc, True)
    Td, rcd, _, _ = _run(desc, False)
    print('optran deselected: KLU rc %d %s, dense rc %d %s' % (rck, Tk.stats(), rcd, Td.stats()))
    assert rck != 0 and rck == rcd
    assert Tk.stats()['iterations'] == Td.stats()['iterations']
    desc = workloads.with_floating_node(chain, node=3)
    Tk, rck, tk, xk = _run(desc, True)
    Td, rcd, td, xd = _run(desc, False)
    assert rck == 0 and rcd == 0 and len(tk) == len(td)
    assert Tk.stats()['iterations'] == Td.stats()['iterations']
    err = np.abs(xk - xd).max()
    print('optran: %d points, %d iterations, %s; waveforms differ by %.2e'
          % (len(tk), Td.stats()['iterations'], Tk.klu_order()[2], err))
    assert err <= 1e-12
    # the nodes with a path to a source settle near the operating point that
    # gshunt = 1e-12 gives; not onto it, since the output nodes held near a
    # rail by an off transistor's leakage relax over times beyond the 10 us
    # of optran (1 fF against picosiemens)
    g = workloads.with_floating_node(chain, node=3, gshunt=1e-12)
    Tg, rcg, tg, xg = _run(g, False)
    assert rcg == 0
    n = chain['nnodes']
    dc = np.abs(xd[0, 1:n + 1] - xg[0, 1:n + 1]).max()
    print('optran operating point against the gshunt one: %.2e V' % dc)
    assert dc <= 1e-3
    Tk, rck, tk, xk = _run(g, True)
    assert rck == 0 and np.abs(xk - xg).max() <= 1e-12
