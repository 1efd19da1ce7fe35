"""The whole-GPU multifrontal LU, host side (ngmi_splu_analyze): static
pivot pairs, the nested-dissection frontal tree and every front's boundary.

The structure is checked by using it: a dense-front factorization in numpy
that follows exactly the fronts, index lists and row exchanges the library
exported must solve the power-grid matrix as SuperLU (scipy) does.  A
missing boundary unknown drops fill and a wrong pair or row exchange puts a
zero on a pivot, so either shows up as a wrong solution or a failed pivot.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from engine_path import ngmi, workloads


def _structure(s, n):
    info = s.info()
    fronts = [s.front(f) for f in range(info['fronts'])]
    return info, fronts, s.rowsrc()


def frontal_solve(s, n, Ap, Ai, Ax, b, stats=None):
    """factor + solve with the library's fronts, dense per front; stats
    (a dict), when given, receives the largest |l| multiplier"""
    info, fronts, rowsrc = _structure(s, n)
    pos = np.full(n, -1)
    owner = np.full(n, -1)
    p = 0
    for f, (ns, nb, par, idx) in enumerate(fronts):
        for u in idx[:ns]:
            assert pos[u] < 0, 'unknown %d pivots twice' % u
            pos[u] = p
            owner[u] = f
            p += 1
    assert p == n and (pos >= 0).all()
    loc = [{int(u): k for k, u in enumerate(fr[3])} for fr in fronts]
    F = [np.zeros((fr[0] + fr[1],) * 2) for fr in fronts]
    for c in range(n):
        for e in range(Ap[c], Ap[c + 1]):
            r = rowsrc[Ai[e]]
            v = r if pos[r] < pos[c] else c
            f = owner[v]
            F[f][loc[f][r], loc[f][c]] += Ax[e]
    kids = [[] for _ in fronts]
    for f, fr in enumerate(fronts):
        if fr[2] >= 0:
            assert fr[2] > f, 'parent before child in postorder'
            kids[fr[2]].append(f)
    for f, (ns, nb, par, idx) in enumerate(fronts):
        for c in kids[f]:
            nsc = fronts[c][0]
            m = [loc[f][int(u)] for u in fronts[c][3][nsc:]]
            F[f][np.ix_(m, m)] += F[c][nsc:, nsc:]
        G = F[f]
        for k in range(ns):
            assert G[k, k] != 0.0, 'zero pivot in front %d' % f
            G[k + 1:, k] /= G[k, k]
            if stats is not None and k + 1 < ns + nb:
                stats['lmax'] = max(stats.get('lmax', 0.0), float(np.abs(G[k + 1:, k]).max()))
            G[k + 1:, k + 1:] -= np.outer(G[k + 1:, k], G[k, k + 1:])
    bp = np.asarray(b)[rowsrc]
    y = np.zeros(n)
    upd = {}
    for f, (ns, nb, par, idx) in enumerate(fronts):
        r = np.zeros(ns + nb)
        r[:ns] = bp[idx[:ns]]
        for c in kids[f]:
            nsc = fronts[c][0]
            r[[loc[f][int(u)] for u in fronts[c][3][nsc:]]] += upd[c]
        G = F[f]
        for k in range(ns):
            r[k + 1:] -= G[k + 1:, k] * r[k]
        y[idx[:ns]] = r[:ns]
        upd[f] = r[ns:]
    x = np.zeros(n)
    for f in reversed(range(len(fronts))):
        ns, nb, par, idx = fronts[f]
        G = F[f]
        r = y[idx[:ns]] - G[:ns, ns:] @ x[idx[ns:]]
        for k in reversed(range(ns)):
            r[k] /= G[k, k]
            r[:k] -= G[:k, k] * r[k]
        x[idx[:ns]] = r
    return x


@pytest.mark.parametrize('side,loads,pad', [(32, 16, 8), (40, 64, 16)])
def test_frontal_structure_solves_the_power_grid(side, loads, pad):
    n, Ap, Ai, Ax = workloads.power_grid_jacobian(side, loads, pad, seed=side)
    s = ngmi.SparseLU(n, Ap, Ai)
    b = workloads.rng(5, side).uniform(-1, 1, n)
    x = frontal_solve(s, n, Ap, Ai, Ax, b)
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    ref = spla.spsolve(A, b)
    err = np.abs(x - ref).max() / np.abs(ref).max()
    print('n %d, %s, max relative difference to SuperLU %.2e' % (n, s.info(), err))
    assert err < 1e-10


def test_boundaries_are_ancestor_pivots():
    n, Ap, Ai, Ax = workloads.power_grid_jacobian(64, 128, 16, seed=3)
    s = ngmi.SparseLU(n, Ap, Ai)
    info, fronts, rowsrc = _structure(s, n)
    owner = np.full(n, -1)
    for f, (ns, nb, par, idx) in enumerate(fronts):
        owner[idx[:ns]] = f
    for f, (ns, nb, par, idx) in enumerate(fronts):
        anc = set()
        a = par
        while a >= 0:
            anc.add(a)
            a = fronts[a][2]
        assert {int(owner[u]) for u in idx[ns:]} <= anc
    assert info['heights'] > 3 and info['max_front'] < 400


def test_branch_rows_are_exchanged_with_their_node():
    n, Ap, Ai, Ax = workloads.power_grid_jacobian(32, 16, 8)
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n)).tocsr()
    rowsrc = ngmi.SparseLU(n, Ap, Ai).rowsrc()
    assert (rowsrc[rowsrc] == np.arange(n)).all()
    nodiag = np.flatnonzero(np.asarray(A.diagonal() == 0))
    moved = np.flatnonzero(rowsrc != np.arange(n))
    assert len(moved) == 2 * len(nodiag) > 0
    B = A[rowsrc]
    assert (B.diagonal() != 0).all()


def test_unknown_without_a_pivot_is_rejected():
    # column 1 empty: no diagonal and no node to pair with
    with pytest.raises(ngmi.EngineError):
        ngmi.SparseLU(2, np.array([0, 2, 2]), np.array([0, 1]))
    assert 'no node to pair with' in ngmi.last_error()


def test_bad_patterns_are_rejected():
    with pytest.raises(ngmi.EngineError) as e:
        ngmi.SparseLU(2, np.array([0, 1, 2]), np.array([0, 5]))
    assert e.value.code == 11 and 'out of range' in ngmi.last_error()
    with pytest.raises(ngmi.EngineError) as e:
        ngmi.SparseLU(2, np.array([0, 2, 1]), np.array([0, 1]))
    assert e.value.code == 11 and 'decrease' in ngmi.last_error()
    # klu_analyze_given.c:55-95 also rejects Ap[0] != 0 and a row twice in a column
    with pytest.raises(ngmi.EngineError) as e:
        ngmi.SparseLU(2, np.array([1, 2, 3]), np.array([0, 0, 1]))
    assert e.value.code == 11 and 'Ap[0]' in ngmi.last_error()
    with pytest.raises(ngmi.EngineError) as e:
        ngmi.SparseLU(2, np.array([0, 2, 3]), np.array([0, 0, 1]))
    assert e.value.code == 11 and 'twice' in ngmi.last_error()


@pytest.mark.parametrize('seed', [0, 1])
def test_small_dense_and_source_systems(seed):
    """a 5x5 dense matrix (one leaf front) and a 3-node circuit with two
    voltage sources (two row exchanges), against numpy's dense solve"""
    g = np.random.default_rng(seed)
    A = g.uniform(-1, 1, (5, 5)) + 5 * np.eye(5)
    S = sp.csc_matrix(A)
    x = frontal_solve(ngmi.SparseLU(5, S.indptr, S.indices), 5, S.indptr, S.indices, S.data,
                      b := g.uniform(-1, 1, 5))
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-12, atol=1e-12)
    # nodes 0..2, branches 3 (source on node 0) and 4 (source from node 2 to node 1)
    M = np.zeros((5, 5))
    for a, c, gv in ((0, 1, 1e-3), (1, 2, 2e-3)):
        M[a, a] += gv; M[c, c] += gv; M[a, c] -= gv; M[c, a] -= gv
    M[2, 2] += 1e-4
    M[0, 3] = M[3, 0] = 1.0
    M[2, 4] = M[4, 2] = 1.0
    M[1, 4] = M[4, 1] = -1.0
    S = sp.csc_matrix(M)
    lu = ngmi.SparseLU(5, S.indptr, S.indices)
    # each branch pairs with the first node in its column: 3 with 0, 4 with 1
    assert list(np.flatnonzero(lu.rowsrc() != np.arange(5))) == [0, 1, 3, 4]
    b = g.uniform(-1, 1, 5)
    assert np.allclose(frontal_solve(lu, 5, S.indptr, S.indices, S.data, b), np.linalg.solve(M, b),
                       rtol=1e-10, atol=1e-12)
