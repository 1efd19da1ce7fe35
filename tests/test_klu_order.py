"""The engine's restatement of KLU's symbolic analysis (ngmi_klu_analyze:
BTF -- maximum transversal and strongly connected components -- then AMD on
each block) against the reference's own klu_analyze, compiled from its
sources into oracle/_ref (oracle/Makefile `ref`; without it these tests
skip).  CPU only: the analysis is host code.

Bar: identical output, element for element -- Symbolic->P, Q, R, the block
count, the off-block entry count and the structural rank -- on
* seeded random patterns: sorted and jumbled columns, structurally singular
  ones (BTF's flipped matching), many small blocks (natural order for
  blocks of three or fewer), denser ones (AMD's dense-row rule);
* 2-D grid Laplacians up to 60 x 60, plain, perturbed and permuted, whose
  fill makes AMD compress its workspace during the elimination;
* the circuit matrices of the ring oscillators, the adder and a power-grid
  mesh as the oracle builds them for KLU (the pattern SMPpreOrder passes).
Then the engine's own pivot sequence with use_klu_order equals the pivot
sequence the reference KLU used in a transient (Numeric->Pnum, Symbolic->Q):
on these circuits klu_factor keeps the analysis' row order; and the static
LU on it has KLU's structure: nnz(L) and nnz(U) of the diagonal blocks plus
the off-block entries, with no fill outside the blocks."""
import os
import sys

import numpy as np
import pytest

from engine_path import ngmi, workloads

sys.path.insert(0, os.path.join(os.path.dirname(__file__), '..', 'oracle'))
import pyoracle as po  # noqa: E402

sp = pytest.importorskip('scipy.sparse')


def _ref_or_skip():
    try:
        po.klu_ref()
    except OSError:
        pytest.skip('reference KLU (oracle/_ref) not built')


def _same(n, Ap, Ai):
    ref = po.klu_symbolic(n, Ap, Ai)
    got = ngmi.klu_analyze(n, Ap, Ai)
    for name, a, b in zip(('P', 'Q', 'R'), ref[:3], got[:3]):
        assert np.array_equal(a, b), name
    assert ref[3:] == got[3:], (ref[3:], got[3:])
    return got


def _random_pattern(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 300))
    dens = float(rng.uniform(0.002, 0.05)) if seed % 4 else float(rng.uniform(0.05, 0.3))
    A = sp.random(n, n, density=dens, format='csc', random_state=seed)
    if rng.random() < 0.7:
        A = A + sp.eye(n)
    A = A.tocsc()
    A.sum_duplicates()
    A.sort_indices()
    if rng.random() < 0.4:   # rows out of order inside the columns
        for j in range(n):
            a, b = A.indptr[j], A.indptr[j + 1]
            A.indices[a:b] = rng.permutation(A.indices[a:b])
    return n, A.indptr, A.indices


def test_random_patterns_match_the_reference():
    _ref_or_skip()
    singular = blocky = 0
    for seed in range(120):
        n, Ap, Ai = _random_pattern(seed)
        P, Q, R, nb, nzoff, rank = _same(n, Ap, Ai)
        singular += rank < n
        blocky += nb > 1
    print('120 patterns: %d structurally singular, %d with several blocks' % (singular, blocky))
    assert singular > 10 and blocky > 10


@pytest.mark.parametrize('side', [10, 30, 60])
def test_grid_laplacians_match_the_reference(side):
    _ref_or_skip()
    n = side * side
    eye = sp.eye(side)
    T = sp.diags([1, 1], [-1, 1], shape=(side, side))
    A = (sp.kron(eye, T) + sp.kron(T, eye) + 4 * sp.eye(n)).tocsc()
    rng = np.random.default_rng(side)
    variants = [A, A + sp.random(n, n, density=2.0 / n, format='csc', random_state=side)]
    p = rng.permutation(n)
    variants.append(A[p][:, p])
    for B in variants:
        B = B.tocsc()
        B.sum_duplicates()
        B.sort_indices()
        _same(n, B.indptr, B.indices)


CIRCUITS = [
    ('ring11', lambda: workloads.ring_oscillator(stages=11, tstop=2e-9)),
    ('ring101', lambda: workloads.ring_oscillator()),
    ('adder', lambda: workloads.mc_adder(0, tstop=0.2e-9)),
    ('mesh24', lambda: workloads.power_grid(side=24, pad_every=8, tstop=0.05e-9)),
]


@pytest.mark.parametrize('name,make', CIRCUITS, ids=[c[0] for c in CIRCUITS])
def test_circuit_patterns_and_pivot_sequence(name, make):
    _ref_or_skip()
    desc = make()
    T = po.Transient(desc)
    assert T.use_reference_klu()
    Ap, Ai = T.csc()
    n = T.n
    P, Q, R, nb, nzoff, rank = _same(n, Ap, Ai)
    rc, _, _ = T.run(desc['tstep'], desc['tstep'] * 5)
    assert rc == 0
    rows, cols, cnt = T.klu_order()
    assert cnt['nblocks'] == nb and cnt['nzoff'] == nzoff
    c = ngmi.Circuit(desc)
    c.use_klu_order()
    a = c.analyze()
    print('%s: n %d, %d blocks, %d off-block entries; engine pivot sequence = KLU Pnum/Q: %s'
          % (name, n, nb, nzoff, bool((a['rows'] == rows).all() and (a['cols'] == cols).all())))
    assert (a['cols'] == Q + 1).all() and (a['rows'] == P + 1).all()
    assert (a['rows'] == rows).all() and (a['cols'] == cols).all()
    # the factor keeps to KLU's structure: fill only inside the diagonal
    # blocks, the off-block entries as they are (klu_factor.c:15-260)
    total = cnt['lnz'] - n + cnt['unz'] + cnt['nzoff']
    print('  engine L %d (KLU lnz - n %d), L+U+off %d (KLU %d), levels %s'
          % (a['l_nonzeros'], cnt['lnz'] - n, a['lu_nonzeros'], total, a['levels']))
    assert a['l_nonzeros'] == cnt['lnz'] - n
    assert a['lu_nonzeros'] == total
    c.use_klu_order(False)
    assert not (c.analyze()['rows'] == rows).all() or name == 'ring11'


def test_bad_patterns_are_rejected():
    with pytest.raises(ngmi.EngineError):
        ngmi.klu_analyze(3, [1, 2, 3, 4], [0, 1, 2])
    with pytest.raises(ngmi.EngineError):
        ngmi.klu_analyze(2, [0, 1, 2], [0, 5])
