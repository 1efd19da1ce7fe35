/*
 * ORACLE (test infrastructure only): a four-call wrapper around the
 * reference's own KLU, compiled from /root/reference/src/maths/KLU into
 * _ref/libklu_ref.so together with this file (see Makefile `ref`).  It lets
 * the oracle's transient restatement solve with the reference's solver the
 * way klusmp.c drives it: klu_analyze once (SMPpreOrder, klusmp.c:1255-1285),
 * klu_factor on SMPreorder (klusmp.c:752-806), klu_refactor on SMPluFac
 * (klusmp.c:573-624, falling back to a fresh klu_factor when the refactor
 * reports a singular matrix, niiter.c:160-190), klu_solve on SMPsolve
 * (klusmp.c:950-1011).
 */
#include <stdlib.h>
#include "ngspice/klu.h"

typedef struct {
    int n;
    int *Ap, *Ai;
    klu_common common;
    klu_symbolic *sym;
    klu_numeric *num;
    /* klu_factor calls, those that chose an off-diagonal pivot, klu_refactor
     * calls and those that met a zero pivot */
    int nfactor, noffdiag, nrefactor, nrefail;
} kg_t;

void *kg_new(int n, const int *Ap, const int *Ai)
{
    kg_t *h = (kg_t *)calloc(1, sizeof(kg_t));
    int k, nz = Ap[n];
    if (!h) return NULL;
    h->n = n;
    h->Ap = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    h->Ai = (int *)malloc(sizeof(int) * (size_t)(nz > 0 ? nz : 1));
    for (k = 0; k <= n; k++) h->Ap[k] = Ap[k];
    for (k = 0; k < nz; k++) h->Ai[k] = Ai[k];
    klu_defaults(&h->common);
    h->sym = klu_analyze(n, h->Ap, h->Ai, &h->common);
    if (!h->sym) {
        free(h->Ap); free(h->Ai); free(h);
        return NULL;
    }
    return h;
}

/* full = 1: klu_factor (pivoting); full = 0: klu_refactor, then klu_factor
 * if that fails.  Returns 0, 2 (the refactor failed and the fresh factor
 * succeeded), or 102 (E_SINGULAR) / 1 on failure. */
int kg_factor(void *hp, double *Ax, int full, double pivrel)
{
    kg_t *h = (kg_t *)hp;
    int refailed = 0;
    h->common.tol = pivrel;
    if (!full && h->num) {
        h->nrefactor++;
        if (klu_refactor(h->Ap, h->Ai, Ax, h->sym, h->num, &h->common)) return 0;
        h->nrefail++;
        refailed = 1;
    }
    if (h->num) klu_free_numeric(&h->num, &h->common);
    h->nfactor++;
    h->num = klu_factor(h->Ap, h->Ai, Ax, h->sym, &h->common);
    if (!h->num) return h->common.status == KLU_SINGULAR ? 102 : 1;
    if (h->common.noffdiag > 0) h->noffdiag++;
    if (h->common.status == KLU_SINGULAR) return 102;
    return refailed ? 2 : 0;
}

int kg_solve(void *hp, double *b)
{
    kg_t *h = (kg_t *)hp;
    return klu_solve(h->sym, h->num, h->n, 1, b, &h->common) ? 0 : 1;
}

void kg_free(void *hp)
{
    kg_t *h = (kg_t *)hp;
    if (!h) return;
    if (h->num) klu_free_numeric(&h->num, &h->common);
    if (h->sym) klu_free_symbolic(&h->sym, &h->common);
    free(h->Ap); free(h->Ai); free(h);
}

/* The pivot sequence of the last klu_factor: rows[k] = Numeric->Pnum[k],
 * cols[k] = Symbolic->Q[k] (0-based), and counts = {lnz, unz, nzoff,
 * nblocks} (klu.h: nnz of L and U including their diagonals, entries of
 * the off-diagonal blocks, BTF blocks), then the klu_factor calls, those
 * that chose an off-diagonal pivot (Common->noffdiag), the klu_refactor
 * calls and those that met a zero pivot (8 ints).  Returns 0, or 1 before
 * a factor. */
int kg_order(void *hp, int *rows, int *cols, int *counts)
{
    kg_t *h = (kg_t *)hp;
    int k;
    if (!h || !h->num) return 1;
    for (k = 0; k < h->n; k++) {
        rows[k] = h->num->Pnum[k];
        cols[k] = h->sym->Q[k];
    }
    counts[0] = (int)h->num->lnz;
    counts[1] = (int)h->num->unz;
    counts[2] = (int)h->num->nzoff;
    counts[3] = (int)h->sym->nblocks;
    counts[4] = h->nfactor;
    counts[5] = h->noffdiag;
    counts[6] = h->nrefactor;
    counts[7] = h->nrefail;
    return 0;
}

/* The symbolic analysis of klu_analyze alone (SMPpreOrder,
 * klusmp.c:1255-1285): P, Q (0-based, size n) and the block boundaries
 * R[0..nblocks]; info = {nblocks, nzoff, structural rank}.  Returns 0, or 1
 * when klu_analyze fails. */
int kg_symbolic(int n, const int *Ap, const int *Ai, int *P, int *Q, int *R, int *info)
{
    void *hp = kg_new(n, Ap, Ai);
    kg_t *h = (kg_t *)hp;
    int k;
    if (!h) return 1;
    for (k = 0; k < n; k++) {
        P[k] = h->sym->P[k];
        Q[k] = h->sym->Q[k];
    }
    for (k = 0; k <= h->sym->nblocks; k++) R[k] = h->sym->R[k];
    info[0] = h->sym->nblocks;
    info[1] = h->sym->nzoff;
    info[2] = h->sym->structural_rank;
    kg_free(hp);
    return 0;
}
