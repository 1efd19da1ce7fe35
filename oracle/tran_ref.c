/*
 * ORACLE (test infrastructure only).  Scalar C restatement of ngspice 46's
 * transient path for circuits of BSIM4 MOSFETs, capacitors, resistors and
 * independent voltage sources (DC or PULSE), used by tests/ and by the CPU
 * baseline of bench.py to check the engine's waveforms.
 *
 * Follows:
 *   src/spicelib/analysis/traninit.c:17-40   TMAX / delmin
 *   src/spicelib/analysis/dctran.c:66-971    DCtran (the XSPICE build, no
 *                                             event-driven instances)
 *   src/spicelib/analysis/cktop.c:26-720     CKTop: direct Newton, then
 *                                             dynamic_gmin / new_gmin /
 *                                             spice3_gmin and gillespie_src /
 *                                             spice3_src, then OPtran
 *   src/spicelib/analysis/optran.c:285-851   OPtran (defaults of init.c:
 *                                             77-93; no supply ramping)
 *   src/maths/ni/niiter.c:28-367             NIiter (INITF state machine,
 *                                             >= 2 iterations, maxIter >= 100)
 *   src/maths/ni/nicomcof.c:17-130           NIcomCof (trapezoidal, Gear)
 *   src/maths/ni/niconv.c:21-90              NIconvTest
 *   src/spicelib/analysis/cktload.c:31-180   CKTload incl. .ic rows in TRANOP
 *   src/spicelib/analysis/ckttrunc.c:20-56   CKTtrunc (min(2 dt, LTE))
 *   src/spicelib/analysis/cktsetbk.c, cktclrbk.c   breakpoint table
 *   src/spicelib/devices/cap/capload.c:19-86, captrunc.c:17-30
 *   src/spicelib/devices/res/resload.c:16-38
 *   src/spicelib/devices/vsrc/vsrcload.c:40-170 (PULSE), vsrcacct.c:20-160
 *   src/maths/misc/equality.c:22-60          AlmostEqualUlps
 *   src/spicelib/analysis/cktic.c:19-80, devices/bsim4/b4getic.c:32-55,
 *   devices/cap/capgetic.c:19-35, niiter.c:41-46   .tran ... uic (CKTic,
 *                                             DEVsetic, the single load)
 *
 * The linear solve is a dense LU with partial pivoting on the MNA matrix,
 * or the reference's own KLU (oracle/_ref, or_tran_use_klu) driven as
 * klusmp.c and niiter.c:107-195 drive it (fresh factorizations while
 * NISHOULDREORDER is set, refactors otherwise); above or_sparse_above
 * unknowns the matrix is held in CSC only.  or_klu_pivot restates
 * klu_factor's pivot choice (klu_kernel.c:370-530, klu_scale.c:19-159) on
 * its own, for checking against the reference's.  Device loads sum into the matrix in device-type order CAP,
 * BSIM4, RES, VSRC as CKTload does.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <dlfcn.h>
#include "bsim4_ref.h"

int or_slot_rc(int slot, int *row_role, int *col_role);

/* breakpoint table size (CKTsetBreak grows its table; one pending
 * breakpoint per PULSE source plus the run's ends fits here) */
#define MAXBRK 65536

typedef struct {
    int p, n, br;
    double dc;
    int pulse;          /* 1: PULSE(v1 v2 td tr tf pw per) */
    int order;          /* number of PULSE coefficients given */
    double c[7];
    double break_time;
} ORVsrc;

typedef struct {
    /* topology */
    int nnodes, nbr, n;       /* unknowns 1..n, 0 = ground */
    int nmos, ncap, nres, nvs;
    ORInst **mos;
    int *cap_p, *cap_n, *cap_q;
    double *cap_c;
    double *cap_ic;           /* CAPinitCond: the .ic difference under UIC (CAPgetic), else 0 */
    int *res_p, *res_n;
    double *res_g;
    ORVsrc *vs;
    int *ic_given;
    double *ic;
    /* options */
    double reltol, abstol, vntol, chgtol, trtol, gmin, temp;
    int maxOrder, method, uic;
    /* operating point: noopiter, gminsteps, srcsteps, gminfactor, gshunt,
     * itl2 (cktntask.c:94-132 defaults 0, 1, 1, 10, 0, 50) */
    int noOpIter, numGminSteps, numSrcSteps, dcTrcvMaxIter;
    /* optran's step size and final time (0: optran deselected) */
    double opstep, opfinal;
    double gminFactor, gshunt;
    /* CKTdiagGmin, CKTsrcFact; hasdiag[u]: A(u, u) exists */
    double diagGmin, srcFact;
    char *hasdiag;
    /* state */
    int nstates;
    double *stbuf;            /* OR_NSTATE x nstates */
    double *states[8];
    double *A;                /* (n+1) x (n+1), row 0 / col 0 are trash */
    double *rhs, *rhsOld;
    double *lu;
    int *piv;
    ORCkt ck;
    double time, delta, saveDelta, finalTime, step, maxStep, delmin, minBreak;
    double *breaks;
    int nbreaks;
    /* output */
    int64_t npts, cap_pts;
    double *out_t, *out_x;
    int64_t niter, accepted, rejected;
    int noncon_dc_fail;
    int iterno;
    int reorder;     /* NIiter's NISHOULDREORDER (KLU mode) */
    double pivrel;   /* CKTpivotRelTol */
    /* optional: the reference's KLU (oracle/_ref) instead of the dense LU */
    void *klu;
    int *Ap, *Ai;
    double *Ax, *kb;
    void *(*kg_new)(int, const int *, const int *);
    int (*kg_factor)(void *, double *, int, double);
    int (*kg_solve)(void *, double *);
    void (*kg_free)(void *);
    int (*kg_order)(void *, int *, int *, int *);
    /* optional: a drop-in BSIM4 load and linear solver behind the same
     * DCtran / NIiter (the engine's per-circuit cores, through function
     * pointers the caller supplies; the oracle links nothing of it) */
    /* sparse mode (large circuits): the matrix lives only in Ax, addressed
     * by a hash of (row, column); ground rows/columns go to a trash cell */
    int sparse;
    int64_t *hkey;
    int *hval;
    int64_t hsize;
    double trash, miss;
    void *dl_ctx;
    int (*dl_run)(void *, const void *, const double *, double *const *, double *, double *, int *);
    void *smp_ctx;
    int (*smp_factor)(void *, const double *, double);
    int (*smp_solve)(void *, double *);
} ORTran;

/* the scalar block the drop-in load receives (the layout of the engine's
 * ngmi_ckt_scalars: CKTmode, order, method, ag[7], delta / deltaOld[1],
 * gmin, reltol, abstol) */
typedef struct {
    int mode, order, method;
    double ag[7];
    double xfact;
    double gmin, reltol, abstol;
} ORDropInScalars;

static int build_csc(ORTran *T);

/* circuits with more unknowns than this keep their matrix sparse */
static int or_sparse_above = 3000;

/* set the sparse-mode threshold for circuits set up afterwards; returns the
 * previous value */
int or_set_sparse_above(int n)
{
    const int old = or_sparse_above;
    or_sparse_above = n;
    return old;
}

/* equality.c:22-60 */
static int almost_equal_ulps(double A, double B, int maxUlps)
{
    int64_t aInt, bInt, intDiff;
    union { double d; int64_t i; } uA, uB;
    if (A == B) return 1;
    uA.d = A;
    aInt = uA.i;
    if (aInt < 0) aInt = INT64_MIN - aInt;
    uB.d = B;
    bInt = uB.i;
    if (bInt < 0) bInt = INT64_MIN - bInt;
    intDiff = llabs(aInt - bInt);
    return intDiff <= maxUlps;
}

/* cktsetbk.c:19-100 */
static int set_break(ORTran *T, double time)
{
    int i, j;
    if (almost_equal_ulps(time, T->time, 3)) return 0;
    if (T->time > time) return -1;
    for (i = 0; i < T->nbreaks; i++) {
        if (T->breaks[i] > time) {
            if ((T->breaks[i] - time) <= T->minBreak) {
                T->breaks[i] = time;
                return 0;
            }
            if (i > 0 && time - T->breaks[i - 1] <= T->minBreak) return 0;
            if (T->nbreaks >= MAXBRK) return -2;
            for (j = T->nbreaks; j > i; j--) T->breaks[j] = T->breaks[j - 1];
            T->breaks[i] = time;
            T->nbreaks++;
            return 0;
        }
    }
    if (T->nbreaks && time - T->breaks[T->nbreaks - 1] <= T->minBreak) return 0;
    if (T->nbreaks >= MAXBRK) return -2;
    T->breaks[T->nbreaks++] = time;
    return 0;
}

/* cktclrbk.c:15-40 */
static void clr_break(ORTran *T)
{
    int j;
    if (T->nbreaks > 2) {
        for (j = 1; j < T->nbreaks; j++) T->breaks[j - 1] = T->breaks[j];
        T->nbreaks--;
    } else {
        T->breaks[0] = T->breaks[1];
        T->breaks[1] = T->finalTime;
    }
}

/* PULSE parameter defaults, vsrcload.c:96-125 */
static void pulse_params(const ORTran *T, const ORVsrc *v, double *V1, double *V2, double *TD,
                         double *TR, double *TF, double *PW, double *PER)
{
    *V1 = v->c[0];
    *V2 = v->c[1];
    *TD = v->order > 2 ? v->c[2] : 0.0;
    *TR = v->order > 3 && v->c[3] > 0.0 ? v->c[3] : T->step;
    *TF = v->order > 4 && v->c[4] > 0.0 ? v->c[4] : T->step;
    *PW = v->order > 5 && v->c[5] >= 0.0 ? v->c[5] : 0.0;
    *PER = v->order > 6 && v->c[6] > 0.0 ? v->c[6] : *TR + *TF + *PW;
}

/* vsrcload.c:78-170 for DC and PULSE */
static double vsrc_value(const ORTran *T, const ORVsrc *v, int mode)
{
    double time, V1, V2, TD, TR, TF, PW, PER, basetime;
    if ((mode & (OR_MODEDCOP | OR_MODEDCTRANCURVE)) && !v->pulse) return v->dc;
    if (!v->pulse) return v->dc;
    time = (mode & (OR_MODEDCOP | OR_MODETRANOP | OR_MODEDCTRANCURVE)) ? 0.0 : T->time;
    pulse_params(T, v, &V1, &V2, &TD, &TR, &TF, &PW, &PER);
    time -= TD;
    if (time > PER) {
        basetime = PER * floor(time / PER);
        time -= basetime;
    }
    if (time <= 0 || time >= TR + PW + TF) return V1;
    if (time >= TR && time <= TR + PW) return V2;
    if (time > 0 && time < TR) return V1 + (V2 - V1) * (time) / TR;
    return V2 + (V1 - V2) * (time - (TR + PW)) / TF;
}

/* vsrcacct.c:20-160 for PULSE: next corner as a breakpoint */
static int vsrc_accept(ORTran *T, int mode)
{
    int k;
    if (!(mode & (OR_MODETRAN | OR_MODETRANOP))) return 0;
    for (k = 0; k < T->nvs; k++) {
        ORVsrc *v = &T->vs[k];
        double V1, V2, TD, TR, TF, PW, PER, time, basetime, wait, atime;
        if (!v->pulse) continue;
        pulse_params(T, v, &V1, &V2, &TD, &TR, &TF, &PW, &PER);
        time = T->time - TD;
        if (T->time >= v->break_time) {
            if (time >= PER) {
                basetime = PER * floor(time / PER);
                time -= basetime;
            }
            atime = time + T->minBreak;
            if (atime < 0.0) wait = -time;
            else if (atime < TR) wait = TR - time;
            else if (atime < TR + PW) wait = TR + PW - time;
            else if (atime < TR + PW + TF) wait = TR + PW + TF - time;
            else wait = PER - time;
            v->break_time = T->time + wait;
            if (set_break(T, v->break_time)) return -1;
            v->break_time -= T->minBreak;
        }
    }
    return 0;
}

static double *ax_sparse(ORTran *T, int r, int c)
{
    int64_t key, h;
    if (r == 0 || c == 0) return &T->trash;
    key = (int64_t)r * (T->n + 1) + c;
    h = (key * 0x9E3779B97F4A7C15ULL) >> 11;
    for (h %= T->hsize;; h = (h + 1) % T->hsize) {
        if (T->hkey[h] == key) return &T->Ax[T->hval[h]];
        if (T->hkey[h] < 0) break;
    }
    T->miss = 0.0;   /* an element the pattern does not hold reads as 0 */
    return &T->miss;
}

#define AX(r, c) (*(T->sparse ? ax_sparse(T, (r), (c)) : &T->A[(size_t)(r) * (T->n + 1) + (c)]))

/* CKTload (cktload.c:31-180): device types in registry order BSIM4, CAP,
 * RES, VSRC (dev.c:142-206), instances newest first as the instance lists
 * are built by prepending (cktcrte.c:62-64), then the .ic rows */
static int ckt_load(ORTran *T)
{
    ORCkt *ck = &T->ck;
    const int mode = ck->mode, n = T->n;
    int k, s;
    if (T->sparse) memset(T->Ax, 0, sizeof(double) * (size_t)T->Ap[n]);
    else memset(T->A, 0, sizeof(double) * (size_t)(n + 1) * (n + 1));
    memset(T->rhs, 0, sizeof(double) * (size_t)(n + 1));
    ck->rhsOld = T->rhsOld;
    if (T->dl_run) {
        /* DEVload through the drop-in: same place in CKTload, same vectors */
        ORDropInScalars sc;
        double *st[3];
        int nc = 0, i;
        sc.mode = mode; sc.order = ck->order; sc.method = ck->method;
        for (i = 0; i < 7; i++) sc.ag[i] = ck->ag[i];
        sc.xfact = ck->deltaOld[1] != 0.0 ? ck->delta / ck->deltaOld[1] : 0.0;
        sc.gmin = ck->gmin; sc.reltol = ck->reltol; sc.abstol = ck->abstol;
        st[0] = ck->states[0]; st[1] = ck->states[1]; st[2] = ck->states[2];
        if (T->sparse) return -1;   /* the drop-in adds at dense positions */
        if (T->dl_run(T->dl_ctx, &sc, T->rhsOld, st, T->A, T->rhs, &nc)) return -1;
        ck->noncon += nc;
    }
    /* b4ld.c:69-5433 */
    for (k = T->dl_run ? -1 : T->nmos - 1; k >= 0; k--) {   /* instances load newest first */
        ORInst *h = T->mos[k];
        if (or_bsim4_load(h, ck)) return -1;
        for (s = 0; s < P__COUNT; s++) {
            int rr, cc, r, c;
            or_slot_rc(s, &rr, &cc);
            r = h->node[rr];
            c = h->node[cc];
            if (r && c) AX(r, c) += h->stamp[s];
        }
        for (s = 0; s < N__COUNT; s++) {
            int r = h->node[s];
            if (r) T->rhs[r] += h->rhs[s];
        }
    }
    /* capload.c:19-86 */
    if (mode & (OR_MODETRAN | OR_MODEAC | OR_MODETRANOP)) {
        const int cond1 = ((mode & (OR_MODEDCOP | OR_MODETRANOP | OR_MODEDCTRANCURVE)) &&
                           (mode & OR_MODEINITJCT)) ||
                          ((mode & OR_MODEUIC) && (mode & OR_MODEINITTRAN));
        for (k = T->ncap - 1; k >= 0; k--) {
            const int p = T->cap_p[k], q = T->cap_n[k], qc = T->cap_q[k];
            double vcap, geq, ceq;
            vcap = cond1 ? T->cap_ic[k] : T->rhsOld[p] - T->rhsOld[q];
            if (mode & (OR_MODETRAN | OR_MODEAC)) {
                if (mode & OR_MODEINITPRED) {
                    ck->states[0][qc] = ck->states[1][qc];
                } else {
                    ck->states[0][qc] = T->cap_c[k] * vcap;
                    if (mode & OR_MODEINITTRAN) ck->states[1][qc] = ck->states[0][qc];
                }
                if (or_integrate(ck, &geq, &ceq, T->cap_c[k], qc)) return -1;
                if (mode & OR_MODEINITTRAN) ck->states[1][qc + 1] = ck->states[0][qc + 1];
                AX(p, p) += geq;
                AX(q, q) += geq;
                AX(p, q) -= geq;
                AX(q, p) -= geq;
                T->rhs[p] -= ceq;
                T->rhs[q] += ceq;
            } else {
                ck->states[0][qc] = T->cap_c[k] * vcap;
            }
        }
    }
    /* resload.c:16-38 */
    for (k = T->nres - 1; k >= 0; k--) {
        const int p = T->res_p[k], q = T->res_n[k];
        const double g = T->res_g[k];
        AX(p, p) += g;
        AX(q, q) += g;
        AX(p, q) -= g;
        AX(q, p) -= g;
    }
    /* vsrcload.c:40-60 + value */
    for (k = T->nvs - 1; k >= 0; k--) {
        const ORVsrc *v = &T->vs[k];
        AX(v->p, v->br) += 1.0;
        AX(v->n, v->br) -= 1.0;
        AX(v->br, v->p) += 1.0;
        AX(v->br, v->n) -= 1.0;
        {   /* vsrcload.c:453-457: TRANOP scales by CKTsrcFact */
            double value = vsrc_value(T, v, mode);
            if (mode & OR_MODETRANOP) value *= T->srcFact;
            T->rhs[v->br] += value;
        }
    }
    /* .ic rows in the transient operating point (cktload.c:146-173) */
    if ((mode & (OR_MODEDCOP | OR_MODETRANOP | OR_MODEDCTRANCURVE)) &&
        (mode & OR_MODETRANOP) && !(mode & OR_MODEUIC)) {
        int node;
        for (node = 1; node <= T->nnodes; node++) {
            int c, currents = 0;
            if (!T->ic_given[node]) continue;
            for (c = 1; c <= n; c++) {
                if (c > T->nnodes) {
                    if (AX(node, c) != 0.0) currents = 1;
                } else {
                    AX(node, c) = 0.0;
                }
            }
            if (currents) {
                T->rhs[node] = 1.0e10 * T->ic[node] * T->srcFact;
                AX(node, node) += 1.0e10;
            } else {
                T->rhs[node] = T->ic[node] * T->srcFact;   /* cktload.c:160 */
                AX(node, node) = 1.0;
            }
        }
    }
    return 0;
}

/* the reference's KLU on the structural pattern.  NIiter factors afresh
 * (klu_factor) while NISHOULDREORDER is set -- it is set by NIreinit, on
 * INITJCT and the first INITTRAN iteration, after either (niiter.c:335-347)
 * and by a failed factorization -- and clears it after a successful one;
 * otherwise it refactors, and a refactor that meets a zero pivot is
 * factored afresh in the same iteration with the flag left set
 * (niiter.c:107-195) */
static int solve_klu(ORTran *T)
{
    const int n = T->n;
    int c, p, full;
    if (!T->sparse)
        for (c = 0; c < n; c++)
            for (p = T->Ap[c]; p < T->Ap[c + 1]; p++) T->Ax[p] = AX(T->Ai[p] + 1, c + 1);
    if ((T->ck.mode & OR_MODEINITJCT) || ((T->ck.mode & OR_MODEINITTRAN) && T->iterno == 1)) T->reorder = 1;
    full = T->reorder;
    c = T->kg_factor(T->klu, T->Ax, full, T->pivrel);
    if (c == 0 || c == 2) T->reorder = (c == 2);
    if (c == 2) c = 0;
    if (c) {
        T->reorder = 1;
        return c;
    }
    for (c = 0; c < n; c++) T->kb[c] = T->rhs[c + 1];
    if (T->kg_solve(T->klu, T->kb)) return 1;
    for (c = 0; c < n; c++) T->rhs[c + 1] = T->kb[c];
    T->rhs[0] = 0.0;
    return 0;
}

/* dense LU with partial pivoting on rows/cols 1..n; solution into rhs */
static int solve(ORTran *T)
{
    if (T->smp_factor) {
        /* SMPluFac / SMPsolve through the drop-in: it adds CKTdiagGmin to
         * the diagonal itself (LoadGmin_CSC) */
        const int n = T->n;
        int c, p, rc;
        if (!T->sparse)
            for (c = 0; c < n; c++)
                for (p = T->Ap[c]; p < T->Ap[c + 1]; p++) T->Ax[p] = AX(T->Ai[p] + 1, c + 1);
        rc = T->smp_factor(T->smp_ctx, T->Ax, T->diagGmin);
        if (rc) return rc;
        if (T->smp_solve(T->smp_ctx, T->rhs)) return 1;
        T->rhs[0] = 0.0;
        return 0;
    }
    /* LoadGmin_CSC (klusmp.c:1765-1777): CKTdiagGmin onto every diagonal
     * element the matrix has, before the (re)factorisation */
    if (T->diagGmin != 0.0) {
        int u;
        for (u = 1; u <= T->n; u++)
            if (T->hasdiag[u]) AX(u, u) += T->diagGmin;
    }
    if (T->klu) return solve_klu(T);
    if (T->sparse) return 1;   /* sparse mode solves with the reference KLU only */
    const int n = T->n, ld = n + 1;
    double *a = T->lu;
    int i, j, k;
    memcpy(a, T->A, sizeof(double) * (size_t)ld * ld);
    for (k = 1; k <= n; k++) {
        int p = k;
        double mx = fabs(a[(size_t)k * ld + k]);
        for (i = k + 1; i <= n; i++) {
            double v = fabs(a[(size_t)i * ld + k]);
            if (v > mx) { mx = v; p = i; }
        }
        if (mx == 0.0) return 102;
        T->piv[k] = p;
        if (p != k) {
            for (j = 0; j <= n; j++) {
                double t = a[(size_t)k * ld + j];
                a[(size_t)k * ld + j] = a[(size_t)p * ld + j];
                a[(size_t)p * ld + j] = t;
            }
            { double t = T->rhs[k]; T->rhs[k] = T->rhs[p]; T->rhs[p] = t; }
        }
        for (i = k + 1; i <= n; i++) {
            double l = a[(size_t)i * ld + k] / a[(size_t)k * ld + k];
            if (l == 0.0) continue;
            a[(size_t)i * ld + k] = l;
            for (j = k + 1; j <= n; j++) a[(size_t)i * ld + j] -= l * a[(size_t)k * ld + j];
            T->rhs[i] -= l * T->rhs[k];
        }
    }
    for (k = n; k >= 1; k--) {
        double s = T->rhs[k];
        for (j = k + 1; j <= n; j++) s -= a[(size_t)k * ld + j] * T->rhs[j];
        T->rhs[k] = s / a[(size_t)k * ld + k];
    }
    T->rhs[0] = 0.0;
    return 0;
}

/* niconv.c:21-90 */
static int conv_test(const ORTran *T)
{
    int i;
    for (i = 1; i <= T->n; i++) {
        const double nw = T->rhs[i], old = T->rhsOld[i];
        double tol;
        if (isnan(nw)) return 1;
        if (i <= T->nnodes) tol = T->reltol * fmax(fabs(old), fabs(nw)) + T->vntol;
        else tol = T->reltol * fmax(fabs(old), fabs(nw)) + T->abstol;
        if (fabs(nw - old) > tol) return 1;
    }
    return 0;
}

/* niiter.c:28-367 (no nodesets, no node damping) */
static int ni_iter(ORTran *T, int maxIter)
{
    ORCkt *ck = &T->ck;
    int iterno = 0;
    if (maxIter < 100) maxIter = 100;
    if ((ck->mode & OR_MODETRANOP) && (ck->mode & OR_MODEUIC)) {
        /* niiter.c:41-46: under UIC the operating point is one load on the
         * initial conditions, no solve */
        double *t = T->rhs; T->rhs = T->rhsOld; T->rhsOld = t;
        return ckt_load(T) ? 1 : 0;
    }
    for (;;) {
        int rc;
        ck->noncon = 0;
        if (ckt_load(T)) return 1;
        iterno++;
        T->iterno = iterno;
        rc = solve(T);
        T->niter++;
        if (rc) return rc;
        if (iterno > maxIter) return 103;
        if (ck->noncon == 0 && iterno != 1) ck->noncon = conv_test(T);
        else ck->noncon = 1;
        if (ck->mode & OR_MODEINITFLOAT) {
            /* converged: CKTrhs holds the last solve, CKTrhsOld the iterate
             * before it (niiter.c:300-305 return before the swap); CKTdump
             * writes CKTrhsOld (cktdump.c:43) */
            if (ck->noncon == 0) return 0;
        } else if (ck->mode & OR_MODEINITJCT) {
            ck->mode = (ck->mode & ~0x3f00) | OR_MODEINITFIX;
            T->reorder = 1;
        } else if (ck->mode & OR_MODEINITFIX) {
            if (ck->noncon == 0) ck->mode = (ck->mode & ~0x3f00) | OR_MODEINITFLOAT;
        } else if (ck->mode & OR_MODEINITSMSIG) {
            ck->mode = (ck->mode & ~0x3f00) | OR_MODEINITFLOAT;
        } else if (ck->mode & OR_MODEINITTRAN) {
            if (iterno <= 1) T->reorder = 1;
            ck->mode = (ck->mode & ~0x3f00) | OR_MODEINITFLOAT;
        } else if (ck->mode & OR_MODEINITPRED) {
            ck->mode = (ck->mode & ~0x3f00) | OR_MODEINITFLOAT;
        } else {
            return 1;
        }
        { double *t = T->rhs; T->rhs = T->rhsOld; T->rhsOld = t; }
    }
}

/* ---- CKTop (cktop.c:26-116) with its stepping fallbacks ---- */

static void op_zero(ORTran *T)
{
    memset(T->rhsOld, 0, sizeof(double) * (T->n + 1));
    memset(T->ck.states[0], 0, sizeof(double) * T->nstates);
}

static void op_save(ORTran *T, double *x, double *s)
{
    memcpy(x, T->rhsOld, sizeof(double) * (T->n + 1));
    memcpy(s, T->ck.states[0], sizeof(double) * T->nstates);
}

static void op_restore(ORTran *T, const double *x, const double *s)
{
    memcpy(T->rhsOld, x, sizeof(double) * (T->n + 1));
    memcpy(T->ck.states[0], s, sizeof(double) * T->nstates);
}

/* NIiter and the iterations it used (STATnumIter difference) */
static int op_niiter(ORTran *T, int maxIter, int *iters)
{
    const int64_t n0 = T->niter;
    const int rc = ni_iter(T, maxIter);
    *iters = (int)(T->niter - n0);
    return rc;
}

/* dynamic_gmin (cktop.c:161-287): CKTdiagGmin from 1e-2 down to gmin */
static int dynamic_gmin(ORTran *T, int firstmode, int continuemode, int iterlim,
                        double *ox, double *os)
{
    ORCkt *ck = &T->ck;
    double OldGmin, gtarget, factor;
    int converged, iters;
    ck->mode = firstmode;
    op_zero(T);
    factor = T->gminFactor;
    OldGmin = 1e-2;
    T->diagGmin = OldGmin / factor;
    gtarget = fmax(T->gmin, T->gshunt);
    for (;;) {
        converged = op_niiter(T, T->dcTrcvMaxIter, &iters);
        if (converged == 0) {
            ck->mode = continuemode;
            if (T->diagGmin <= gtarget) break;
            op_save(T, ox, os);
            if (iters <= (T->dcTrcvMaxIter / 4)) {
                factor *= sqrt(factor);
                if (factor > T->gminFactor) factor = T->gminFactor;
            }
            if (iters > (3 * T->dcTrcvMaxIter / 4)) factor = fmax(sqrt(factor), 1.00005);
            OldGmin = T->diagGmin;
            if (T->diagGmin < factor * gtarget) {
                factor = T->diagGmin / gtarget;
                T->diagGmin = gtarget;
            } else {
                T->diagGmin /= factor;
            }
        } else {
            if (factor < 1.00005) break;
            factor = sqrt(sqrt(factor));
            T->diagGmin = OldGmin / factor;
            op_restore(T, ox, os);
        }
    }
    T->diagGmin = T->gshunt;
    return ni_iter(T, iterlim);
}

/* new_gmin (cktop.c:349-478): the devices' CKTgmin from 1e-2 down */
static int new_gmin(ORTran *T, int firstmode, int continuemode, int iterlim, double *ox, double *os)
{
    ORCkt *ck = &T->ck;
    double OldGmin, gtarget, factor, startgmin;
    int converged, iters;
    ck->mode = firstmode;
    op_zero(T);
    startgmin = ck->gmin;
    factor = T->gminFactor;
    OldGmin = 1e-2;
    ck->gmin = OldGmin / factor;
    gtarget = fmax(startgmin, T->gshunt);
    for (;;) {
        converged = op_niiter(T, T->dcTrcvMaxIter, &iters);
        if (converged == 0) {
            ck->mode = continuemode;
            if (ck->gmin <= gtarget) break;
            op_save(T, ox, os);
            if (iters <= (T->dcTrcvMaxIter / 4)) {
                factor *= sqrt(factor);
                if (factor > T->gminFactor) factor = T->gminFactor;
            }
            if (iters > (3 * T->dcTrcvMaxIter / 4)) factor = fmax(sqrt(factor), 3);
            OldGmin = ck->gmin;
            if (ck->gmin < factor * gtarget) {
                factor = ck->gmin / gtarget;
                ck->gmin = gtarget;
            } else {
                ck->gmin /= factor;
            }
        } else {
            if (factor < 1.00005) break;
            factor = sqrt(sqrt(factor));
            ck->gmin = OldGmin / factor;
            op_restore(T, ox, os);
        }
    }
    ck->gmin = fmax(startgmin, T->gshunt);
    return ni_iter(T, iterlim);
}

/* gillespie_src (cktop.c:480-670): sources ramped from 0 */
static int gillespie_src(ORTran *T, int firstmode, int continuemode, double *ox, double *os)
{
    ORCkt *ck = &T->ck;
    int converged, i, iters;
    double ConvFact = 0;
    const double gminstart = ck->gmin;
    ck->mode = firstmode;
    T->srcFact = 0;
    op_zero(T);
    converged = op_niiter(T, T->dcTrcvMaxIter, &iters);
    if (converged != 0) {
        T->diagGmin = (T->gshunt <= 0) ? ck->gmin : T->gshunt;
        for (i = 0; i < 10; i++) T->diagGmin *= 10;
        for (i = 0; i <= 10; i++) {
            converged = op_niiter(T, T->dcTrcvMaxIter, &iters);
            if (converged != 0) {
                T->diagGmin = T->gshunt;
                break;
            }
            T->diagGmin /= 10;
            ck->mode = continuemode;
        }
        T->diagGmin = T->gshunt;
    }
    if (converged == 0) {
        double raise = 0.001;
        op_save(T, ox, os);
        T->srcFact = ConvFact + raise;
        do {
            converged = op_niiter(T, T->dcTrcvMaxIter, &iters);
            ck->mode = continuemode;
            if (converged == 0) {
                ConvFact = T->srcFact;
                op_save(T, ox, os);
                T->srcFact = ConvFact + raise;
                if (iters <= (T->dcTrcvMaxIter / 4)) raise *= 1.5;
                if (iters > (3 * T->dcTrcvMaxIter / 4)) raise *= 0.5;
            } else {
                if (T->srcFact - ConvFact < 1e-8) break;
                raise /= 10;
                if (raise > 0.01) raise = 0.01;
                T->srcFact = ConvFact;
                op_restore(T, ox, os);
            }
            if (T->srcFact > 1) T->srcFact = 1;
        } while ((raise >= 1e-7) && (ConvFact < 1));
        T->diagGmin = ck->gmin = gminstart;
    }
    T->srcFact = 1;
    return (ConvFact != 1) ? 103 : 0;
}

/* spice3_gmin (cktop.c:289-347): CKTdiagGmin from gmin * gminFactor^steps
 * down by gminFactor per converged step, then a plain NIiter */
static int spice3_gmin(ORTran *T, int firstmode, int continuemode, int iterlim)
{
    ORCkt *ck = &T->ck;
    int converged, i;
    ck->mode = firstmode;
    T->diagGmin = (T->gshunt == 0) ? ck->gmin : T->gshunt;
    for (i = 0; i < T->numGminSteps; i++) T->diagGmin *= T->gminFactor;
    for (i = 0; i <= T->numGminSteps; i++) {
        converged = ni_iter(T, T->dcTrcvMaxIter);
        if (converged != 0) {
            T->diagGmin = T->gshunt;
            break;
        }
        T->diagGmin /= T->gminFactor;
        ck->mode = continuemode;
    }
    T->diagGmin = T->gshunt;
    return ni_iter(T, iterlim);
}

/* spice3_src (cktop.c:672-720): CKTsrcFact = i / steps for i = 0..steps */
static int spice3_src(ORTran *T, int firstmode, int continuemode)
{
    ORCkt *ck = &T->ck;
    int converged, i;
    ck->mode = firstmode;
    for (i = 0; i <= T->numSrcSteps; i++) {
        T->srcFact = ((double)i) / ((double)T->numSrcSteps);
        converged = ni_iter(T, T->dcTrcvMaxIter);
        ck->mode = continuemode;
        if (converged != 0) {
            T->srcFact = 1;
            return converged;
        }
    }
    T->srcFact = 1;
    return 0;
}

static int op_tran(ORTran *T, int oldconverged);

static int ckt_op(ORTran *T, int firstmode, int continuemode, int iterlim)
{
    int converged = 1;
    double *ox = (double *)malloc(sizeof(double) * (T->n + 1));
    double *os = (double *)malloc(sizeof(double) * (T->nstates + 1));
    T->ck.mode = firstmode;
    /* under UIC noopiter is not taken: the stepping's NIiter calls would
     * each swap the previous load's right-hand side in as the iterate */
    if (!T->noOpIter || (firstmode & OR_MODEUIC)) converged = ni_iter(T, iterlim);
    /* cktop.c:54-91 (the frontend's dyngmin variable unset) */
    if (converged != 0 && T->numGminSteps >= 1) {
        if (T->numGminSteps == 1) {
            converged = dynamic_gmin(T, firstmode, continuemode, iterlim, ox, os);
            if (converged != 0) converged = new_gmin(T, firstmode, continuemode, iterlim, ox, os);
        } else {
            converged = spice3_gmin(T, firstmode, continuemode, iterlim);
        }
    }
    if (converged != 0 && T->numSrcSteps >= 1) {
        if (T->numSrcSteps == 1) converged = gillespie_src(T, firstmode, continuemode, ox, os);
        else converged = spice3_src(T, firstmode, continuemode);
    }
    free(ox);
    free(os);
    /* cktop.c:93-105: the transient operating point, last */
    if (converged != 0) converged = op_tran(T, converged);
    return converged;
}

/* operating-point options (cktsopt.c): noopiter, gminsteps (0: off, 1:
 * dynamic then true gmin stepping, above 1: spice3 gmin stepping in that
 * many steps), srcsteps (0: off, 1: Gillespie source stepping, above 1:
 * spice3 source stepping), gminfactor, gshunt, itl2.  Returns 0. */
int or_tran_set_op_options(ORTran *T, int noopiter, int gminsteps, int srcsteps, double gminfactor,
                           double gshunt, int itl2)
{
    T->noOpIter = noopiter;
    T->numGminSteps = gminsteps;
    T->numSrcSteps = srcsteps;
    T->gminFactor = gminfactor;
    T->gshunt = gshunt;
    T->dcTrcvMaxIter = itl2;
    return 0;
}

/* nicomcof.c:17-130 */
static int com_cof(ORTran *T)
{
    ORCkt *ck = &T->ck;
    double mat[8][8], arg, arg1;
    int i, j, k;
    if (ck->method == OR_TRAPEZOIDAL) {
        if (ck->order == 1) {
            ck->ag[0] = 1 / ck->delta;
            ck->ag[1] = -1 / ck->delta;
        } else if (ck->order == 2) {
            const double xmu = 0.5;
            ck->ag[0] = 1.0 / ck->delta / (1.0 - xmu);
            ck->ag[1] = xmu / (1.0 - xmu);
        } else return -1;
        return 0;
    }
    if (ck->order < 1 || ck->order > 6) return -1;
    memset(ck->ag, 0, 7 * sizeof(double));
    ck->ag[1] = -1 / ck->delta;
    arg = 0;
    for (i = 0; i <= ck->order; i++) mat[0][i] = 1;
    for (i = 1; i <= ck->order; i++) mat[i][0] = 0;
    for (i = 1; i <= ck->order; i++) {
        arg += ck->deltaOld[i - 1];
        arg1 = 1;
        for (j = 1; j <= ck->order; j++) {
            arg1 *= arg / ck->delta;
            mat[j][i] = arg1;
        }
    }
    for (i = 1; i <= ck->order; i++)
        for (j = i + 1; j <= ck->order; j++) {
            mat[j][i] /= mat[i][i];
            for (k = i + 1; k <= ck->order; k++) mat[j][k] -= mat[j][i] * mat[i][k];
        }
    for (i = 1; i <= ck->order; i++)
        for (j = i + 1; j <= ck->order; j++) ck->ag[j] = ck->ag[j] - mat[j][i] * ck->ag[i];
    ck->ag[ck->order] /= mat[ck->order][ck->order];
    for (i = ck->order - 1; i >= 0; i--) {
        for (j = i + 1; j <= ck->order; j++) ck->ag[i] = ck->ag[i] - mat[i][j] * ck->ag[j];
        ck->ag[i] /= mat[i][i];
    }
    return 0;
}

/* NIcomCof alone: ag[0..6] for method (OR_TRAPEZOIDAL / OR_GEAR), order,
 * the step delta and deltaOld[0..6].  Returns 0, or -1 for an order the
 * method does not have. */
int or_comcof(int method, int order, double delta, const double *deltaOld, double *ag)
{
    ORTran *T = (ORTran *)calloc(1, sizeof(ORTran));
    int k, rc;
    if (!T) return -1;
    T->ck.method = method;
    T->ck.order = order;
    T->ck.delta = delta;
    for (k = 0; k < 7; k++) T->ck.deltaOld[k] = deltaOld[k];
    rc = com_cof(T);
    for (k = 0; k < 7; k++) ag[k] = T->ck.ag[k];
    free(T);
    return rc;
}

/* ckttrunc.c:20-56 with CAPtrunc and BSIM4trunc */
static void ckt_trunc(ORTran *T, double *timeStep)
{
    double timetemp = 1e308;
    int k;
    for (k = 0; k < T->ncap; k++) or_terr(T->cap_q[k], &T->ck, &timetemp, T->trtol, T->chgtol);
    for (k = 0; k < T->nmos; k++) or_bsim4_trunc(T->mos[k], &T->ck, &timetemp, T->trtol, T->chgtol);
    *timeStep = fmin(2 * *timeStep, timetemp);
}

static int dump(ORTran *T)
{
    if (T->npts == T->cap_pts) {
        int64_t nc = T->cap_pts ? 2 * T->cap_pts : 1024;
        double *t = (double *)realloc(T->out_t, sizeof(double) * nc);
        double *x = (double *)realloc(T->out_x, sizeof(double) * nc * (T->n + 1));
        if (!t || !x) return -1;
        T->out_t = t;
        T->out_x = x;
        T->cap_pts = nc;
    }
    T->out_t[T->npts] = T->time;
    memcpy(T->out_x + T->npts * (T->n + 1), T->rhsOld, sizeof(double) * (T->n + 1));
    T->npts++;
    return 0;
}

static void rotate_states(ORTran *T)
{
    double *temp = T->ck.states[T->maxOrder + 1];
    int i;
    for (i = T->maxOrder; i >= 0; i--) T->ck.states[i + 1] = T->ck.states[i];
    T->ck.states[0] = temp;
}

/* OPtran (optran.c:285-851, the XSPICE build with no event instances): the
 * operating point as a transient from the state the failed stepping left,
 * sources held at CKTtime (0: the transient has not begun), steps of
 * opstep (CKTstep = CKTmaxStep during it) up to opfinal on breakpoints of
 * its own, without output.  Its accepts are CKTaccept's (VSRCaccept at
 * CKTtime 0 into the transient's breakpoints).  Unlike DCtran's, its
 * resume block has no 1.9-step rule, and its loop that throws out passed
 * breakpoints has no final-time guard: with the table down to the final
 * time and the time within CKTminBreak of it, OPclrBreak changes nothing
 * and the reference would loop; here the loop ends.  Returns 0, 106
 * (timestep too small) or oldconverged when deselected. */
static int op_tran(ORTran *T, int oldconverged)
{
    ORCkt *ck = &T->ck;
    const size_t ns = (size_t)T->nstates;
    const int uic = ck->mode & OR_MODEUIC;
    const double prevmax = T->maxStep, prevstep = T->step;
    double optime = 0, ob[2], olddelta, newdelta;
    int firsttime = 1, converged, i;
    if (T->opstep == 0) return oldconverged;
    T->maxStep = T->step = T->opstep;
    T->delta = fmin(T->opfinal / 100, T->step) / 10;
    ob[0] = 0;
    ob[1] = T->opfinal;
    ck->order = 1;
    for (i = 0; i < 7; i++) ck->deltaOld[i] = T->maxStep;
    ck->delta = T->delta;
    T->saveDelta = T->opfinal / 50;
    ck->mode = uic | OR_MODETRAN | OR_MODEINITTRAN;
    ck->ag[0] = ck->ag[1] = 0;
    memcpy(ck->states[1], ck->states[0], ns * sizeof(double));
    for (;;) {
        /* nextTime: */
        if (vsrc_accept(T, ck->mode)) return 1;
        if (optime > ob[0]) {
            ob[0] = ob[1];
            ob[1] = T->opfinal;
        }
        if (almost_equal_ulps(optime, T->opfinal, 100)) {
            T->maxStep = prevmax;
            T->step = prevstep;
            return 0;
        }
        /* resume: */
        T->delta = fmin(T->delta, T->maxStep);
        if (almost_equal_ulps(optime, ob[0], 100) || ob[0] - optime <= T->delmin) {
            ck->order = 1;
            T->delta = fmin(T->delta, .1 * fmin(T->saveDelta, ob[1] - ob[0]));
            if (firsttime) T->delta /= 10;
        }
        while ((almost_equal_ulps(ob[0], optime, 100) || ob[0] <= optime + T->minBreak) &&
               !(ob[0] == T->opfinal && ob[1] == T->opfinal)) {
            ob[0] = ob[1];
            ob[1] = T->opfinal;
        }
        if (optime + T->delta > ob[0]) {
            T->saveDelta = T->delta;
            T->delta = ob[0] - optime;
        }
        for (i = 5; i >= 0; i--) ck->deltaOld[i + 1] = ck->deltaOld[i];
        ck->deltaOld[0] = T->delta;
        rotate_states(T);
        for (;;) {
            int accept = 0;
            olddelta = T->delta;
            optime += T->delta;
            ck->delta = T->delta;
            ck->deltaOld[0] = T->delta;
            if (com_cof(T)) return 1;
            converged = ni_iter(T, 10);
            ck->mode = uic | OR_MODETRAN | OR_MODEINITPRED;
            if (firsttime) {
                memcpy(ck->states[2], ck->states[1], ns * sizeof(double));
                memcpy(ck->states[3], ck->states[1], ns * sizeof(double));
            }
            if (converged != 0) {
                optime -= T->delta;
                T->delta = T->delta / 8;
                if (firsttime) ck->mode = uic | OR_MODETRAN | OR_MODEINITTRAN;
                ck->order = 1;
            } else if (firsttime) {
                firsttime = 0;
                accept = 1;
            } else {
                newdelta = T->delta;
                ckt_trunc(T, &newdelta);
                if (newdelta > .9 * T->delta) {
                    if (ck->order == 1 && T->maxOrder > 1) {
                        newdelta = T->delta;
                        ck->order = 2;
                        ckt_trunc(T, &newdelta);
                        if (newdelta <= 1.05 * T->delta) ck->order = 1;
                    }
                    T->delta = newdelta;
                    accept = 1;
                } else {
                    optime -= T->delta;
                    T->delta = newdelta;
                }
            }
            if (accept) break;
            if (T->delta <= T->delmin) {
                if (olddelta > T->delmin) T->delta = T->delmin;
                else return 106;
            }
        }
    }
}

/* optran's step size and final time (com_optran, optran.c:75-190: a step
 * above final / 50 becomes final / 50; step 0 deselects it).  Defaults
 * 100n and 10u, ngspice's own (init.c:77-93).  Returns 0, or -1 when the
 * step exceeds the final time. */
int or_tran_set_optran(ORTran *T, double opstep, double opfinal)
{
    if (opstep < 0 || opfinal <= 0 || opstep > opfinal) return -1;
    if (opstep > opfinal / 50.) opstep = opfinal / 50.;
    T->opstep = opstep;
    T->opfinal = opfinal;
    return 0;
}

/* traninit.c + dctran.c (XSPICE build, no event instances) */
/* CKTic (cktic.c:19-80) with DEVsetic under UIC: CKTrhs = 0 but for the
 * .ic nodes, where CKTrhs = CKTrhsOld = the value; BSIM4getic
 * (b4getic.c:32-55) takes an instance's icvds/icvgs/icvbs not given from
 * its terminals' CKTrhs, CAPgetic (capgetic.c:19-35) each capacitor's
 * initial voltage */
static void or_uic_setic(ORTran *T)
{
    int u, k;
    memset(T->rhs, 0, sizeof(double) * (T->n + 1));
    for (u = 1; u <= T->nnodes; u++)
        if (T->ic_given[u]) T->rhs[u] = T->rhsOld[u] = T->ic[u];
    for (k = 0; k < T->nmos; k++) {
        ORInst *h = T->mos[k];
        const double vs = T->rhs[h->node[N_s]];
        if (!h->igiven[BI_icVDS]) h->ip[BI_icVDS] = T->rhs[h->node[N_d]] - vs;
        if (!h->igiven[BI_icVGS]) h->ip[BI_icVGS] = T->rhs[h->node[N_g]] - vs;
        if (!h->igiven[BI_icVBS]) h->ip[BI_icVBS] = T->rhs[h->node[N_b]] - vs;
    }
    for (k = 0; k < T->ncap; k++) T->cap_ic[k] = T->rhs[T->cap_p[k]] - T->rhs[T->cap_n[k]];
}

int or_tran_run(ORTran *T, double tstep, double tstop, double tmax)
{
    ORCkt *ck = &T->ck;
    double delta, olddelta, newdelta;
    int firsttime, converged, i;
    const size_t ns = (size_t)T->nstates;
    T->finalTime = tstop;
    T->step = tstep;
    T->maxStep = tmax;
    if (T->maxStep == 0) {
        if (T->step < T->finalTime / 50.0) T->maxStep = T->step;
        else T->maxStep = T->finalTime / 50.0;
    }
    T->delmin = 1e-11 * T->maxStep;
    delta = fmin(T->finalTime / 100, T->step) / 10;
    T->breaks[0] = 0;
    T->breaks[1] = T->finalTime;
    T->nbreaks = 2;
    T->minBreak = 10.0 * T->delmin;
    T->time = 0;
    ck->delta = 0;
    firsttime = 1;
    ck->method = T->method;
    ck->order = 1;
    for (i = 0; i < T->nvs; i++) T->vs[i].break_time = -1.0;
    memset(T->rhsOld, 0, sizeof(double) * (T->n + 1));
    if (T->uic) or_uic_setic(T);
    converged = ckt_op(T, (T->uic ? OR_MODEUIC : 0) | OR_MODETRANOP | OR_MODEINITJCT,
                       (T->uic ? OR_MODEUIC : 0) | OR_MODETRANOP | OR_MODEINITFLOAT, 100);
    if (converged) { T->noncon_dc_fail = 1; return converged; }
    ck->order = 1;
    for (i = 0; i < 7; i++) ck->deltaOld[i] = T->maxStep;
    ck->delta = delta;
    T->delta = delta;
    T->saveDelta = T->finalTime / 50;
    ck->mode = (T->uic ? OR_MODEUIC : 0) | OR_MODETRAN | OR_MODEINITTRAN;
    ck->ag[0] = ck->ag[1] = 0;
    memcpy(ck->states[1], ck->states[0], ns * sizeof(double));
    for (;;) {
        /* nextTime: */
        if (vsrc_accept(T, ck->mode)) return 1;
        if (T->time > T->breaks[0]) clr_break(T);
        T->accepted++;
        /* dctran.c:416-419: no time-0 point under UIC */
        if (!(T->uic && T->time <= 0) && dump(T)) return 8;
        if (almost_equal_ulps(T->time, T->finalTime, 100)) return 0;
        /* resume: */
        T->delta = fmin(T->delta, T->maxStep);
        if (almost_equal_ulps(T->time, T->breaks[0], 100) || T->breaks[0] - T->time <= T->delmin) {
            ck->order = 1;
            T->delta = fmin(T->delta, .1 * fmin(T->saveDelta, T->breaks[1] - T->breaks[0]));
            if (firsttime) {
                /* dctran.c:509-512 */
                if (T->uic && set_break(T, T->step)) return 1;
                T->delta /= 10;
            }
        }
        while ((T->breaks[0] <= T->time + T->minBreak ||
                almost_equal_ulps(T->breaks[0], T->time, 100)) &&
               T->breaks[0] < T->finalTime)
            clr_break(T);
        if (T->time + T->delta > T->breaks[0]) {
            T->saveDelta = T->delta;
            T->delta = T->breaks[0] - T->time;
        } else if (T->time + 1.9 * T->delta > T->breaks[0]) {
            T->saveDelta = T->delta;
            T->delta = (T->breaks[0] - T->time) / 2.;
        }
        for (i = 5; i >= 0; i--) ck->deltaOld[i + 1] = ck->deltaOld[i];
        ck->deltaOld[0] = T->delta;
        rotate_states(T);
        for (;;) {
            int accept = 0;
            olddelta = T->delta;
            T->time += T->delta;
            ck->delta = T->delta;
            ck->deltaOld[0] = T->delta;
            if (com_cof(T)) return 1;
            converged = ni_iter(T, 10);
            ck->mode = (T->uic ? OR_MODEUIC : 0) | OR_MODETRAN | OR_MODEINITPRED;
            if (firsttime) {
                memcpy(ck->states[2], ck->states[1], ns * sizeof(double));
                memcpy(ck->states[3], ck->states[1], ns * sizeof(double));
            }
            if (converged != 0) {
                T->time = T->time - T->delta;
                T->rejected++;
                T->delta = T->delta / 8;
                if (firsttime) ck->mode = (T->uic ? OR_MODEUIC : 0) | OR_MODETRAN | OR_MODEINITTRAN;
                ck->order = 1;
            } else {
                if (firsttime) {
                    firsttime = 0;
                    accept = 1;
                } else {
                    newdelta = T->delta;
                    ckt_trunc(T, &newdelta);
                    if (newdelta > .9 * T->delta) {
                        if (ck->order == 1 && T->maxOrder > 1) {
                            newdelta = T->delta;
                            ck->order = 2;
                            ckt_trunc(T, &newdelta);
                            if (newdelta <= 1.05 * T->delta) ck->order = 1;
                        }
                        T->delta = newdelta;
                        accept = 1;
                    } else {
                        T->time = T->time - T->delta;
                        T->rejected++;
                        T->delta = newdelta;
                    }
                }
            }
            if (accept) break;
            if (T->delta <= T->delmin) {
                if (olddelta > T->delmin) T->delta = T->delmin;
                else return 106;
            }
        }
    }
}

/* ---- construction API (ctypes) ---- */

ORTran *or_tran_new(int nnodes, int nmos, int ncap, int nres, int nvs)
{
    ORTran *T = (ORTran *)calloc(1, sizeof(ORTran));
    if (!T) return NULL;
    T->nnodes = nnodes;
    T->mos = (ORInst **)calloc((size_t)nmos + 1, sizeof(ORInst *));
    T->cap_p = (int *)calloc((size_t)ncap + 1, sizeof(int));
    T->cap_n = (int *)calloc((size_t)ncap + 1, sizeof(int));
    T->cap_q = (int *)calloc((size_t)ncap + 1, sizeof(int));
    T->cap_c = (double *)calloc((size_t)ncap + 1, sizeof(double));
    T->cap_ic = (double *)calloc((size_t)ncap + 1, sizeof(double));
    T->res_p = (int *)calloc((size_t)nres + 1, sizeof(int));
    T->res_n = (int *)calloc((size_t)nres + 1, sizeof(int));
    T->res_g = (double *)calloc((size_t)nres + 1, sizeof(double));
    T->vs = (ORVsrc *)calloc((size_t)nvs + 1, sizeof(ORVsrc));
    T->ic_given = (int *)calloc((size_t)nnodes + 1, sizeof(int));
    T->ic = (double *)calloc((size_t)nnodes + 1, sizeof(double));
    T->reltol = 1e-3; T->abstol = 1e-12; T->vntol = 1e-6; T->chgtol = 1e-14;
    T->trtol = 7; T->gmin = 1e-12; T->temp = 300.15; T->maxOrder = 2;
    T->noOpIter = 0; T->numGminSteps = 1; T->numSrcSteps = 1; T->dcTrcvMaxIter = 50;
    T->gminFactor = 10; T->gshunt = 0; T->srcFact = 1; T->diagGmin = 0;
    T->opstep = 1e-7; T->opfinal = 1e-5;
    T->method = OR_TRAPEZOIDAL;
    T->breaks = (double *)calloc(MAXBRK, sizeof(double));
    return T;
}

int or_tran_set_options(ORTran *T, double reltol, double abstol, double vntol, double chgtol,
                        double trtol, double gmin, int method, int maxord, int uic)
{
    T->reltol = reltol; T->abstol = abstol; T->vntol = vntol; T->chgtol = chgtol;
    /* cktsopt.c:123-133: maxord within 1..6 */
    if (maxord < 1) maxord = 1;
    if (maxord > 6) maxord = 6;
    T->trtol = trtol; T->gmin = gmin; T->method = method; T->maxOrder = maxord; T->uic = uic;
    return 0;
}

/* CKTpivotRelTol, KLU's pivot tolerance (default 1e-3) */
int or_tran_set_pivrel(ORTran *T, double pivrel)
{
    if (!(pivrel > 0 && pivrel <= 1)) return -1;
    T->pivrel = pivrel;
    return 0;
}

/* KLU's numeric pivot choice, restated: klu_factor's partial pivoting with
 * diagonal preference (lpivot, klu_kernel.c:370-530; the displaced diagonal
 * re-logged, klu_kernel.c:930-945) on the row-scaled matrix (klu_scale.c:
 * 19-159, scale = 2: each row divided by its largest magnitude, a zero row
 * by 1), columns taken in the order Q with P[k] the initial diagonal
 * candidate of column Q[k] -- Symbolic->P and Q.  A is dense, row-major
 * n x n in original (0-based) coordinates.  The elimination is dense and
 * right-looking: the candidates of column k are its Schur-complement
 * entries in the rows not yet pivotal, which is what KLU's left-looking
 * kernel sees for that column.  Among equal largest candidates the lowest
 * row wins: KLU takes the first in its column pattern, which lists the
 * column's own entries by ascending row among the rows its depth-first
 * search adds, so the two differ only on some exact ties.  Pnum[k] gets the
 * pivot row of column Q[k], x (may be NULL) the solution of A x = b.
 * Returns 0, or 102 on a zero or non-finite pivot (KLU_SINGULAR with
 * halt_if_singular). */
int or_klu_pivot(int n, const double *A, const int *P, const int *Q, double tol, const double *b, int *Pnum,
                 double *x)
{
    double *W = (double *)malloc(sizeof(double) * (size_t)n * (n + 1) + 1);
    double *Rs = (double *)malloc(sizeof(double) * (size_t)n + 1);
    int *Pinv = (int *)malloc(sizeof(int) * (size_t)n + 1);   /* row -> its step, else -1 - candidate step */
    int *Pd = (int *)malloc(sizeof(int) * (size_t)n + 1);     /* step -> diagonal candidate row */
    const int ld = n + 1;
    int i, j, k, rc = 0;
    for (i = 0; i < n; i++) {
        double m = 0;
        for (j = 0; j < n; j++) {
            const double v = A[(size_t)i * n + j];
            W[(size_t)i * ld + j] = v;
            if (fabs(v) > m) m = fabs(v);
        }
        W[(size_t)i * ld + n] = b ? b[i] : 0.0;
        Rs[i] = m == 0.0 ? 1.0 : m;
    }
    for (k = 0; k < n; k++) {
        Pd[k] = P[k];
        Pinv[P[k]] = -1 - k;
    }
    for (k = 0; k < n; k++) {
        const int col = Q[k];
        const int d = Pd[k];
        double best = -1.0, piv;
        int p = -1;
        for (i = 0; i < n; i++) {
            double v;
            if (Pinv[i] >= 0) continue;
            v = fabs(W[(size_t)i * ld + col]) / Rs[i];
            if (v > best) {
                best = v;
                p = i;
            }
        }
        if (p < 0) p = d;
        if (Pinv[d] < 0 && fabs(W[(size_t)d * ld + col]) / Rs[d] >= tol * best) p = d;
        piv = W[(size_t)p * ld + col];
        if (piv == 0.0 || !isfinite(piv)) {
            rc = 102;
            break;
        }
        if (p != d && Pinv[d] < 0) {
            const int kbar = -1 - Pinv[p];
            Pd[kbar] = d;
            Pinv[d] = -1 - kbar;
        }
        Pnum[k] = p;
        Pinv[p] = k;
        for (i = 0; i < n; i++) {
            double l;
            if (Pinv[i] >= 0) continue;
            l = W[(size_t)i * ld + col];
            if (l == 0.0) continue;
            l /= piv;
            /* the columns still to come, and the right-hand side */
            for (j = k + 1; j < n; j++) {
                const int cj = Q[j];
                W[(size_t)i * ld + cj] -= l * W[(size_t)p * ld + cj];
            }
            W[(size_t)i * ld + n] -= l * W[(size_t)p * ld + n];
        }
    }
    if (rc == 0 && x) {
        for (k = n - 1; k >= 0; k--) {
            const int p = Pnum[k];
            double acc = W[(size_t)p * ld + n];
            for (j = k + 1; j < n; j++) acc -= W[(size_t)p * ld + Q[j]] * x[Q[j]];
            x[Q[k]] = acc / W[(size_t)p * ld + Q[k]];
        }
    }
    free(W);
    free(Rs);
    free(Pinv);
    free(Pd);
    return rc;
}

/* the instance must be set up (defaults) but not yet temperature-resolved */
int or_tran_add_mos(ORTran *T, ORInst *h, int d, int g, int s, int b)
{
    int bits = or_bsim4_internal_nodes(h);
    if (bits) return -1;
    h->node[N_d] = d; h->node[N_g] = g; h->node[N_s] = s; h->node[N_b] = b;
    h->node[N_dp] = d; h->node[N_gp] = g; h->node[N_gm] = g; h->node[N_sp] = s;
    h->node[N_db] = h->node[N_bp] = h->node[N_sb] = b;
    h->node[N_q] = 0;
    h->state = T->nstates;
    T->nstates += S__COUNT;
    T->mos[T->nmos++] = h;
    return 0;
}

int or_tran_add_cap(ORTran *T, int p, int n, double c)
{
    T->cap_p[T->ncap] = p; T->cap_n[T->ncap] = n; T->cap_c[T->ncap] = c;
    T->cap_q[T->ncap] = T->nstates;
    T->nstates += 2;
    T->ncap++;
    return 0;
}

int or_tran_add_res(ORTran *T, int p, int n, double r)
{
    T->res_p[T->nres] = p; T->res_n[T->nres] = n; T->res_g[T->nres] = 1.0 / r;
    T->nres++;
    return 0;
}

/* pulse: NULL for DC, else npulse coefficients (v1 v2 td tr tf pw per) */
int or_tran_add_vsrc(ORTran *T, int p, int n, double dc, const double *pulse, int npulse)
{
    ORVsrc *v = &T->vs[T->nvs];
    int k;
    v->p = p; v->n = n; v->dc = dc;
    v->pulse = pulse != NULL;
    v->order = npulse;
    for (k = 0; k < npulse && k < 7; k++) v->c[k] = pulse[k];
    v->br = -1;
    T->nvs++;
    return 0;
}

int or_tran_set_ic(ORTran *T, int node, double v)
{
    if (node < 1 || node > T->nnodes) return -1;
    T->ic_given[node] = 1;
    T->ic[node] = v;
    return 0;
}

/* branch numbering, state arrays, instance temperature; returns 0 */
int or_tran_setup(ORTran *T)
{
    int k;
    T->reorder = 1;   /* NIreinit (nireinit.c:49) */
    if (!(T->pivrel > 0)) T->pivrel = 1e-3;
    T->nbr = T->nvs;
    T->n = T->nnodes + T->nbr;
    for (k = 0; k < T->nvs; k++) T->vs[k].br = T->nnodes + 1 + k;
    /* CKTmaxOrder + 2 state vectors (cktsetup.c:192-194), at most 8 */
    T->stbuf = (double *)calloc((size_t)8 * (T->nstates + 1), sizeof(double));
    for (k = 0; k < 8; k++) T->states[k] = T->stbuf + (size_t)k * (T->nstates + 1);
    for (k = 0; k < 8; k++) T->ck.states[k] = T->states[k];
    /* large circuits keep the matrix sparse (KLU only); small ones dense */
    T->sparse = T->n > or_sparse_above;
    if (!T->sparse) {
        T->A = (double *)calloc((size_t)(T->n + 1) * (T->n + 1), sizeof(double));
        T->lu = (double *)calloc((size_t)(T->n + 1) * (T->n + 1), sizeof(double));
    }
    T->piv = (int *)calloc((size_t)T->n + 1, sizeof(int));
    T->rhs = (double *)calloc((size_t)T->n + 1, sizeof(double));
    T->rhsOld = (double *)calloc((size_t)T->n + 1, sizeof(double));
    T->ck.gmin = T->gmin; T->ck.reltol = T->reltol; T->ck.abstol = T->abstol;
    /* which diagonal elements exist (SMPmakeElt calls of the devices and of
     * the .ic rows, CKTic) */
    T->hasdiag = (char *)calloc((size_t)T->n + 1, 1);
    for (k = 0; k < T->nmos; k++) {
        int r;
        for (r = 0; r < 4; r++) T->hasdiag[T->mos[k]->node[r]] = 1;
    }
    for (k = 0; k < T->ncap; k++) { T->hasdiag[T->cap_p[k]] = 1; T->hasdiag[T->cap_n[k]] = 1; }
    for (k = 0; k < T->nres; k++) { T->hasdiag[T->res_p[k]] = 1; T->hasdiag[T->res_n[k]] = 1; }
    for (k = 1; k <= T->nnodes; k++)
        if (T->ic_given[k]) T->hasdiag[k] = 1;
    T->hasdiag[0] = 0;
    T->ck.voltTol = T->vntol; T->ck.temp = T->temp;
    for (k = 0; k < T->nmos; k++)
        if (or_bsim4_inst_temp(T->mos[k], T->temp)) return -1;
    if (T->sparse && build_csc(T)) return -1;
    return ((T->sparse || (T->A && T->lu)) && T->rhs && T->rhsOld && T->stbuf) ? 0 : -1;
}

/* switch the linear solve to the reference's KLU in the library at path
 * (oracle/_ref/libklu_ref.so); call after or_tran_setup.  Returns 0. */
/* the structural CSC pattern (0-based) of the circuit's matrix: every
 * element the devices' SMPmakeElt calls create */
static int cmp_i64(const void *a, const void *b)
{
    const int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return x < y ? -1 : x > y;
}

static int build_csc(ORTran *T)
{
    int n = T->n, k, s, c;
    int64_t ne = 0, cap, i, nz;
    int64_t *E;
    if (T->Ap) return 0;
    cap = (int64_t)T->nmos * P__COUNT + 4 * (int64_t)(T->ncap + T->nres + T->nvs) + T->nnodes + 1;
    E = (int64_t *)malloc(sizeof(int64_t) * (size_t)cap);
    if (!E) return -1;
    /* key = column-major position (c, r), both 1-based */
#define MARK(a, b) do { if ((a) && (b)) E[ne++] = (int64_t)(b) * (n + 1) + (a); } while (0)
    for (k = 0; k < T->nmos; k++)
        for (s = 0; s < P__COUNT; s++) {
            int rr, cc;
            or_slot_rc(s, &rr, &cc);
            MARK(T->mos[k]->node[rr], T->mos[k]->node[cc]);
        }
    for (k = 0; k < T->ncap; k++) {
        MARK(T->cap_p[k], T->cap_p[k]); MARK(T->cap_n[k], T->cap_n[k]);
        MARK(T->cap_p[k], T->cap_n[k]); MARK(T->cap_n[k], T->cap_p[k]);
    }
    for (k = 0; k < T->nres; k++) {
        MARK(T->res_p[k], T->res_p[k]); MARK(T->res_n[k], T->res_n[k]);
        MARK(T->res_p[k], T->res_n[k]); MARK(T->res_n[k], T->res_p[k]);
    }
    for (k = 0; k < T->nvs; k++) {
        MARK(T->vs[k].p, T->vs[k].br); MARK(T->vs[k].n, T->vs[k].br);
        MARK(T->vs[k].br, T->vs[k].p); MARK(T->vs[k].br, T->vs[k].n);
    }
    for (k = 1; k <= T->nnodes; k++)
        if (T->ic_given[k]) MARK(k, k);
#undef MARK
    qsort(E, (size_t)ne, sizeof(int64_t), cmp_i64);
    nz = 0;
    for (i = 0; i < ne; i++)
        if (i == 0 || E[i] != E[i - 1]) E[nz++] = E[i];
    T->Ap = (int *)calloc((size_t)n + 1, sizeof(int));
    T->Ai = (int *)calloc((size_t)nz + 1, sizeof(int));
    T->Ax = (double *)calloc((size_t)nz + 1, sizeof(double));
    T->kb = (double *)calloc((size_t)n + 1, sizeof(double));
    for (i = 0, c = 1; c <= n; c++) {
        T->Ap[c - 1] = (int)i;
        while (i < nz && E[i] / (n + 1) == c) {
            T->Ai[i] = (int)(E[i] % (n + 1)) - 1;
            i++;
        }
    }
    T->Ap[n] = (int)nz;
    if (T->sparse) {
        T->hsize = 2 * nz + 7;
        T->hkey = (int64_t *)malloc(sizeof(int64_t) * (size_t)T->hsize);
        T->hval = (int *)malloc(sizeof(int) * (size_t)T->hsize);
        for (i = 0; i < T->hsize; i++) T->hkey[i] = -1;
        for (c = 0; c < n; c++)
            for (k = T->Ap[c]; k < T->Ap[c + 1]; k++) {
                const int64_t key = (int64_t)(T->Ai[k] + 1) * (n + 1) + (c + 1);
                int64_t h = ((key * 0x9E3779B97F4A7C15ULL) >> 11) % T->hsize;
                while (T->hkey[h] >= 0) h = (h + 1) % T->hsize;
                T->hkey[h] = key;
                T->hval[h] = k;
            }
    }
    free(E);
    return 0;
}

int or_tran_use_klu(ORTran *T, const char *path)
{
    void *dl = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!dl) return -1;
    T->kg_new = (void *(*)(int, const int *, const int *))dlsym(dl, "kg_new");
    T->kg_factor = (int (*)(void *, double *, int, double))dlsym(dl, "kg_factor");
    T->kg_solve = (int (*)(void *, double *))dlsym(dl, "kg_solve");
    T->kg_free = (void (*)(void *))dlsym(dl, "kg_free");
    T->kg_order = (int (*)(void *, int *, int *, int *))dlsym(dl, "kg_order");
    if (!T->kg_new || !T->kg_factor || !T->kg_solve || !T->kg_free) return -2;
    if (build_csc(T)) return -3;
    T->klu = T->kg_new(T->n, T->Ap, T->Ai);
    return T->klu ? 0 : -3;
}

/* The CSC pattern for a caller's own solver: Ap[n+1] and Ai[nnz] (0-based)
 * into the given arrays when non-NULL; returns nnz.  Call after
 * or_tran_setup. */
int64_t or_tran_csc(ORTran *T, int64_t *Ap, int *Ai)
{
    int k;
    if (build_csc(T)) return -1;
    if (Ap) for (k = 0; k <= T->n; k++) Ap[k] = T->Ap[k];
    if (Ai) for (k = 0; k < T->Ap[T->n]; k++) Ai[k] = T->Ai[k];
    return T->Ap[T->n];
}

/* State offset and terminals (d g s b) of MOSFET k, in insertion order. */
int or_tran_mos_info(const ORTran *T, int k, int *state, int *nodes)
{
    if (k < 0 || k >= T->nmos) return -1;
    if (state) *state = T->mos[k]->state;
    if (nodes) { nodes[0] = T->mos[k]->node[N_d]; nodes[1] = T->mos[k]->node[N_g];
                 nodes[2] = T->mos[k]->node[N_s]; nodes[3] = T->mos[k]->node[N_b]; }
    return 0;
}

/* Drop-in BSIM4 load and linear solver (NULL functions restore the
 * oracle's own): run(ctx, scalars, rhsOld, states[3], A, rhs, &noncon)
 * replaces the MOSFET loads of CKTload, adding into the dense matrix A
 * ((n+1) x (n+1), row-major, entry (r, c) at r (n+1) + c) and rhs;
 * factor(ctx, Ax, gmin) and solve(ctx, rhs) replace the linear solve, Ax in
 * the order of or_tran_csc's pattern, rhs 1-based.  Call after
 * or_tran_setup. */
int or_tran_use_dropin(ORTran *T, void *run, void *dl_ctx, void *factor, void *solve_fn, void *smp_ctx)
{
    T->dl_run = (int (*)(void *, const void *, const double *, double *const *, double *, double *, int *))run;
    T->dl_ctx = dl_ctx;
    T->smp_factor = (int (*)(void *, const double *, double))factor;
    T->smp_solve = (int (*)(void *, double *))solve_fn;
    T->smp_ctx = smp_ctx;
    if (T->smp_factor && build_csc(T)) return -1;
    if (T->smp_factor && !T->Ax) T->Ax = (double *)calloc((size_t)T->Ap[T->n] + 1, sizeof(double));
    return 0;
}

/* The reference KLU's pivot sequence from its last klu_factor, as 1-based
 * unknowns (rows[k] = Pnum[k] + 1, cols[k] = Q[k] + 1), and counts = {lnz,
 * unz, nzoff, nblocks, klu_factor calls, those with an off-diagonal pivot,
 * klu_refactor calls, those that met a zero pivot} (8 ints).  Returns 0, or
 * nonzero without a KLU factor. */
int or_tran_klu_order(const ORTran *T, int *rows, int *cols, int *counts)
{
    int k, rc;
    if (!T->klu || !T->kg_order) return -1;
    rc = T->kg_order(T->klu, rows, cols, counts);
    if (rc) return rc;
    for (k = 0; k < T->n; k++) {
        rows[k] += 1;
        cols[k] += 1;
    }
    return 0;
}

int64_t or_tran_npts(const ORTran *T) { return T->npts; }
int or_tran_n(const ORTran *T) { return T->n; }
void or_tran_stats(const ORTran *T, int64_t *acc, int64_t *rej, int64_t *iters)
{
    *acc = T->accepted; *rej = T->rejected; *iters = T->niter;
}
void or_tran_output(const ORTran *T, double *t, double *x)
{
    memcpy(t, T->out_t, sizeof(double) * T->npts);
    memcpy(x, T->out_x, sizeof(double) * T->npts * (T->n + 1));
}

void or_tran_free(ORTran *T)
{
    if (!T) return;
    free(T->mos); free(T->cap_p); free(T->cap_n); free(T->cap_q); free(T->cap_c); free(T->cap_ic);
    free(T->res_p); free(T->res_n); free(T->res_g); free(T->vs); free(T->ic_given);
    free(T->ic); free(T->stbuf); free(T->hasdiag); free(T->A); free(T->lu); free(T->piv); free(T->rhs);
    free(T->rhsOld); free(T->out_t); free(T->out_x);
    if (T->klu) T->kg_free(T->klu);
    free(T->Ap); free(T->Ai); free(T->Ax); free(T->kb);
    free(T->hkey); free(T->hval); free(T->breaks);
    free(T);
}
