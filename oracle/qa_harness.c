/*
 * ORACLE (test infrastructure only): single-device operating-point harness
 * used to check the BSIM4 restatement against the CMC QA goldens in
 * tests/bsim4/{nmos,pmos}/reference/*.standard.
 *
 * The QA netlists (tests/bin/ngspice.pm:443-540) drive the four terminals of
 * one instance with voltage sources and print the source branch currents of
 * a .dc sweep.  Here the external terminals are held at the given voltages,
 * the internal nodes (gate resistance, body network, S/D series resistance)
 * are solved by Newton with the restated load in MODEDCTRANCURVE, and the
 * terminal currents are the KCL residuals of the converged linearisation.
 * Each bias point is solved from its own initial guess to machine precision.
 */
#include <math.h>
#include <string.h>
#include <stdlib.h>
#include "bsim4_ref.h"

#define MAXN 16

/* node numbering: 0 ground, 1..4 = d g s b, internal nodes after */
static int assign_nodes(ORInst *h, int *nint)
{
    int bits = or_bsim4_internal_nodes(h);
    int next = 5;
    h->node[N_d] = 1; h->node[N_g] = 2; h->node[N_s] = 3; h->node[N_b] = 4;
    h->node[N_dp] = (bits & 1) ? next++ : 1;
    h->node[N_sp] = (bits & 2) ? next++ : 3;
    h->node[N_gp] = (bits & 4) ? next++ : 2;
    h->node[N_gm] = (bits & 8) ? next++ : 2;
    if (bits & 16) {
        h->node[N_db] = next++;
        h->node[N_bp] = next++;
        h->node[N_sb] = next++;
    } else {
        h->node[N_db] = h->node[N_bp] = h->node[N_sb] = 4;
    }
    h->node[N_q] = 0;
    *nint = next - 5;
    return next;   /* number of node slots incl. ground */
}

static const int slot_rc[P__COUNT][2] = {
    {N_dp, N_bp}, {N_gp, N_bp}, {N_sp, N_bp}, {N_bp, N_dp}, {N_bp, N_gp}, {N_bp, N_sp},
    {N_bp, N_bp}, {N_d, N_d}, {N_gp, N_gp}, {N_s, N_s}, {N_dp, N_dp}, {N_sp, N_sp},
    {N_d, N_dp}, {N_gp, N_dp}, {N_gp, N_sp}, {N_s, N_sp}, {N_dp, N_sp}, {N_dp, N_d},
    {N_dp, N_gp}, {N_sp, N_gp}, {N_sp, N_s}, {N_sp, N_dp}, {N_q, N_q}, {N_q, N_bp},
    {N_q, N_dp}, {N_q, N_sp}, {N_q, N_gp}, {N_dp, N_q}, {N_sp, N_q}, {N_gp, N_q},
    {N_g, N_g}, {N_g, N_gp}, {N_gp, N_g}, {N_g, N_dp}, {N_g, N_sp}, {N_g, N_bp},
    {N_gm, N_dp}, {N_gm, N_gp}, {N_gm, N_gm}, {N_gm, N_g}, {N_gm, N_sp}, {N_gm, N_bp},
    {N_dp, N_gm}, {N_gp, N_gm}, {N_g, N_gm}, {N_sp, N_gm}, {N_bp, N_gm},
    {N_dp, N_db}, {N_sp, N_sb}, {N_db, N_dp}, {N_db, N_db}, {N_db, N_bp}, {N_db, N_b},
    {N_bp, N_db}, {N_bp, N_b}, {N_bp, N_sb}, {N_sb, N_sp}, {N_sb, N_bp}, {N_sb, N_b},
    {N_sb, N_sb}, {N_b, N_db}, {N_b, N_bp}, {N_b, N_sb}, {N_b, N_b},
    {N_d, N_gp}, {N_d, N_sp}, {N_d, N_bp}, {N_s, N_dp}, {N_s, N_gp}, {N_s, N_bp},
};

int or_slot_rc(int slot, int *row_role, int *col_role)
{
    if (slot < 0 || slot >= P__COUNT) return -1;
    *row_role = slot_rc[slot][0];
    *col_role = slot_rc[slot][1];
    return 0;
}

static int solve_dense(double *A, double *b, int n)
{
    int i, j, k, p;
    for (k = 0; k < n; k++) {
        double best = fabs(A[k * n + k]);
        p = k;
        for (i = k + 1; i < n; i++) if (fabs(A[i * n + k]) > best) { best = fabs(A[i * n + k]); p = i; }
        if (best == 0.0) return -1;
        if (p != k) {
            for (j = 0; j < n; j++) { double t = A[k * n + j]; A[k * n + j] = A[p * n + j]; A[p * n + j] = t; }
            { double t = b[k]; b[k] = b[p]; b[p] = t; }
        }
        for (i = k + 1; i < n; i++) {
            double f = A[i * n + k] / A[k * n + k];
            for (j = k; j < n; j++) A[i * n + j] -= f * A[k * n + j];
            b[i] -= f * b[k];
        }
    }
    for (i = n - 1; i >= 0; i--) {
        double s = b[i];
        for (j = i + 1; j < n; j++) s -= A[i * n + j] * b[j];
        b[i] = s / A[i * n + i];
    }
    return 0;
}

/* Assemble the linearised system of one instance at the node voltages v.
 * G[r][c] = sum of stamps, I[r] = RHS. */
static int assemble(ORInst *h, ORCkt *ckt, double *v, int nn, double *G, double *I)
{
    int s, r, c;
    ckt->rhsOld = v;
    ckt->noncon = 0;
    if (or_bsim4_load(h, ckt)) return -1;
    memset(G, 0, sizeof(double) * nn * nn);
    memset(I, 0, sizeof(double) * nn);
    for (s = 0; s < P__COUNT; s++) {
        r = h->node[slot_rc[s][0]];
        c = h->node[slot_rc[s][1]];
        if (r == 0 || c == 0) continue;
        G[r * nn + c] += h->stamp[s];
    }
    for (s = 0; s < N__COUNT; s++) {
        r = h->node[s];
        if (r == 0) continue;
        I[r] += h->rhs[s];
    }
    return 0;
}

/* Solve one bias point; returns the 4 terminal currents (into d, g, s, b). */
int or_qa_dc_point(ORInst *h, double temp, double vd, double vg, double vs, double vb,
                   double gmin, double *cur, double *vout)
{
    double v[MAXN], G[MAXN * MAXN], I[MAXN], A[MAXN * MAXN], b[MAXN];
    double states[4][S__COUNT];
    ORCkt ckt;
    int nint, nn, i, j, it;
    nn = assign_nodes(h, &nint);
    memset(&ckt, 0, sizeof(ckt));
    memset(states, 0, sizeof(states));
    ckt.mode = OR_MODEDCTRANCURVE | OR_MODEINITFLOAT;
    ckt.order = 1; ckt.method = OR_TRAPEZOIDAL;
    ckt.gmin = gmin; ckt.reltol = 1e-5; ckt.abstol = 1e-15; ckt.voltTol = 1e-6;
    ckt.temp = temp;
    ckt.states[0] = states[0]; ckt.states[1] = states[1];
    ckt.states[2] = states[2]; ckt.states[3] = states[3];
    h->state = 0;
    memset(v, 0, sizeof(v));
    v[1] = vd; v[2] = vg; v[3] = vs; v[4] = vb;
    for (i = 5; i < nn; i++) v[i] = 0.0;
    /* internal nodes start at their external terminal's value */
    if (h->node[N_dp] >= 5) v[h->node[N_dp]] = vd;
    if (h->node[N_sp] >= 5) v[h->node[N_sp]] = vs;
    if (h->node[N_gp] >= 5) v[h->node[N_gp]] = vg;
    if (h->node[N_gm] >= 5) v[h->node[N_gm]] = vg;
    if (h->node[N_db] >= 5) { v[h->node[N_db]] = vb; v[h->node[N_bp]] = vb; v[h->node[N_sb]] = vb; }
    /* seed the state vector with the current bias so the limiting starts there */
    states[0][S_vds] = h->model->type * (vd - vs);
    states[0][S_vgs] = h->model->type * (vg - vs);
    states[0][S_vbs] = h->model->type * (vb - vs);
    states[0][S_vbd] = states[0][S_vbs] - states[0][S_vds];
    states[0][S_vges] = states[0][S_vgms] = states[0][S_vgs];
    states[0][S_vdbs] = states[0][S_vbs]; states[0][S_vsbs] = states[0][S_vbs];
    states[0][S_vdbd] = states[0][S_vbd];
    states[0][S_vdes] = states[0][S_vds];
    h->von = 0.0;
    for (it = 0; it < 200; it++) {
        double maxd = 0.0;
        if (assemble(h, &ckt, v, nn, G, I)) return -1;
        if (nint > 0) {
            for (i = 0; i < nint; i++) {
                b[i] = I[5 + i];
                for (j = 1; j <= 4; j++) b[i] -= G[(5 + i) * nn + j] * v[j];
                for (j = 0; j < nint; j++) A[i * nint + j] = G[(5 + i) * nn + 5 + j];
            }
            if (solve_dense(A, b, nint)) return -2;
            for (i = 0; i < nint; i++) {
                double d = fabs(b[i] - v[5 + i]);
                if (d > maxd) maxd = d;
                v[5 + i] = b[i];
            }
        }
        if (it > 2 && maxd < 1e-15 && ckt.noncon == 0) break;
    }
    /* terminal currents from the converged linearisation */
    if (assemble(h, &ckt, v, nn, G, I)) return -1;
    for (i = 1; i <= 4; i++) {
        double s = -I[i];
        for (j = 1; j < nn; j++) s += G[i * nn + j] * v[j];
        cur[i - 1] = s;
    }
    if (vout) for (i = 0; i < MAXN; i++) vout[i] = (i < nn) ? v[i] : 0.0;
    return it;
}
