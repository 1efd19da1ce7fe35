/*
 * ORACLE (test infrastructure only): flat C entry points over the BSIM4
 * restatement for ctypes callers in tests/ and bench.py's CPU baseline.
 */
#include <stdlib.h>
#include <string.h>
#include "bsim4_ref.h"

int or_qa_dc_point(ORInst *h, double temp, double vd, double vg, double vs, double vb,
                   double gmin, double *cur, double *vout);
int or_slot_rc(int slot, int *row_role, int *col_role);

void *or_model_new(void)
{
    ORModel *m = (ORModel *)calloc(1, sizeof(ORModel));
    if (m) or_bsim4_model_init(m);
    return m;
}

void or_model_free(void *m) { free(m); }

int or_model_param(void *m, const char *name, double value)
{
    return or_bsim4_model_set((ORModel *)m, name, value);
}

/* defaults + model temperature; temperatures in kelvin */
int or_model_finish(void *mp, double nomTemp, double temp)
{
    ORModel *m = (ORModel *)mp;
    int rc = or_bsim4_model_defaults(m, nomTemp);
    if (rc) return rc;
    return or_bsim4_model_temp(m, temp);
}

double or_model_get(void *mp, const char *what)
{
    ORModel *m = (ORModel *)mp;
    if (!strcmp(what, "coxe")) return m->coxe;
    if (!strcmp(what, "vtm")) return m->vtm;
    if (!strcmp(what, "vcrit")) return m->vcrit;
    if (!strcmp(what, "type")) return m->type;
    return 0.0;
}

void *or_inst_new(void *m)
{
    ORInst *h = (ORInst *)calloc(1, sizeof(ORInst));
    if (h) or_bsim4_inst_init(h, (ORModel *)m);
    return h;
}

void or_inst_free(void *h) { free(h); }

int or_inst_param(void *h, const char *name, double value)
{
    return or_bsim4_inst_set((ORInst *)h, name, value);
}

/* defaults; returns the internal-node bits (>= 0) or a negative error */
int or_inst_defaults(void *hp)
{
    ORInst *h = (ORInst *)hp;
    int rc = or_bsim4_inst_defaults(h);
    if (rc) return rc;
    return or_bsim4_internal_nodes(h);
}

/* node roles d g s b dp gp gm sp db bp sb q, then temperature */
int or_inst_finish(void *hp, const int *nodes, int state, double temp)
{
    ORInst *h = (ORInst *)hp;
    int k;
    for (k = 0; k < N__COUNT; k++) h->node[k] = nodes[k];
    h->state = state;
    return or_bsim4_inst_temp(h, temp);
}

int or_qa_dc(void *hp, double temp, double vd, double vg, double vs, double vb,
             double gmin, double *cur, double *vout)
{
    ORInst *h = (ORInst *)hp;
    int bits = or_bsim4_internal_nodes(h), next = 5, rc;
    h->node[N_d] = 1; h->node[N_g] = 2; h->node[N_s] = 3; h->node[N_b] = 4;
    h->node[N_dp] = (bits & 1) ? next++ : 1;
    h->node[N_sp] = (bits & 2) ? next++ : 3;
    h->node[N_gp] = (bits & 4) ? next++ : 2;
    h->node[N_gm] = (bits & 8) ? next++ : 2;
    if (bits & 16) { h->node[N_db] = next++; h->node[N_bp] = next++; h->node[N_sb] = next++; }
    else h->node[N_db] = h->node[N_bp] = h->node[N_sb] = 4;
    h->node[N_q] = 0;
    rc = or_bsim4_inst_temp(h, temp);
    if (rc) return rc;
    return or_qa_dc_point(h, temp, vd, vg, vs, vb, gmin, cur, vout);
}

/* Evaluate one instance at given node voltages in a caller-provided circuit
 * context.  states: 4 arrays laid out [4][nstate]; rhsOld: node voltages.
 * out_stamp[P__COUNT], out_rhs[N__COUNT]; returns noncon increment. */
int or_inst_load(void *hp, int mode, double temp, double gmin, double reltol,
                 double abstol, double voltTol, const double *ag, int order,
                 int method, double delta, const double *deltaOld,
                 double *rhsOld, double *s0, double *s1, double *s2, double *s3,
                 double *out_stamp, double *out_rhs)
{
    ORInst *h = (ORInst *)hp;
    ORCkt ckt;
    int k;
    memset(&ckt, 0, sizeof(ckt));
    ckt.mode = mode; ckt.temp = temp; ckt.gmin = gmin; ckt.reltol = reltol;
    ckt.abstol = abstol; ckt.voltTol = voltTol; ckt.order = order; ckt.method = method;
    for (k = 0; k < 7; k++) { ckt.ag[k] = ag ? ag[k] : 0.0; ckt.deltaOld[k] = deltaOld ? deltaOld[k] : 0.0; }
    ckt.delta = delta;
    ckt.rhsOld = rhsOld;
    ckt.states[0] = s0; ckt.states[1] = s1; ckt.states[2] = s2; ckt.states[3] = s3;
    if (or_bsim4_load(h, &ckt)) return -1;
    memcpy(out_stamp, h->stamp, sizeof(h->stamp));
    memcpy(out_rhs, h->rhs, sizeof(h->rhs));
    return ckt.noncon;
}

/* named per-instance quantities after a load, for parity checks */
int or_inst_node(void *hp, int role) { return ((ORInst *)hp)->node[role]; }

double or_inst_get(void *hp, const char *what)
{
    ORInst *h = (ORInst *)hp;
#define G(f) if (!strcmp(what, #f)) return h->f;
    G(cd) G(csub) G(cbd) G(cbs) G(gbd) G(gbs) G(gm) G(gds) G(gmbs) G(gbbs) G(gbgs) G(gbds)
    G(Igidl) G(Igisl) G(ggidld) G(ggidlg) G(ggidlb) G(ggisls) G(ggislg) G(ggislb)
    G(Igs) G(Igd) G(Igcs) G(Igcd) G(Igb) G(von) G(vdsat) G(Vgsteff) G(Vdseff) G(ueff)
    G(Abulk) G(cggb) G(cgsb) G(cgdb) G(cdgb) G(cdsb) G(cddb) G(cbgb) G(cbsb) G(cbdb)
    G(capbs) G(capbd) G(qgate) G(qbulk) G(qdrn) G(qsrc) G(leff) G(weff) G(vth0) G(u0temp)
    G(vsattemp) G(phi) G(k2) G(vfbzb) G(Aseff) G(Adeff) G(Pseff) G(Pdeff) G(grgeltd)
    G(sourceConductance) G(drainConductance) G(rds0) G(abulkCVfactor) G(litl) G(vbsc)
#undef G
    if (!strcmp(what, "mode")) return h->mode;
    return 0.0;
}

int or_counts(int *nslots, int *nroles, int *nstates)
{
    *nslots = P__COUNT;
    *nroles = N__COUNT;
    *nstates = S__COUNT;
    return 0;
}
