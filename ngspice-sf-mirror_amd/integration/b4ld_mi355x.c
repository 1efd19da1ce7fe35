/*
 * BSIM4load on the MI355X engine -- the file a maintainer compiles into
 * ngspice's BSIM4 device (src/spicelib/devices/bsim4/) in place of b4ld.c,
 * with -I<engine>/include -I<engine>/ngspice-sf-mirror_amd/csrc and
 * -lngspice_mi355x.  It replaces DEVload (devdefs.h:57) for BSIM4:
 *
 *   int BSIM4load(GENmodel *inModel, CKTcircuit *ckt)      (b4ld.c:69)
 *
 * with the same signature, return convention (OK / E_* numbers,
 * iferrmsg.h:17, sperror.h:15-32) and non-convergence signalling (votes
 * added to ckt->CKTnoncon, b4ld.c:4072-4110).
 *
 * On the first call for a circuit it hands the engine what BSIM4setup,
 * BSIM4temp and BSIM4bindCSC established:
 *   - every model card's given parameters (the Given flags of b4mpar.c,
 *     through the engine's generated parameter table) and every instance's
 *     given parameters (b4par.c), as an ngmi_bsim4_set;
 *   - each instance's terminal unknowns, its GENstate offset (b4set.c:2318)
 *     and the positions its 16 bound matrix pointers take in KLU's Ax
 *     (b4bindCSC.c:12-111), as an ngmi_devload.
 * Every call then runs the load of all instances on the GPU
 * (ngmi_devload_run): CKTrhsOld and CKTstate0..2 in, CKTstate0/1 back,
 * stamps added into KLU's Ax and CKTrhs instance by instance.
 *
 * The engine builds BSIM4's default selector set (capMod 2, no rgate,
 * rbody, S/D resistance nodes, NQS); an instance with internal nodes makes
 * this load return E_UNSUPP, and the circuit must run with the reference's
 * b4ld.c.  KLU must be on (.options klu): the matrix positions come from
 * the KLU binding.
 */
#include "ngspice/ngspice.h"
#include "ngspice/cktdefs.h"
#include "ngspice/smpdefs.h"
#include "ngspice/sperror.h"
#include "ngspice/klu.h"
#include "bsim4def.h"

#include "ngspice_mi355x.h"
#include "bsim4_tables.h"   /* the engine's generated BSIM4 parameter table */

typedef struct mi_ckt {
    CKTcircuit *ckt;
    double *ax_base;             /* KLU's Ax when the positions were taken */
    int nmodels;
    ngmi_bsim4_model **models;
    ngmi_bsim4_set *set;
    ngmi_devload *dl;
    struct mi_ckt *next;
} mi_ckt;

static mi_ckt *mi_circuits = NULL;

/* --- model cards: the given parameters, by name ------------------------ */

#define MI_MODEL_r(name, field) \
    if (model->BSIM4##field##Given) rc |= ngmi_bsim4_model_param(m, name, model->BSIM4##field);
#define MI_MODEL_i(name, field) \
    if (model->BSIM4##field##Given) rc |= ngmi_bsim4_model_param(m, name, (double)model->BSIM4##field);
#define MI_MODEL_f(name, field) MI_MODEL_i(name, field)
#define MI_MODEL_n(name, field)
#define MI_MODEL_p(name, field)
#define MI_MODEL(name, field, kind) MI_MODEL_##kind(name, field)

static ngmi_bsim4_model *mi_model(BSIM4model *model, CKTcircuit *ckt)
{
    ngmi_bsim4_model *m = ngmi_bsim4_model_new();
    int rc = 0;
    if (!m) return NULL;
    rc |= ngmi_bsim4_model_param(m, model->BSIM4type == NMOS ? "nmos" : "pmos", 1.0);
    /* tnom is stored in kelvin (b4mpar.c adds CONSTCtoK): it goes to the
     * setup call below, not through the table */
    model->BSIM4tnomGiven = 0;
    BSIM4_MODEL_PARAMS(MI_MODEL)
    model->BSIM4tnomGiven = 1;
    if (rc == 0) rc = ngmi_bsim4_model_setup(m, model->BSIM4tnom, ckt->CKTtemp);
    if (rc) {
        fprintf(stderr, "Error: BSIM4 model %s on the MI355X engine: %s\n", model->BSIM4modName,
                ngmi_last_error());
        ngmi_bsim4_model_free(m);
        return NULL;
    }
    return m;
}

/* --- instances: the given parameters of one instance ------------------- */

#define MI_INST_r(name, field) \
    if (here->BSIM4##field##Given) { names[np] = name; vals[np++] = here->BSIM4##field; }
#define MI_INST_i(name, field) \
    if (here->BSIM4##field##Given) { names[np] = name; vals[np++] = (double)here->BSIM4##field; }
#define MI_INST_f(name, field) MI_INST_i(name, field)
#define MI_INST(name, field, kind) MI_INST_##kind(name, field)

/* the 16 Jacobian entries of one instance, row-major over d g s b; with
 * no internal nodes the primed nodes are the terminals themselves */
static double *mi_ptr(BSIM4instance *here, int k)
{
    double *p[16] = {
        here->BSIM4DPdpPtr, here->BSIM4DPgpPtr, here->BSIM4DPspPtr, here->BSIM4DPbpPtr,
        here->BSIM4GPdpPtr, here->BSIM4GPgpPtr, here->BSIM4GPspPtr, here->BSIM4GPbpPtr,
        here->BSIM4SPdpPtr, here->BSIM4SPgpPtr, here->BSIM4SPspPtr, here->BSIM4SPbpPtr,
        here->BSIM4BPdpPtr, here->BSIM4BPgpPtr, here->BSIM4BPspPtr, here->BSIM4BPbpPtr};
    return p[k];
}

static mi_ckt *mi_build(BSIM4model *first, CKTcircuit *ckt)
{
    KLUmatrix *K = ckt->CKTmatrix->SMPkluMatrix;
    BSIM4model *model;
    BSIM4instance *here;
    mi_ckt *c;
    int n = 0, k, mi = 0, rc = 0;
    int *midx, *nodes, *soff;
    int64_t *pos;

    if (!ckt->CKTmatrix->CKTkluMODE || !K || !K->KLUmatrixAx) {
        fprintf(stderr, "Error: the MI355X BSIM4 load needs the KLU matrix (.options klu)\n");
        return NULL;
    }
    c = TMALLOC(mi_ckt, 1);
    c->ckt = ckt;
    c->ax_base = K->KLUmatrixAx;
    c->nmodels = 0;
    for (model = first; model; model = BSIM4nextModel(model)) {
        c->nmodels++;
        for (here = BSIM4instances(model); here; here = BSIM4nextInstance(here)) {
            if (here->BSIM4dNodePrime != here->BSIM4dNode || here->BSIM4sNodePrime != here->BSIM4sNode ||
                here->BSIM4gNodePrime != here->BSIM4gNodeExt || here->BSIM4bNodePrime != here->BSIM4bNode) {
                fprintf(stderr, "Error: BSIM4 instance %s has internal nodes; the MI355X engine "
                        "builds the default selectors only\n", here->BSIM4name);
                tfree(c);
                return NULL;
            }
            n++;
        }
    }
    c->models = TMALLOC(ngmi_bsim4_model *, c->nmodels);
    midx = TMALLOC(int, n);
    nodes = TMALLOC(int, 4 * n);
    soff = TMALLOC(int, n);
    pos = TMALLOC(int64_t, 16 * n);
    c->set = ngmi_bsim4_set_new();
    k = 0;
    for (model = first; model; model = BSIM4nextModel(model), mi++) {
        c->models[mi] = mi_model(model, ckt);
        if (!c->models[mi]) { rc = 1; break; }
        for (here = BSIM4instances(model); here; here = BSIM4nextInstance(here), k++) {
            const char *names[64];
            double vals[64];
            int np = 0, e, q;
            BSIM4_INST_PARAMS(MI_INST)
            midx[k] = 0;
            if (ngmi_bsim4_set_add(c->set, &c->models[mi], 1, &midx[k], 1, names, np, vals) < 0) {
                fprintf(stderr, "Error: BSIM4 instance %s: %s\n", here->BSIM4name, ngmi_last_error());
                rc = 1;
                break;
            }
            nodes[4 * k + 0] = here->BSIM4dNode;
            nodes[4 * k + 1] = here->BSIM4gNodeExt;
            nodes[4 * k + 2] = here->BSIM4sNode;
            nodes[4 * k + 3] = here->BSIM4bNode;
            soff[k] = here->BSIM4states;
            for (e = 0; e < 16; e++) {
                double *p = mi_ptr(here, e);
                pos[16 * k + e] = (p >= K->KLUmatrixAx && p < K->KLUmatrixAx + K->KLUmatrixNZ)
                                      ? (int64_t)(p - K->KLUmatrixAx) : -1;
            }
            (void)q;
        }
        if (rc) break;
    }
    if (rc == 0 && ngmi_bsim4_set_finalize(c->set) != NGMI_OK) {
        fprintf(stderr, "Error: BSIM4 on the MI355X engine: %s\n", ngmi_last_error());
        rc = 1;
    }
    if (rc == 0) {
        c->dl = ngmi_devload_new(c->set, ckt->CKTmaxEqNum - 1, ckt->CKTnumStates, nodes, soff, pos);
        if (!c->dl) {
            fprintf(stderr, "Error: BSIM4 on the MI355X engine: %s\n", ngmi_last_error());
            rc = 1;
        }
    }
    tfree(midx);
    tfree(nodes);
    tfree(soff);
    tfree(pos);
    if (rc) {
        ngmi_bsim4_set_free(c->set);
        for (mi = 0; mi < c->nmodels; mi++)
            if (c->models[mi]) ngmi_bsim4_model_free(c->models[mi]);
        tfree(c->models);
        tfree(c);
        return NULL;
    }
    c->next = mi_circuits;
    mi_circuits = c;
    return c;
}

int BSIM4load(GENmodel *inModel, CKTcircuit *ckt)
{
    BSIM4model *first = (BSIM4model *)inModel;
    mi_ckt *c;
    ngmi_ckt_scalars sc;
    double *st[3];
    int i, nc = 0, rc;

    for (c = mi_circuits; c; c = c->next)
        if (c->ckt == ckt && c->ax_base == ckt->CKTmatrix->SMPkluMatrix->KLUmatrixAx) break;
    if (!c && !(c = mi_build(first, ckt))) return E_UNSUPP;

    sc.mode = (int)ckt->CKTmode;
    sc.order = ckt->CKTorder;
    sc.method = ckt->CKTintegrateMethod == GEAR ? 2 : 1;
    for (i = 0; i < 7; i++) sc.ag[i] = ckt->CKTag[i];
    sc.xfact = ckt->CKTdeltaOld[1] != 0.0 ? ckt->CKTdelta / ckt->CKTdeltaOld[1] : 0.0;
    sc.gmin = ckt->CKTgmin;
    sc.reltol = ckt->CKTreltol;
    sc.abstol = ckt->CKTabstol;
    st[0] = ckt->CKTstate0;
    st[1] = ckt->CKTstate1;
    st[2] = ckt->CKTstate2;
    rc = ngmi_devload_run(c->dl, &sc, ckt->CKTrhsOld, st, ckt->CKTmatrix->SMPkluMatrix->KLUmatrixAx,
                          ckt->CKTrhs, &nc);
    if (rc != NGMI_OK) {
        SPfrontEnd->IFerrorf(ERR_FATAL, "BSIM4load on the MI355X engine: %s", ngmi_last_error());
        return rc == NGMI_E_UNSUPP ? E_UNSUPP : E_PANIC;
    }
    ckt->CKTnoncon += nc;
    return OK;
}
