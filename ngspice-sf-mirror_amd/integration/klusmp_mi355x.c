/*
 * SMP* on the MI355X engine -- the KLU branch of klusmp.c's four hot
 * calls, for a maintainer to compile into ngspice (src/maths/KLU/) in
 * place of those functions, with -I<engine>/include and -lngspice_mi355x.
 * The signatures and return conventions are smpdefs.h:61-86's:
 *
 *   int  SMPpreOrder(SMPmatrix *)                             klusmp.c:1255-1285
 *   int  SMPreorder(SMPmatrix *, double PivTol, double PivRel, double Gmin)
 *                                                             klusmp.c:752-806
 *   int  SMPluFac(SMPmatrix *, double PivTol, double Gmin)    klusmp.c:573-624
 *   void SMPsolve(SMPmatrix *, double RHS[], double Spare[])  klusmp.c:950-1011
 *
 * The matrix stays where KLU's binding put it (KLUmatrixAp/Ai/Ax, 0-based
 * CSC, the devices' pointers into Ax): SMPpreOrder hands its pattern to the
 * engine's GPU sparse LU (ngmi_smp_new), SMPluFac and SMPreorder factor a
 * device copy of Ax after LoadGmin_CSC (ngmi_smp_factor adds Gmin to every
 * diagonal entry, as LoadGmin_CSC does, when KLUloadDiagGmin is set), and
 * SMPsolve gathers the RHS through KLU's node collapsing
 * (KLUmatrixNodeCollapsingNewToOld), solves on the GPU and scatters back.
 * The engine fixes its pivot sequence at SMPpreOrder (voltage-source rows
 * paired with their nodes, nested dissection); SMPreorder does not search
 * new pivots, and a zero pivot returns E_SINGULAR, which NIiter answers as
 * it answers KLU's (niiter.c:162-195).  The other SMP functions are KLU's
 * own, unchanged.
 */
#include "ngspice/ngspice.h"
#include "ngspice/smpdefs.h"
#include "ngspice/sperror.h"
#include "ngspice/klu.h"

#include "ngspice_mi355x.h"

typedef struct mi_smp {
    SMPmatrix *matrix;
    ngmi_smp *lu;
    double *rhs;            /* 1-based work vector in KLU's numbering */
    struct mi_smp *next;
} mi_smp;

static mi_smp *mi_matrices = NULL;

static mi_smp *mi_find(SMPmatrix *Matrix)
{
    mi_smp *m;
    for (m = mi_matrices; m; m = m->next)
        if (m->matrix == Matrix) return m;
    return NULL;
}

int SMPpreOrder(SMPmatrix *Matrix)
{
    KLUmatrix *K = Matrix->SMPkluMatrix;
    mi_smp *m;
    int64_t *Ap;
    unsigned int c;
    int rc;
    if (!Matrix->CKTkluMODE) {
        spMNA_Preorder(Matrix->SPmatrix);
        return spError(Matrix->SPmatrix);
    }
    if (K->KLUmatrixN == 0) return 0;
    if (!(m = mi_find(Matrix))) {
        m = TMALLOC(mi_smp, 1);
        m->matrix = Matrix;
        m->next = mi_matrices;
        mi_matrices = m;
    } else {
        ngmi_smp_free(m->lu);
        tfree(m->rhs);
    }
    Ap = TMALLOC(int64_t, K->KLUmatrixN + 1);
    for (c = 0; c <= K->KLUmatrixN; c++) Ap[c] = K->KLUmatrixAp[c];
    rc = ngmi_smp_new((int)K->KLUmatrixN, Ap, K->KLUmatrixAi, &m->lu);
    tfree(Ap);
    m->rhs = TMALLOC(double, K->KLUmatrixN + 1);
    if (rc != NGMI_OK) {
        fprintf(stderr, "Error (PreOrder): MI355X sparse LU: %s\n", ngmi_last_error());
        return 1;
    }
    return 0;
}

static int mi_factor(SMPmatrix *Matrix, double Gmin)
{
    KLUmatrix *K = Matrix->SMPkluMatrix;
    mi_smp *m = mi_find(Matrix);
    int rc;
    if (K->KLUmatrixN == 0) return 0;
    if (!m || !m->lu) return 1;
    rc = ngmi_smp_factor(m->lu, K->KLUmatrixAx, K->KLUloadDiagGmin ? Gmin : 0.0);
    if (rc == NGMI_E_SINGULAR) return E_SINGULAR;
    if (rc != NGMI_OK) {
        fprintf(stderr, "Error (Factor): MI355X sparse LU: %s\n", ngmi_last_error());
        return 1;
    }
    return 0;
}

int SMPreorder(SMPmatrix *Matrix, double PivTol, double PivRel, double Gmin)
{
    NG_IGNORE(PivTol);
    NG_IGNORE(PivRel);
    if (!Matrix->CKTkluMODE) {
        spSetReal(Matrix->SPmatrix);
        LoadGmin(Matrix, Gmin);
        return spOrderAndFactor(Matrix->SPmatrix, NULL, PivRel, PivTol, YES);
    }
    return mi_factor(Matrix, Gmin);
}

int SMPluFac(SMPmatrix *Matrix, double PivTol, double Gmin)
{
    NG_IGNORE(PivTol);
    if (!Matrix->CKTkluMODE) {
        spSetReal(Matrix->SPmatrix);
        LoadGmin(Matrix, Gmin);
        return spFactor(Matrix->SPmatrix);
    }
    return mi_factor(Matrix, Gmin);
}

void SMPsolve(SMPmatrix *Matrix, double RHS[], double Spare[])
{
    KLUmatrix *K = Matrix->SMPkluMatrix;
    mi_smp *m;
    unsigned int i;
    NG_IGNORE(Spare);
    if (!Matrix->CKTkluMODE) {
        spSolve(Matrix->SPmatrix, RHS, RHS, NULL, NULL);
        return;
    }
    if (K->KLUmatrixN == 0) return;
    m = mi_find(Matrix);
    if (!m || !m->lu) return;
    m->rhs[0] = 0.0;
    for (i = 0; i < K->KLUmatrixN; i++)
        m->rhs[i + 1] = K->KLUmatrixNodeCollapsingNewToOld[i + 1] != 0
                            ? RHS[K->KLUmatrixNodeCollapsingNewToOld[i + 1]] : 0.0;
    if (ngmi_smp_solve(m->lu, m->rhs) != NGMI_OK)
        fprintf(stderr, "Error (Solve): MI355X sparse LU: %s\n", ngmi_last_error());
    for (i = 0; i < K->KLUmatrixNrhs; i++) RHS[i] = 0;
    for (i = 0; i < K->KLUmatrixN; i++)
        if (K->KLUmatrixNodeCollapsingNewToOld[i + 1] != 0)
            RHS[K->KLUmatrixNodeCollapsingNewToOld[i + 1]] = m->rhs[i + 1];
}
