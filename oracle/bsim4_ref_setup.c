/*
 * ORACLE (test infrastructure only): BSIM4 parameter handling, defaults and
 * temperature/size-dependent parameters, restated from ngspice 46:
 *   b4mpar.c / b4par.c (setters), b4set.c:79-2716 (defaults, node creation),
 *   b4temp.c:69-2404 (temperature and size dependence), b4geo.c (geometry),
 *   b4check.c:355-560 (value clamps).
 */
#include <math.h>
#include <string.h>
#include <stdio.h>
#include "bsim4_ref.h"

#define Kb 1.3806226e-23
#define KboQ 8.617087e-5
#define EPS0 8.85418e-12
#define EPSSI 1.03594e-10
#define PI 3.141592654
#define MAX_EXP 5.834617425e14
#define MIN_EXP 1.713908431e-15
#define EXP_THRESHOLD 34.0
#define Charge_q 1.60219e-19
#define DELTA 1.0E-9
#define CONSTCtoK 273.15
#define CONSTboltz 1.38064852e-23
#define CHARGE 1.6021766208e-19
#define REFTEMP (27.0 + CONSTCtoK)
#define OR_MAX(a, b) ((a) > (b) ? (a) : (b))

static double dexp_clamped(double a)
{   /* b4temp.c:42-50 DEXP */
    if (a > EXP_THRESHOLD) return MAX_EXP * (1.0 + a - EXP_THRESHOLD);
    if (a < -EXP_THRESHOLD) return MIN_EXP;
    return exp(a);
}

#define MP(f) (m->p[BM_##f])
#define MG(f) (m->given[BM_##f])

/* ---------------------------------------------------------------- setters */

typedef struct { const char *name; int field; char kind; } ParDesc;

static const ParDesc model_pars[] = {
#define X(nm, fld, k) { nm, BM_##fld, (#k)[0] },
    BSIM4_MODEL_PARAMS(X)
#undef X
};

int or_bsim4_model_init(ORModel *m)
{
    memset(m, 0, sizeof(*m));
    return 0;
}

int or_bsim4_model_set(ORModel *m, const char *name, double value)
{
    size_t i;
    for (i = 0; i < sizeof(model_pars) / sizeof(model_pars[0]); i++) {
        if (strcmp(model_pars[i].name, name) == 0) {
            int f = model_pars[i].field;
            if (model_pars[i].kind == 'n') {
                if (value != 0.0) { m->p[f] = 1; m->given[f] = 1; }
            } else if (model_pars[i].kind == 'p') {
                if (value != 0.0) { m->p[f] = -1; m->given[f] = 1; }
            } else if (model_pars[i].kind == 'i') {
                m->p[f] = (double)(int)value; m->given[f] = 1;
            } else {
                /* b4mpar.c:3368: tnom is given in Celsius */
                m->p[f] = (f == BM_tnom) ? value + CONSTCtoK : value;
                m->given[f] = 1;
            }
            return 0;
        }
    }
    return -1;
}

static const struct { const char *name; int field; char kind; } inst_pars[] = {
    {"l", BI_l, 'r'}, {"w", BI_w, 'r'}, {"m", BI_m, 'r'}, {"nf", BI_nf, 'r'},
    {"sa", BI_sa, 'r'}, {"sb", BI_sb, 'r'}, {"sd", BI_sd, 'r'},
    {"sca", BI_sca, 'r'}, {"scb", BI_scb, 'r'}, {"scc", BI_scc, 'r'},
    {"sc", BI_sc, 'r'}, {"min", BI_min, 'i'}, {"ad", BI_drainArea, 'r'},
    {"as", BI_sourceArea, 'r'}, {"pd", BI_drainPerimeter, 'r'},
    {"ps", BI_sourcePerimeter, 'r'}, {"nrd", BI_drainSquares, 'r'},
    {"nrs", BI_sourceSquares, 'r'}, {"off", BI_off, 'i'},
    {"rbdb", BI_rbdb, 'r'}, {"rbsb", BI_rbsb, 'r'}, {"rbpb", BI_rbpb, 'r'},
    {"rbps", BI_rbps, 'r'}, {"rbpd", BI_rbpd, 'r'}, {"delvto", BI_delvto, 'r'},
    {"delvt0", BI_delvto, 'r'}, {"mulu0", BI_mulu0, 'r'}, {"xgw", BI_xgw, 'r'},
    {"ngcon", BI_ngcon, 'r'}, {"wnflag", BI_wnflag, 'i'},
    {"trnqsmod", BI_trnqsMod, 'i'}, {"acnqsmod", BI_acnqsMod, 'i'},
    {"rbodymod", BI_rbodyMod, 'i'}, {"rgatemod", BI_rgateMod, 'i'},
    {"geomod", BI_geoMod, 'i'}, {"rgeomod", BI_rgeoMod, 'i'},
    {"mult_i", BI_mult_i, 'r'}, {"mult_q", BI_mult_q, 'r'},
    {"mult_fn", BI_mult_fn, 'r'}, {"icvgs", BI_icVGS, 'r'},
    {"icvds", BI_icVDS, 'r'}, {"icvbs", BI_icVBS, 'r'},
};

int or_bsim4_inst_init(ORInst *h, const ORModel *m)
{
    memset(h, 0, sizeof(*h));
    h->model = m;
    return 0;
}

int or_bsim4_inst_set(ORInst *h, const char *name, double value)
{
    size_t i;
    for (i = 0; i < sizeof(inst_pars) / sizeof(inst_pars[0]); i++) {
        if (strcmp(inst_pars[i].name, name) == 0) {
            h->ip[inst_pars[i].field] = inst_pars[i].kind == 'i' ? (double)(int)value : value;
            h->igiven[inst_pars[i].field] = 1;
            return 0;
        }
    }
    return -1;
}

/* --------------------------------------------------------------- defaults */

static double type_pick(const ORModel *m, double n, double p)
{
    return (m->p[BM_type] >= 0) ? n : p;   /* (type == NMOS) ? n : p */
}

int or_bsim4_model_defaults(ORModel *m, double nomTemp)
{
    if (!MG(type)) MP(type) = 1;   /* b4set.c:79 */

    /* table-driven defaults, in b4set.c order */
#define C_(f, v) if (!MG(f)) MP(f) = (v);
#define Y_(f, src) if (!MG(f)) MP(f) = MP(src);
#define T_(f, n, pv) if (!MG(f)) MP(f) = type_pick(m, n, pv);
#define M_(f, a, b, v1, v2) if (!MG(f)) MP(f) = ((int)MP(mobMod) == a || (int)MP(mobMod) == b) ? (v1) : (v2);
#define N_(f, z) if (!MG(f)) MP(f) = nomTemp;
    BSIM4_MODEL_DEFAULTS(C_, Y_, T_, M_, N_)
#undef C_
#undef Y_
#undef T_
#undef M_
#undef N_

    /* gate-to-S/D tunnelling families (b4set.c:548-580, 1055-1093, 1434-1472,
     * 1814-1852): *sd overrides the s/d pair when given. */
#define SDFAM(q, nv, pv) \
    if (MG(q##igsd)) { MP(q##igs) = MP(q##igd) = MP(q##igsd); } \
    else { MP(q##igsd) = type_pick(m, nv, pv); \
           if (!MG(q##igs)) MP(q##igs) = type_pick(m, nv, pv); \
           if (!MG(q##igd)) MP(q##igd) = type_pick(m, nv, pv); }
    SDFAM(a, 1.36e-2, 9.80e-3)
    SDFAM(b, 1.71e-3, 7.59e-4)
    SDFAM(c, 0.075, 0.03)
#undef SDFAM
#define SDFAMLWP(pre, q) \
    if (!MG(q##igsd) && (MG(q##igs) || MG(q##igd))) { \
        if (!MG(pre##q##igs)) MP(pre##q##igs) = 0.0; \
        if (!MG(pre##q##igd)) MP(pre##q##igd) = 0.0; } \
    else { if (!MG(pre##q##igsd)) MP(pre##q##igsd) = 0.0; \
           MP(pre##q##igs) = MP(pre##q##igd) = MP(pre##q##igsd); }
    SDFAMLWP(l, a) SDFAMLWP(l, b) SDFAMLWP(l, c)
    SDFAMLWP(w, a) SDFAMLWP(w, b) SDFAMLWP(w, c)
    SDFAMLWP(p, a) SDFAMLWP(p, b) SDFAMLWP(p, c)
#undef SDFAMLWP

    /* b4set.c:2013-2019 */
    if (!MG(dlcigd)) MP(dlcigd) = MG(dlcig) ? MP(dlcig) : MP(Lint);
    /* b4set.c:2103-2169 */
    if (!MG(njtsd)) MP(njtsd) = MG(njts) ? MP(njts) : 20.0;
    if (!MG(njtsswd)) MP(njtsswd) = MG(njtssw) ? MP(njtssw) : 20.0;
    if (!MG(njtsswgd)) MP(njtsswgd) = MG(njtsswg) ? MP(njtsswg) : 20.0;
    if (!MG(tnjtsd)) MP(tnjtsd) = MG(tnjts) ? MP(tnjts) : 0.0;
    if (!MG(tnjtsswd)) MP(tnjtsswd) = MG(tnjtssw) ? MP(tnjtssw) : 0.0;
    if (!MG(tnjtsswgd)) MP(tnjtsswgd) = MG(tnjtsswg) ? MP(tnjtsswg) : 0.0;

    m->type = MP(type) >= 0 ? 1 : -1;

    /* selectors this restatement implements */
    if ((int)MP(capMod) != 2) return -2;
    if ((int)MP(rdsMod) != 0) return -3;
    if ((int)MP(trnqsMod) != 0 || (int)MP(acnqsMod) != 0) return -4;
    if ((int)MP(mtrlMod) != 0) return -5;
    if ((int)MP(rgateMod) > 1) return -6;
    if ((int)MP(rbodyMod) > 1) return -7;
    return 0;
}

int or_bsim4_inst_defaults(ORInst *h)
{
    const ORModel *m = h->model;
#define ID(f, v) if (!h->igiven[BI_##f]) h->ip[BI_##f] = (v);
    /* b4set.c:2317-2420 */
    ID(l, 5.0e-6) ID(w, 5.0e-6) ID(m, 1.0) ID(mult_i, 1.0) ID(mult_q, 1.0)
    ID(mult_fn, h->ip[BI_mult_i]) ID(nf, 1.0) ID(min, 0) ID(icVDS, 0.0)
    ID(icVGS, 0.0) ID(icVBS, 0.0) ID(drainArea, 0.0) ID(drainPerimeter, 0.0)
    ID(drainSquares, 1.0) ID(sourceArea, 0.0) ID(sourcePerimeter, 0.0)
    ID(sourceSquares, 1.0) ID(rbdb, MP(rbdb)) ID(rbsb, MP(rbsb))
    ID(rbpb, MP(rbpb)) ID(rbps, MP(rbps)) ID(rbpd, MP(rbpd)) ID(delvto, 0.0)
    ID(mulu0, 1.0) ID(xgw, MP(xgw)) ID(ngcon, MP(ngcon)) ID(dtemp, 0.0)
    ID(rbodyMod, MP(rbodyMod)) ID(rgateMod, MP(rgateMod)) ID(geoMod, MP(geoMod))
    ID(rgeoMod, 0) ID(trnqsMod, MP(trnqsMod)) ID(acnqsMod, MP(acnqsMod))
    ID(sa, 0.0) ID(sb, 0.0) ID(sd, 2 * MP(dmcg)) ID(sca, 0.0) ID(scb, 0.0)
    ID(scc, 0.0) ID(sc, 0.0) ID(off, 0)
#undef ID
    h->rgateMod = (int)h->ip[BI_rgateMod];
    h->rbodyMod = (int)h->ip[BI_rbodyMod];
    h->trnqsMod = (int)h->ip[BI_trnqsMod];
    h->acnqsMod = (int)h->ip[BI_acnqsMod];
    h->geoMod = (int)h->ip[BI_geoMod];
    h->rgeoMod = (int)h->ip[BI_rgeoMod];
    h->off = (int)h->ip[BI_off];
    h->min = (int)h->ip[BI_min];
    h->rbdb = h->ip[BI_rbdb]; h->rbsb = h->ip[BI_rbsb]; h->rbpb = h->ip[BI_rbpb];
    h->rbps = h->ip[BI_rbps]; h->rbpd = h->ip[BI_rbpd];
    if (h->rgateMod > 1 || h->rbodyMod > 1 || h->trnqsMod || h->acnqsMod) return -6;
    return 0;
}

/* -------------------------------------------------------------- geometry */

static void num_finger_diff(double nf, int minSD, double *nuIntD, double *nuEndD,
                            double *nuIntS, double *nuEndS)
{   /* b4geo.c:26-58 */
    int NF = (int)nf;
    if ((NF % 2) != 0) {
        *nuEndD = *nuEndS = 1.0;
        *nuIntD = *nuIntS = 2.0 * OR_MAX((nf - 1.0) / 2.0, 0.0);
    } else if (minSD == 1) {
        *nuEndD = 2.0; *nuIntD = 2.0 * OR_MAX((nf / 2.0 - 1.0), 0.0);
        *nuEndS = 0.0; *nuIntS = nf;
    } else {
        *nuEndD = 0.0; *nuIntD = nf;
        *nuEndS = 2.0; *nuIntS = 2.0 * OR_MAX((nf / 2.0 - 1.0), 0.0);
    }
}

static void pa_eff_geo(double nf, int geo, int minSD, double Weffcj, double DMCG,
                       double DMCI, double DMDG, double *Ps, double *Pd,
                       double *As, double *Ad)
{   /* b4geo.c:61-147 */
    double T0, T1, T2, ADiso, ADsha, ADmer, ASiso, ASsha, ASmer;
    double PDiso, PDsha, PDmer, PSiso, PSsha, PSmer;
    double nuIntD = 0.0, nuEndD = 0.0, nuIntS = 0.0, nuEndS = 0.0;
    if (geo < 9) num_finger_diff(nf, minSD, &nuIntD, &nuEndD, &nuIntS, &nuEndS);
    T0 = DMCG + DMCI; T1 = DMCG + DMCG; T2 = DMDG + DMDG;
    PSiso = PDiso = T0 + T0 + Weffcj; PSsha = PDsha = T1; PSmer = PDmer = T2;
    ASiso = ADiso = T0 * Weffcj; ASsha = ADsha = DMCG * Weffcj; ASmer = ADmer = DMDG * Weffcj;
    switch (geo) {
    case 0: *Ps = nuEndS * PSiso + nuIntS * PSsha; *Pd = nuEndD * PDiso + nuIntD * PDsha;
            *As = nuEndS * ASiso + nuIntS * ASsha; *Ad = nuEndD * ADiso + nuIntD * ADsha; break;
    case 1: *Ps = nuEndS * PSiso + nuIntS * PSsha; *Pd = (nuEndD + nuIntD) * PDsha;
            *As = nuEndS * ASiso + nuIntS * ASsha; *Ad = (nuEndD + nuIntD) * ADsha; break;
    case 2: *Ps = (nuEndS + nuIntS) * PSsha; *Pd = nuEndD * PDiso + nuIntD * PDsha;
            *As = (nuEndS + nuIntS) * ASsha; *Ad = nuEndD * ADiso + nuIntD * ADsha; break;
    case 3: *Ps = (nuEndS + nuIntS) * PSsha; *Pd = (nuEndD + nuIntD) * PDsha;
            *As = (nuEndS + nuIntS) * ASsha; *Ad = (nuEndD + nuIntD) * ADsha; break;
    case 4: *Ps = nuEndS * PSiso + nuIntS * PSsha; *Pd = nuEndD * PDmer + nuIntD * PDsha;
            *As = nuEndS * ASiso + nuIntS * ASsha; *Ad = nuEndD * ADmer + nuIntD * ADsha; break;
    case 5: *Ps = (nuEndS + nuIntS) * PSsha; *Pd = nuEndD * PDmer + nuIntD * PDsha;
            *As = (nuEndS + nuIntS) * ASsha; *Ad = nuEndD * ADmer + nuIntD * ADsha; break;
    case 6: *Ps = nuEndS * PSmer + nuIntS * PSsha; *Pd = nuEndD * PDiso + nuIntD * PDsha;
            *As = nuEndS * ASmer + nuIntS * ASsha; *Ad = nuEndD * ADiso + nuIntD * ADsha; break;
    case 7: *Ps = nuEndS * PSmer + nuIntS * PSsha; *Pd = (nuEndD + nuIntD) * PDsha;
            *As = nuEndS * ASmer + nuIntS * ASsha; *Ad = (nuEndD + nuIntD) * ADsha; break;
    case 8: *Ps = nuEndS * PSmer + nuIntS * PSsha; *Pd = nuEndD * PDmer + nuIntD * PDsha;
            *As = nuEndS * ASmer + nuIntS * ASsha; *Ad = nuEndD * ADmer + nuIntD * ADsha; break;
    case 9: *Ps = PSiso + (nf - 1.0) * PSsha; *Pd = nf * PDsha;
            *As = ASiso + (nf - 1.0) * ASsha; *Ad = nf * ADsha; break;
    case 10: *Ps = nf * PSsha; *Pd = PDiso + (nf - 1.0) * PDsha;
             *As = nf * ASsha; *Ad = ADiso + (nf - 1.0) * ADsha; break;
    default: break;
    }
}

static void rds_end_iso(double Weffcj, double Rsh, double DMCG, double DMCI,
                        double nuEnd, int rgeo, int Type, double *Rend)
{   /* b4geo.c:249-300 */
    if (Type == 1) {
        switch (rgeo) {
        case 1: case 2: case 5:
            *Rend = (nuEnd == 0.0) ? 0.0 : Rsh * DMCG / (Weffcj * nuEnd); break;
        case 3: case 4: case 6:
            *Rend = ((nuEnd == 0.0) || ((DMCG + DMCI) == 0.0)) ? 0.0
                  : Rsh * Weffcj / (3.0 * nuEnd * (DMCG + DMCI)); break;
        default: break;
        }
    } else {
        switch (rgeo) {
        case 1: case 3: case 7:
            *Rend = (nuEnd == 0.0) ? 0.0 : Rsh * DMCG / (Weffcj * nuEnd); break;
        case 2: case 4: case 8:
            *Rend = ((nuEnd == 0.0) || ((DMCG + DMCI) == 0.0)) ? 0.0
                  : Rsh * Weffcj / (3.0 * nuEnd * (DMCG + DMCI)); break;
        default: break;
        }
    }
}

static void rds_end_sha(double Weffcj, double Rsh, double DMCG, double nuEnd,
                        int rgeo, int Type, double *Rend)
{   /* b4geo.c:303-355 */
    if (Type == 1) {
        switch (rgeo) {
        case 1: case 2: case 5:
            *Rend = (nuEnd == 0.0) ? 0.0 : Rsh * DMCG / (Weffcj * nuEnd); break;
        case 3: case 4: case 6:
            *Rend = (nuEnd == 0.0) ? 0.0 : Rsh * Weffcj / (6.0 * nuEnd * DMCG); break;
        default: break;
        }
    } else {
        switch (rgeo) {
        case 1: case 3: case 7:
            *Rend = (nuEnd == 0.0) ? 0.0 : Rsh * DMCG / (Weffcj * nuEnd); break;
        case 2: case 4: case 8:
            *Rend = (nuEnd == 0.0) ? 0.0 : Rsh * Weffcj / (6.0 * nuEnd * DMCG); break;
        default: break;
        }
    }
}

static void rdseff_geo(double nf, int geo, int rgeo, int minSD, double Weffcj,
                       double Rsh, double DMCG, double DMCI, double DMDG, int Type,
                       double *Rtot)
{   /* b4geo.c:150-246 */
    double Rint = 0.0, Rend = 0.0;
    double nuIntD = 0.0, nuEndD = 0.0, nuIntS = 0.0, nuEndS = 0.0;
    if (geo < 9) {
        num_finger_diff(nf, minSD, &nuIntD, &nuEndD, &nuIntS, &nuEndS);
        if (Type == 1) Rint = (nuIntS == 0.0) ? 0.0 : Rsh * DMCG / (Weffcj * nuIntS);
        else Rint = (nuIntD == 0.0) ? 0.0 : Rsh * DMCG / (Weffcj * nuIntD);
    }
    switch (geo) {
    case 0: if (Type == 1) rds_end_iso(Weffcj, Rsh, DMCG, DMCI, nuEndS, rgeo, 1, &Rend);
            else rds_end_iso(Weffcj, Rsh, DMCG, DMCI, nuEndD, rgeo, 0, &Rend); break;
    case 1: if (Type == 1) rds_end_iso(Weffcj, Rsh, DMCG, DMCI, nuEndS, rgeo, 1, &Rend);
            else rds_end_sha(Weffcj, Rsh, DMCG, nuEndD, rgeo, 0, &Rend); break;
    case 2: if (Type == 1) rds_end_sha(Weffcj, Rsh, DMCG, nuEndS, rgeo, 1, &Rend);
            else rds_end_iso(Weffcj, Rsh, DMCG, DMCI, nuEndD, rgeo, 0, &Rend); break;
    case 3: if (Type == 1) rds_end_sha(Weffcj, Rsh, DMCG, nuEndS, rgeo, 1, &Rend);
            else rds_end_sha(Weffcj, Rsh, DMCG, nuEndD, rgeo, 0, &Rend); break;
    case 4: if (Type == 1) rds_end_iso(Weffcj, Rsh, DMCG, DMCI, nuEndS, rgeo, 1, &Rend);
            else Rend = Rsh * DMDG / Weffcj; break;
    case 5: if (Type == 1) rds_end_sha(Weffcj, Rsh, DMCG, nuEndS, rgeo, 1, &Rend);
            else Rend = Rsh * DMDG / (Weffcj * nuEndD); break;
    case 6: if (Type == 1) Rend = Rsh * DMDG / Weffcj;
            else rds_end_iso(Weffcj, Rsh, DMCG, DMCI, nuEndD, rgeo, 0, &Rend); break;
    case 7: if (Type == 1) Rend = Rsh * DMDG / (Weffcj * nuEndS);
            else rds_end_sha(Weffcj, Rsh, DMCG, nuEndD, rgeo, 0, &Rend); break;
    case 8: Rend = Rsh * DMDG / Weffcj; break;
    case 9:
        if (Type == 1) { Rend = 0.5 * Rsh * DMCG / Weffcj;
                         Rint = (nf == 2.0) ? 0.0 : Rsh * DMCG / (Weffcj * (nf - 2.0)); }
        else { Rend = 0.0; Rint = Rsh * DMCG / (Weffcj * nf); }
        break;
    case 10:
        if (Type == 1) { Rend = 0.0; Rint = Rsh * DMCG / (Weffcj * nf); }
        else { Rend = 0.5 * Rsh * DMCG / Weffcj;
               Rint = (nf == 2.0) ? 0.0 : Rsh * DMCG / (Weffcj * (nf - 2.0)); }
        break;
    default: break;
    }
    if (Rint <= 0.0) *Rtot = Rend;
    else if (Rend <= 0.0) *Rtot = Rint;
    else *Rtot = Rint * Rend / (Rint + Rend);
}

int or_bsim4_internal_nodes(const ORInst *h)
{   /* b4set.c:2434-2583 (noise analysis never requested here) */
    const ORModel *m = h->model;
    int bits = 0;
    double Rtot;
    double DMCGeff = MP(dmcg) - MP(dmcgt), DMCIeff = MP(dmci), DMDGeff = MP(dmdg) - MP(dmcgt);
    if ((int)MP(rdsMod) != 0) bits |= 3;
    else if (MP(sheetResistance) > 0) {
        if (h->igiven[BI_drainSquares] && h->ip[BI_drainSquares] > 0) bits |= 1;
        else if (!h->igiven[BI_drainSquares] && h->rgeoMod != 0) {
            rdseff_geo(h->ip[BI_nf], h->geoMod, h->rgeoMod, h->min, h->ip[BI_w],
                       MP(sheetResistance), DMCGeff, DMCIeff, DMDGeff, 0, &Rtot);
            if (Rtot > 0) bits |= 1;
        }
        if (h->igiven[BI_sourceSquares] && h->ip[BI_sourceSquares] > 0) bits |= 2;
        else if (!h->igiven[BI_sourceSquares] && h->rgeoMod != 0) {
            rdseff_geo(h->ip[BI_nf], h->geoMod, h->rgeoMod, h->min, h->ip[BI_w],
                       MP(sheetResistance), DMCGeff, DMCIeff, DMDGeff, 1, &Rtot);
            if (Rtot > 0) bits |= 2;
        }
    }
    if (h->rgateMod > 0) bits |= 4;
    if (h->rgateMod == 3) bits |= 8;
    if (h->rbodyMod == 1 || h->rbodyMod == 2) bits |= 16;
    return bits;
}

/* ------------------------------------------------------------ model temp */

static void diode_vjm(double Nvtm, double Ijth, double Isb, double XExpBV, double *Vjm)
{   /* b4temp.c:53-66 */
    double Tc = XExpBV;
    double Tb = 1.0 + Ijth / Isb - Tc;
    double EVjmovNv = 0.5 * (Tb + sqrt(Tb * Tb + 4.0 * Tc));
    *Vjm = Nvtm * log(EVjmovNv);
}

int or_bsim4_model_temp(ORModel *m, double Temp)
{   /* b4temp.c:90-417 */
    double T0, T1, T2, T3, Eg, Eg0, ni, Tnom, TRatio, Vtm0, delTemp, epssub, toxe, epsrox;
    double CONSTvt0 = CONSTboltz * REFTEMP / CHARGE;
    if (MP(SbulkJctPotential) < 0.1) MP(SbulkJctPotential) = 0.1;
    if (MP(SsidewallJctPotential) < 0.1) MP(SsidewallJctPotential) = 0.1;
    if (MP(SGatesidewallJctPotential) < 0.1) MP(SGatesidewallJctPotential) = 0.1;
    if (MP(DbulkJctPotential) < 0.1) MP(DbulkJctPotential) = 0.1;
    if (MP(DsidewallJctPotential) < 0.1) MP(DsidewallJctPotential) = 0.1;
    if (MP(DGatesidewallJctPotential) < 0.1) MP(DGatesidewallJctPotential) = 0.1;

    {   double chktol = fabs(MP(toxe));
        if (fabs(MP(toxp)) > chktol) chktol = fabs(MP(toxp));
        if (fabs(MP(dtox)) > chktol) chktol = fabs(MP(dtox));
        chktol = chktol * 1e-14;
        if (MG(toxe) && MG(toxp) && MG(dtox) && (fabs(MP(toxe) - (MP(toxp) + MP(dtox))) > chktol)) {
            /* dtox ignored */
        } else if (MG(toxe) && !MG(toxp)) {
            MP(toxp) = MP(toxe) - MP(dtox);
        } else if (!MG(toxe) && MG(toxp)) {
            MP(toxe) = MP(toxp) + MP(dtox);
            if (!MG(toxm)) MP(toxm) = MP(toxe);
        }
    }
    epsrox = MP(epsrox); toxe = MP(toxe); epssub = EPSSI;
    if (!MG(cf)) MP(cf) = 2.0 * epsrox * EPS0 / PI * log(1.0 + 0.4e-6 / toxe);
    m->coxe = epsrox * EPS0 / toxe;
    m->coxp = MP(epsrox) * EPS0 / MP(toxp);
    if (!MG(cgdo)) {
        if (MG(dlc) && (MP(dlc) > 0.0)) MP(cgdo) = MP(dlc) * m->coxe - MP(cgdl);
        else MP(cgdo) = 0.6 * MP(xj) * m->coxe;
    }
    if (!MG(cgso)) {
        if (MG(dlc) && (MP(dlc) > 0.0)) MP(cgso) = MP(dlc) * m->coxe - MP(cgsl);
        else MP(cgso) = 0.6 * MP(xj) * m->coxe;
    }
    if (!MG(cgbo)) MP(cgbo) = 2.0 * MP(dwc) * m->coxe;

    Tnom = MP(tnom);
    TRatio = Temp / Tnom;
    m->vcrit = CONSTvt0 * log(CONSTvt0 / (sqrt(2.0) * 1.0e-14));
    m->factor1 = sqrt(epssub / (epsrox * EPS0) * toxe);
    Vtm0 = m->vtm0 = KboQ * Tnom;
    Eg0 = 1.16 - 7.02e-4 * Tnom * Tnom / (Tnom + 1108.0);
    ni = 1.45e10 * (Tnom / 300.15) * sqrt(Tnom / 300.15) * exp(21.5565981 - Eg0 / (2.0 * Vtm0));
    m->Eg0 = Eg0; m->ni = ni;
    m->vtm = KboQ * Temp;
    Eg = 1.16 - 7.02e-4 * Temp * Temp / (Temp + 1108.0);
    if (Temp != Tnom) {
        T0 = Eg0 / Vtm0 - Eg / m->vtm;
        T1 = log(Temp / Tnom);
        T2 = T0 + MP(SjctTempExponent) * T1;
        T3 = exp(T2 / MP(SjctEmissionCoeff));
        m->SjctTempSatCurDensity = MP(SjctSatCurDensity) * T3;
        m->SjctSidewallTempSatCurDensity = MP(SjctSidewallSatCurDensity) * T3;
        m->SjctGateSidewallTempSatCurDensity = MP(SjctGateSidewallSatCurDensity) * T3;
        T2 = T0 + MP(DjctTempExponent) * T1;
        T3 = exp(T2 / MP(DjctEmissionCoeff));
        m->DjctTempSatCurDensity = MP(DjctSatCurDensity) * T3;
        m->DjctSidewallTempSatCurDensity = MP(DjctSidewallSatCurDensity) * T3;
        m->DjctGateSidewallTempSatCurDensity = MP(DjctGateSidewallSatCurDensity) * T3;
    } else {
        m->SjctTempSatCurDensity = MP(SjctSatCurDensity);
        m->SjctSidewallTempSatCurDensity = MP(SjctSidewallSatCurDensity);
        m->SjctGateSidewallTempSatCurDensity = MP(SjctGateSidewallSatCurDensity);
        m->DjctTempSatCurDensity = MP(DjctSatCurDensity);
        m->DjctSidewallTempSatCurDensity = MP(DjctSidewallSatCurDensity);
        m->DjctGateSidewallTempSatCurDensity = MP(DjctGateSidewallSatCurDensity);
    }
    if (m->SjctTempSatCurDensity < 0.0) m->SjctTempSatCurDensity = 0.0;
    if (m->SjctSidewallTempSatCurDensity < 0.0) m->SjctSidewallTempSatCurDensity = 0.0;
    if (m->SjctGateSidewallTempSatCurDensity < 0.0) m->SjctGateSidewallTempSatCurDensity = 0.0;
    if (m->DjctTempSatCurDensity < 0.0) m->DjctTempSatCurDensity = 0.0;
    if (m->DjctSidewallTempSatCurDensity < 0.0) m->DjctSidewallTempSatCurDensity = 0.0;
    if (m->DjctGateSidewallTempSatCurDensity < 0.0) m->DjctGateSidewallTempSatCurDensity = 0.0;

    delTemp = Temp - MP(tnom);
    T0 = MP(tcj) * delTemp;
    if (T0 >= -1.0) {
        m->SunitAreaTempJctCap = MP(SunitAreaJctCap) * (1.0 + T0);
        m->DunitAreaTempJctCap = MP(DunitAreaJctCap) * (1.0 + T0);
    } else {
        if (MP(SunitAreaJctCap) > 0.0) m->SunitAreaTempJctCap = 0.0;
        if (MP(DunitAreaJctCap) > 0.0) m->DunitAreaTempJctCap = 0.0;
    }
    T0 = MP(tcjsw) * delTemp;
    if (MP(SunitLengthSidewallJctCap) < 0.0) MP(SunitLengthSidewallJctCap) = 0.0;
    if (MP(DunitLengthSidewallJctCap) < 0.0) MP(DunitLengthSidewallJctCap) = 0.0;
    if (T0 >= -1.0) {
        m->SunitLengthSidewallTempJctCap = MP(SunitLengthSidewallJctCap) * (1.0 + T0);
        m->DunitLengthSidewallTempJctCap = MP(DunitLengthSidewallJctCap) * (1.0 + T0);
    } else {
        if (MP(SunitLengthSidewallJctCap) > 0.0) m->SunitLengthSidewallTempJctCap = 0.0;
        if (MP(DunitLengthSidewallJctCap) > 0.0) m->DunitLengthSidewallTempJctCap = 0.0;
    }
    T0 = MP(tcjswg) * delTemp;
    if (T0 >= -1.0) {
        m->SunitLengthGateSidewallTempJctCap = MP(SunitLengthGateSidewallJctCap) * (1.0 + T0);
        m->DunitLengthGateSidewallTempJctCap = MP(DunitLengthGateSidewallJctCap) * (1.0 + T0);
    } else {
        if (MP(SunitLengthGateSidewallJctCap) > 0.0) m->SunitLengthGateSidewallTempJctCap = 0.0;
        if (MP(DunitLengthGateSidewallJctCap) > 0.0) m->DunitLengthGateSidewallTempJctCap = 0.0;
    }
    m->PhiBS = MP(SbulkJctPotential) - MP(tpb) * delTemp;
    if (m->PhiBS < 0.01) m->PhiBS = 0.01;
    m->PhiBD = MP(DbulkJctPotential) - MP(tpb) * delTemp;
    if (m->PhiBD < 0.01) m->PhiBD = 0.01;
    m->PhiBSWS = MP(SsidewallJctPotential) - MP(tpbsw) * delTemp;
    if (m->PhiBSWS < 0.01) m->PhiBSWS = 0.01;
    m->PhiBSWD = MP(DsidewallJctPotential) - MP(tpbsw) * delTemp;
    if (m->PhiBSWD < 0.01) m->PhiBSWD = 0.01;
    m->PhiBSWGS = MP(SGatesidewallJctPotential) - MP(tpbswg) * delTemp;
    if (m->PhiBSWGS < 0.01) m->PhiBSWGS = 0.01;
    m->PhiBSWGD = MP(DGatesidewallJctPotential) - MP(tpbswg) * delTemp;
    if (m->PhiBSWGD < 0.01) m->PhiBSWGD = 0.01;

    if (MP(ijthdfwd) <= 0.0) MP(ijthdfwd) = 0.0;
    if (MP(ijthsfwd) <= 0.0) MP(ijthsfwd) = 0.0;
    if (MP(ijthdrev) <= 0.0) MP(ijthdrev) = 0.0;
    if (MP(ijthsrev) <= 0.0) MP(ijthsrev) = 0.0;
    if ((MP(xjbvd) <= 0.0) && ((int)MP(dioMod) == 2)) MP(xjbvd) = 0.0;
    else if ((MP(xjbvd) < 0.0) && ((int)MP(dioMod) == 0)) MP(xjbvd) = 0.0;
    if (MP(bvd) <= 0.0) MP(bvd) = 0.0;
    if ((MP(xjbvs) <= 0.0) && ((int)MP(dioMod) == 2)) MP(xjbvs) = 0.0;
    else if ((MP(xjbvs) < 0.0) && ((int)MP(dioMod) == 0)) MP(xjbvs) = 0.0;
    if (MP(bvs) <= 0.0) MP(bvs) = 0.0;

    /* b4temp.c:2242-2247: reverse-bias tunnelling emission coefficients */
    T0 = (TRatio - 1.0);
    m->njtsstemp = MP(njts) * (1.0 + MP(tnjts) * T0);
    m->njtsswstemp = MP(njtssw) * (1.0 + MP(tnjtssw) * T0);
    m->njtsswgstemp = MP(njtsswg) * (1.0 + MP(tnjtsswg) * T0);
    m->njtsdtemp = MP(njtsd) * (1.0 + MP(tnjtsd) * T0);
    m->njtsswdtemp = MP(njtsswd) * (1.0 + MP(tnjtsswd) * T0);
    m->njtsswgdtemp = MP(njtsswgd) * (1.0 + MP(tnjtsswgd) * T0);
    if (MP(jtweff) < 0.0) MP(jtweff) = 0.0;
    if (MP(dlcig) < 0.0) MP(dlcig) = 0.0;
    if (MP(dlcigd) < 0.0) MP(dlcigd) = 0.0;
    m->checked = 0;
    return 0;
}

/* --------------------------------------------------------- instance temp */

#define BP(f) (h->bin[BB_##f])

int or_bsim4_inst_temp(ORInst *h, double Temp)
{
    ORModel *m = (ORModel *)h->model;
    double T0, T1, T2, T3, T4, T5, T8, T9, T10, T11, tmp, tmp1, tmp2, Lnew, Wnew;
    double delTemp, TRatio, Inv_L, Inv_W, Inv_LW, Vtm0, Tnom, PowWeffWr;
    double DMCGeff, DMCIeff, DMDGeff, dumPs, dumPd, dumAs, dumAd;
    double Nvtms, Nvtmd, SourceSatCurrent, DrainSatCurrent;
    double Inv_saref, Inv_sbref, Inv_sa, Inv_sb, rho, Ldrn, Wdrn, dvth0_lod;
    double W_tmp, Inv_ODeff, OD_offset, dk2_lod, deta0_lod, kvsat, wlod, sceff;
    double epssub = EPSSI, epsrox = MP(epsrox), toxe = MP(toxe), Eg0 = m->Eg0, ni = m->ni;
    int i;
    const double l = h->ip[BI_l], w = h->ip[BI_w], nf = h->ip[BI_nf];

    Tnom = MP(tnom);
    TRatio = Temp / Tnom;
    delTemp = Temp - MP(tnom);
    Vtm0 = m->vtm0;
    Ldrn = l;
    Wdrn = w / nf;

    /* size-dependent parameters (b4temp.c:436-1692) */
    Lnew = l + MP(xl);
    Wnew = w / nf + MP(xw);
    T0 = pow(Lnew, MP(Lln));
    T1 = pow(Wnew, MP(Lwn));
    tmp1 = MP(Ll) / T0 + MP(Lw) / T1 + MP(Lwl) / (T0 * T1);
    h->dl = MP(Lint) + tmp1;
    tmp2 = MP(Llc) / T0 + MP(Lwc) / T1 + MP(Lwlc) / (T0 * T1);
    h->dlc = MP(dlc) + tmp2;
    T2 = pow(Lnew, MP(Wln));
    T3 = pow(Wnew, MP(Wwn));
    tmp1 = MP(Wl) / T2 + MP(Ww) / T3 + MP(Wwl) / (T2 * T3);
    h->dw = MP(Wint) + tmp1;
    tmp2 = MP(Wlc) / T2 + MP(Wwc) / T3 + MP(Wwlc) / (T2 * T3);
    h->dwc = MP(dwc) + tmp2;
    h->dwj = MP(dwj) + tmp2;
    h->leff = Lnew - 2.0 * h->dl;
    if (h->leff <= 0.0) return -10;
    h->weff = Wnew - 2.0 * h->dw;
    if (h->weff <= 0.0) return -11;
    h->leffCV = Lnew - 2.0 * h->dlc;
    if (h->leffCV <= 0.0) return -12;
    h->weffCV = Wnew - 2.0 * h->dwc;
    if (h->weffCV <= 0.0) return -13;
    h->weffCJ = Wnew - 2.0 * h->dwj;
    if (h->weffCJ <= 0.0) return -14;

    if ((int)MP(binUnit) == 1) {
        Inv_L = 1.0e-6 / h->leff;
        Inv_W = 1.0e-6 / h->weff;
        Inv_LW = 1.0e-12 / (h->leff * h->weff);
    } else {
        Inv_L = 1.0 / h->leff;
        Inv_W = 1.0 / h->weff;
        Inv_LW = 1.0 / (h->leff * h->weff);
    }
#define BINIT(f) h->bin[BB_##f] = MP(f) + MP(l##f) * Inv_L + MP(w##f) * Inv_W + MP(p##f) * Inv_LW;
    BSIM4_BINNED(BINIT)
#undef BINIT
    h->abulkCVfactor = 1.0 + pow((BP(clc) / h->leffCV), BP(cle));

    T0 = (TRatio - 1.0);
    PowWeffWr = pow(h->weffCJ * 1.0e6, BP(wr)) * nf;
    T1 = T2 = T3 = T4 = 0.0;
    BP(ucs) = BP(ucs) * pow(TRatio, BP(ucste));
    if ((int)MP(tempMod) == 0) {
        BP(ua) = BP(ua) + BP(ua1) * T0;
        BP(ub) = BP(ub) + BP(ub1) * T0;
        BP(uc) = BP(uc) + BP(uc1) * T0;
        BP(ud) = BP(ud) + BP(ud1) * T0;
        h->vsattempP = BP(vsat) - BP(at) * T0;
        T10 = BP(prt) * T0;
        h->rds0 = (BP(rdsw) + T10) * nf / PowWeffWr;
        h->rdswminP = (MP(rdswmin) + T10) * nf / PowWeffWr;
    } else {
        if ((int)MP(tempMod) == 3) {
            BP(ua) = BP(ua) * pow(TRatio, BP(ua1));
            BP(ub) = BP(ub) * pow(TRatio, BP(ub1));
            BP(uc) = BP(uc) * pow(TRatio, BP(uc1));
            BP(ud) = BP(ud) * pow(TRatio, BP(ud1));
        } else {
            BP(ua) = BP(ua) * (1.0 + BP(ua1) * delTemp);
            BP(ub) = BP(ub) * (1.0 + BP(ub1) * delTemp);
            BP(uc) = BP(uc) * (1.0 + BP(uc1) * delTemp);
            BP(ud) = BP(ud) * (1.0 + BP(ud1) * delTemp);
        }
        h->vsattempP = BP(vsat) * (1.0 - BP(at) * delTemp);
        T10 = 1.0 + BP(prt) * delTemp;
        h->rds0 = BP(rdsw) * T10 * nf / PowWeffWr;
        h->rdswminP = MP(rdswmin) * T10 * nf / PowWeffWr;
    }
    h->rd0 = T1 / PowWeffWr; h->rdwminP = T2 / PowWeffWr;
    h->rs0 = T3 / PowWeffWr; h->rswminP = T4 / PowWeffWr;
    if (BP(u0) > 1.0) BP(u0) = BP(u0) / 1.0e4;
    T5 = 1.0 - BP(up) * exp(-h->leff / BP(lp));
    h->u0tempP = BP(u0) * T5 * pow(TRatio, BP(ute));
    if (BP(eu) < 0.0) BP(eu) = 0.0;
    if (BP(ucs) < 0.0) BP(ucs) = 0.0;
    BP(vfbsdoff) = BP(vfbsdoff) * (1.0 + BP(tvfbsdoff) * delTemp);
    BP(voff) = BP(voff) * (1.0 + BP(tvoff) * delTemp);
    BP(nfactor) = BP(nfactor) + BP(tnfactor) * delTemp / Tnom;
    BP(voffcv) = BP(voffcv) * (1.0 + BP(tvoffcv) * delTemp);
    BP(eta0) = BP(eta0) + BP(teta0) * delTemp / Tnom;
    if (MG(vtl) && (MP(vtl) > 0.0)) {
        h->lc = (MP(lc) < 0.0) ? 0.0 : MP(lc);
        T0 = h->leff / (BP(xn) * h->leff + h->lc);
        h->tfactor = (1.0 - T0) / (1.0 + T0);
    }
    h->cgdoP = (MP(cgdo) + BP(cf)) * h->weffCV;
    h->cgsoP = (MP(cgso) + BP(cf)) * h->weffCV;
    h->cgboP = MP(cgbo) * h->leffCV * nf;
    if (!MG(ndep) && MG(gamma1)) {
        T0 = BP(gamma1) * m->coxe;
        BP(ndep) = 3.01248e22 * T0 * T0;
    }
    h->phi = Vtm0 * log(BP(ndep) / ni) + BP(phin) + 0.4;
    if (h->phi <= 0.0) return -15;
    h->sqrtPhi = sqrt(h->phi);
    h->phis3 = h->sqrtPhi * h->phi;
    h->Xdep0 = sqrt(2.0 * epssub / (Charge_q * BP(ndep) * 1.0e6)) * h->sqrtPhi;
    h->sqrtXdep0 = sqrt(h->Xdep0);
    h->litl = sqrt(3.0 * 3.9 / epsrox * BP(xj) * toxe);
    h->vbi = Vtm0 * log(BP(nsd) * BP(ndep) / (ni * ni));
    if (BP(ngate) > 0.0) h->vfbsd = Vtm0 * log(BP(ngate) / BP(nsd));
    else h->vfbsd = 0.0;
    h->cdep0 = sqrt(Charge_q * epssub * BP(ndep) * 1.0e6 / 2.0 / h->phi);
    h->ToxRatio = exp(BP(ntox) * log(MP(toxref) / toxe)) / toxe / toxe;
    h->ToxRatioEdge = exp(BP(ntox) * log(MP(toxref) / (toxe * BP(poxedge))))
                    / toxe / toxe / BP(poxedge) / BP(poxedge);
    h->Aechvb = (m->type == 1) ? 4.97232e-7 : 3.42537e-7;
    h->Bechvb = (m->type == 1) ? 7.45669e11 : 1.16645e12;
    /* version >= 4.8.1 branch (b4temp.c:1384-1399) */
    h->AechvbEdgeS = h->Aechvb * h->weff * MP(dlcig) * h->ToxRatioEdge;
    h->AechvbEdgeD = h->Aechvb * h->weff * MP(dlcigd) * h->ToxRatioEdge;
    h->BechvbEdge = -h->Bechvb * toxe * BP(poxedge);
    h->Aechvb *= h->weff * h->leff * h->ToxRatio;
    h->Bechvb *= -toxe;
    h->mstar = 0.5 + atan(BP(minv)) / PI;
    h->mstarcv = 0.5 + atan(BP(minvcv)) / PI;
    h->voffcbn = BP(voff) + MP(voffl) / h->leff;
    h->voffcbncv = BP(voffcv) + MP(voffcvl) / h->leff;
    h->ldeb = sqrt(epssub * Vtm0 / (Charge_q * BP(ndep) * 1.0e6)) / 3.0;
    BP(acde) *= pow((BP(ndep) / 2.0e16), -0.25);

    if (MG(k1) || MG(k2)) {
        if (!MG(k1)) BP(k1) = 0.53;
        if (!MG(k2)) BP(k2) = -0.0186;
    } else {
        if (!MG(vbx)) BP(vbx) = h->phi - 7.7348e-4 * BP(ndep) * BP(xt) * BP(xt);
        if (BP(vbx) > 0.0) BP(vbx) = -BP(vbx);
        if (BP(vbm) > 0.0) BP(vbm) = -BP(vbm);
        if (!MG(gamma1)) BP(gamma1) = 5.753e-12 * sqrt(BP(ndep)) / m->coxe;
        if (!MG(gamma2)) BP(gamma2) = 5.753e-12 * sqrt(BP(nsub)) / m->coxe;
        T0 = BP(gamma1) - BP(gamma2);
        T1 = sqrt(h->phi - BP(vbx)) - h->sqrtPhi;
        T2 = sqrt(h->phi * (h->phi - BP(vbm))) - h->phi;
        BP(k2) = T0 * T1 / (2.0 * T2 + BP(vbm));
        BP(k1) = BP(gamma2) - 2.0 * BP(k2) * sqrt(h->phi - BP(vbm));
    }
    if (!MG(vfb)) {
        if (MG(vth0)) BP(vfb) = m->type * BP(vth0) - h->phi - BP(k1) * h->sqrtPhi;
        else BP(vfb) = -1.0;
    }
    if (!MG(vth0)) BP(vth0) = m->type * (BP(vfb) + h->phi + BP(k1) * h->sqrtPhi);
    h->k1ox = BP(k1) * toxe / MP(toxm);

    tmp = sqrt(epssub / (epsrox * EPS0) * toxe * h->Xdep0);
    T0 = BP(dsub) * h->leff / tmp;
    if (T0 < EXP_THRESHOLD) {
        T1 = exp(T0); T2 = T1 - 1.0; T3 = T2 * T2; T4 = T3 + 2.0 * T1 * MIN_EXP;
        h->theta0vb0 = T1 / T4;
    } else h->theta0vb0 = 1.0 / (MAX_EXP - 2.0);
    T0 = BP(drout) * h->leff / tmp;
    if (T0 < EXP_THRESHOLD) {
        T1 = exp(T0); T2 = T1 - 1.0; T3 = T2 * T2; T4 = T3 + 2.0 * T1 * MIN_EXP;
        T5 = T1 / T4;
    } else T5 = 1.0 / (MAX_EXP - 2.0);
    h->thetaRout = BP(pdibl1) * T5 + BP(pdibl2);

    tmp = sqrt(h->Xdep0);
    tmp1 = h->vbi - h->phi;
    tmp2 = m->factor1 * tmp;
    T0 = BP(dvt1w) * h->weff * h->leff / tmp2;
    if (T0 < EXP_THRESHOLD) {
        T1 = exp(T0); T2 = T1 - 1.0; T3 = T2 * T2; T4 = T3 + 2.0 * T1 * MIN_EXP;
        T8 = T1 / T4;
    } else T8 = 1.0 / (MAX_EXP - 2.0);
    T0 = BP(dvt0w) * T8;
    T8 = T0 * tmp1;
    T0 = BP(dvt1) * h->leff / tmp2;
    if (T0 < EXP_THRESHOLD) {
        T1 = exp(T0); T2 = T1 - 1.0; T3 = T2 * T2; T4 = T3 + 2.0 * T1 * MIN_EXP;
        T9 = T1 / T4;
    } else T9 = 1.0 / (MAX_EXP - 2.0);
    T9 = BP(dvt0) * T9 * tmp1;
    T4 = toxe * h->phi / (h->weff + BP(w0));
    T0 = sqrt(1.0 + BP(lpe0) / h->leff);
    T3 = 0.0;
    if (((int)MP(tempMod) == 1) || ((int)MP(tempMod) == 0))
        T3 = (BP(kt1) + BP(kt1l) / h->leff) * (TRatio - 1.0);
    if (((int)MP(tempMod) == 2) || ((int)MP(tempMod) == 3))
        T3 = -BP(kt1) * (TRatio - 1.0);
    T5 = h->k1ox * (T0 - 1.0) * h->sqrtPhi + T3;
    h->vfbzbfactor = -T8 - T9 + BP(k3) * T4 + T5 - h->phi - BP(k1) * h->sqrtPhi;

    /* stress effect (b4temp.c:1577-1604) */
    wlod = MP(wlod);
    if (MP(wlod) < 0.0) wlod = 0.0;
    T0 = pow(Lnew, MP(llodku0));
    W_tmp = Wnew + wlod;
    T1 = pow(W_tmp, MP(wlodku0));
    tmp1 = MP(lku0) / T0 + MP(wku0) / T1 + MP(pku0) / (T0 * T1);
    h->ku0 = 1.0 + tmp1;
    T0 = pow(Lnew, MP(llodvth));
    T1 = pow(W_tmp, MP(wlodvth));
    tmp1 = MP(lkvth0) / T0 + MP(wkvth0) / T1 + MP(pkvth0) / (T0 * T1);
    h->kvth0 = 1.0 + tmp1;
    h->kvth0 = sqrt(h->kvth0 * h->kvth0 + DELTA);
    T0 = (TRatio - 1.0);
    h->ku0temp = h->ku0 * (1.0 + MP(tku0) * T0) + DELTA;
    Inv_saref = 1.0 / (MP(saref) + 0.5 * Ldrn);
    Inv_sbref = 1.0 / (MP(sbref) + 0.5 * Ldrn);
    h->inv_od_ref = Inv_saref + Inv_sbref;
    h->rho_ref = MP(ku0) / h->ku0temp * h->inv_od_ref;

    if ((int)MP(mobMod) == 3) {   /* b4temp.c:1606-1652 */
        double lt1, Theta0, n0, tmp3;
        lt1 = m->factor1 * h->sqrtXdep0;
        T0 = BP(dvt1) * h->leff / lt1;
        if (T0 < EXP_THRESHOLD) {
            T1 = exp(T0); T2 = T1 - 1.0; T3 = T2 * T2; T4 = T3 + 2.0 * T1 * MIN_EXP;
            Theta0 = T1 / T4;
        } else Theta0 = 1.0 / (MAX_EXP - 2.0);
        tmp1 = epssub / h->Xdep0;
        tmp2 = BP(nfactor) * tmp1;
        tmp3 = (tmp2 + BP(cdsc) * Theta0 + BP(cit)) / m->coxe;
        if (tmp3 >= -0.5) n0 = 1.0 + tmp3;
        else { T0 = 1.0 / (3.0 + 8.0 * tmp3); n0 = (1.0 + 3.0 * tmp3) * T0; }
        T0 = n0 * m->vtm;
        T1 = h->voffcbn;
        T2 = T1 / T0;
        if (T2 < -EXP_THRESHOLD) { T3 = m->coxe * MIN_EXP / h->cdep0; T4 = h->mstar + T3 * n0; }
        else if (T2 > EXP_THRESHOLD) { T3 = m->coxe * MAX_EXP / h->cdep0; T4 = h->mstar + T3 * n0; }
        else { T3 = exp(T2) * m->coxe / h->cdep0; T4 = h->mstar + T3 * n0; }
        h->VgsteffVth = T0 * log(2.0) / T4;
    }
    T0 = -BP(dvtp3) * log(h->leff);
    T1 = dexp_clamped(T0);
    h->dvtp2factor = BP(dvtp5) + BP(dvtp2) * T1;

    /* b4check.c value clamps (the first instance of a model is the one checked) */
    if (!m->checked) {
        m->checked = 1;
        if (BP(ckappas) < 0.02) BP(ckappas) = 0.02;
        if (BP(ckappad) < 0.02) BP(ckappad) = 0.02;
        if ((int)MP(paramChk) == 1) {
            if (BP(a2) < 0.01) BP(a2) = 0.01;
            else if (BP(a2) > 1.0) { BP(a2) = 1.0; BP(a1) = 0.0; }
            if (BP(prwg) < 0.0) BP(prwg) = 0.0;
            if (BP(rdsw) < 0.0) { BP(rdsw) = 0.0; h->rds0 = 0.0; }
            if (h->rds0 < 0.0) h->rds0 = 0.0;
            if (h->rdswminP < 0.0) h->rdswminP = 0.0;
            if (MP(cgdo) < 0.0) MP(cgdo) = 0.0;
            if (MP(cgso) < 0.0) MP(cgso) = 0.0;
            if (MP(cgbo) < 0.0) MP(cgbo) = 0.0;
        }
    }

    /* stress effect, instance part (b4temp.c:1694-1738) */
    if ((h->ip[BI_sa] > 0.0) && (h->ip[BI_sb] > 0.0) &&
        ((nf == 1.0) || ((nf > 1.0) && (h->ip[BI_sd] > 0.0)))) {
        Inv_sa = 0; Inv_sb = 0;
        kvsat = MP(kvsat);
        if (MP(kvsat) < -1.0) kvsat = -1.0;
        if (MP(kvsat) > 1.0) kvsat = 1.0;
        for (i = 0; i < nf; i++) {
            T0 = 1.0 / nf / (h->ip[BI_sa] + 0.5 * Ldrn + i * (h->ip[BI_sd] + Ldrn));
            T1 = 1.0 / nf / (h->ip[BI_sb] + 0.5 * Ldrn + i * (h->ip[BI_sd] + Ldrn));
            Inv_sa += T0;
            Inv_sb += T1;
        }
        Inv_ODeff = Inv_sa + Inv_sb;
        rho = MP(ku0) / h->ku0temp * Inv_ODeff;
        T0 = (1.0 + rho) / (1.0 + h->rho_ref);
        h->u0temp = h->u0tempP * T0;
        T1 = (1.0 + kvsat * rho) / (1.0 + kvsat * h->rho_ref);
        h->vsattemp = h->vsattempP * T1;
        OD_offset = Inv_ODeff - h->inv_od_ref;
        dvth0_lod = MP(kvth0) / h->kvth0 * OD_offset;
        dk2_lod = MP(stk2) / pow(h->kvth0, MP(lodk2)) * OD_offset;
        deta0_lod = MP(steta0) / pow(h->kvth0, MP(lodeta0)) * OD_offset;
        h->vth0 = BP(vth0) + dvth0_lod;
        h->eta0 = BP(eta0) + deta0_lod;
        h->k2 = BP(k2) + dk2_lod;
    } else {
        h->u0temp = h->u0tempP;
        h->vth0 = BP(vth0);
        h->vsattemp = h->vsattempP;
        h->eta0 = BP(eta0);
        h->k2 = BP(k2);
    }

    /* well proximity effect (b4temp.c:1741-1790) */
    if ((int)MP(wpemod)) {
        if (!h->igiven[BI_sca] && !h->igiven[BI_scb] && !h->igiven[BI_scc]) {
            if (h->igiven[BI_sc] && (h->ip[BI_sc] > 0.0)) {
                double sc = h->ip[BI_sc];
                T1 = sc + Wdrn;
                T2 = 1.0 / MP(scref);
                h->ip[BI_sca] = MP(scref) * MP(scref) / (sc * T1);
                h->ip[BI_scb] = ((0.1 * sc + 0.01 * MP(scref)) * exp(-10.0 * sc * T2)
                               - (0.1 * T1 + 0.01 * MP(scref)) * exp(-10.0 * T1 * T2)) / Wdrn;
                h->ip[BI_scc] = ((0.05 * sc + 0.0025 * MP(scref)) * exp(-20.0 * sc * T2)
                               - (0.05 * T1 + 0.0025 * MP(scref)) * exp(-20.0 * T1 * T2)) / Wdrn;
            }
        }
        if (h->ip[BI_sca] < 0.0) h->ip[BI_sca] = 0.0;
        if (h->ip[BI_scb] < 0.0) h->ip[BI_scb] = 0.0;
        if (h->ip[BI_scc] < 0.0) h->ip[BI_scc] = 0.0;
        if (h->ip[BI_sc] < 0.0) h->ip[BI_sc] = 0.0;
        sceff = h->ip[BI_sca] + MP(web) * h->ip[BI_scb] + MP(wec) * h->ip[BI_scc];
        h->vth0 += BP(kvth0we) * sceff;
        h->k2 += BP(k2we) * sceff;
        T3 = 1.0 + BP(ku0we) * sceff;
        if (T3 <= 0.0) T3 = 0.0;
        h->u0temp *= T3;
    }

    /* delvto and mulu0 (b4temp.c:1792-1797); u0temp is re-derived from the
     * size-dependent value, as the reference does */
    h->vth0 += h->ip[BI_delvto];
    h->vfb = BP(vfb) + m->type * h->ip[BI_delvto];
    h->u0temp = h->u0tempP * h->ip[BI_mulu0];

    T3 = m->type * h->vth0 - h->vfb - h->phi;
    T4 = T3 + T3;
    T5 = 2.5 * T3;
    h->vtfbphi1 = (m->type == 1) ? T4 : T5;
    if (h->vtfbphi1 < 0.0) h->vtfbphi1 = 0.0;
    h->vtfbphi2 = 4.0 * T3;
    if (h->vtfbphi2 < 0.0) h->vtfbphi2 = 0.0;
    if (h->k2 < 0.0) {
        T0 = 0.5 * BP(k1) / h->k2;
        h->vbsc = 0.9 * (h->phi - T0 * T0);
        if (h->vbsc > -3.0) h->vbsc = -3.0;
        else if (h->vbsc < -30.0) h->vbsc = -30.0;
    } else h->vbsc = -30.0;
    if (h->vbsc > BP(vbm)) h->vbsc = BP(vbm);
    h->k2ox = h->k2 * toxe / MP(toxm);
    h->vfbzb = h->vfbzbfactor + m->type * h->vth0;
    h->cgso = h->cgsoP;
    h->cgdo = h->cgdoP;

    /* body resistances (b4temp.c:1805-1920); rbodyMod 2 is not supported */
    if (h->rbodyMod == 1) {
        h->grbdb = (h->rbdb < 1.0e-3) ? 1.0e3 : MP(gbmin) + 1.0 / h->rbdb;
        h->grbpb = (h->rbpb < 1.0e-3) ? 1.0e3 : MP(gbmin) + 1.0 / h->rbpb;
        h->grbps = (h->rbps < 1.0e-3) ? 1.0e3 : MP(gbmin) + 1.0 / h->rbps;
        h->grbsb = (h->rbsb < 1.0e-3) ? 1.0e3 : MP(gbmin) + 1.0 / h->rbsb;
        h->grbpd = (h->rbpd < 1.0e-3) ? 1.0e3 : MP(gbmin) + 1.0 / h->rbpd;
    }

    /* gate resistance (b4temp.c:1926-1936) */
    h->grgeltd = MP(rshg) * (h->ip[BI_xgw] + h->weffCJ / 3.0 / h->ip[BI_ngcon])
               / (h->ip[BI_ngcon] * nf * (Lnew - MP(xgl)));
    if (h->grgeltd > 0.0) h->grgeltd = 1.0 / h->grgeltd;
    else h->grgeltd = 1.0e3;

    /* S/D perimeters and areas (b4temp.c:1938-2049) */
    DMCGeff = MP(dmcg) - MP(dmcgt);
    DMCIeff = MP(dmci);
    DMDGeff = MP(dmdg) - MP(dmcgt);
    if (h->igiven[BI_sourcePerimeter]) {
        if (h->ip[BI_sourcePerimeter] == 0.0) h->Pseff = 0.0;
        else if (h->ip[BI_sourcePerimeter] < 0.0) h->Pseff = 0.0;
        else if ((int)MP(perMod) == 0) h->Pseff = h->ip[BI_sourcePerimeter];
        else h->Pseff = h->ip[BI_sourcePerimeter] - h->weffCJ * nf;
    } else
        pa_eff_geo(nf, h->geoMod, h->min, h->weffCJ, DMCGeff, DMCIeff, DMDGeff,
                   &h->Pseff, &dumPd, &dumAs, &dumAd);
    if (h->Pseff < 0.0) h->Pseff = 0.0;
    if (h->igiven[BI_drainPerimeter]) {
        if (h->ip[BI_drainPerimeter] == 0.0) h->Pdeff = 0.0;
        else if (h->ip[BI_drainPerimeter] < 0.0) h->Pdeff = 0.0;
        else if ((int)MP(perMod) == 0) h->Pdeff = h->ip[BI_drainPerimeter];
        else h->Pdeff = h->ip[BI_drainPerimeter] - h->weffCJ * nf;
    } else
        pa_eff_geo(nf, h->geoMod, h->min, h->weffCJ, DMCGeff, DMCIeff, DMDGeff,
                   &dumPs, &h->Pdeff, &dumAs, &dumAd);
    if (h->Pdeff < 0.0) h->Pdeff = 0.0;
    if (h->igiven[BI_sourceArea]) h->Aseff = h->ip[BI_sourceArea];
    else pa_eff_geo(nf, h->geoMod, h->min, h->weffCJ, DMCGeff, DMCIeff, DMDGeff,
                    &dumPs, &dumPd, &h->Aseff, &dumAd);
    if (h->Aseff < 0.0) h->Aseff = 0.0;
    if (h->igiven[BI_drainArea]) h->Adeff = h->ip[BI_drainArea];
    else pa_eff_geo(nf, h->geoMod, h->min, h->weffCJ, DMCGeff, DMCIeff, DMDGeff,
                    &dumPs, &dumPd, &dumAs, &h->Adeff);
    if (h->Adeff < 0.0) h->Adeff = 0.0;

    /* S/D series conductances (b4temp.c:2051-2104) */
    if (h->node[N_sp] != h->node[N_s]) {
        h->sourceConductance = 0.0;
        if (h->igiven[BI_sourceSquares])
            h->sourceConductance = MP(sheetResistance) * h->ip[BI_sourceSquares];
        else if (h->rgeoMod > 0)
            rdseff_geo(nf, h->geoMod, h->rgeoMod, h->min, h->weffCJ, MP(sheetResistance),
                       DMCGeff, DMCIeff, DMDGeff, 1, &h->sourceConductance);
        if (h->sourceConductance > 0.0) h->sourceConductance = 1.0 / h->sourceConductance;
        else h->sourceConductance = 1.0e3;
    } else h->sourceConductance = 0.0;
    if (h->node[N_dp] != h->node[N_d]) {
        h->drainConductance = 0.0;
        if (h->igiven[BI_drainSquares])
            h->drainConductance = MP(sheetResistance) * h->ip[BI_drainSquares];
        else if (h->rgeoMod > 0)
            rdseff_geo(nf, h->geoMod, h->rgeoMod, h->min, h->weffCJ, MP(sheetResistance),
                       DMCGeff, DMCIeff, DMDGeff, 0, &h->drainConductance);
        if (h->drainConductance > 0.0) h->drainConductance = 1.0 / h->drainConductance;
        else h->drainConductance = 1.0e3;
    } else h->drainConductance = 0.0;

    /* junction diode limits (b4temp.c:2107-2236) */
    Nvtms = m->vtm * MP(SjctEmissionCoeff);
    if ((h->Aseff <= 0.0) && (h->Pseff <= 0.0)) SourceSatCurrent = 0.0;
    else SourceSatCurrent = h->Aseff * m->SjctTempSatCurDensity
                          + h->Pseff * m->SjctSidewallTempSatCurDensity
                          + h->weffCJ * nf * m->SjctGateSidewallTempSatCurDensity;
    if (SourceSatCurrent > 0.0) {
        switch ((int)MP(dioMod)) {
        case 0:
            if ((MP(bvs) / Nvtms) > EXP_THRESHOLD) h->XExpBVS = MP(xjbvs) * MIN_EXP;
            else h->XExpBVS = MP(xjbvs) * exp(-MP(bvs) / Nvtms);
            break;
        case 1:
            diode_vjm(Nvtms, MP(ijthsfwd), SourceSatCurrent, 0.0, &h->vjsmFwd);
            h->IVjsmFwd = SourceSatCurrent * exp(h->vjsmFwd / Nvtms);
            break;
        case 2:
            if ((MP(bvs) / Nvtms) > EXP_THRESHOLD) { h->XExpBVS = MP(xjbvs) * MIN_EXP; tmp = MIN_EXP; }
            else { h->XExpBVS = exp(-MP(bvs) / Nvtms); tmp = h->XExpBVS; h->XExpBVS *= MP(xjbvs); }
            diode_vjm(Nvtms, MP(ijthsfwd), SourceSatCurrent, h->XExpBVS, &h->vjsmFwd);
            T0 = exp(h->vjsmFwd / Nvtms);
            h->IVjsmFwd = SourceSatCurrent * (T0 - h->XExpBVS / T0 + h->XExpBVS - 1.0);
            h->SslpFwd = SourceSatCurrent * (T0 + h->XExpBVS / T0) / Nvtms;
            T2 = MP(ijthsrev) / SourceSatCurrent;
            if (T2 < 1.0) T2 = 10.0;
            h->vjsmRev = -MP(bvs) - Nvtms * log((T2 - 1.0) / MP(xjbvs));
            T1 = MP(xjbvs) * exp(-(MP(bvs) + h->vjsmRev) / Nvtms);
            h->IVjsmRev = SourceSatCurrent * (1.0 + T1);
            h->SslpRev = -SourceSatCurrent * T1 / Nvtms;
            break;
        default: break;
        }
    }
    Nvtmd = m->vtm * MP(DjctEmissionCoeff);
    if ((h->Adeff <= 0.0) && (h->Pdeff <= 0.0)) DrainSatCurrent = 0.0;
    else DrainSatCurrent = h->Adeff * m->DjctTempSatCurDensity
                         + h->Pdeff * m->DjctSidewallTempSatCurDensity
                         + h->weffCJ * nf * m->DjctGateSidewallTempSatCurDensity;
    if (DrainSatCurrent > 0.0) {
        switch ((int)MP(dioMod)) {
        case 0:
            if ((MP(bvd) / Nvtmd) > EXP_THRESHOLD) h->XExpBVD = MP(xjbvd) * MIN_EXP;
            else h->XExpBVD = MP(xjbvd) * exp(-MP(bvd) / Nvtmd);
            break;
        case 1:
            diode_vjm(Nvtmd, MP(ijthdfwd), DrainSatCurrent, 0.0, &h->vjdmFwd);
            h->IVjdmFwd = DrainSatCurrent * exp(h->vjdmFwd / Nvtmd);
            break;
        case 2:
            if ((MP(bvd) / Nvtmd) > EXP_THRESHOLD) { h->XExpBVD = MP(xjbvd) * MIN_EXP; tmp = MIN_EXP; }
            else { h->XExpBVD = exp(-MP(bvd) / Nvtmd); tmp = h->XExpBVD; h->XExpBVD *= MP(xjbvd); }
            diode_vjm(Nvtmd, MP(ijthdfwd), DrainSatCurrent, h->XExpBVD, &h->vjdmFwd);
            T0 = exp(h->vjdmFwd / Nvtmd);
            h->IVjdmFwd = DrainSatCurrent * (T0 - h->XExpBVD / T0 + h->XExpBVD - 1.0);
            h->DslpFwd = DrainSatCurrent * (T0 + h->XExpBVD / T0) / Nvtmd;
            T2 = MP(ijthdrev) / DrainSatCurrent;
            if (T2 < 1.0) T2 = 10.0;
            h->vjdmRev = -MP(bvd) - Nvtmd * log((T2 - 1.0) / MP(xjbvd));
            T1 = MP(xjbvd) * exp(-(MP(bvd) + h->vjdmRev) / Nvtmd);
            h->IVjdmRev = DrainSatCurrent * (1.0 + T1);
            h->DslpRev = -DrainSatCurrent * T1 / Nvtmd;
            break;
        default: break;
        }
    }
    (void)tmp;

    /* trap-assisted tunnelling saturation currents (b4temp.c:2239-2275) */
    T0 = (TRatio - 1.0);
    {
        double T7 = Eg0 / m->vtm * T0, Ta, Tb, Tc, Td, Te, Tf;
        Ta = dexp_clamped(MP(xtss) * T7);
        Tb = dexp_clamped(MP(xtsd) * T7);
        Tc = dexp_clamped(MP(xtssws) * T7);
        Td = dexp_clamped(MP(xtsswd) * T7);
        Te = dexp_clamped(MP(xtsswgs) * T7);
        Tf = dexp_clamped(MP(xtsswgd) * T7);
        T11 = sqrt(MP(jtweff) / h->weffCJ) + 1.0;
        T10 = h->weffCJ * nf;
        h->SjctTempRevSatCur = Ta * h->Aseff * MP(jtss);
        h->DjctTempRevSatCur = Tb * h->Adeff * MP(jtsd);
        h->SswTempRevSatCur = Tc * h->Pseff * MP(jtssws);
        h->DswTempRevSatCur = Td * h->Pdeff * MP(jtsswd);
        h->SswgTempRevSatCur = Te * T10 * T11 * MP(jtsswgs);
        h->DswgTempRevSatCur = Tf * T10 * T11 * MP(jtsswgd);
    }
    h->toxp = MP(toxp);
    h->coxp = m->coxp;
    (void)Eg0;
    return 0;
}
