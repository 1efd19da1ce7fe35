/*
 * Host side of the BSIM4 4.8.3 device: model cards, defaults, temperature and
 * size dependence, resolved once per instance into the evaluation table the
 * HIP kernels read (bsim4_fields.h).
 *
 * Follows src/spicelib/devices/bsim4/b4set.c:79-2420 (defaults),
 * b4temp.c:69-2404 (temperature / size dependence), b4geo.c (S/D geometry)
 * and the value clamps of b4check.c:355-560.  The engine supports the
 * selectors of the standard-cell transient path: capMod 2, rdsMod 0,
 * rgateMod 0, rbodyMod 0, trnqsMod/acnqsMod 0, mtrlMod 0, and no S/D
 * series-resistance nodes; other cards are rejected by ngmi_b4_check().
 */
#include <cmath>
#include <cstring>
#include <cstdio>
#include <initializer_list>
#include "bsim4_setup.h"

namespace ngmi {

static const double KboQ = 8.617087e-5;
static const double EPS0 = 8.85418e-12;
static const double EPSSI = 1.03594e-10;
static const double PI = 3.141592654;
static const double MAX_EXP = 5.834617425e14;
static const double MIN_EXP = 1.713908431e-15;
static const double EXP_THRESHOLD = 34.0;
static const double Charge_q = 1.60219e-19;
static const double DELTA = 1.0E-9;
static const double CtoK = 273.15;

static double dexp_clamped(double a)
{
    if (a > EXP_THRESHOLD) return MAX_EXP * (1.0 + a - EXP_THRESHOLD);
    if (a < -EXP_THRESHOLD) return MIN_EXP;
    return std::exp(a);
}

struct ParDesc { const char *name; int field; char kind; };
static const ParDesc kModelPars[] = {
#define X(nm, fld, k) { nm, BM_##fld, (#k)[0] },
    BSIM4_MODEL_PARAMS(X)
#undef X
};

#define MP(f) (m.p[BM_##f])
#define MG(f) (m.given[BM_##f])

int B4Model::set(const char *name, double value)
{
    for (const ParDesc &d : kModelPars) {
        if (std::strcmp(d.name, name) != 0) continue;
        if (d.kind == 'n') { if (value != 0.0) { p[d.field] = 1; given[d.field] = 1; } }
        else if (d.kind == 'p') { if (value != 0.0) { p[d.field] = -1; given[d.field] = 1; } }
        else if (d.kind == 'i') { p[d.field] = (double)(int)value; given[d.field] = 1; }
        else { p[d.field] = (d.field == BM_tnom) ? value + CtoK : value; given[d.field] = 1; }
        return 0;
    }
    return -1;
}

/* b4set.c:79-2300 */
static void apply_defaults(B4Model &m, double nomTemp)
{
    if (!MG(type)) MP(type) = 1;
    auto tp = [&](double n, double pv) { return MP(type) >= 0 ? n : pv; };
#define C_(f, v) if (!MG(f)) MP(f) = (v);
#define Y_(f, src) if (!MG(f)) MP(f) = MP(src);
#define T_(f, n, pv) if (!MG(f)) MP(f) = tp(n, pv);
#define M_(f, a, b, v1, v2) if (!MG(f)) MP(f) = ((int)MP(mobMod) == a || (int)MP(mobMod) == b) ? (v1) : (v2);
#define N_(f, z) if (!MG(f)) MP(f) = nomTemp;
    BSIM4_MODEL_DEFAULTS(C_, Y_, T_, M_, N_)
#undef C_
#undef Y_
#undef T_
#undef M_
#undef N_
    /* S/D edge tunnelling parameter families, b4set.c:548-580 and their
     * length/width/product variants */
#define SDFAM(q, nv, pv) \
    if (MG(q##igsd)) { MP(q##igs) = MP(q##igd) = MP(q##igsd); } \
    else { MP(q##igsd) = tp(nv, pv); \
           if (!MG(q##igs)) MP(q##igs) = tp(nv, pv); \
           if (!MG(q##igd)) MP(q##igd) = tp(nv, pv); }
    SDFAM(a, 1.36e-2, 9.80e-3)
    SDFAM(b, 1.71e-3, 7.59e-4)
    SDFAM(c, 0.075, 0.03)
#undef SDFAM
#define SDLWP(pre, q) \
    if (!MG(q##igsd) && (MG(q##igs) || MG(q##igd))) { \
        if (!MG(pre##q##igs)) MP(pre##q##igs) = 0.0; \
        if (!MG(pre##q##igd)) MP(pre##q##igd) = 0.0; } \
    else { if (!MG(pre##q##igsd)) MP(pre##q##igsd) = 0.0; \
           MP(pre##q##igs) = MP(pre##q##igd) = MP(pre##q##igsd); }
    SDLWP(l, a) SDLWP(l, b) SDLWP(l, c) SDLWP(w, a) SDLWP(w, b) SDLWP(w, c)
    SDLWP(p, a) SDLWP(p, b) SDLWP(p, c)
#undef SDLWP
    if (!MG(dlcigd)) MP(dlcigd) = MG(dlcig) ? MP(dlcig) : MP(Lint);
    if (!MG(njtsd)) MP(njtsd) = MG(njts) ? MP(njts) : 20.0;
    if (!MG(njtsswd)) MP(njtsswd) = MG(njtssw) ? MP(njtssw) : 20.0;
    if (!MG(njtsswgd)) MP(njtsswgd) = MG(njtsswg) ? MP(njtsswg) : 20.0;
    if (!MG(tnjtsd)) MP(tnjtsd) = MG(tnjts) ? MP(tnjts) : 0.0;
    if (!MG(tnjtsswd)) MP(tnjtsswd) = MG(tnjtssw) ? MP(tnjtssw) : 0.0;
    if (!MG(tnjtsswgd)) MP(tnjtsswgd) = MG(tnjtsswg) ? MP(tnjtsswg) : 0.0;
}

const char *B4Model::check() const
{
    const B4Model &m = *this;
    if ((int)MP(capMod) != 2) return "capMod other than 2";
    if ((int)MP(rdsMod) != 0) return "rdsMod 1 (bias-dependent S/D resistance nodes)";
    if ((int)MP(rgateMod) != 0) return "rgateMod other than 0 (gate resistance node)";
    if ((int)MP(rbodyMod) != 0) return "rbodyMod other than 0 (substrate resistance network)";
    if ((int)MP(trnqsMod) != 0 || (int)MP(acnqsMod) != 0) return "NQS models";
    if ((int)MP(mtrlMod) != 0) return "mtrlMod 1";
    if ((int)MP(mobMod) < 0 || (int)MP(mobMod) > 6) return "mobMod out of range";
    if ((int)MP(dioMod) < 0 || (int)MP(dioMod) > 2) return "dioMod out of range";
    if ((int)MP(igcMod) < 0 || (int)MP(igcMod) > 2) return "igcMod out of range";
    return nullptr;
}

/* b4temp.c:90-417 */
int B4Model::finish(double nomTemp, double T)
{
    B4Model &m = *this;
    apply_defaults(m, nomTemp);
    type = MP(type) >= 0 ? 1 : -1;
    if (check()) return -1;
    temp = T;
    const double CONSTvt0 = 1.38064852e-23 * (27.0 + CtoK) / 1.6021766208e-19;
    if (MP(SbulkJctPotential) < 0.1) MP(SbulkJctPotential) = 0.1;
    if (MP(SsidewallJctPotential) < 0.1) MP(SsidewallJctPotential) = 0.1;
    if (MP(SGatesidewallJctPotential) < 0.1) MP(SGatesidewallJctPotential) = 0.1;
    if (MP(DbulkJctPotential) < 0.1) MP(DbulkJctPotential) = 0.1;
    if (MP(DsidewallJctPotential) < 0.1) MP(DsidewallJctPotential) = 0.1;
    if (MP(DGatesidewallJctPotential) < 0.1) MP(DGatesidewallJctPotential) = 0.1;
    {
        double chktol = std::fabs(MP(toxe));
        if (std::fabs(MP(toxp)) > chktol) chktol = std::fabs(MP(toxp));
        if (std::fabs(MP(dtox)) > chktol) chktol = std::fabs(MP(dtox));
        chktol *= 1e-14;
        if (MG(toxe) && MG(toxp) && MG(dtox) && std::fabs(MP(toxe) - (MP(toxp) + MP(dtox))) > chktol) {
        } else if (MG(toxe) && !MG(toxp)) {
            MP(toxp) = MP(toxe) - MP(dtox);
        } else if (!MG(toxe) && MG(toxp)) {
            MP(toxe) = MP(toxp) + MP(dtox);
            if (!MG(toxm)) MP(toxm) = MP(toxe);
        }
    }
    const double epsrox = MP(epsrox), toxe = MP(toxe), epssub = EPSSI;
    if (!MG(cf)) MP(cf) = 2.0 * epsrox * EPS0 / PI * std::log(1.0 + 0.4e-6 / toxe);
    coxe = epsrox * EPS0 / toxe;
    coxp = MP(epsrox) * EPS0 / MP(toxp);
    if (!MG(cgdo)) MP(cgdo) = (MG(dlc) && MP(dlc) > 0.0) ? MP(dlc) * coxe - MP(cgdl) : 0.6 * MP(xj) * coxe;
    if (!MG(cgso)) MP(cgso) = (MG(dlc) && MP(dlc) > 0.0) ? MP(dlc) * coxe - MP(cgsl) : 0.6 * MP(xj) * coxe;
    if (!MG(cgbo)) MP(cgbo) = 2.0 * MP(dwc) * coxe;
    const double Tnom = MP(tnom);
    vcrit = CONSTvt0 * std::log(CONSTvt0 / (std::sqrt(2.0) * 1.0e-14));
    factor1 = std::sqrt(epssub / (epsrox * EPS0) * toxe);
    vtm0 = KboQ * Tnom;
    Eg0 = 1.16 - 7.02e-4 * Tnom * Tnom / (Tnom + 1108.0);
    ni = 1.45e10 * (Tnom / 300.15) * std::sqrt(Tnom / 300.15) * std::exp(21.5565981 - Eg0 / (2.0 * vtm0));
    vtm = KboQ * T;
    const double Eg = 1.16 - 7.02e-4 * T * T / (T + 1108.0);
    if (T != Tnom) {
        double T0 = Eg0 / vtm0 - Eg / vtm, T1 = std::log(T / Tnom);
        double T3 = std::exp((T0 + MP(SjctTempExponent) * T1) / MP(SjctEmissionCoeff));
        SjSat = MP(SjctSatCurDensity) * T3;
        SjSwSat = MP(SjctSidewallSatCurDensity) * T3;
        SjSwgSat = MP(SjctGateSidewallSatCurDensity) * T3;
        T3 = std::exp((T0 + MP(DjctTempExponent) * T1) / MP(DjctEmissionCoeff));
        DjSat = MP(DjctSatCurDensity) * T3;
        DjSwSat = MP(DjctSidewallSatCurDensity) * T3;
        DjSwgSat = MP(DjctGateSidewallSatCurDensity) * T3;
    } else {
        SjSat = MP(SjctSatCurDensity); SjSwSat = MP(SjctSidewallSatCurDensity);
        SjSwgSat = MP(SjctGateSidewallSatCurDensity); DjSat = MP(DjctSatCurDensity);
        DjSwSat = MP(DjctSidewallSatCurDensity); DjSwgSat = MP(DjctGateSidewallSatCurDensity);
    }
    if (SjSat < 0.0) SjSat = 0.0;
    if (SjSwSat < 0.0) SjSwSat = 0.0;
    if (SjSwgSat < 0.0) SjSwgSat = 0.0;
    if (DjSat < 0.0) DjSat = 0.0;
    if (DjSwSat < 0.0) DjSwSat = 0.0;
    if (DjSwgSat < 0.0) DjSwgSat = 0.0;
    const double delTemp = T - MP(tnom);
    double T0 = MP(tcj) * delTemp;
    if (T0 >= -1.0) {
        SAreaCap = MP(SunitAreaJctCap) * (1.0 + T0);
        DAreaCap = MP(DunitAreaJctCap) * (1.0 + T0);
    } else {
        if (MP(SunitAreaJctCap) > 0.0) SAreaCap = 0.0;
        if (MP(DunitAreaJctCap) > 0.0) DAreaCap = 0.0;
    }
    T0 = MP(tcjsw) * delTemp;
    if (MP(SunitLengthSidewallJctCap) < 0.0) MP(SunitLengthSidewallJctCap) = 0.0;
    if (MP(DunitLengthSidewallJctCap) < 0.0) MP(DunitLengthSidewallJctCap) = 0.0;
    if (T0 >= -1.0) {
        SSwCap = MP(SunitLengthSidewallJctCap) * (1.0 + T0);
        DSwCap = MP(DunitLengthSidewallJctCap) * (1.0 + T0);
    } else {
        if (MP(SunitLengthSidewallJctCap) > 0.0) SSwCap = 0.0;
        if (MP(DunitLengthSidewallJctCap) > 0.0) DSwCap = 0.0;
    }
    T0 = MP(tcjswg) * delTemp;
    if (T0 >= -1.0) {
        SSwgCap = MP(SunitLengthGateSidewallJctCap) * (1.0 + T0);
        DSwgCap = MP(DunitLengthGateSidewallJctCap) * (1.0 + T0);
    } else {
        if (MP(SunitLengthGateSidewallJctCap) > 0.0) SSwgCap = 0.0;
        if (MP(DunitLengthGateSidewallJctCap) > 0.0) DSwgCap = 0.0;
    }
    auto clamp01 = [](double v) { return v < 0.01 ? 0.01 : v; };
    PhiBS = clamp01(MP(SbulkJctPotential) - MP(tpb) * delTemp);
    PhiBD = clamp01(MP(DbulkJctPotential) - MP(tpb) * delTemp);
    PhiBSWS = clamp01(MP(SsidewallJctPotential) - MP(tpbsw) * delTemp);
    PhiBSWD = clamp01(MP(DsidewallJctPotential) - MP(tpbsw) * delTemp);
    PhiBSWGS = clamp01(MP(SGatesidewallJctPotential) - MP(tpbswg) * delTemp);
    PhiBSWGD = clamp01(MP(DGatesidewallJctPotential) - MP(tpbswg) * delTemp);
    if (MP(ijthdfwd) <= 0.0) MP(ijthdfwd) = 0.0;
    if (MP(ijthsfwd) <= 0.0) MP(ijthsfwd) = 0.0;
    if (MP(ijthdrev) <= 0.0) MP(ijthdrev) = 0.0;
    if (MP(ijthsrev) <= 0.0) MP(ijthsrev) = 0.0;
    if (MP(xjbvd) <= 0.0 && (int)MP(dioMod) == 2) MP(xjbvd) = 0.0;
    else if (MP(xjbvd) < 0.0 && (int)MP(dioMod) == 0) MP(xjbvd) = 0.0;
    if (MP(bvd) <= 0.0) MP(bvd) = 0.0;
    if (MP(xjbvs) <= 0.0 && (int)MP(dioMod) == 2) MP(xjbvs) = 0.0;
    else if (MP(xjbvs) < 0.0 && (int)MP(dioMod) == 0) MP(xjbvs) = 0.0;
    if (MP(bvs) <= 0.0) MP(bvs) = 0.0;
    T0 = T / Tnom - 1.0;
    njtss = MP(njts) * (1.0 + MP(tnjts) * T0);
    njtssws = MP(njtssw) * (1.0 + MP(tnjtssw) * T0);
    njtsswgs = MP(njtsswg) * (1.0 + MP(tnjtsswg) * T0);
    njtsd = MP(njtsd) * (1.0 + MP(tnjtsd) * T0);
    njtsswd = MP(njtsswd) * (1.0 + MP(tnjtsswd) * T0);
    njtsswgd = MP(njtsswgd) * (1.0 + MP(tnjtsswgd) * T0);
    if (MP(jtweff) < 0.0) MP(jtweff) = 0.0;
    if (MP(dlcig) < 0.0) MP(dlcig) = 0.0;
    if (MP(dlcigd) < 0.0) MP(dlcigd) = 0.0;
    checked = false;
    return 0;
}

/* instance parameters */
static const struct { const char *name; int field; bool integer; } kInstPars[] = {
    {"l", BI_l, false}, {"w", BI_w, false}, {"m", BI_m, false}, {"nf", BI_nf, false},
    {"sa", BI_sa, false}, {"sb", BI_sb, false}, {"sd", BI_sd, false}, {"sca", BI_sca, false},
    {"scb", BI_scb, false}, {"scc", BI_scc, false}, {"sc", BI_sc, false}, {"min", BI_min, true},
    {"ad", BI_ad, false}, {"as", BI_as, false}, {"pd", BI_pd, false}, {"ps", BI_ps, false},
    {"nrd", BI_nrd, false}, {"nrs", BI_nrs, false}, {"off", BI_off, true},
    {"delvto", BI_delvto, false}, {"delvt0", BI_delvto, false}, {"mulu0", BI_mulu0, false},
    {"xgw", BI_xgw, false}, {"ngcon", BI_ngcon, false}, {"geomod", BI_geoMod, true},
    {"rgeomod", BI_rgeoMod, true}, {"mult_i", BI_mult_i, false}, {"mult_q", BI_mult_q, false},
    {"icvgs", BI_icVGS, false}, {"icvds", BI_icVDS, false}, {"icvbs", BI_icVBS, false},
    {"rgatemod", BI_rgateMod, true}, {"rbodymod", BI_rbodyMod, true},
    {"trnqsmod", BI_trnqsMod, true}, {"acnqsmod", BI_acnqsMod, true},
};

int B4Inst::set(const char *name, double value)
{
    for (auto &d : kInstPars) {
        if (std::strcmp(d.name, name) == 0) {
            ip[d.field] = d.integer ? (double)(int)value : value;
            igiven[d.field] = 1;
            return 0;
        }
    }
    return -1;
}

#define BP(f) (bin[BB_##f])
#define IP(f) (ip[BI_##f])
#define IG(f) (igiven[BI_##f])

static void finger_diff(double nf, int minSD, double *nuIntD, double *nuEndD,
                        double *nuIntS, double *nuEndS)
{   /* b4geo.c:26-58 */
    int NF = (int)nf;
    if ((NF % 2) != 0) {
        *nuEndD = *nuEndS = 1.0;
        *nuIntD = *nuIntS = 2.0 * std::fmax((nf - 1.0) / 2.0, 0.0);
    } else if (minSD == 1) {
        *nuEndD = 2.0; *nuIntD = 2.0 * std::fmax((nf / 2.0 - 1.0), 0.0);
        *nuEndS = 0.0; *nuIntS = nf;
    } else {
        *nuEndD = 0.0; *nuIntD = nf;
        *nuEndS = 2.0; *nuIntS = 2.0 * std::fmax((nf / 2.0 - 1.0), 0.0);
    }
}

static void paeff(double nf, int geo, int minSD, double W, double DMCG, double DMCI,
                  double DMDG, double *Ps, double *Pd, double *As, double *Ad)
{   /* b4geo.c:61-147 */
    double nuIntD = 0, nuEndD = 0, nuIntS = 0, nuEndS = 0;
    if (geo < 9) finger_diff(nf, minSD, &nuIntD, &nuEndD, &nuIntS, &nuEndS);
    double T0 = DMCG + DMCI, T1 = DMCG + DMCG, T2 = DMDG + DMDG;
    double Piso = T0 + T0 + W, Psha = T1, Pmer = T2;
    double Aiso = T0 * W, Asha = DMCG * W, Amer = DMDG * W;
    switch (geo) {
    case 0: *Ps = nuEndS * Piso + nuIntS * Psha; *Pd = nuEndD * Piso + nuIntD * Psha;
            *As = nuEndS * Aiso + nuIntS * Asha; *Ad = nuEndD * Aiso + nuIntD * Asha; break;
    case 1: *Ps = nuEndS * Piso + nuIntS * Psha; *Pd = (nuEndD + nuIntD) * Psha;
            *As = nuEndS * Aiso + nuIntS * Asha; *Ad = (nuEndD + nuIntD) * Asha; break;
    case 2: *Ps = (nuEndS + nuIntS) * Psha; *Pd = nuEndD * Piso + nuIntD * Psha;
            *As = (nuEndS + nuIntS) * Asha; *Ad = nuEndD * Aiso + nuIntD * Asha; break;
    case 3: *Ps = (nuEndS + nuIntS) * Psha; *Pd = (nuEndD + nuIntD) * Psha;
            *As = (nuEndS + nuIntS) * Asha; *Ad = (nuEndD + nuIntD) * Asha; break;
    case 4: *Ps = nuEndS * Piso + nuIntS * Psha; *Pd = nuEndD * Pmer + nuIntD * Psha;
            *As = nuEndS * Aiso + nuIntS * Asha; *Ad = nuEndD * Amer + nuIntD * Asha; break;
    case 5: *Ps = (nuEndS + nuIntS) * Psha; *Pd = nuEndD * Pmer + nuIntD * Psha;
            *As = (nuEndS + nuIntS) * Asha; *Ad = nuEndD * Amer + nuIntD * Asha; break;
    case 6: *Ps = nuEndS * Pmer + nuIntS * Psha; *Pd = nuEndD * Piso + nuIntD * Psha;
            *As = nuEndS * Amer + nuIntS * Asha; *Ad = nuEndD * Aiso + nuIntD * Asha; break;
    case 7: *Ps = nuEndS * Pmer + nuIntS * Psha; *Pd = (nuEndD + nuIntD) * Psha;
            *As = nuEndS * Amer + nuIntS * Asha; *Ad = (nuEndD + nuIntD) * Asha; break;
    case 8: *Ps = nuEndS * Pmer + nuIntS * Psha; *Pd = nuEndD * Pmer + nuIntD * Psha;
            *As = nuEndS * Amer + nuIntS * Asha; *Ad = nuEndD * Amer + nuIntD * Asha; break;
    case 9: *Ps = Piso + (nf - 1.0) * Psha; *Pd = nf * Psha;
            *As = Aiso + (nf - 1.0) * Asha; *Ad = nf * Asha; break;
    case 10: *Ps = nf * Psha; *Pd = Piso + (nf - 1.0) * Psha;
             *As = nf * Asha; *Ad = Aiso + (nf - 1.0) * Asha; break;
    default: break;
    }
}

const char *B4Inst::check() const
{
    if ((int)IP(rgateMod) != 0 || (int)IP(rbodyMod) != 0) return "instance rgateMod/rbodyMod";
    if ((int)IP(trnqsMod) != 0 || (int)IP(acnqsMod) != 0) return "instance NQS";
    if (IG(nrs) && IP(nrs) > 0 && model->p[BM_sheetResistance] > 0) return "source series resistance node";
    if (IG(nrd) && IP(nrd) > 0 && model->p[BM_sheetResistance] > 0) return "drain series resistance node";
    if ((int)IP(rgeoMod) != 0 && model->p[BM_sheetResistance] > 0) return "rgeoMod series resistance";
    return nullptr;
}

void B4Inst::defaults()
{
    const B4Model &m = *model;
    auto D = [&](int f, double v) { if (!igiven[f]) ip[f] = v; };
    D(BI_l, 5.0e-6); D(BI_w, 5.0e-6); D(BI_m, 1.0); D(BI_mult_i, 1.0); D(BI_mult_q, 1.0);
    D(BI_nf, 1.0); D(BI_min, 0); D(BI_icVDS, 0.0); D(BI_icVGS, 0.0); D(BI_icVBS, 0.0);
    D(BI_ad, 0.0); D(BI_pd, 0.0); D(BI_nrd, 1.0); D(BI_as, 0.0); D(BI_ps, 0.0);
    D(BI_nrs, 1.0); D(BI_delvto, 0.0); D(BI_mulu0, 1.0); D(BI_xgw, MP(xgw));
    D(BI_ngcon, MP(ngcon)); D(BI_geoMod, MP(geoMod)); D(BI_rgeoMod, 0);
    D(BI_sa, 0.0); D(BI_sb, 0.0); D(BI_sd, 2 * MP(dmcg)); D(BI_sca, 0.0); D(BI_scb, 0.0);
    D(BI_scc, 0.0); D(BI_sc, 0.0); D(BI_off, 0);
    D(BI_rgateMod, MP(rgateMod)); D(BI_rbodyMod, MP(rbodyMod));
    D(BI_trnqsMod, MP(trnqsMod)); D(BI_acnqsMod, MP(acnqsMod));
}

/* b4temp.c:418-2400 for one instance; fills one row of the evaluation table */
int B4Inst::resolve(double *row)
{
    B4Model &m = *model;
    double bin[BB__COUNT];
    const double T = m.temp, Tnom = MP(tnom), TRatio = T / Tnom, delTemp = T - MP(tnom);
    const double toxe = MP(toxe), epsrox = MP(epsrox), epssub = EPSSI, Vtm0 = m.vtm0, ni = m.ni;
    const double l = IP(l), w = IP(w), nf = IP(nf);
    double T0, T1, T2, T3, T4, T5, T8, T9, T10, T11, tmp, tmp1, tmp2;
    double R[B4F__COUNT];
    std::memset(R, 0, sizeof(R));
#define F(f) R[B4F_##f]

    const double Lnew = l + MP(xl), Wnew = w / nf + MP(xw);
    T0 = std::pow(Lnew, MP(Lln)); T1 = std::pow(Wnew, MP(Lwn));
    tmp1 = MP(Ll) / T0 + MP(Lw) / T1 + MP(Lwl) / (T0 * T1);
    const double dl = MP(Lint) + tmp1;
    tmp2 = MP(Llc) / T0 + MP(Lwc) / T1 + MP(Lwlc) / (T0 * T1);
    const double dlc = MP(dlc) + tmp2;
    T2 = std::pow(Lnew, MP(Wln)); T3 = std::pow(Wnew, MP(Wwn));
    tmp1 = MP(Wl) / T2 + MP(Ww) / T3 + MP(Wwl) / (T2 * T3);
    const double dw = MP(Wint) + tmp1;
    tmp2 = MP(Wlc) / T2 + MP(Wwc) / T3 + MP(Wwlc) / (T2 * T3);
    const double dwc = MP(dwc) + tmp2, dwj = MP(dwj) + tmp2;
    const double leff = Lnew - 2.0 * dl, weff = Wnew - 2.0 * dw;
    const double leffCV = Lnew - 2.0 * dlc, weffCV = Wnew - 2.0 * dwc, weffCJ = Wnew - 2.0 * dwj;
    if (leff <= 0.0 || weff <= 0.0 || leffCV <= 0.0 || weffCV <= 0.0 || weffCJ <= 0.0) return -1;
    double Inv_L, Inv_W, Inv_LW;
    if ((int)MP(binUnit) == 1) { Inv_L = 1.0e-6 / leff; Inv_W = 1.0e-6 / weff; Inv_LW = 1.0e-12 / (leff * weff); }
    else { Inv_L = 1.0 / leff; Inv_W = 1.0 / weff; Inv_LW = 1.0 / (leff * weff); }
#define BINIT(f) bin[BB_##f] = MP(f) + MP(l##f) * Inv_L + MP(w##f) * Inv_W + MP(p##f) * Inv_LW;
    BSIM4_BINNED(BINIT)
#undef BINIT
    const double abulkCVfactor = 1.0 + std::pow(BP(clc) / leffCV, BP(cle));
    T0 = TRatio - 1.0;
    const double PowWeffWr = std::pow(weffCJ * 1.0e6, BP(wr)) * nf;
    double vsattempP, rds0, rdswmin;
    BP(ucs) = BP(ucs) * std::pow(TRatio, BP(ucste));
    if ((int)MP(tempMod) == 0) {
        BP(ua) = BP(ua) + BP(ua1) * T0; BP(ub) = BP(ub) + BP(ub1) * T0;
        BP(uc) = BP(uc) + BP(uc1) * T0; BP(ud) = BP(ud) + BP(ud1) * T0;
        vsattempP = BP(vsat) - BP(at) * T0;
        T10 = BP(prt) * T0;
        rds0 = (BP(rdsw) + T10) * nf / PowWeffWr;
        rdswmin = (MP(rdswmin) + T10) * nf / PowWeffWr;
    } else {
        if ((int)MP(tempMod) == 3) {
            BP(ua) *= std::pow(TRatio, BP(ua1)); BP(ub) *= std::pow(TRatio, BP(ub1));
            BP(uc) *= std::pow(TRatio, BP(uc1)); BP(ud) *= std::pow(TRatio, BP(ud1));
        } else {
            BP(ua) *= (1.0 + BP(ua1) * delTemp); BP(ub) *= (1.0 + BP(ub1) * delTemp);
            BP(uc) *= (1.0 + BP(uc1) * delTemp); BP(ud) *= (1.0 + BP(ud1) * delTemp);
        }
        vsattempP = BP(vsat) * (1.0 - BP(at) * delTemp);
        T10 = 1.0 + BP(prt) * delTemp;
        rds0 = BP(rdsw) * T10 * nf / PowWeffWr;
        rdswmin = MP(rdswmin) * T10 * nf / PowWeffWr;
    }
    if (BP(u0) > 1.0) BP(u0) = BP(u0) / 1.0e4;
    T5 = 1.0 - BP(up) * std::exp(-leff / BP(lp));
    const double u0tempP = BP(u0) * T5 * std::pow(TRatio, BP(ute));
    if (BP(eu) < 0.0) BP(eu) = 0.0;
    if (BP(ucs) < 0.0) BP(ucs) = 0.0;
    BP(vfbsdoff) = BP(vfbsdoff) * (1.0 + BP(tvfbsdoff) * delTemp);
    BP(voff) = BP(voff) * (1.0 + BP(tvoff) * delTemp);
    BP(nfactor) = BP(nfactor) + BP(tnfactor) * delTemp / Tnom;
    BP(voffcv) = BP(voffcv) * (1.0 + BP(tvoffcv) * delTemp);
    BP(eta0) = BP(eta0) + BP(teta0) * delTemp / Tnom;
    double tfactor = 0.0;
    if (MG(vtl) && MP(vtl) > 0.0) {
        double lc = MP(lc) < 0.0 ? 0.0 : MP(lc);
        T0 = leff / (BP(xn) * leff + lc);
        tfactor = (1.0 - T0) / (1.0 + T0);
    }
    const double cgdoP = (MP(cgdo) + BP(cf)) * weffCV;
    const double cgsoP = (MP(cgso) + BP(cf)) * weffCV;
    const double cgboP = MP(cgbo) * leffCV * nf;
    if (!MG(ndep) && MG(gamma1)) { T0 = BP(gamma1) * m.coxe; BP(ndep) = 3.01248e22 * T0 * T0; }
    const double phi = Vtm0 * std::log(BP(ndep) / ni) + BP(phin) + 0.4;
    if (phi <= 0.0) return -2;
    const double sqrtPhi = std::sqrt(phi);
    const double Xdep0 = std::sqrt(2.0 * epssub / (Charge_q * BP(ndep) * 1.0e6)) * sqrtPhi;
    const double sqrtXdep0 = std::sqrt(Xdep0);
    const double litl = std::sqrt(3.0 * 3.9 / epsrox * BP(xj) * toxe);
    const double vbi = Vtm0 * std::log(BP(nsd) * BP(ndep) / (ni * ni));
    const double vfbsd = BP(ngate) > 0.0 ? Vtm0 * std::log(BP(ngate) / BP(nsd)) : 0.0;
    const double cdep0 = std::sqrt(Charge_q * epssub * BP(ndep) * 1.0e6 / 2.0 / phi);
    const double ToxRatio = std::exp(BP(ntox) * std::log(MP(toxref) / toxe)) / toxe / toxe;
    const double ToxRatioEdge = std::exp(BP(ntox) * std::log(MP(toxref) / (toxe * BP(poxedge))))
                              / toxe / toxe / BP(poxedge) / BP(poxedge);
    double Aechvb = (m.type == 1) ? 4.97232e-7 : 3.42537e-7;
    double Bechvb = (m.type == 1) ? 7.45669e11 : 1.16645e12;
    const double AechvbEdgeS = Aechvb * weff * MP(dlcig) * ToxRatioEdge;
    const double AechvbEdgeD = Aechvb * weff * MP(dlcigd) * ToxRatioEdge;
    const double BechvbEdge = -Bechvb * toxe * BP(poxedge);
    Aechvb *= weff * leff * ToxRatio;
    Bechvb *= -toxe;
    const double mstar = 0.5 + std::atan(BP(minv)) / PI;
    const double mstarcv = 0.5 + std::atan(BP(minvcv)) / PI;
    const double voffcbn = BP(voff) + MP(voffl) / leff;
    const double voffcbncv = BP(voffcv) + MP(voffcvl) / leff;
    const double ldeb = std::sqrt(epssub * Vtm0 / (Charge_q * BP(ndep) * 1.0e6)) / 3.0;
    BP(acde) *= std::pow(BP(ndep) / 2.0e16, -0.25);
    if (MG(k1) || MG(k2)) {
        if (!MG(k1)) BP(k1) = 0.53;
        if (!MG(k2)) BP(k2) = -0.0186;
    } else {
        if (!MG(vbx)) BP(vbx) = phi - 7.7348e-4 * BP(ndep) * BP(xt) * BP(xt);
        if (BP(vbx) > 0.0) BP(vbx) = -BP(vbx);
        if (BP(vbm) > 0.0) BP(vbm) = -BP(vbm);
        if (!MG(gamma1)) BP(gamma1) = 5.753e-12 * std::sqrt(BP(ndep)) / m.coxe;
        if (!MG(gamma2)) BP(gamma2) = 5.753e-12 * std::sqrt(BP(nsub)) / m.coxe;
        T0 = BP(gamma1) - BP(gamma2);
        T1 = std::sqrt(phi - BP(vbx)) - sqrtPhi;
        T2 = std::sqrt(phi * (phi - BP(vbm))) - phi;
        BP(k2) = T0 * T1 / (2.0 * T2 + BP(vbm));
        BP(k1) = BP(gamma2) - 2.0 * BP(k2) * std::sqrt(phi - BP(vbm));
    }
    if (!MG(vfb)) BP(vfb) = MG(vth0) ? m.type * BP(vth0) - phi - BP(k1) * sqrtPhi : -1.0;
    if (!MG(vth0)) BP(vth0) = m.type * (BP(vfb) + phi + BP(k1) * sqrtPhi);
    const double k1ox = BP(k1) * toxe / MP(toxm);
    tmp = std::sqrt(epssub / (epsrox * EPS0) * toxe * Xdep0);
    auto theta_of = [](double t0) {
        if (t0 < EXP_THRESHOLD) {
            double a = std::exp(t0), b = a - 1.0, c = b * b, d = c + 2.0 * a * MIN_EXP;
            return a / d;
        }
        return 1.0 / (MAX_EXP - 2.0);
    };
    const double theta0vb0 = theta_of(BP(dsub) * leff / tmp);
    const double thetaRout = BP(pdibl1) * theta_of(BP(drout) * leff / tmp) + BP(pdibl2);
    tmp = std::sqrt(Xdep0);
    tmp1 = vbi - phi;
    tmp2 = m.factor1 * tmp;
    T8 = BP(dvt0w) * theta_of(BP(dvt1w) * weff * leff / tmp2) * tmp1;
    T9 = BP(dvt0) * theta_of(BP(dvt1) * leff / tmp2) * tmp1;
    T4 = toxe * phi / (weff + BP(w0));
    T0 = std::sqrt(1.0 + BP(lpe0) / leff);
    T3 = 0.0;
    if ((int)MP(tempMod) == 1 || (int)MP(tempMod) == 0) T3 = (BP(kt1) + BP(kt1l) / leff) * (TRatio - 1.0);
    if ((int)MP(tempMod) == 2 || (int)MP(tempMod) == 3) T3 = -BP(kt1) * (TRatio - 1.0);
    T5 = k1ox * (T0 - 1.0) * sqrtPhi + T3;
    const double vfbzbfactor = -T8 - T9 + BP(k3) * T4 + T5 - phi - BP(k1) * sqrtPhi;
    double wlod = MP(wlod) < 0.0 ? 0.0 : MP(wlod);
    T0 = std::pow(Lnew, MP(llodku0));
    const double W_tmp = Wnew + wlod;
    T1 = std::pow(W_tmp, MP(wlodku0));
    const double ku0 = 1.0 + MP(lku0) / T0 + MP(wku0) / T1 + MP(pku0) / (T0 * T1);
    T0 = std::pow(Lnew, MP(llodvth));
    T1 = std::pow(W_tmp, MP(wlodvth));
    double kvth0 = 1.0 + MP(lkvth0) / T0 + MP(wkvth0) / T1 + MP(pkvth0) / (T0 * T1);
    kvth0 = std::sqrt(kvth0 * kvth0 + DELTA);
    const double ku0temp = ku0 * (1.0 + MP(tku0) * (TRatio - 1.0)) + DELTA;
    const double inv_od_ref = 1.0 / (MP(saref) + 0.5 * l) + 1.0 / (MP(sbref) + 0.5 * l);
    const double rho_ref = MP(ku0) / ku0temp * inv_od_ref;
    double VgsteffVth = 0.0;
    if ((int)MP(mobMod) == 3) {
        double lt1 = m.factor1 * sqrtXdep0, Theta0 = theta_of(BP(dvt1) * leff / lt1);
        double tmp1b = epssub / Xdep0, tmp3 = (BP(nfactor) * tmp1b + BP(cdsc) * Theta0 + BP(cit)) / m.coxe;
        double n0 = tmp3 >= -0.5 ? 1.0 + tmp3 : (1.0 + 3.0 * tmp3) / (3.0 + 8.0 * tmp3);
        T0 = n0 * m.vtm;
        T2 = voffcbn / T0;
        if (T2 < -EXP_THRESHOLD) T4 = mstar + m.coxe * MIN_EXP / cdep0 * n0;
        else if (T2 > EXP_THRESHOLD) T4 = mstar + m.coxe * MAX_EXP / cdep0 * n0;
        else T4 = mstar + std::exp(T2) * m.coxe / cdep0 * n0;
        VgsteffVth = T0 * std::log(2.0) / T4;
    }
    const double dvtp2factor = BP(dvtp5) + BP(dvtp2) * dexp_clamped(-BP(dvtp3) * std::log(leff));
    /* b4check.c clamps, applied to the model's first instance as the reference does */
    if (!m.checked) {
        m.checked = true;
        if (BP(ckappas) < 0.02) BP(ckappas) = 0.02;
        if (BP(ckappad) < 0.02) BP(ckappad) = 0.02;
        if ((int)MP(paramChk) == 1) {
            if (BP(a2) < 0.01) BP(a2) = 0.01;
            else if (BP(a2) > 1.0) { BP(a2) = 1.0; BP(a1) = 0.0; }
            if (BP(prwg) < 0.0) BP(prwg) = 0.0;
            if (BP(rdsw) < 0.0) { BP(rdsw) = 0.0; rds0 = 0.0; }
            if (rds0 < 0.0) rds0 = 0.0;
            if (rdswmin < 0.0) rdswmin = 0.0;
            if (MP(cgdo) < 0.0) MP(cgdo) = 0.0;
            if (MP(cgso) < 0.0) MP(cgso) = 0.0;
            if (MP(cgbo) < 0.0) MP(cgbo) = 0.0;
        }
    }
    double u0temp, vsattemp, vth0, eta0, k2;
    if (IP(sa) > 0.0 && IP(sb) > 0.0 && (nf == 1.0 || (nf > 1.0 && IP(sd) > 0.0))) {
        double Inv_sa = 0, Inv_sb = 0, kvsat = MP(kvsat);
        if (kvsat < -1.0) kvsat = -1.0;
        if (kvsat > 1.0) kvsat = 1.0;
        for (int i = 0; i < nf; i++) {
            Inv_sa += 1.0 / nf / (IP(sa) + 0.5 * l + i * (IP(sd) + l));
            Inv_sb += 1.0 / nf / (IP(sb) + 0.5 * l + i * (IP(sd) + l));
        }
        double Inv_ODeff = Inv_sa + Inv_sb, rho = MP(ku0) / ku0temp * Inv_ODeff;
        u0temp = u0tempP * (1.0 + rho) / (1.0 + rho_ref);
        vsattemp = vsattempP * (1.0 + kvsat * rho) / (1.0 + kvsat * rho_ref);
        double OD_offset = Inv_ODeff - inv_od_ref;
        vth0 = BP(vth0) + MP(kvth0) / kvth0 * OD_offset;
        eta0 = BP(eta0) + MP(steta0) / std::pow(kvth0, MP(lodeta0)) * OD_offset;
        k2 = BP(k2) + MP(stk2) / std::pow(kvth0, MP(lodk2)) * OD_offset;
    } else {
        u0temp = u0tempP; vth0 = BP(vth0); vsattemp = vsattempP; eta0 = BP(eta0); k2 = BP(k2);
    }
    if ((int)MP(wpemod)) {
        if (!IG(sca) && !IG(scb) && !IG(scc) && IG(sc) && IP(sc) > 0.0) {
            double sc = IP(sc), Wdrn = w / nf;
            T1 = sc + Wdrn; T2 = 1.0 / MP(scref);
            IP(sca) = MP(scref) * MP(scref) / (sc * T1);
            IP(scb) = ((0.1 * sc + 0.01 * MP(scref)) * std::exp(-10.0 * sc * T2)
                     - (0.1 * T1 + 0.01 * MP(scref)) * std::exp(-10.0 * T1 * T2)) / Wdrn;
            IP(scc) = ((0.05 * sc + 0.0025 * MP(scref)) * std::exp(-20.0 * sc * T2)
                     - (0.05 * T1 + 0.0025 * MP(scref)) * std::exp(-20.0 * T1 * T2)) / Wdrn;
        }
        for (int f : {BI_sca, BI_scb, BI_scc, BI_sc}) if (ip[f] < 0.0) ip[f] = 0.0;
        double sceff = IP(sca) + MP(web) * IP(scb) + MP(wec) * IP(scc);
        vth0 += BP(kvth0we) * sceff;
        k2 += BP(k2we) * sceff;
        T3 = 1.0 + BP(ku0we) * sceff;
        if (T3 <= 0.0) T3 = 0.0;
        u0temp *= T3;
    }
    vth0 += IP(delvto);
    const double vfb = BP(vfb) + m.type * IP(delvto);
    u0temp = u0tempP * IP(mulu0);   /* b4temp.c:1797 re-derives from the size value */
    T3 = m.type * vth0 - vfb - phi;
    double vtfbphi1 = (m.type == 1) ? T3 + T3 : 2.5 * T3;
    if (vtfbphi1 < 0.0) vtfbphi1 = 0.0;
    double vtfbphi2 = 4.0 * T3;
    if (vtfbphi2 < 0.0) vtfbphi2 = 0.0;
    double vbsc;
    if (k2 < 0.0) {
        T0 = 0.5 * BP(k1) / k2;
        vbsc = 0.9 * (phi - T0 * T0);
        if (vbsc > -3.0) vbsc = -3.0;
        else if (vbsc < -30.0) vbsc = -30.0;
    } else vbsc = -30.0;
    if (vbsc > BP(vbm)) vbsc = BP(vbm);
    const double k2ox = k2 * toxe / MP(toxm);
    const double vfbzb = vfbzbfactor + m.type * vth0;

    /* S/D areas and perimeters (b4temp.c:1938-2049) */
    const double DMCGeff = MP(dmcg) - MP(dmcgt), DMCIeff = MP(dmci), DMDGeff = MP(dmdg) - MP(dmcgt);
    double Pseff, Pdeff, Aseff, Adeff, dum1, dum2, dum3;
    const int geo = (int)IP(geoMod), minSD = (int)IP(min);
    if (IG(ps)) Pseff = IP(ps) <= 0.0 ? 0.0 : ((int)MP(perMod) == 0 ? IP(ps) : IP(ps) - weffCJ * nf);
    else paeff(nf, geo, minSD, weffCJ, DMCGeff, DMCIeff, DMDGeff, &Pseff, &dum1, &dum2, &dum3);
    if (Pseff < 0.0) Pseff = 0.0;
    if (IG(pd)) Pdeff = IP(pd) <= 0.0 ? 0.0 : ((int)MP(perMod) == 0 ? IP(pd) : IP(pd) - weffCJ * nf);
    else paeff(nf, geo, minSD, weffCJ, DMCGeff, DMCIeff, DMDGeff, &dum1, &Pdeff, &dum2, &dum3);
    if (Pdeff < 0.0) Pdeff = 0.0;
    if (IG(as)) Aseff = IP(as);
    else paeff(nf, geo, minSD, weffCJ, DMCGeff, DMCIeff, DMDGeff, &dum1, &dum2, &Aseff, &dum3);
    if (Aseff < 0.0) Aseff = 0.0;
    if (IG(ad)) Adeff = IP(ad);
    else paeff(nf, geo, minSD, weffCJ, DMCGeff, DMCIeff, DMDGeff, &dum1, &dum2, &dum3, &Adeff);
    if (Adeff < 0.0) Adeff = 0.0;

    /* junction saturation currents and diode limits (b4temp.c:2107-2236) */
    const double Nvtms = m.vtm * MP(SjctEmissionCoeff), Nvtmd = m.vtm * MP(DjctEmissionCoeff);
    const double IsatS = (Aseff <= 0.0 && Pseff <= 0.0) ? 0.0
        : Aseff * m.SjSat + Pseff * m.SjSwSat + weffCJ * nf * m.SjSwgSat;
    const double IsatD = (Adeff <= 0.0 && Pdeff <= 0.0) ? 0.0
        : Adeff * m.DjSat + Pdeff * m.DjSwSat + weffCJ * nf * m.DjSwgSat;
    auto vjm = [](double Nvtm, double Ijth, double Isb, double XExpBV) {
        double Tc = XExpBV, Tb = 1.0 + Ijth / Isb - Tc;
        return Nvtm * std::log(0.5 * (Tb + std::sqrt(Tb * Tb + 4.0 * Tc)));
    };
    auto diode = [&](double Isat, double Nvtm, double bv, double xjbv, double ijthfwd,
                     double ijthrev, double *XExp, double *vFwd, double *IFwd, double *vRev,
                     double *IRev, double *slpF, double *slpR) {
        *XExp = *vFwd = *IFwd = *vRev = *IRev = *slpF = *slpR = 0.0;
        if (Isat <= 0.0) return;
        switch ((int)MP(dioMod)) {
        case 0:
            *XExp = (bv / Nvtm > EXP_THRESHOLD) ? xjbv * MIN_EXP : xjbv * std::exp(-bv / Nvtm);
            break;
        case 1:
            *vFwd = vjm(Nvtm, ijthfwd, Isat, 0.0);
            *IFwd = Isat * std::exp(*vFwd / Nvtm);
            break;
        case 2: {
            if (bv / Nvtm > EXP_THRESHOLD) *XExp = xjbv * MIN_EXP;
            else *XExp = std::exp(-bv / Nvtm) * xjbv;
            *vFwd = vjm(Nvtm, ijthfwd, Isat, *XExp);
            double e = std::exp(*vFwd / Nvtm);
            *IFwd = Isat * (e - *XExp / e + *XExp - 1.0);
            *slpF = Isat * (e + *XExp / e) / Nvtm;
            double r = ijthrev / Isat;
            if (r < 1.0) r = 10.0;
            *vRev = -bv - Nvtm * std::log((r - 1.0) / xjbv);
            double t1 = xjbv * std::exp(-(bv + *vRev) / Nvtm);
            *IRev = Isat * (1.0 + t1);
            *slpR = -Isat * t1 / Nvtm;
            break; }
        default: break;
        }
    };
    diode(IsatS, Nvtms, MP(bvs), MP(xjbvs), MP(ijthsfwd), MP(ijthsrev), &F(XExpBVS), &F(vjsmFwd),
          &F(IVjsmFwd), &F(vjsmRev), &F(IVjsmRev), &F(SslpFwd), &F(SslpRev));
    diode(IsatD, Nvtmd, MP(bvd), MP(xjbvd), MP(ijthdfwd), MP(ijthdrev), &F(XExpBVD), &F(vjdmFwd),
          &F(IVjdmFwd), &F(vjdmRev), &F(IVjdmRev), &F(DslpFwd), &F(DslpRev));
    {   /* trap-assisted tunnelling (b4temp.c:2239-2275) */
        double T7 = m.Eg0 / m.vtm * (TRatio - 1.0);
        T11 = std::sqrt(MP(jtweff) / weffCJ) + 1.0;
        T10 = weffCJ * nf;
        F(SjctTempRevSatCur) = dexp_clamped(MP(xtss) * T7) * Aseff * MP(jtss);
        F(DjctTempRevSatCur) = dexp_clamped(MP(xtsd) * T7) * Adeff * MP(jtsd);
        F(SswTempRevSatCur) = dexp_clamped(MP(xtssws) * T7) * Pseff * MP(jtssws);
        F(DswTempRevSatCur) = dexp_clamped(MP(xtsswd) * T7) * Pdeff * MP(jtsswd);
        F(SswgTempRevSatCur) = dexp_clamped(MP(xtsswgs) * T7) * T10 * T11 * MP(jtsswgs);
        F(DswgTempRevSatCur) = dexp_clamped(MP(xtsswgd) * T7) * T10 * T11 * MP(jtsswgd);
    }

    F(type) = m.type; F(tvth0) = m.type * vth0;
    F(icVDS) = IP(icVDS); F(icVGS) = IP(icVGS); F(icVBS) = IP(icVBS); F(off) = IP(off);
    F(vcrit) = m.vcrit; F(vtm) = m.vtm; F(vtm0) = m.vtm0; F(coxe) = m.coxe; F(factor1) = m.factor1;
    F(toxe) = toxe; F(TempRatio) = T / MP(tnom) - 1.0;
    F(Nvtms) = Nvtms; F(Nvtmd) = Nvtmd; F(IsatS) = IsatS; F(IsatD) = IsatD;
    F(xjbvs) = MP(xjbvs); F(bvs) = MP(bvs); F(xjbvd) = MP(xjbvd); F(bvd) = MP(bvd);
    F(Nvtmrss) = m.vtm0 * m.njtss; F(Nvtmrsd) = m.vtm0 * m.njtsd;
    F(Nvtmrssws) = m.vtm0 * m.njtssws; F(Nvtmrsswd) = m.vtm0 * m.njtsswd;
    F(Nvtmrsswgs) = m.vtm0 * m.njtsswgs; F(Nvtmrsswgd) = m.vtm0 * m.njtsswgd;
    F(vtss) = MP(vtss); F(vtsd) = MP(vtsd); F(vtssws) = MP(vtssws); F(vtsswd) = MP(vtsswd);
    F(vtsswgs) = MP(vtsswgs); F(vtsswgd) = MP(vtsswgd);
    F(vbsc) = vbsc; F(phi) = phi; F(sqrtPhi) = sqrtPhi; F(Xdep0) = Xdep0; F(vbi) = vbi;
    F(dvt2) = BP(dvt2); F(dvt2w) = BP(dvt2w); F(dvt1) = BP(dvt1); F(dvt0) = BP(dvt0);
    F(dvt1w) = BP(dvt1w); F(dvt0w) = BP(dvt0w); F(weff) = weff; F(leff) = leff;
    F(lpe0) = BP(lpe0); F(lpeb) = BP(lpeb); F(k1ox) = k1ox; F(k1) = BP(k1); F(kt1) = BP(kt1);
    F(kt1l) = BP(kt1l); F(kt2) = BP(kt2); F(w0) = BP(w0); F(eta0) = eta0; F(etab) = BP(etab);
    F(theta0vb0) = theta0vb0; F(k3) = BP(k3); F(k3b) = BP(k3b); F(k2ox) = k2ox; F(cit) = BP(cit);
    F(nfactor) = BP(nfactor); F(cdsc) = BP(cdsc); F(cdscb) = BP(cdscb); F(cdscd) = BP(cdscd);
    F(dvtp0) = BP(dvtp0); F(dvtp1) = BP(dvtp1); F(dvtp4) = BP(dvtp4); F(dvtp2factor) = dvtp2factor;
    F(vfb) = vfb; F(ngate) = BP(ngate); F(mstar) = mstar; F(voffcbn) = voffcbn; F(cdep0) = cdep0;
    F(dwg) = BP(dwg); F(dwb) = BP(dwb); F(prwg) = BP(prwg); F(prwb) = BP(prwb);
    F(rds0) = rds0; F(rdswmin) = rdswmin; F(xj) = BP(xj); F(a0) = BP(a0); F(ags) = BP(ags);
    F(b0) = BP(b0); F(b1) = BP(b1); F(keta) = BP(keta); F(ketac) = BP(ketac);
    F(ua) = BP(ua); F(ub) = BP(ub); F(uc) = BP(uc); F(ud) = BP(ud); F(u0temp) = u0temp;
    F(vtfbphi1) = vtfbphi1; F(eu) = BP(eu); F(ucs) = BP(ucs); F(VgsteffVth) = VgsteffVth;
    F(vsattemp) = vsattemp; F(a1) = BP(a1); F(a2) = BP(a2); F(delta) = BP(delta);
    F(lambda) = BP(lambda); F(litl) = litl; F(vtfbphi2) = vtfbphi2; F(toxp) = MP(toxp);
    F(coxp) = m.coxp; F(bdos) = MP(bdos); F(ados) = MP(ados); F(fprout) = BP(fprout);
    F(pvag) = BP(pvag); F(pclm) = BP(pclm); F(thetaRout) = thetaRout; F(pdiblb) = BP(pdiblb);
    F(pditsd) = BP(pditsd); F(pdits) = BP(pdits); F(pditsl) = MP(pditsl); F(pscbe1) = BP(pscbe1);
    F(pscbe2) = BP(pscbe2); F(alpha0) = BP(alpha0); F(alpha1) = BP(alpha1); F(beta0) = BP(beta0);
    F(vtl) = BP(vtl); F(tfactor) = tfactor; F(idovvdsc) = MP(idovvdsc);
    F(agidl) = BP(agidl); F(bgidl) = BP(bgidl); F(cgidl) = BP(cgidl); F(egidl) = BP(egidl);
    F(agisl) = BP(agisl); F(bgisl) = BP(bgisl); F(cgisl) = BP(cgisl); F(egisl) = BP(egisl);
    F(rgidl) = BP(rgidl); F(kgidl) = BP(kgidl); F(fgidl) = BP(fgidl); F(rgisl) = BP(rgisl);
    F(kgisl) = BP(kgisl); F(fgisl) = BP(fgisl); F(gidlclamp) = MP(gidlclamp);
    F(weffCJ) = weffCJ; F(nf) = nf; F(vfbzb) = vfbzb; F(aigc) = BP(aigc); F(bigc) = BP(bigc);
    F(cigc) = BP(cigc); F(nigc) = BP(nigc); F(Aechvb) = Aechvb; F(Bechvb) = Bechvb;
    F(pigcd) = BP(pigcd); F(vfbsd) = vfbsd; F(vfbsdoff) = BP(vfbsdoff);
    F(AechvbEdgeS) = AechvbEdgeS; F(AechvbEdgeD) = AechvbEdgeD; F(BechvbEdge) = BechvbEdge;
    F(aigs) = BP(aigs); F(bigs) = BP(bigs); F(cigs) = BP(cigs); F(aigd) = BP(aigd);
    F(bigd) = BP(bigd); F(cigd) = BP(cigd); F(nigbacc) = BP(nigbacc);
    F(Tigb) = 4.97232e-7 * weff * leff * ToxRatio;
    F(aigbacc) = BP(aigbacc); F(bigbacc) = BP(bigbacc); F(cigbacc) = BP(cigbacc);
    F(nigbinv) = BP(nigbinv); F(eigbinv) = BP(eigbinv); F(aigbinv) = BP(aigbinv);
    F(bigbinv) = BP(bigbinv); F(cigbinv) = BP(cigbinv);
    F(abulkCVfactor) = abulkCVfactor; F(weffCV) = weffCV; F(leffCV) = leffCV;
    F(noff) = BP(noff); F(voffcv) = BP(voffcv); F(mstarcv) = mstarcv; F(voffcbncv) = voffcbncv;
    F(ldeb) = ldeb; F(acde) = BP(acde); F(moin) = BP(moin); F(xpart) = MP(xpart);
    F(czbs) = m.SAreaCap * Aseff; F(czbssw) = m.SSwCap * Pseff; F(czbsswg) = m.SSwgCap * weffCJ * nf;
    F(czbd) = m.DAreaCap * Adeff; F(czbdsw) = m.DSwCap * Pdeff; F(czbdswg) = m.DSwgCap * weffCJ * nf;
    F(MJS) = MP(SbulkJctBotGradingCoeff); F(MJSWS) = MP(SbulkJctSideGradingCoeff);
    F(MJSWGS) = MP(SbulkJctGateSideGradingCoeff); F(MJD) = MP(DbulkJctBotGradingCoeff);
    F(MJSWD) = MP(DbulkJctSideGradingCoeff); F(MJSWGD) = MP(DbulkJctGateSideGradingCoeff);
    F(PhiBS) = m.PhiBS; F(PhiBSWS) = m.PhiBSWS; F(PhiBSWGS) = m.PhiBSWGS;
    F(PhiBD) = m.PhiBD; F(PhiBSWD) = m.PhiBSWD; F(PhiBSWGD) = m.PhiBSWGD;
    F(cgdl) = BP(cgdl); F(cgsl) = BP(cgsl); F(ckappad) = BP(ckappad); F(ckappas) = BP(ckappas);
    F(cgdoP) = cgdoP; F(cgsoP) = cgsoP; F(cgboP) = cgboP;
    F(mult_i) = IP(mult_i) * IP(m); F(mult_q) = IP(mult_q) * IP(m);
    int flags = ((int)MP(mobMod) & B4FL_MOBMOD_MASK) | (((int)MP(dioMod) & 3) << B4FL_DIOMOD_SHIFT)
              | (((int)MP(igcMod) & 3) << B4FL_IGCMOD_SHIFT) | (((int)MP(tempMod) & 3) << B4FL_TEMPMOD_SHIFT);
    if ((int)MP(igbMod)) flags |= B4FL_IGBMOD;
    if ((int)MP(gidlMod)) flags |= B4FL_GIDLMOD;
    if ((int)MP(cvchargeMod)) flags |= B4FL_CVCHARGE;
    if (MG(pigcd)) flags |= B4FL_PIGCD_GIVEN;
    if (MG(lambda) && MP(lambda) > 0.0) flags |= B4FL_LAMBDA;
    if (MG(vtl) && MP(vtl) > 0.0) flags |= B4FL_VTL;
    F(flags) = (double)flags;
#undef F
    std::memcpy(row, R, sizeof(R));
    (void)sqrtXdep0; (void)dum1;
    return 0;
}

}  // namespace ngmi
