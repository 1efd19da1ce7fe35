/*
 * ORACLE (test infrastructure only): one BSIM4 4.8.3 instance evaluation,
 * restated from ngspice 46 src/spicelib/devices/bsim4/b4ld.c:69-5433 for the
 * selectors listed in bsim4_ref.h.  The matrix contributions go to
 * h->stamp[slot] (slot = TSTALLOC pointer, b4set.c:2587-2676) and the RHS
 * contributions to h->rhs[node role], mirroring the OpenMP build's split of
 * BSIM4LoadOMP / BSIM4LoadRhsMat (b4ld.c:5435-5623).  Also the limiters
 * (devsup.c:21-151), NIintegrate (niinteg.c:18-80), CKTterr (cktterr.c:13-78)
 * and BSIM4trunc (b4trunc.c:32-71).
 */
#include <math.h>
#include <string.h>
#include "bsim4_ref.h"

#define MAX_EXPL 2.688117142e+43
#define MIN_EXPL 3.720075976e-44
#define EXPL_THRESHOLD 100.0
#define MAX_EXP 5.834617425e14
#define MIN_EXP 1.713908431e-15
#define EXP_THRESHOLD 34.0
#define EPS0 8.85418e-12
#define EPSSI 1.03594e-10
#define Charge_q 1.60219e-19
#define DELTA_1 0.02
#define DELTA_3 0.02
#define DELTA_4 0.02
#define MM 3
#define CHARGE 1.6021766208e-19
#define CONSTboltz 1.38064852e-23
#define REFTEMP (27.0 + 273.15)
#define OR_MAX(a, b) ((a) > (b) ? (a) : (b))
#define OR_MIN(a, b) ((a) < (b) ? (a) : (b))

#define MP(f) (m->p[BM_##f])
#define MG(f) (m->given[BM_##f])
#define BP(f) (h->bin[BB_##f])

static void dexp3(double A, double *B, double *C)
{   /* b4ld.c:48-59 DEXP */
    if (A > EXP_THRESHOLD) { *B = MAX_EXP * (1.0 + A - EXP_THRESHOLD); *C = MAX_EXP; }
    else if (A < -EXP_THRESHOLD) { *B = MIN_EXP; *C = 0; }
    else { *B = exp(A); *C = *B; }
}

double or_limvds(double vnew, double vold)
{   /* devsup.c:21-40 */
    if (vold >= 3.5) {
        if (vnew > vold) vnew = OR_MIN(vnew, (3 * vold) + 2);
        else if (vnew < 3.5) vnew = OR_MAX(vnew, 2);
    } else {
        if (vnew > vold) vnew = OR_MIN(vnew, 4);
        else vnew = OR_MAX(vnew, -.5);
    }
    return vnew;
}

double or_pnjlim(double vnew, double vold, double vt, double vcrit, int *icheck)
{   /* devsup.c:50-84 */
    double arg;
    if ((vnew > vcrit) && (fabs(vnew - vold) > (vt + vt))) {
        if (vold > 0) {
            arg = (vnew - vold) / vt;
            if (arg > 0) vnew = vold + vt * (2 + log(arg - 2));
            else vnew = vold - vt * (2 + log(2 - arg));
        } else vnew = vt * log(vnew / vt);
        *icheck = 1;
    } else {
        if (vnew < 0) {
            if (vold > 0) arg = -1 * vold - 1;
            else arg = 2 * vold - 1;
            if (vnew < arg) { vnew = arg; *icheck = 1; }
            else *icheck = 0;
        } else *icheck = 0;
    }
    return vnew;
}

double or_fetlim(double vnew, double vold, double vto)
{   /* devsup.c:93-151 */
    double vtsthi = fabs(2 * (vold - vto)) + 2;
    double vtstlo = fabs(vold - vto) + 1;
    double vtox = vto + 3.5;
    double delv = vnew - vold;
    double vtemp;
    if (vold >= vto) {
        if (vold >= vtox) {
            if (delv <= 0) {
                if (vnew >= vtox) { if (-delv > vtstlo) vnew = vold - vtstlo; }
                else vnew = OR_MAX(vnew, vto + 2);
            } else if (delv >= vtsthi) vnew = vold + vtsthi;
        } else {
            if (delv <= 0) vnew = OR_MAX(vnew, vto - .5);
            else vnew = OR_MIN(vnew, vto + 4);
        }
    } else {
        if (delv <= 0) { if (-delv > vtsthi) vnew = vold - vtsthi; }
        else {
            vtemp = vto + .5;
            if (vnew <= vtemp) { if (delv > vtstlo) vnew = vold + vtstlo; }
            else vnew = vtemp;
        }
    }
    return vnew;
}

int or_integrate(ORCkt *ckt, double *geq, double *ceq, double cap, int qcap)
{   /* niinteg.c:18-80 */
    double *s0 = ckt->states[0], *s1 = ckt->states[1];
    int ccap = qcap + 1, k;
    if (ckt->method == OR_TRAPEZOIDAL) {
        if (ckt->order == 1)
            s0[ccap] = ckt->ag[0] * s0[qcap] + ckt->ag[1] * s1[qcap];
        else if (ckt->order == 2)
            s0[ccap] = -s1[ccap] * ckt->ag[1] + ckt->ag[0] * (s0[qcap] - s1[qcap]);
        else return -1;
    } else {
        s0[ccap] = 0;
        for (k = ckt->order; k >= 2; k--) s0[ccap] += ckt->ag[k] * ckt->states[k][qcap];
        s0[ccap] += ckt->ag[1] * s1[qcap];
        s0[ccap] += ckt->ag[0] * s0[qcap];
    }
    *ceq = s0[ccap] - ckt->ag[0] * s0[qcap];
    *geq = ckt->ag[0] * cap;
    return 0;
}

static void poly_depletion(double phi, double ngate, double epsgate, double coxe,
                           double Vgs, double *Vgs_eff, double *dVgs_eff_dVg)
{   /* b4ld.c:5402-5433 */
    double T1, T2, T3, T4, T5, T6, T7, T8;
    if ((ngate > 1.0e18) && (ngate < 1.0e25) && (Vgs > phi) && (epsgate != 0)) {
        T1 = 1.0e6 * CHARGE * epsgate * ngate / (coxe * coxe);
        T8 = Vgs - phi;
        T4 = sqrt(1.0 + 2.0 * T8 / T1);
        T2 = 2.0 * T8 / (T4 + 1.0);
        T3 = 0.5 * T2 * T2 / T1;
        T7 = 1.12 - T3 - 0.05;
        T6 = sqrt(T7 * T7 + 0.224);
        T5 = 1.12 - 0.5 * (T7 + T6);
        *Vgs_eff = Vgs - T5;
        *dVgs_eff_dVg = 1.0 - (0.5 - 0.5 / T4) * (1.0 + T7 / T6);
    } else {
        *Vgs_eff = Vgs;
        *dVgs_eff_dVg = 1.0;
    }
}

/* junction diode current for one side (b4ld.c:700-860) */
static void junction_dc(int dioMod, double SatCurrent, double vjct, double Nvtm,
                        double gmin, double xjbv, double bv, double XExpBV,
                        double vjmFwd, double IVjmFwd, double vjmRev, double IVjmRev,
                        double slpFwd, double slpRev, double *g, double *c)
{
    double ev, T0, T1, T2, T3, dev_dv;
    if (SatCurrent <= 0.0) { *g = gmin; *c = *g * vjct; return; }
    switch (dioMod) {
    case 0:
        ev = exp(vjct / Nvtm);
        T1 = xjbv * exp(-(bv + vjct) / Nvtm);
        *g = SatCurrent * (ev + T1) / Nvtm + gmin;
        *c = SatCurrent * (ev + XExpBV - T1 - 1.0) + gmin * vjct;
        break;
    case 1:
        T2 = vjct / Nvtm;
        if (T2 < -EXP_THRESHOLD) {
            *g = gmin;
            *c = SatCurrent * (MIN_EXP - 1.0) + gmin * vjct;
        } else if (vjct <= vjmFwd) {
            ev = exp(T2);
            *g = SatCurrent * ev / Nvtm + gmin;
            *c = SatCurrent * (ev - 1.0) + gmin * vjct;
        } else {
            T0 = IVjmFwd / Nvtm;
            *g = T0 + gmin;
            *c = IVjmFwd - SatCurrent + T0 * (vjct - vjmFwd) + gmin * vjct;
        }
        break;
    case 2:
        if (vjct < vjmRev) {
            T0 = vjct / Nvtm;
            if (T0 < -EXP_THRESHOLD) { ev = MIN_EXP; dev_dv = 0.0; }
            else { ev = exp(T0); dev_dv = ev / Nvtm; }
            T1 = ev - 1.0;
            T2 = IVjmRev + slpRev * (vjct - vjmRev);
            *g = dev_dv * T2 + T1 * slpRev + gmin;
            *c = T1 * T2 + gmin * vjct;
        } else if (vjct <= vjmFwd) {
            T0 = vjct / Nvtm;
            if (T0 < -EXP_THRESHOLD) { ev = MIN_EXP; dev_dv = 0.0; }
            else { ev = exp(T0); dev_dv = ev / Nvtm; }
            T1 = (bv + vjct) / Nvtm;
            if (T1 > EXP_THRESHOLD) { T2 = MIN_EXP; T3 = 0.0; }
            else { T2 = exp(-T1); T3 = -T2 / Nvtm; }
            *g = SatCurrent * (dev_dv - xjbv * T3) + gmin;
            *c = SatCurrent * (ev + XExpBV - 1.0 - xjbv * T2) + gmin * vjct;
        } else {
            *g = slpFwd + gmin;
            *c = IVjmFwd + slpFwd * (vjct - vjmFwd) + gmin * vjct;
        }
        break;
    default: break;
    }
}

/* trap-assisted tunnelling factor (b4ld.c:869-988) */
static void tat_factor(double vts, double vjct, double Nvtmr, double *T, double *dT)
{
    double T0, T9, T10, dT0;
    if ((vts - vjct) < (vts * 1e-3)) {
        T9 = 1.0e3;
        T0 = -vjct / Nvtmr * T9;
        dexp3(T0, T, &T10);
        *dT = T10 / Nvtmr * T9;
    } else {
        T9 = 1.0 / (vts - vjct);
        T0 = -vjct / Nvtmr * vts * T9;
        dT0 = vts / Nvtmr * (T9 + vjct * T9 * T9);
        dexp3(T0, T, &T10);
        *dT = T10 * dT0;
    }
}

/* junction depletion charge (b4ld.c:3965-4063) */
static void junction_cv(double vjct, double czb, double czbsw, double czbswg,
                        double MJ, double MJSW, double MJSWG, double PhiB,
                        double PhiBSW, double PhiBSWG, double *q, double *cap)
{
    double arg, sarg, T0, T1;
    if (vjct == 0.0) { *q = 0.0; *cap = czb + czbsw + czbswg; }
    else if (vjct < 0.0) {
        if (czb > 0.0) {
            arg = 1.0 - vjct / PhiB;
            sarg = (MJ == 0.5) ? 1.0 / sqrt(arg) : exp(-MJ * log(arg));
            *q = PhiB * czb * (1.0 - arg * sarg) / (1.0 - MJ);
            *cap = czb * sarg;
        } else { *q = 0.0; *cap = 0.0; }
        if (czbsw > 0.0) {
            arg = 1.0 - vjct / PhiBSW;
            sarg = (MJSW == 0.5) ? 1.0 / sqrt(arg) : exp(-MJSW * log(arg));
            *q += PhiBSW * czbsw * (1.0 - arg * sarg) / (1.0 - MJSW);
            *cap += czbsw * sarg;
        }
        if (czbswg > 0.0) {
            arg = 1.0 - vjct / PhiBSWG;
            sarg = (MJSWG == 0.5) ? 1.0 / sqrt(arg) : exp(-MJSWG * log(arg));
            *q += PhiBSWG * czbswg * (1.0 - arg * sarg) / (1.0 - MJSWG);
            *cap += czbswg * sarg;
        }
    } else {
        T0 = czb + czbsw + czbswg;
        T1 = vjct * (czb * MJ / PhiB + czbsw * MJSW / PhiBSW + czbswg * MJSWG / PhiBSWG);
        *q = vjct * (T0 + 0.5 * T1);
        *cap = T0 + T1;
    }
}

int or_bsim4_load(ORInst *h, ORCkt *ckt)
{
    const ORModel *m = h->model;
    const int type = m->type;
    double *s0 = ckt->states[0], *s1 = ckt->states[1], *s2 = ckt->states[2];
    const int st = h->state;
    const double *rhsOld = ckt->rhsOld;
    const double gmin = ckt->gmin;
    double CONSTvt0 = CONSTboltz * REFTEMP / CHARGE;

    double vds, vgs, vbs, vges, vgms, vdbs, vsbs, vses, vdes, qdef, vbd, vdbd, vgd, vged, vgmd, vgb, vgmb;
    double vgdo, vsbdo, vsbd, vbs_jct, vbd_jct;
    double delvbd, delvdbd, delvgd, delvds, delvgs, delvbs, delvsbs, delvses, delvded, delvdes;
    /* the hat currents are only read when Check == 0, which only the
     * limiting branch can produce */
    double vdedo, Idtot = 0, cdhat = 0, Ibtot = 0, cbhat = 0, Igstot = 0, cgshat = 0;
    double Igdtot = 0, cgdhat = 0, Igbtot = 0, cgbhat = 0;
    double Isestot = 0, cseshat = 0, Idedtot = 0, cdedhat = 0, von, xfact;
    double Nvtms, Nvtmd, SourceSatCurrent, DrainSatCurrent;
    double Nvtmrss, Nvtmrssws, Nvtmrsswgs, Nvtmrsd, Nvtmrsswd, Nvtmrsswgd;
    double T0, T1, T2, T3, T4, T5, T6, T7, T8, T9, T10, T11, T12, T13, T14;
    double dT0_dVg, dT0_dVd, dT0_dVb, dT1_dVg, dT1_dVd, dT1_dVb, dT2_dVg, dT2_dVd, dT2_dVb;
    double dT3_dVg, dT3_dVd, dT3_dVb, dT4_dVd, dT4_dVb, dT5_dVg, dT5_dVd, dT5_dVb;
    double dT6_dVg, dT6_dVd, dT6_dVb, dT7_dVg, dT7_dVd, dT7_dVb, dT8_dVg, dT8_dVd, dT8_dVb;
    double dT9_dVg, dT9_dVd, dT9_dVb, dT10_dVg, dT10_dVd, dT10_dVb, dT11_dVg;
    double dT1b, dT2b, dT3b, dT4b, dT5b, dT6b, Ta, Tb, Tc, Td, Te, Tf;
    double Vds, Vgs, Vbs, Vdb, epssub, toxe, epsrox;
    double Vbseff, dVbseff_dVb, Phis, dPhis_dVb, sqrtPhis, dsqrtPhis_dVb, Xdep, dXdep_dVb;
    double Leff, Vtm, Vtm0, V0, lt1, dlt1_dVb, ltw, dltw_dVb, Theta0, dTheta0_dVb;
    double Delt_vth, dDelt_vth_dVb, TempRatio, Vth_NarrowW, dDIBL_Sft_dVd, DIBL_Sft, Lpe_Vb;
    double Vth, dVth_dVb, dVth_dVd, tmp, tmp1, tmp2, tmp3, tmp4, n, dn_dVb, dn_dVd;
    double dDITS_Sft_dVd, dDITS_Sft_dVb, DITS_Sft2, dDITS_Sft2_dVd;
    double vgs_eff, dvgs_eff_dvg, vgd_eff, dvgd_eff_dvg, Vgs_eff, dVgs_eff_dVg, Vgst;
    double Vgsteff, dVgsteff_dVg, dVgsteff_dVd, dVgsteff_dVb, ExpVgst;
    double Weff, dWeff_dVg, dWeff_dVb, Rds, dRds_dVg, dRds_dVb;
    double Abulk, dAbulk_dVb, dAbulk_dVg, Abulk0, dAbulk0_dVb, Abulk0_Q, dAbulk0_Q_dVb;
    double Denomi, dDenomi_dVg, dDenomi_dVd, dDenomi_dVb, ueff, dueff_dVg, dueff_dVd, dueff_dVb;
    double WVCox, WVCoxRds, Esat, EsatL, dEsatL_dVg, dEsatL_dVd, dEsatL_dVb, a1, Lambda, dLambda_dVg;
    double Vgst2Vtm, Vdsat, dVdsat_dVg, dVdsat_dVd, dVdsat_dVb;
    double Vdseff, dVdseff_dVg, dVdseff_dVd, dVdseff_dVb, diffVds;
    double Vasat, dVasat_dVg, dVasat_dVb, dVasat_dVd, Tcen, dTcen_dVg, dTcen_dVd, dTcen_dVb;
    double Coxeff, dCoxeff_dVg, dCoxeff_dVd, dCoxeff_dVb, CoxeffWovL, beta, dbeta_dVg, dbeta_dVd, dbeta_dVb;
    double fgche1, dfgche1_dVg, dfgche1_dVd, dfgche1_dVb, fgche2, dfgche2_dVg, dfgche2_dVd, dfgche2_dVb;
    double gche, dgche_dVg, dgche_dVd, dgche_dVb, Idl, dIdl_dVg, dIdl_dVd, dIdl_dVb;
    double FP, dFP_dVg, PvagTerm, dPvagTerm_dVg, dPvagTerm_dVd, dPvagTerm_dVb;
    double Cclm, dCclm_dVg, dCclm_dVd, dCclm_dVb, VACLM, dVACLM_dVg, dVACLM_dVd, dVACLM_dVb;
    double VADIBL, dVADIBL_dVg, dVADIBL_dVd, dVADIBL_dVb, Va, dVa_dVd, dVa_dVg, dVa_dVb;
    double VADITS, dVADITS_dVg, dVADITS_dVd, VASCBE, dVASCBE_dVg, dVASCBE_dVd, dVASCBE_dVb;
    double Idsa, dIdsa_dVg, dIdsa_dVd, dIdsa_dVb, Isub, Gbd, Gbg, Gbb, Ids, Gm, Gds, Gmb, cdrain;
    double vs, dvs_dVg, dvs_dVd, dvs_dVb, Fsevl, dFsevl_dVg, dFsevl_dVd, dFsevl_dVb;
    double Igidl, Ggidld, Ggidlg, Ggidlb, Igisl, Ggisls, Ggislg, Ggislb;
    double Vfb = 0.0, V3, V4, Vfbeff, dVfbeff_dVg, dVfbeff_dVb, Voxacc = 0.0, dVoxacc_dVg = 0.0, dVoxacc_dVb = 0.0;
    double Voxdepinv = 0.0, dVoxdepinv_dVg = 0.0, dVoxdepinv_dVd = 0.0, dVoxdepinv_dVb = 0.0;
    double VxNVt = 0.0, ExpVxNVt, Vaux = 0.0, dVaux_dVg = 0.0, dVaux_dVd = 0.0, dVaux_dVb = 0.0;
    double Igc, dIgc_dVg, dIgc_dVd, dIgc_dVb, Igcs, dIgcs_dVg, dIgcs_dVd, dIgcs_dVb;
    double Igcd, dIgcd_dVg, dIgcd_dVd, dIgcd_dVb, Igs, dIgs_dVg, dIgs_dVs, Igd, dIgd_dVg, dIgd_dVd;
    double Igbacc, dIgbacc_dVg, dIgbacc_dVb, Igbinv, dIgbinv_dVg, dIgbinv_dVd, dIgbinv_dVb;
    double Pigcd, dPigcd_dVg, dPigcd_dVd, dPigcd_dVb, nf = h->ip[BI_nf];
    double VbseffCV, dVbseffCV_dVb, CoxWL, noff, dnoff_dVd, dnoff_dVb, voffcv, VgstNVt;
    double Cox, Tox, Ccen, CoxWLcen, Qac0, QovCox, dQac0_dVg, dQac0_dVb, Qsub0;
    double dQsub0_dVg, dQsub0_dVd, dQsub0_dVb, DeltaPhi, dDeltaPhi_dVg, VgDP, dVgDP_dVg;
    double LINK, AbulkCV, dAbulkCV_dVb, VdsatCV, VdseffCV, dVdseffCV_dVg, dVdseffCV_dVd, dVdseffCV_dVb;
    double Cgg1, Cgd1, Cgb1, Cbg1, Cbd1, Cbb1, Csg, Csd, Csb, Cbg, Cbd, Cbb, Cgg, Cgd, Cgb;
    double qgate = 0.0, qbulk = 0.0, qdrn = 0.0, qsrc, qgb, qgdo, qgso, cgdo, cgso;
    double czbd, czbs, czbdsw, czbdswg, czbssw, czbsswg, qbs, qbd;
    double ag0, gcggb, gcgdb, gcgsb, gcgbb, gcdgb, gcddb, gcdsb, gcdbb, gcsgb, gcsdb, gcssb, gcsbb;
    double gcbgb, gcbdb, gcbsb, gcbbb, gcdbdb, gcsbsb, sxpart, dxpart;
    double geq, ceq, ceqqg, ceqqb, ceqqd, ceqqjd = 0.0, ceqqjs = 0.0, cqgate, cqbody, cqdrn;
    double Gmbs, FwdSum, RevSum, ceqdrn, ceqbd, ceqbs, gbbdp, gbbsp, gbdpg, gbdpdp, gbdpb, gbdpsp;
    double gbspg, gbspdp, gbspb, gbspsp;
    double gIstotg, gIstotd, gIstots, gIstotb, Istoteq, gIdtotg, gIdtotd, gIdtots, gIdtotb, Idtoteq;
    double gIbtotg, gIbtotd, gIbtots, gIbtotb, Ibtoteq, gIgtotg, gIgtotd, gIgtots, gIgtotb, Igtoteq;
    double ceqjs, ceqjd, mult_i, mult_q, gjbd, gjbs, gdpr, gspr, geltd;
    int Check = 1, Check1 = 1, Check2 = 1, ChargeComputationNeeded, k;
    double *st0 = s0 + st, *st1 = s1 + st, *st2 = s2 + st;
    double *S = h->stamp, *R = h->rhs;

#define NV(role) (rhsOld[h->node[role]])

    ChargeComputationNeeded =
        ((ckt->mode & (OR_MODEDCTRANCURVE | OR_MODEAC | OR_MODETRAN | OR_MODEINITSMSIG)) ||
         ((ckt->mode & OR_MODETRANOP) && (ckt->mode & OR_MODEUIC))) ? 1 : 0;

    for (k = 0; k < P__COUNT; k++) S[k] = 0.0;
    for (k = 0; k < N__COUNT; k++) R[k] = 0.0;

    /* ---- terminal voltages by mode (b4ld.c:261-407) */
    if (ckt->mode & OR_MODEINITSMSIG) {
        vds = st0[S_vds]; vgs = st0[S_vgs]; vbs = st0[S_vbs]; vges = st0[S_vges];
        vgms = st0[S_vgms]; vdbs = st0[S_vdbs]; vsbs = st0[S_vsbs]; vses = st0[S_vses];
        vdes = st0[S_vdes]; qdef = st0[S_qdef];
    } else if (ckt->mode & OR_MODEINITTRAN) {
        vds = st1[S_vds]; vgs = st1[S_vgs]; vbs = st1[S_vbs]; vges = st1[S_vges];
        vgms = st1[S_vgms]; vdbs = st1[S_vdbs]; vsbs = st1[S_vsbs]; vses = st1[S_vses];
        vdes = st1[S_vdes]; qdef = st1[S_qdef];
    } else if ((ckt->mode & OR_MODEINITJCT) && !h->off) {
        vds = type * h->ip[BI_icVDS];
        vgs = vges = vgms = type * h->ip[BI_icVGS];
        vbs = vdbs = vsbs = type * h->ip[BI_icVBS];
        if (vds > 0.0) { vdes = vds + 0.01; vses = -0.01; }
        else if (vds < 0.0) { vdes = vds - 0.01; vses = 0.01; }
        else vdes = vses = 0.0;
        qdef = 0.0;
        if ((vds == 0.0) && (vgs == 0.0) && (vbs == 0.0) &&
            ((ckt->mode & (OR_MODETRAN | OR_MODEAC | OR_MODEDCOP | OR_MODEDCTRANCURVE)) ||
             (!(ckt->mode & OR_MODEUIC)))) {
            vds = 0.1; vdes = 0.11; vses = -0.01;
            vgs = vges = vgms = type * h->vth0 + 0.1;
            vbs = vdbs = vsbs = 0.0;
        }
    } else if ((ckt->mode & (OR_MODEINITJCT | OR_MODEINITFIX)) && h->off) {
        vds = vgs = vbs = vges = vgms = 0.0;
        vdbs = vsbs = vdes = vses = qdef = 0.0;
    } else {
        if (ckt->mode & OR_MODEINITPRED) {
            xfact = ckt->delta / ckt->deltaOld[1];
#define PRED(x) st0[S_##x] = st1[S_##x]; x = (1.0 + xfact) * st1[S_##x] - (xfact * st2[S_##x]);
            PRED(vds)
            PRED(vgs)
            PRED(vges)
            PRED(vgms)
            PRED(vbs)
            st0[S_vbd] = st0[S_vbs] - st0[S_vds];
            PRED(vdbs)
            st0[S_vdbd] = st0[S_vdbs] - st0[S_vds];
            PRED(vsbs)
            PRED(vses)
            PRED(vdes)
            PRED(qdef)
#undef PRED
        } else {
            vds = type * (NV(N_dp) - NV(N_sp));
            vgs = type * (NV(N_gp) - NV(N_sp));
            vbs = type * (NV(N_bp) - NV(N_sp));
            vges = type * (NV(N_g) - NV(N_sp));
            vgms = type * (NV(N_gm) - NV(N_sp));
            vdbs = type * (NV(N_db) - NV(N_sp));
            vsbs = type * (NV(N_sb) - NV(N_sp));
            vses = type * (NV(N_s) - NV(N_sp));
            vdes = type * (NV(N_d) - NV(N_sp));
            qdef = type * (NV(N_q));
        }
        vgdo = st0[S_vgs] - st0[S_vds];
        vbd = vbs - vds;
        vdbd = vdbs - vds;
        vgd = vgs - vds;
        vged = vges - vds;
        vgmd = vgms - vds;
        delvbd = vbd - st0[S_vbd];
        delvdbd = vdbd - st0[S_vdbd];
        delvgd = vgd - vgdo;
        delvds = vds - st0[S_vds];
        delvgs = vgs - st0[S_vgs];
        delvbs = vbs - st0[S_vbs];
        delvsbs = vsbs - st0[S_vsbs];
        delvses = vses - st0[S_vses];
        vdedo = st0[S_vdes] - st0[S_vds];
        delvdes = vdes - st0[S_vdes];
        delvded = vdes - vds - vdedo;
        (void)delvdes; (void)delvdbd; (void)vged; (void)vgmd;
        {
            double delvbd_jct = (!h->rbodyMod) ? delvbd : delvdbd;
            double delvbs_jct = (!h->rbodyMod) ? delvbs : delvsbs;
            if (h->mode >= 0) {
                Idtot = h->cd + h->csub - h->cbd + h->Igidl;
                cdhat = Idtot - h->gbd * delvbd_jct + (h->gmbs + h->gbbs + h->ggidlb) * delvbs
                      + (h->gm + h->gbgs + h->ggidlg) * delvgs
                      + (h->gds + h->gbds + h->ggidld) * delvds;
                Ibtot = h->cbs + h->cbd - h->Igidl - h->Igisl - h->csub;
                cbhat = Ibtot + h->gbd * delvbd_jct + h->gbs * delvbs_jct
                      - (h->gbbs + h->ggidlb) * delvbs - (h->gbgs + h->ggidlg) * delvgs
                      - (h->gbds + h->ggidld - h->ggisls) * delvds
                      - h->ggislg * delvgd - h->ggislb * delvbd;
                Igstot = h->Igs + h->Igcs;
                cgshat = Igstot + (h->gIgsg + h->gIgcsg) * delvgs + h->gIgcsd * delvds + h->gIgcsb * delvbs;
                Igdtot = h->Igd + h->Igcd;
                cgdhat = Igdtot + h->gIgdg * delvgd + h->gIgcdg * delvgs + h->gIgcdd * delvds + h->gIgcdb * delvbs;
                Igbtot = h->Igb;
                cgbhat = h->Igb + h->gIgbg * delvgs + h->gIgbd * delvds + h->gIgbb * delvbs;
            } else {
                Idtot = h->cd + h->cbd - h->Igidl;
                cdhat = Idtot + h->gbd * delvbd_jct + h->gmbs * delvbd + h->gm * delvgd
                      - (h->gds + h->ggidls) * delvds - h->ggidlg * delvgs - h->ggidlb * delvbs;
                Ibtot = h->cbs + h->cbd - h->Igidl - h->Igisl - h->csub;
                cbhat = Ibtot + h->gbs * delvbs_jct + h->gbd * delvbd_jct
                      - (h->gbbs + h->ggislb) * delvbd - (h->gbgs + h->ggislg) * delvgd
                      + (h->gbds + h->ggisld - h->ggidls) * delvds
                      - h->ggidlg * delvgs - h->ggidlb * delvbs;
                Igstot = h->Igs + h->Igcd;
                cgshat = Igstot + h->gIgsg * delvgs + h->gIgcdg * delvgd - h->gIgcdd * delvds + h->gIgcdb * delvbd;
                Igdtot = h->Igd + h->Igcs;
                cgdhat = Igdtot + (h->gIgdg + h->gIgcsg) * delvgd - h->gIgcsd * delvds + h->gIgcsb * delvbd;
                Igbtot = h->Igb;
                cgbhat = h->Igb + h->gIgbg * delvgd - h->gIgbd * delvds + h->gIgbb * delvbd;
            }
            Isestot = h->gstot * st0[S_vses];
            cseshat = Isestot + h->gstot * delvses + h->gstotd * delvds + h->gstotg * delvgs + h->gstotb * delvbs;
            Idedtot = h->gdtot * vdedo;
            cdedhat = Idedtot + h->gdtot * delvded + h->gdtotd * delvds + h->gdtotg * delvgs + h->gdtotb * delvbs;
        }
        /* CKTbypass is 0 by default (cktntask.c:105): no bypass */

        /* limiting (b4ld.c:605-686) */
        von = h->von;
        if (st0[S_vds] >= 0.0) {
            vgs = or_fetlim(vgs, st0[S_vgs], von);
            vds = vgs - vgd;
            vds = or_limvds(vds, st0[S_vds]);
            vgd = vgs - vds;
            if (h->rgateMod == 1) {
                vges = or_fetlim(vges, st0[S_vges], von);
                vged = vges - vds;
            }
        } else {
            vgd = or_fetlim(vgd, vgdo, von);
            vds = vgs - vgd;
            vds = -or_limvds(-vds, -(st0[S_vds]));
            vgs = vgd + vds;
            if (h->rgateMod == 1) {
                double vgedo = st0[S_vges] - st0[S_vds];
                vged = or_fetlim(vged, vgedo, von);
                vges = vged + vds;
            }
        }
        if (vds >= 0.0) {
            vbs = or_pnjlim(vbs, st0[S_vbs], CONSTvt0, m->vcrit, &Check);
            vbd = vbs - vds;
            if (h->rbodyMod) {
                vdbs = or_pnjlim(vdbs, st0[S_vdbs], CONSTvt0, m->vcrit, &Check1);
                vdbd = vdbs - vds;
                vsbs = or_pnjlim(vsbs, st0[S_vsbs], CONSTvt0, m->vcrit, &Check2);
                Check = ((Check1 == 0) && (Check2 == 0)) ? 0 : 1;
            }
        } else {
            vbd = or_pnjlim(vbd, st0[S_vbd], CONSTvt0, m->vcrit, &Check);
            vbs = vbd + vds;
            if (h->rbodyMod) {
                vdbd = or_pnjlim(vdbd, st0[S_vdbd], CONSTvt0, m->vcrit, &Check1);
                vdbs = vdbd + vds;
                vsbdo = st0[S_vsbs] - st0[S_vds];
                vsbd = vsbs - vds;
                vsbd = or_pnjlim(vsbd, vsbdo, CONSTvt0, m->vcrit, &Check2);
                vsbs = vsbd + vds;
                Check = ((Check1 == 0) && (Check2 == 0)) ? 0 : 1;
            }
        }
    }

    /* ---- DC currents (b4ld.c:688-1002) */
    vbd = vbs - vds;
    vgd = vgs - vds;
    vgb = vgs - vbs;
    vged = vges - vds;
    vgmd = vgms - vds;
    vgmb = vgms - vbs;
    vdbd = vdbs - vds;
    (void)vged; (void)vgmd;
    vbs_jct = (!h->rbodyMod) ? vbs : vsbs;
    vbd_jct = (!h->rbodyMod) ? vbd : vdbd;

    Nvtms = m->vtm * MP(SjctEmissionCoeff);
    if ((h->Aseff <= 0.0) && (h->Pseff <= 0.0)) SourceSatCurrent = 0.0;
    else SourceSatCurrent = h->Aseff * m->SjctTempSatCurDensity
                          + h->Pseff * m->SjctSidewallTempSatCurDensity
                          + h->weffCJ * nf * m->SjctGateSidewallTempSatCurDensity;
    junction_dc((int)MP(dioMod), SourceSatCurrent, vbs_jct, Nvtms, gmin, MP(xjbvs), MP(bvs),
                h->XExpBVS, h->vjsmFwd, h->IVjsmFwd, h->vjsmRev, h->IVjsmRev, h->SslpFwd,
                h->SslpRev, &h->gbs, &h->cbs);
    Nvtmd = m->vtm * MP(DjctEmissionCoeff);
    if ((h->Adeff <= 0.0) && (h->Pdeff <= 0.0)) DrainSatCurrent = 0.0;
    else DrainSatCurrent = h->Adeff * m->DjctTempSatCurDensity
                         + h->Pdeff * m->DjctSidewallTempSatCurDensity
                         + h->weffCJ * nf * m->DjctGateSidewallTempSatCurDensity;
    junction_dc((int)MP(dioMod), DrainSatCurrent, vbd_jct, Nvtmd, gmin, MP(xjbvd), MP(bvd),
                h->XExpBVD, h->vjdmFwd, h->IVjdmFwd, h->vjdmRev, h->IVjdmRev, h->DslpFwd,
                h->DslpRev, &h->gbd, &h->cbd);

    Nvtmrssws = m->vtm0 * m->njtsswstemp;
    Nvtmrsswgs = m->vtm0 * m->njtsswgstemp;
    Nvtmrss = m->vtm0 * m->njtsstemp;
    Nvtmrsswd = m->vtm0 * m->njtsswdtemp;
    Nvtmrsswgd = m->vtm0 * m->njtsswgdtemp;
    Nvtmrsd = m->vtm0 * m->njtsdtemp;
    tat_factor(MP(vtss), vbs_jct, Nvtmrss, &Ta, &dT1b);
    tat_factor(MP(vtsd), vbd_jct, Nvtmrsd, &Tb, &dT2b);
    tat_factor(MP(vtssws), vbs_jct, Nvtmrssws, &Tc, &dT3b);
    tat_factor(MP(vtsswd), vbd_jct, Nvtmrsswd, &Td, &dT4b);
    tat_factor(MP(vtsswgs), vbs_jct, Nvtmrsswgs, &Te, &dT5b);
    tat_factor(MP(vtsswgd), vbd_jct, Nvtmrsswgd, &Tf, &dT6b);
    h->gbs += h->SjctTempRevSatCur * dT1b + h->SswTempRevSatCur * dT3b + h->SswgTempRevSatCur * dT5b;
    h->cbs -= h->SjctTempRevSatCur * (Ta - 1.0) + h->SswTempRevSatCur * (Tc - 1.0)
            + h->SswgTempRevSatCur * (Te - 1.0);
    h->gbd += h->DjctTempRevSatCur * dT2b + h->DswTempRevSatCur * dT4b + h->DswgTempRevSatCur * dT6b;
    h->cbd -= h->DjctTempRevSatCur * (Tb - 1.0) + h->DswTempRevSatCur * (Td - 1.0)
            + h->DswgTempRevSatCur * (Tf - 1.0);

    if (vds >= 0.0) { h->mode = 1; Vds = vds; Vgs = vgs; Vbs = vbs; Vdb = vds - vbs; }
    else { h->mode = -1; Vds = -vds; Vgs = vgd; Vbs = vbd; Vdb = -vbs; }
    (void)Vdb;
    epsrox = MP(epsrox); toxe = MP(toxe); epssub = EPSSI;

    /* ---- Vbseff, Vth (b4ld.c:1036-1245) */
    T0 = Vbs - h->vbsc - 0.001;
    T1 = sqrt(T0 * T0 - 0.004 * h->vbsc);
    if (T0 >= 0.0) {
        Vbseff = h->vbsc + 0.5 * (T0 + T1);
        dVbseff_dVb = 0.5 * (1.0 + T0 / T1);
    } else {
        T2 = -0.002 / (T1 - T0);
        Vbseff = h->vbsc * (1.0 + T2);
        dVbseff_dVb = T2 * h->vbsc / T1;
    }
    T9 = 0.95 * h->phi;
    T0 = T9 - Vbseff - 0.001;
    T1 = sqrt(T0 * T0 + 0.004 * T9);
    Vbseff = T9 - 0.5 * (T0 + T1);
    dVbseff_dVb *= 0.5 * (1.0 + T0 / T1);
    Phis = h->phi - Vbseff;
    dPhis_dVb = -1.0;
    sqrtPhis = sqrt(Phis);
    dsqrtPhis_dVb = -0.5 / sqrtPhis;
    Xdep = h->Xdep0 * sqrtPhis / h->sqrtPhi;
    dXdep_dVb = (h->Xdep0 / h->sqrtPhi) * dsqrtPhis_dVb;
    Leff = h->leff;
    Vtm = m->vtm;
    Vtm0 = m->vtm0;
    T3 = sqrt(Xdep);
    V0 = h->vbi - h->phi;
    T0 = BP(dvt2) * Vbseff;
    if (T0 >= -0.5) { T1 = 1.0 + T0; T2 = BP(dvt2); }
    else { T4 = 1.0 / (3.0 + 8.0 * T0); T1 = (1.0 + 3.0 * T0) * T4; T2 = BP(dvt2) * T4 * T4; }
    lt1 = m->factor1 * T3 * T1;
    dlt1_dVb = m->factor1 * (0.5 / T3 * T1 * dXdep_dVb + T3 * T2);
    T0 = BP(dvt2w) * Vbseff;
    if (T0 >= -0.5) { T1 = 1.0 + T0; T2 = BP(dvt2w); }
    else { T4 = 1.0 / (3.0 + 8.0 * T0); T1 = (1.0 + 3.0 * T0) * T4; T2 = BP(dvt2w) * T4 * T4; }
    ltw = m->factor1 * T3 * T1;
    dltw_dVb = m->factor1 * (0.5 / T3 * T1 * dXdep_dVb + T3 * T2);
    T0 = BP(dvt1) * Leff / lt1;
    if (T0 < EXP_THRESHOLD) {
        T1 = exp(T0); T2 = T1 - 1.0; T3 = T2 * T2; T4 = T3 + 2.0 * T1 * MIN_EXP;
        Theta0 = T1 / T4;
        dT1_dVb = -T0 * T1 * dlt1_dVb / lt1;
        dTheta0_dVb = dT1_dVb * (T4 - 2.0 * T1 * (T2 + MIN_EXP)) / T4 / T4;
    } else { Theta0 = 1.0 / (MAX_EXP - 2.0); dTheta0_dVb = 0.0; }
    h->thetavth = BP(dvt0) * Theta0;
    Delt_vth = h->thetavth * V0;
    dDelt_vth_dVb = BP(dvt0) * dTheta0_dVb * V0;
    T0 = BP(dvt1w) * h->weff * Leff / ltw;
    if (T0 < EXP_THRESHOLD) {
        T1 = exp(T0); T2 = T1 - 1.0; T3 = T2 * T2; T4 = T3 + 2.0 * T1 * MIN_EXP;
        T5 = T1 / T4;
        dT1_dVb = -T0 * T1 * dltw_dVb / ltw;
        dT5_dVb = dT1_dVb * (T4 - 2.0 * T1 * (T2 + MIN_EXP)) / T4 / T4;
    } else { T5 = 1.0 / (MAX_EXP - 2.0); dT5_dVb = 0.0; }
    T0 = BP(dvt0w) * T5;
    T2 = T0 * V0;
    dT2_dVb = BP(dvt0w) * dT5_dVb * V0;
    TempRatio = ckt->temp / MP(tnom) - 1.0;
    T0 = sqrt(1.0 + BP(lpe0) / Leff);
    T1 = h->k1ox * (T0 - 1.0) * h->sqrtPhi + (BP(kt1) + BP(kt1l) / Leff + BP(kt2) * Vbseff) * TempRatio;
    Vth_NarrowW = toxe * h->phi / (h->weff + BP(w0));
    T3 = h->eta0 + BP(etab) * Vbseff;
    if (T3 < 1.0e-4) { T9 = 1.0 / (3.0 - 2.0e4 * T3); T3 = (2.0e-4 - T3) * T9; T4 = T9 * T9; }
    else T4 = 1.0;
    dDIBL_Sft_dVd = T3 * h->theta0vb0;
    DIBL_Sft = dDIBL_Sft_dVd * Vds;
    Lpe_Vb = sqrt(1.0 + BP(lpeb) / Leff);
    Vth = type * h->vth0 + (h->k1ox * sqrtPhis - BP(k1) * h->sqrtPhi) * Lpe_Vb
        - h->k2ox * Vbseff - Delt_vth - T2 + (BP(k3) + BP(k3b) * Vbseff) * Vth_NarrowW + T1 - DIBL_Sft;
    dVth_dVb = Lpe_Vb * h->k1ox * dsqrtPhis_dVb - h->k2ox - dDelt_vth_dVb - dT2_dVb
             + BP(k3b) * Vth_NarrowW - BP(etab) * Vds * h->theta0vb0 * T4 + BP(kt2) * TempRatio;
    dVth_dVd = -dDIBL_Sft_dVd;

    tmp1 = epssub / Xdep;
    h->nstar = m->vtm / Charge_q * (m->coxe + tmp1 + BP(cit));
    tmp2 = BP(nfactor) * tmp1;
    tmp3 = BP(cdsc) + BP(cdscb) * Vbseff + BP(cdscd) * Vds;
    tmp4 = (tmp2 + tmp3 * Theta0 + BP(cit)) / m->coxe;
    if (tmp4 >= -0.5) {
        n = 1.0 + tmp4;
        dn_dVb = (-tmp2 / Xdep * dXdep_dVb + tmp3 * dTheta0_dVb + BP(cdscb) * Theta0) / m->coxe;
        dn_dVd = BP(cdscd) * Theta0 / m->coxe;
    } else {
        T0 = 1.0 / (3.0 + 8.0 * tmp4);
        n = (1.0 + 3.0 * tmp4) * T0;
        T0 *= T0;
        dn_dVb = (-tmp2 / Xdep * dXdep_dVb + tmp3 * dTheta0_dVb + BP(cdscb) * Theta0) / m->coxe * T0;
        dn_dVd = BP(cdscd) * Theta0 / m->coxe * T0;
    }
    if (BP(dvtp0) > 0.0) {
        T0 = -BP(dvtp1) * Vds;
        if (T0 < -EXP_THRESHOLD) { T2 = MIN_EXP; dT2_dVd = 0.0; }
        else { T2 = exp(T0); dT2_dVd = -BP(dvtp1) * T2; }
        T3 = Leff + BP(dvtp0) * (1.0 + T2);
        dT3_dVd = BP(dvtp0) * dT2_dVd;
        if ((int)MP(tempMod) < 2) { T4 = Vtm * log(Leff / T3); dT4_dVd = -Vtm * dT3_dVd / T3; }
        else { T4 = m->vtm0 * log(Leff / T3); dT4_dVd = -m->vtm0 * dT3_dVd / T3; }
        dDITS_Sft_dVd = dn_dVd * T4 + n * dT4_dVd;
        dDITS_Sft_dVb = T4 * dn_dVb;
        Vth -= n * T4;
        dVth_dVd -= dDITS_Sft_dVd;
        dVth_dVb -= dDITS_Sft_dVb;
    }
    if ((BP(dvtp4) == 0.0) || (h->dvtp2factor == 0.0)) {
        T0 = 0.0; DITS_Sft2 = 0.0;
    } else {
        T1 = 2.0 * BP(dvtp4) * Vds;
        dexp3(T1, &T0, &T10);
        DITS_Sft2 = h->dvtp2factor * (T0 - 1) / (T0 + 1);
        dDITS_Sft2_dVd = h->dvtp2factor * BP(dvtp4) * 4.0 * T10 / ((T0 + 1) * (T0 + 1));
        Vth -= DITS_Sft2;
        dVth_dVd -= dDITS_Sft2_dVd;
    }
    h->von = Vth;

    /* ---- poly depletion, Vgsteff (b4ld.c:1245-1332) */
    T0 = h->vfb + h->phi;
    poly_depletion(T0, BP(ngate), EPSSI, m->coxe, vgs, &vgs_eff, &dvgs_eff_dvg);
    poly_depletion(T0, BP(ngate), EPSSI, m->coxe, vgd, &vgd_eff, &dvgd_eff_dvg);
    if (h->mode > 0) { Vgs_eff = vgs_eff; dVgs_eff_dVg = dvgs_eff_dvg; }
    else { Vgs_eff = vgd_eff; dVgs_eff_dVg = dvgd_eff_dvg; }
    h->vgs_eff = vgs_eff; h->vgd_eff = vgd_eff;
    h->dvgs_eff_dvg = dvgs_eff_dvg; h->dvgd_eff_dvg = dvgd_eff_dvg;
    Vgst = Vgs_eff - Vth;
    T0 = n * Vtm;
    T1 = h->mstar * Vgst;
    T2 = T1 / T0;
    if (T2 > EXP_THRESHOLD) {
        T10 = T1;
        dT10_dVg = h->mstar * dVgs_eff_dVg;
        dT10_dVd = -dVth_dVd * h->mstar;
        dT10_dVb = -dVth_dVb * h->mstar;
    } else if (T2 < -EXP_THRESHOLD) {
        T10 = Vtm * log(1.0 + MIN_EXP);
        dT10_dVg = 0.0;
        dT10_dVd = T10 * dn_dVd;
        dT10_dVb = T10 * dn_dVb;
        T10 *= n;
    } else {
        ExpVgst = exp(T2);
        T3 = Vtm * log(1.0 + ExpVgst);
        T10 = n * T3;
        dT10_dVg = h->mstar * ExpVgst / (1.0 + ExpVgst);
        dT10_dVb = T3 * dn_dVb - dT10_dVg * (dVth_dVb + Vgst * dn_dVb / n);
        dT10_dVd = T3 * dn_dVd - dT10_dVg * (dVth_dVd + Vgst * dn_dVd / n);
        dT10_dVg *= dVgs_eff_dVg;
    }
    T1 = h->voffcbn - (1.0 - h->mstar) * Vgst;
    T2 = T1 / T0;
    if (T2 < -EXP_THRESHOLD) {
        T3 = m->coxe * MIN_EXP / h->cdep0;
        T9 = h->mstar + T3 * n;
        dT9_dVg = 0.0; dT9_dVd = dn_dVd * T3; dT9_dVb = dn_dVb * T3;
    } else if (T2 > EXP_THRESHOLD) {
        T3 = m->coxe * MAX_EXP / h->cdep0;
        T9 = h->mstar + T3 * n;
        dT9_dVg = 0.0; dT9_dVd = dn_dVd * T3; dT9_dVb = dn_dVb * T3;
    } else {
        ExpVgst = exp(T2);
        T3 = m->coxe / h->cdep0;
        T4 = T3 * ExpVgst;
        T5 = T1 * T4 / T0;
        T9 = h->mstar + n * T4;
        dT9_dVg = T3 * (h->mstar - 1.0) * ExpVgst / Vtm;
        dT9_dVb = T4 * dn_dVb - dT9_dVg * dVth_dVb - T5 * dn_dVb;
        dT9_dVd = T4 * dn_dVd - dT9_dVg * dVth_dVd - T5 * dn_dVd;
        dT9_dVg *= dVgs_eff_dVg;
    }
    h->Vgsteff = Vgsteff = T10 / T9;
    T11 = T9 * T9;
    dVgsteff_dVg = (T9 * dT10_dVg - T10 * dT9_dVg) / T11;
    dVgsteff_dVd = (T9 * dT10_dVd - T10 * dT9_dVd) / T11;
    dVgsteff_dVb = (T9 * dT10_dVb - T10 * dT9_dVb) / T11;

    /* ---- effective geometry, Rds, Abulk (b4ld.c:1334-1450) */
    T9 = sqrtPhis - h->sqrtPhi;
    Weff = h->weff - 2.0 * (BP(dwg) * Vgsteff + BP(dwb) * T9);
    dWeff_dVg = -2.0 * BP(dwg);
    dWeff_dVb = -2.0 * BP(dwb) * dsqrtPhis_dVb;
    if (Weff < 2.0e-8) {
        T0 = 1.0 / (6.0e-8 - 2.0 * Weff);
        Weff = 2.0e-8 * (4.0e-8 - Weff) * T0;
        T0 *= T0 * 4.0e-16;
        dWeff_dVg *= T0;
        dWeff_dVb *= T0;
    }
    T0 = 1.0 + BP(prwg) * Vgsteff;
    dT0_dVg = -BP(prwg) / T0 / T0;
    T1 = BP(prwb) * T9;
    dT1_dVb = BP(prwb) * dsqrtPhis_dVb;
    T2 = 1.0 / T0 + T1;
    T3 = T2 + sqrt(T2 * T2 + 0.01);
    dT3_dVg = 1.0 + T2 / (T3 - T2);
    dT3_dVb = dT3_dVg * dT1_dVb;
    dT3_dVg *= dT0_dVg;
    T4 = h->rds0 * 0.5;
    Rds = h->rdswminP + T3 * T4;
    dRds_dVg = T4 * dT3_dVg;
    dRds_dVb = T4 * dT3_dVb;
    h->grdsw = (Rds > 0.0) ? 1.0 / Rds * nf : 0.0;

    T9 = 0.5 * h->k1ox * Lpe_Vb / sqrtPhis;
    T1 = T9 + h->k2ox - BP(k3b) * Vth_NarrowW;
    dT1_dVb = -T9 / sqrtPhis * dsqrtPhis_dVb;
    T9 = sqrt(BP(xj) * Xdep);
    tmp1 = Leff + 2.0 * T9;
    T5 = Leff / tmp1;
    tmp2 = BP(a0) * T5;
    tmp3 = h->weff + BP(b1);
    tmp4 = BP(b0) / tmp3;
    T2 = tmp2 + tmp4;
    dT2_dVb = -T9 / tmp1 / Xdep * dXdep_dVb;
    T6 = T5 * T5;
    T7 = T5 * T6;
    Abulk0 = 1.0 + T1 * T2;
    dAbulk0_dVb = T1 * tmp2 * dT2_dVb + T2 * dT1_dVb;
    T8 = BP(ags) * BP(a0) * T7;
    dAbulk_dVg = -T1 * T8;
    Abulk = Abulk0 + dAbulk_dVg * Vgsteff;
    dAbulk_dVb = dAbulk0_dVb - T8 * Vgsteff * (dT1_dVb + 3.0 * T1 * dT2_dVb);
    if (Abulk0 < 0.1) {
        T9 = 1.0 / (3.0 - 20.0 * Abulk0);
        Abulk0 = (0.2 - Abulk0) * T9;
        dAbulk0_dVb *= T9 * T9;
    }
    if (Abulk < 0.1) {
        T9 = 1.0 / (3.0 - 20.0 * Abulk);
        Abulk = (0.2 - Abulk) * T9;
        T10 = T9 * T9;
        dAbulk_dVb *= T10;
        dAbulk_dVg *= T10;
    }
    h->Abulk = Abulk;
    T2 = BP(keta) * Vbseff;
    if (T2 >= -0.9) { T0 = 1.0 / (1.0 + T2); dT0_dVb = -BP(keta) * T0 * T0; }
    else { T1 = 1.0 / (0.8 + T2); T0 = (17.0 + 20.0 * T2) * T1; dT0_dVb = -BP(keta) * T1 * T1; }
    dAbulk_dVg *= T0;
    dAbulk_dVb = dAbulk_dVb * T0 + Abulk * dT0_dVb;
    dAbulk0_Q_dVb = dAbulk0_dVb;
    dAbulk0_dVb = dAbulk0_dVb * T0 + Abulk0 * dT0_dVb;
    Abulk *= T0;
    Abulk0_Q = Abulk0;
    Abulk0 *= T0;
    if (BP(ketac) != BP(keta)) {
        T2 = BP(ketac) * Vbseff;
        if (T2 >= -0.9) { T0 = 1.0 / (1.0 + T2); dT0_dVb = -BP(ketac) * T0 * T0; }
        else { T1 = 1.0 / (0.8 + T2); T0 = (17.0 + 20.0 * T2) * T1; dT0_dVb = -BP(ketac) * T1 * T1; }
        dAbulk0_Q_dVb = dAbulk0_Q_dVb * T0 + Abulk0_Q * dT0_dVb;
        Abulk0_Q *= T0;
    } else {
        dAbulk0_Q_dVb = dAbulk0_dVb;
        Abulk0_Q = Abulk0;
    }

    /* ---- mobility (b4ld.c:1450-1614) */
    T14 = 0.0;
    switch ((int)MP(mobMod)) {
    case 0:
        T0 = Vgsteff + Vth + Vth - T14;
        T2 = BP(ua) + BP(uc) * Vbseff;
        T3 = T0 / toxe;
        T12 = sqrt(Vth * Vth + 0.0001);
        T9 = 1.0 / (Vgsteff + 2 * T12);
        T10 = T9 * toxe;
        T8 = BP(ud) * T10 * T10 * Vth;
        T6 = T8 * Vth;
        T5 = T3 * (T2 + BP(ub) * T3) + T6;
        T7 = -2.0 * T6 * T9;
        T11 = T7 * Vth / T12;
        dDenomi_dVg = (T2 + 2.0 * BP(ub) * T3) / toxe;
        T13 = 2.0 * (dDenomi_dVg + T11 + T8);
        dDenomi_dVd = T13 * dVth_dVd;
        dDenomi_dVb = T13 * dVth_dVb + BP(uc) * T3;
        dDenomi_dVg += T7;
        break;
    case 1:
        T0 = Vgsteff + Vth + Vth - T14;
        T2 = 1.0 + BP(uc) * Vbseff;
        T3 = T0 / toxe;
        T4 = T3 * (BP(ua) + BP(ub) * T3);
        T12 = sqrt(Vth * Vth + 0.0001);
        T9 = 1.0 / (Vgsteff + 2 * T12);
        T10 = T9 * toxe;
        T8 = BP(ud) * T10 * T10 * Vth;
        T6 = T8 * Vth;
        T5 = T4 * T2 + T6;
        T7 = -2.0 * T6 * T9;
        T11 = T7 * Vth / T12;
        dDenomi_dVg = (BP(ua) + 2.0 * BP(ub) * T3) * T2 / toxe;
        T13 = 2.0 * (dDenomi_dVg + T11 + T8);
        dDenomi_dVd = T13 * dVth_dVd;
        dDenomi_dVb = T13 * dVth_dVb + BP(uc) * T4;
        dDenomi_dVg += T7;
        break;
    case 2:
        T0 = (Vgsteff + h->vtfbphi1) / toxe;
        T1 = exp(BP(eu) * log(T0));
        dT1_dVg = T1 * BP(eu) / T0 / toxe;
        T2 = BP(ua) + BP(uc) * Vbseff;
        T12 = sqrt(Vth * Vth + 0.0001);
        T9 = 1.0 / (Vgsteff + 2 * T12);
        T10 = T9 * toxe;
        T8 = BP(ud) * T10 * T10 * Vth;
        T6 = T8 * Vth;
        T5 = T1 * T2 + T6;
        T7 = -2.0 * T6 * T9;
        T11 = T7 * Vth / T12;
        dDenomi_dVg = T2 * dT1_dVg + T7;
        T13 = 2.0 * (T11 + T8);
        dDenomi_dVd = T13 * dVth_dVd;
        dDenomi_dVb = T13 * dVth_dVb + T1 * BP(uc);
        break;
    case 4:
        T0 = Vgsteff + h->vtfbphi1 - T14;
        T2 = BP(ua) + BP(uc) * Vbseff;
        T3 = T0 / toxe;
        T12 = sqrt(h->vtfbphi1 * h->vtfbphi1 + 0.0001);
        T9 = 1.0 / (Vgsteff + 2 * T12);
        T10 = T9 * toxe;
        T8 = BP(ud) * T10 * T10 * h->vtfbphi1;
        T6 = T8 * h->vtfbphi1;
        T5 = T3 * (T2 + BP(ub) * T3) + T6;
        T7 = -2.0 * T6 * T9;
        dDenomi_dVg = (T2 + 2.0 * BP(ub) * T3) / toxe;
        dDenomi_dVd = 0.0;
        dDenomi_dVb = BP(uc) * T3;
        dDenomi_dVg += T7;
        break;
    case 5:
        T0 = Vgsteff + h->vtfbphi1 - T14;
        T2 = 1.0 + BP(uc) * Vbseff;
        T3 = T0 / toxe;
        T4 = T3 * (BP(ua) + BP(ub) * T3);
        T12 = sqrt(h->vtfbphi1 * h->vtfbphi1 + 0.0001);
        T9 = 1.0 / (Vgsteff + 2 * T12);
        T10 = T9 * toxe;
        T8 = BP(ud) * T10 * T10 * h->vtfbphi1;
        T6 = T8 * h->vtfbphi1;
        T5 = T4 * T2 + T6;
        T7 = -2.0 * T6 * T9;
        dDenomi_dVg = (BP(ua) + 2.0 * BP(ub) * T3) * T2 / toxe;
        dDenomi_dVd = 0.0;
        dDenomi_dVb = BP(uc) * T4;
        dDenomi_dVg += T7;
        break;
    case 6:
        T0 = (Vgsteff + h->vtfbphi1) / toxe;
        T1 = exp(BP(eu) * log(T0));
        dT1_dVg = T1 * BP(eu) / T0 / toxe;
        T2 = BP(ua) + BP(uc) * Vbseff;
        T12 = sqrt(h->vtfbphi1 * h->vtfbphi1 + 0.0001);
        T9 = 1.0 / (Vgsteff + 2 * T12);
        T10 = T9 * toxe;
        T8 = BP(ud) * T10 * T10 * h->vtfbphi1;
        T6 = T8 * h->vtfbphi1;
        T5 = T1 * T2 + T6;
        T7 = -2.0 * T6 * T9;
        dDenomi_dVg = T2 * dT1_dVg + T7;
        dDenomi_dVd = 0;
        dDenomi_dVb = T1 * BP(uc);
        break;
    default: {   /* mobMod 3 */
        double VgsteffVth = h->VgsteffVth;
        T0 = (Vgsteff + h->vtfbphi1) * 1.0e-8 / toxe / 6.0;
        T1 = exp(BP(eu) * log(T0));
        dT1_dVg = T1 * BP(eu) * 1.0e-8 / T0 / toxe / 6.0;
        T2 = BP(ua) + BP(uc) * Vbseff;
        T10 = exp(BP(ucs) * log(0.5 + 0.5 * Vgsteff / VgsteffVth));
        T11 = BP(ud) / T10;
        dT11_dVg = -0.5 * BP(ucs) * T11 / (0.5 + 0.5 * Vgsteff / VgsteffVth) / VgsteffVth;
        dDenomi_dVg = T2 * dT1_dVg + dT11_dVg;
        dDenomi_dVd = 0.0;
        dDenomi_dVb = T1 * BP(uc);
        T5 = T1 * T2 + T11;
        break; }
    }
    if (T5 >= -0.8) Denomi = 1.0 + T5;
    else {
        T9 = 1.0 / (7.0 + 10.0 * T5);
        Denomi = (0.6 + T5) * T9;
        T9 *= T9;
        dDenomi_dVg *= T9; dDenomi_dVd *= T9; dDenomi_dVb *= T9;
    }
    h->ueff = ueff = h->u0temp / Denomi;
    T9 = -ueff / Denomi;
    dueff_dVg = T9 * dDenomi_dVg;
    dueff_dVd = T9 * dDenomi_dVd;
    dueff_dVb = T9 * dDenomi_dVb;

    /* ---- Vdsat (b4ld.c:1614-1700) */
    WVCox = Weff * h->vsattemp * m->coxe;
    WVCoxRds = WVCox * Rds;
    Esat = 2.0 * h->vsattemp / ueff;
    h->EsatL = EsatL = Esat * Leff;
    T0 = -EsatL / ueff;
    dEsatL_dVg = T0 * dueff_dVg;
    dEsatL_dVd = T0 * dueff_dVd;
    dEsatL_dVb = T0 * dueff_dVb;
    a1 = BP(a1);
    if (a1 == 0.0) { Lambda = BP(a2); dLambda_dVg = 0.0; }
    else if (a1 > 0.0) {
        T0 = 1.0 - BP(a2);
        T1 = T0 - BP(a1) * Vgsteff - 0.0001;
        T2 = sqrt(T1 * T1 + 0.0004 * T0);
        Lambda = BP(a2) + T0 - 0.5 * (T1 + T2);
        dLambda_dVg = 0.5 * BP(a1) * (1.0 + T1 / T2);
    } else {
        T1 = BP(a2) + BP(a1) * Vgsteff - 0.0001;
        T2 = sqrt(T1 * T1 + 0.0004 * BP(a2));
        Lambda = 0.5 * (T1 + T2);
        dLambda_dVg = 0.5 * BP(a1) * (1.0 + T1 / T2);
    }
    Vgst2Vtm = Vgsteff + 2.0 * Vtm;
    if (Rds > 0) { tmp2 = dRds_dVg / Rds + dWeff_dVg / Weff; tmp3 = dRds_dVb / Rds + dWeff_dVb / Weff; }
    else { tmp2 = dWeff_dVg / Weff; tmp3 = dWeff_dVb / Weff; }
    if ((Rds == 0.0) && (Lambda == 1.0)) {
        T0 = 1.0 / (Abulk * EsatL + Vgst2Vtm);
        tmp1 = 0.0;
        T1 = T0 * T0;
        T2 = Vgst2Vtm * T0;
        T3 = EsatL * Vgst2Vtm;
        Vdsat = T3 * T0;
        dT0_dVg = -(Abulk * dEsatL_dVg + EsatL * dAbulk_dVg + 1.0) * T1;
        dT0_dVd = -(Abulk * dEsatL_dVd) * T1;
        dT0_dVb = -(Abulk * dEsatL_dVb + dAbulk_dVb * EsatL) * T1;
        dVdsat_dVg = T3 * dT0_dVg + T2 * dEsatL_dVg + EsatL * T0;
        dVdsat_dVd = T3 * dT0_dVd + T2 * dEsatL_dVd;
        dVdsat_dVb = T3 * dT0_dVb + T2 * dEsatL_dVb;
    } else {
        tmp1 = dLambda_dVg / (Lambda * Lambda);
        T9 = Abulk * WVCoxRds;
        T8 = Abulk * T9;
        T7 = Vgst2Vtm * T9;
        T6 = Vgst2Vtm * WVCoxRds;
        T0 = 2.0 * Abulk * (T9 - 1.0 + 1.0 / Lambda);
        dT0_dVg = 2.0 * (T8 * tmp2 - Abulk * tmp1 + (2.0 * T9 + 1.0 / Lambda - 1.0) * dAbulk_dVg);
        dT0_dVb = 2.0 * (T8 * (2.0 / Abulk * dAbulk_dVb + tmp3) + (1.0 / Lambda - 1.0) * dAbulk_dVb);
        dT0_dVd = 0.0;
        T1 = Vgst2Vtm * (2.0 / Lambda - 1.0) + Abulk * EsatL + 3.0 * T7;
        dT1_dVg = (2.0 / Lambda - 1.0) - 2.0 * Vgst2Vtm * tmp1 + Abulk * dEsatL_dVg
                + EsatL * dAbulk_dVg + 3.0 * (T9 + T7 * tmp2 + T6 * dAbulk_dVg);
        dT1_dVb = Abulk * dEsatL_dVb + EsatL * dAbulk_dVb + 3.0 * (T6 * dAbulk_dVb + T7 * tmp3);
        dT1_dVd = Abulk * dEsatL_dVd;
        T2 = Vgst2Vtm * (EsatL + 2.0 * T6);
        dT2_dVg = EsatL + Vgst2Vtm * dEsatL_dVg + T6 * (4.0 + 2.0 * Vgst2Vtm * tmp2);
        dT2_dVb = Vgst2Vtm * (dEsatL_dVb + 2.0 * T6 * tmp3);
        dT2_dVd = Vgst2Vtm * dEsatL_dVd;
        T3 = sqrt(T1 * T1 - 2.0 * T0 * T2);
        Vdsat = (T1 - T3) / T0;
        dT3_dVg = (T1 * dT1_dVg - 2.0 * (T0 * dT2_dVg + T2 * dT0_dVg)) / T3;
        dT3_dVd = (T1 * dT1_dVd - 2.0 * (T0 * dT2_dVd + T2 * dT0_dVd)) / T3;
        dT3_dVb = (T1 * dT1_dVb - 2.0 * (T0 * dT2_dVb + T2 * dT0_dVb)) / T3;
        (void)dT3_dVg; (void)dT3_dVd; (void)dT3_dVb;
        dVdsat_dVg = (dT1_dVg - (T1 * dT1_dVg - dT0_dVg * T2 - T0 * dT2_dVg) / T3 - Vdsat * dT0_dVg) / T0;
        dVdsat_dVb = (dT1_dVb - (T1 * dT1_dVb - dT0_dVb * T2 - T0 * dT2_dVb) / T3 - Vdsat * dT0_dVb) / T0;
        dVdsat_dVd = (dT1_dVd - (T1 * dT1_dVd - T0 * dT2_dVd) / T3) / T0;
    }
    h->vdsat = Vdsat;

    /* ---- Vdseff (b4ld.c:1702-1757) */
    T1 = Vdsat - Vds - BP(delta);
    dT1_dVg = dVdsat_dVg;
    dT1_dVd = dVdsat_dVd - 1.0;
    dT1_dVb = dVdsat_dVb;
    T2 = sqrt(T1 * T1 + 4.0 * BP(delta) * Vdsat);
    T0 = T1 / T2;
    T9 = 2.0 * BP(delta);
    T3 = T9 / T2;
    dT2_dVg = T0 * dT1_dVg + T3 * dVdsat_dVg;
    dT2_dVd = T0 * dT1_dVd + T3 * dVdsat_dVd;
    dT2_dVb = T0 * dT1_dVb + T3 * dVdsat_dVb;
    if (T1 >= 0.0) {
        Vdseff = Vdsat - 0.5 * (T1 + T2);
        dVdseff_dVg = dVdsat_dVg - 0.5 * (dT1_dVg + dT2_dVg);
        dVdseff_dVd = dVdsat_dVd - 0.5 * (dT1_dVd + dT2_dVd);
        dVdseff_dVb = dVdsat_dVb - 0.5 * (dT1_dVb + dT2_dVb);
    } else {
        T4 = T9 / (T2 - T1);
        T5 = 1.0 - T4;
        T6 = Vdsat * T4 / (T2 - T1);
        Vdseff = Vdsat * T5;
        dVdseff_dVg = dVdsat_dVg * T5 + T6 * (dT2_dVg - dT1_dVg);
        dVdseff_dVd = dVdsat_dVd * T5 + T6 * (dT2_dVd - dT1_dVd);
        dVdseff_dVb = dVdsat_dVb * T5 + T6 * (dT2_dVb - dT1_dVb);
    }
    if (Vds == 0.0) { Vdseff = 0.0; dVdseff_dVg = 0.0; dVdseff_dVb = 0.0; }
    if (Vdseff > Vds) Vdseff = Vds;
    diffVds = Vds - Vdseff;
    h->Vdseff = Vdseff;

    /* velocity overshoot (b4ld.c:1759-1808) */
    if (MG(lambda) && (MP(lambda) > 0.0)) {
        T1 = Leff * ueff;
        T2 = BP(lambda) / T1;
        T3 = -T2 / T1 * Leff;
        dT2_dVd = T3 * dueff_dVd;
        dT2_dVg = T3 * dueff_dVg;
        dT2_dVb = T3 * dueff_dVb;
        T5 = 1.0 / (Esat * h->litl);
        T4 = -T5 / EsatL;
        dT5_dVg = dEsatL_dVg * T4;
        dT5_dVd = dEsatL_dVd * T4;
        dT5_dVb = dEsatL_dVb * T4;
        T6 = 1.0 + diffVds * T5;
        dT6_dVg = dT5_dVg * diffVds - dVdseff_dVg * T5;
        dT6_dVd = dT5_dVd * diffVds + (1.0 - dVdseff_dVd) * T5;
        dT6_dVb = dT5_dVb * diffVds - dVdseff_dVb * T5;
        T7 = 2.0 / (T6 * T6 + 1.0);
        T8 = 1.0 - T7;
        T9 = T6 * T7 * T7;
        dT8_dVg = T9 * dT6_dVg;
        dT8_dVd = T9 * dT6_dVd;
        dT8_dVb = T9 * dT6_dVb;
        T10 = 1.0 + T2 * T8;
        dT10_dVg = dT2_dVg * T8 + T2 * dT8_dVg;
        dT10_dVd = dT2_dVd * T8 + T2 * dT8_dVd;
        dT10_dVb = dT2_dVb * T8 + T2 * dT8_dVb;
        if (T10 == 1.0) dT10_dVg = dT10_dVd = dT10_dVb = 0.0;
        dEsatL_dVg *= T10; dEsatL_dVg += EsatL * dT10_dVg;
        dEsatL_dVd *= T10; dEsatL_dVd += EsatL * dT10_dVd;
        dEsatL_dVb *= T10; dEsatL_dVb += EsatL * dT10_dVb;
        EsatL *= T10;
        Esat = EsatL / Leff;
        h->EsatL = EsatL;
    }

    /* ---- Vasat (b4ld.c:1810-1838) */
    tmp4 = 1.0 - 0.5 * Abulk * Vdsat / Vgst2Vtm;
    T9 = WVCoxRds * Vgsteff;
    T8 = T9 / Vgst2Vtm;
    T0 = EsatL + Vdsat + 2.0 * T9 * tmp4;
    T7 = 2.0 * WVCoxRds * tmp4;
    dT0_dVg = dEsatL_dVg + dVdsat_dVg + T7 * (1.0 + tmp2 * Vgsteff)
            - T8 * (Abulk * dVdsat_dVg - Abulk * Vdsat / Vgst2Vtm + Vdsat * dAbulk_dVg);
    dT0_dVb = dEsatL_dVb + dVdsat_dVb + T7 * tmp3 * Vgsteff
            - T8 * (dAbulk_dVb * Vdsat + Abulk * dVdsat_dVb);
    dT0_dVd = dEsatL_dVd + dVdsat_dVd - T8 * Abulk * dVdsat_dVd;
    T9 = WVCoxRds * Abulk;
    T1 = 2.0 / Lambda - 1.0 + T9;
    dT1_dVg = -2.0 * tmp1 + WVCoxRds * (Abulk * tmp2 + dAbulk_dVg);
    dT1_dVb = dAbulk_dVb * WVCoxRds + T9 * tmp3;
    Vasat = T0 / T1;
    dVasat_dVg = (dT0_dVg - Vasat * dT1_dVg) / T1;
    dVasat_dVb = (dT0_dVb - Vasat * dT1_dVb) / T1;
    dVasat_dVd = dT0_dVd / T1;

    /* ---- Idl (b4ld.c:1840-1893) */
    tmp1 = h->vtfbphi2;
    tmp2 = 2.0e8 * h->toxp;
    dT0_dVg = 1.0 / tmp2;
    T0 = (Vgsteff + tmp1) * dT0_dVg;
    tmp3 = exp(MP(bdos) * 0.7 * log(T0));
    T1 = 1.0 + tmp3;
    T2 = MP(bdos) * 0.7 * tmp3 / T0;
    Tcen = MP(ados) * 1.9e-9 / T1;
    dTcen_dVg = -Tcen * T2 * dT0_dVg / T1;
    Coxeff = epssub * h->coxp / (epssub + h->coxp * Tcen);
    h->Coxeff = Coxeff;
    dCoxeff_dVg = -Coxeff * Coxeff * dTcen_dVg / epssub;
    CoxeffWovL = Coxeff * Weff / Leff;
    beta = ueff * CoxeffWovL;
    T3 = ueff / Leff;
    dbeta_dVg = CoxeffWovL * dueff_dVg + T3 * (Weff * dCoxeff_dVg + Coxeff * dWeff_dVg);
    dbeta_dVd = CoxeffWovL * dueff_dVd;
    dbeta_dVb = CoxeffWovL * dueff_dVb + T3 * Coxeff * dWeff_dVb;
    h->AbovVgst2Vtm = Abulk / Vgst2Vtm;
    T0 = 1.0 - 0.5 * Vdseff * h->AbovVgst2Vtm;
    dT0_dVg = -0.5 * (Abulk * dVdseff_dVg - Abulk * Vdseff / Vgst2Vtm + Vdseff * dAbulk_dVg) / Vgst2Vtm;
    dT0_dVd = -0.5 * Abulk * dVdseff_dVd / Vgst2Vtm;
    dT0_dVb = -0.5 * (Abulk * dVdseff_dVb + dAbulk_dVb * Vdseff) / Vgst2Vtm;
    fgche1 = Vgsteff * T0;
    dfgche1_dVg = Vgsteff * dT0_dVg + T0;
    dfgche1_dVd = Vgsteff * dT0_dVd;
    dfgche1_dVb = Vgsteff * dT0_dVb;
    T9 = Vdseff / EsatL;
    fgche2 = 1.0 + T9;
    dfgche2_dVg = (dVdseff_dVg - T9 * dEsatL_dVg) / EsatL;
    dfgche2_dVd = (dVdseff_dVd - T9 * dEsatL_dVd) / EsatL;
    dfgche2_dVb = (dVdseff_dVb - T9 * dEsatL_dVb) / EsatL;
    gche = beta * fgche1 / fgche2;
    dgche_dVg = (beta * dfgche1_dVg + fgche1 * dbeta_dVg - gche * dfgche2_dVg) / fgche2;
    dgche_dVd = (beta * dfgche1_dVd + fgche1 * dbeta_dVd - gche * dfgche2_dVd) / fgche2;
    dgche_dVb = (beta * dfgche1_dVb + fgche1 * dbeta_dVb - gche * dfgche2_dVb) / fgche2;
    T0 = 1.0 + gche * Rds;
    Idl = gche / T0;
    T1 = (1.0 - Idl * Rds) / T0;
    T2 = Idl * Idl;
    dIdl_dVg = T1 * dgche_dVg - T2 * dRds_dVg;
    dIdl_dVd = T1 * dgche_dVd;
    dIdl_dVb = T1 * dgche_dVb - T2 * dRds_dVb;

    /* ---- output resistance terms (b4ld.c:1895-2044) */
    if (BP(fprout) <= 0.0) { FP = 1.0; dFP_dVg = 0.0; }
    else {
        T9 = BP(fprout) * sqrt(Leff) / Vgst2Vtm;
        FP = 1.0 / (1.0 + T9);
        dFP_dVg = FP * FP * T9 / Vgst2Vtm;
    }
    T8 = BP(pvag) / EsatL;
    T9 = T8 * Vgsteff;
    if (T9 > -0.9) {
        PvagTerm = 1.0 + T9;
        dPvagTerm_dVg = T8 * (1.0 - Vgsteff * dEsatL_dVg / EsatL);
        dPvagTerm_dVb = -T9 * dEsatL_dVb / EsatL;
        dPvagTerm_dVd = -T9 * dEsatL_dVd / EsatL;
    } else {
        T4 = 1.0 / (17.0 + 20.0 * T9);
        PvagTerm = (0.8 + T9) * T4;
        T4 *= T4;
        dPvagTerm_dVg = T8 * (1.0 - Vgsteff * dEsatL_dVg / EsatL) * T4;
        T9 *= T4 / EsatL;
        dPvagTerm_dVb = -T9 * dEsatL_dVb;
        dPvagTerm_dVd = -T9 * dEsatL_dVd;
    }
    if ((BP(pclm) > MIN_EXP) && (diffVds > 1.0e-10)) {
        T0 = 1.0 + Rds * Idl;
        dT0_dVg = dRds_dVg * Idl + Rds * dIdl_dVg;
        dT0_dVd = Rds * dIdl_dVd;
        dT0_dVb = dRds_dVb * Idl + Rds * dIdl_dVb;
        T2 = Vdsat / Esat;
        T1 = Leff + T2;
        dT1_dVg = (dVdsat_dVg - T2 * dEsatL_dVg / Leff) / Esat;
        dT1_dVd = (dVdsat_dVd - T2 * dEsatL_dVd / Leff) / Esat;
        dT1_dVb = (dVdsat_dVb - T2 * dEsatL_dVb / Leff) / Esat;
        Cclm = FP * PvagTerm * T0 * T1 / (BP(pclm) * h->litl);
        dCclm_dVg = Cclm * (dFP_dVg / FP + dPvagTerm_dVg / PvagTerm + dT0_dVg / T0 + dT1_dVg / T1);
        dCclm_dVb = Cclm * (dPvagTerm_dVb / PvagTerm + dT0_dVb / T0 + dT1_dVb / T1);
        dCclm_dVd = Cclm * (dPvagTerm_dVd / PvagTerm + dT0_dVd / T0 + dT1_dVd / T1);
        VACLM = Cclm * diffVds;
        dVACLM_dVg = dCclm_dVg * diffVds - dVdseff_dVg * Cclm;
        dVACLM_dVb = dCclm_dVb * diffVds - dVdseff_dVb * Cclm;
        dVACLM_dVd = dCclm_dVd * diffVds + (1.0 - dVdseff_dVd) * Cclm;
    } else {
        VACLM = Cclm = MAX_EXP;
        dVACLM_dVd = dVACLM_dVg = dVACLM_dVb = 0.0;
        dCclm_dVd = dCclm_dVg = dCclm_dVb = 0.0;
    }
    if (h->thetaRout > MIN_EXP) {
        T8 = Abulk * Vdsat;
        T0 = Vgst2Vtm * T8;
        dT0_dVg = Vgst2Vtm * Abulk * dVdsat_dVg + T8 + Vgst2Vtm * Vdsat * dAbulk_dVg;
        dT0_dVb = Vgst2Vtm * (dAbulk_dVb * Vdsat + Abulk * dVdsat_dVb);
        dT0_dVd = Vgst2Vtm * Abulk * dVdsat_dVd;
        T1 = Vgst2Vtm + T8;
        dT1_dVg = 1.0 + Abulk * dVdsat_dVg + Vdsat * dAbulk_dVg;
        dT1_dVb = Abulk * dVdsat_dVb + dAbulk_dVb * Vdsat;
        dT1_dVd = Abulk * dVdsat_dVd;
        T9 = T1 * T1;
        T2 = h->thetaRout;
        VADIBL = (Vgst2Vtm - T0 / T1) / T2;
        dVADIBL_dVg = (1.0 - dT0_dVg / T1 + T0 * dT1_dVg / T9) / T2;
        dVADIBL_dVb = (-dT0_dVb / T1 + T0 * dT1_dVb / T9) / T2;
        dVADIBL_dVd = (-dT0_dVd / T1 + T0 * dT1_dVd / T9) / T2;
        T7 = BP(pdiblb) * Vbseff;
        if (T7 >= -0.9) {
            T3 = 1.0 / (1.0 + T7);
            VADIBL *= T3;
            dVADIBL_dVg *= T3;
            dVADIBL_dVb = (dVADIBL_dVb - VADIBL * BP(pdiblb)) * T3;
            dVADIBL_dVd *= T3;
        } else {
            T4 = 1.0 / (0.8 + T7);
            T3 = (17.0 + 20.0 * T7) * T4;
            dVADIBL_dVg *= T3;
            dVADIBL_dVb = dVADIBL_dVb * T3 - VADIBL * BP(pdiblb) * T4 * T4;
            dVADIBL_dVd *= T3;
            VADIBL *= T3;
        }
        dVADIBL_dVg = dVADIBL_dVg * PvagTerm + VADIBL * dPvagTerm_dVg;
        dVADIBL_dVb = dVADIBL_dVb * PvagTerm + VADIBL * dPvagTerm_dVb;
        dVADIBL_dVd = dVADIBL_dVd * PvagTerm + VADIBL * dPvagTerm_dVd;
        VADIBL *= PvagTerm;
    } else {
        VADIBL = MAX_EXP;
        dVADIBL_dVd = dVADIBL_dVg = dVADIBL_dVb = 0.0;
    }
    Va = Vasat + VACLM;
    dVa_dVg = dVasat_dVg + dVACLM_dVg;
    dVa_dVb = dVasat_dVb + dVACLM_dVb;
    dVa_dVd = dVasat_dVd + dVACLM_dVd;
    T0 = BP(pditsd) * Vds;
    if (T0 > EXP_THRESHOLD) { T1 = MAX_EXP; dT1_dVd = 0; }
    else { T1 = exp(T0); dT1_dVd = T1 * BP(pditsd); }
    if (BP(pdits) > MIN_EXP) {
        T2 = 1.0 + MP(pditsl) * Leff;
        VADITS = (1.0 + T2 * T1) / BP(pdits);
        dVADITS_dVg = VADITS * dFP_dVg;
        dVADITS_dVd = FP * T2 * dT1_dVd / BP(pdits);
        VADITS *= FP;
    } else {
        VADITS = MAX_EXP;
        dVADITS_dVg = dVADITS_dVd = 0;
    }
    if ((BP(pscbe2) > 0.0) && (BP(pscbe1) >= 0.0)) {
        if (diffVds > BP(pscbe1) * h->litl / EXP_THRESHOLD) {
            T0 = BP(pscbe1) * h->litl / diffVds;
            VASCBE = Leff * exp(T0) / BP(pscbe2);
            T1 = T0 * VASCBE / diffVds;
            dVASCBE_dVg = T1 * dVdseff_dVg;
            dVASCBE_dVd = -T1 * (1.0 - dVdseff_dVd);
            dVASCBE_dVb = T1 * dVdseff_dVb;
        } else {
            VASCBE = MAX_EXP * Leff / BP(pscbe2);
            dVASCBE_dVg = dVASCBE_dVd = dVASCBE_dVb = 0.0;
        }
    } else {
        VASCBE = MAX_EXP;
        dVASCBE_dVg = dVASCBE_dVd = dVASCBE_dVb = 0.0;
    }

    /* ---- Ids with DIBL/DITS/CLM, Isub, SCBE (b4ld.c:2046-2152) */
    T9 = diffVds / VADIBL;
    T0 = 1.0 + T9;
    Idsa = Idl * T0;
    dIdsa_dVg = T0 * dIdl_dVg - Idl * (dVdseff_dVg + T9 * dVADIBL_dVg) / VADIBL;
    dIdsa_dVd = T0 * dIdl_dVd + Idl * (1.0 - dVdseff_dVd - T9 * dVADIBL_dVd) / VADIBL;
    dIdsa_dVb = T0 * dIdl_dVb - Idl * (dVdseff_dVb + T9 * dVADIBL_dVb) / VADIBL;
    T9 = diffVds / VADITS;
    T0 = 1.0 + T9;
    dIdsa_dVg = T0 * dIdsa_dVg - Idsa * (dVdseff_dVg + T9 * dVADITS_dVg) / VADITS;
    dIdsa_dVd = T0 * dIdsa_dVd + Idsa * (1.0 - dVdseff_dVd - T9 * dVADITS_dVd) / VADITS;
    dIdsa_dVb = T0 * dIdsa_dVb - Idsa * dVdseff_dVb / VADITS;
    Idsa *= T0;
    T0 = log(Va / Vasat);
    dT0_dVg = dVa_dVg / Va - dVasat_dVg / Vasat;
    dT0_dVb = dVa_dVb / Va - dVasat_dVb / Vasat;
    dT0_dVd = dVa_dVd / Va - dVasat_dVd / Vasat;
    T1 = T0 / Cclm;
    T9 = 1.0 + T1;
    dT9_dVg = (dT0_dVg - T1 * dCclm_dVg) / Cclm;
    dT9_dVb = (dT0_dVb - T1 * dCclm_dVb) / Cclm;
    dT9_dVd = (dT0_dVd - T1 * dCclm_dVd) / Cclm;
    dIdsa_dVg = dIdsa_dVg * T9 + Idsa * dT9_dVg;
    dIdsa_dVb = dIdsa_dVb * T9 + Idsa * dT9_dVb;
    dIdsa_dVd = dIdsa_dVd * T9 + Idsa * dT9_dVd;
    Idsa *= T9;

    tmp = BP(alpha0) + BP(alpha1) * Leff;
    if ((tmp <= 0.0) || (BP(beta0) <= 0.0)) {
        Isub = Gbd = Gbb = Gbg = 0.0;
    } else {
        T2 = tmp / Leff;
        if (diffVds > BP(beta0) / EXP_THRESHOLD) {
            T0 = -BP(beta0) / diffVds;
            T1 = T2 * diffVds * exp(T0);
            T3 = T1 / diffVds * (T0 - 1.0);
            dT1_dVg = T3 * dVdseff_dVg;
            dT1_dVd = T3 * (dVdseff_dVd - 1.0);
            dT1_dVb = T3 * dVdseff_dVb;
        } else {
            T3 = T2 * MIN_EXP;
            T1 = T3 * diffVds;
            dT1_dVg = -T3 * dVdseff_dVg;
            dT1_dVd = T3 * (1.0 - dVdseff_dVd);
            dT1_dVb = -T3 * dVdseff_dVb;
        }
        T4 = Idsa * Vdseff;
        Isub = T1 * T4;
        Gbg = T1 * (dIdsa_dVg * Vdseff + Idsa * dVdseff_dVg) + T4 * dT1_dVg;
        Gbd = T1 * (dIdsa_dVd * Vdseff + Idsa * dVdseff_dVd) + T4 * dT1_dVd;
        Gbb = T1 * (dIdsa_dVb * Vdseff + Idsa * dVdseff_dVb) + T4 * dT1_dVb;
        Gbd += Gbg * dVgsteff_dVd;
        Gbb += Gbg * dVgsteff_dVb;
        Gbg *= dVgsteff_dVg;
        Gbb *= dVbseff_dVb;
    }
    h->csub = Isub; h->gbbs = Gbb; h->gbgs = Gbg; h->gbds = Gbd;

    T9 = diffVds / VASCBE;
    T0 = 1.0 + T9;
    Ids = Idsa * T0;
    Gm = T0 * dIdsa_dVg - Idsa * (dVdseff_dVg + T9 * dVASCBE_dVg) / VASCBE;
    Gds = T0 * dIdsa_dVd + Idsa * (1.0 - dVdseff_dVd - T9 * dVASCBE_dVd) / VASCBE;
    Gmb = T0 * dIdsa_dVb - Idsa * (dVdseff_dVb + T9 * dVASCBE_dVb) / VASCBE;
    tmp1 = Gds + Gm * dVgsteff_dVd;
    tmp2 = Gmb + Gm * dVgsteff_dVb;
    tmp3 = Gm;
    Gm = (Ids * dVdseff_dVg + Vdseff * tmp3) * dVgsteff_dVg;
    Gds = Ids * (dVdseff_dVd + dVdseff_dVg * dVgsteff_dVd) + Vdseff * tmp1;
    Gmb = (Ids * (dVdseff_dVb + dVdseff_dVg * dVgsteff_dVb) + Vdseff * tmp2) * dVbseff_dVb;
    cdrain = Ids * Vdseff;

    if (MG(vtl) && (MP(vtl) > 0.0)) {   /* b4ld.c:2154-2188 */
        T12 = 1.0 / Leff / CoxeffWovL;
        T11 = T12 / Vgsteff;
        T10 = -T11 / Vgsteff;
        vs = cdrain * T11;
        dvs_dVg = Gm * T11 + cdrain * T10 * dVgsteff_dVg;
        dvs_dVd = Gds * T11 + cdrain * T10 * dVgsteff_dVd;
        dvs_dVb = Gmb * T11 + cdrain * T10 * dVgsteff_dVb;
        T0 = 2 * MM;
        T1 = vs / (BP(vtl) * h->tfactor);
        if (T1 > 0.0) {
            T2 = 1.0 + exp(T0 * log(T1));
            T3 = (T2 - 1.0) * T0 / vs;
            Fsevl = 1.0 / exp(log(T2) / T0);
            dT2_dVg = T3 * dvs_dVg;
            dT2_dVd = T3 * dvs_dVd;
            dT2_dVb = T3 * dvs_dVb;
            T4 = -1.0 / T0 * Fsevl / T2;
            dFsevl_dVg = T4 * dT2_dVg;
            dFsevl_dVd = T4 * dT2_dVd;
            dFsevl_dVb = T4 * dT2_dVb;
        } else {
            Fsevl = 1.0;
            dFsevl_dVg = dFsevl_dVd = dFsevl_dVb = 0.0;
        }
        Gm *= Fsevl; Gm += cdrain * dFsevl_dVg;
        Gmb *= Fsevl; Gmb += cdrain * dFsevl_dVb;
        Gds *= Fsevl; Gds += cdrain * dFsevl_dVd;
        cdrain *= Fsevl;
    }
    h->gds = Gds; h->gm = Gm; h->gmbs = Gmb;
    h->IdovVds = Ids;
    if (h->IdovVds <= MP(idovvdsc)) h->IdovVds = MP(idovvdsc);
    (void)beta; (void)dbeta_dVd; (void)dbeta_dVb;
    h->gstot = h->gstotd = h->gstotg = h->gstots = h->gstotb = 0.0;
    h->gdtot = h->gdtotd = h->gdtotg = h->gdtots = h->gdtotb = 0.0;

    /* ---- GIDL / GISL (b4ld.c:2308-2526) */
    T0 = 3.0 * toxe;
    vgs_eff = h->vgs_eff; dvgs_eff_dvg = h->dvgs_eff_dvg;
    vgd_eff = h->vgd_eff; dvgd_eff_dvg = h->dvgd_eff_dvg;
    if ((int)MP(gidlMod) == 0) {
        T1 = (vds - vgs_eff - BP(egidl)) / T0;
        if ((BP(agidl) <= 0.0) || (BP(bgidl) <= 0.0) || (T1 <= 0.0) || (BP(cgidl) <= 0.0) || (vbd > 0.0))
            Igidl = Ggidld = Ggidlg = Ggidlb = 0.0;
        else {
            dT1_dVd = 1.0 / T0;
            dT1_dVg = -dvgs_eff_dvg * dT1_dVd;
            T2 = BP(bgidl) / T1;
            if (T2 < 100.0) {
                Igidl = BP(agidl) * h->weffCJ * T1 * exp(-T2);
                T3 = Igidl * (1.0 + T2) / T1;
                Ggidld = T3 * dT1_dVd;
                Ggidlg = T3 * dT1_dVg;
            } else {
                Igidl = BP(agidl) * h->weffCJ * 3.720075976e-44;
                Ggidld = Igidl * dT1_dVd;
                Ggidlg = Igidl * dT1_dVg;
                Igidl *= T1;
            }
            T4 = vbd * vbd;
            T5 = -vbd * T4;
            T6 = BP(cgidl) + T5;
            T7 = T5 / T6;
            T8 = 3.0 * BP(cgidl) * T4 / T6 / T6;
            Ggidld = Ggidld * T7 + Igidl * T8;
            Ggidlg = Ggidlg * T7;
            Ggidlb = -Igidl * T8;
            Igidl *= T7;
        }
        T1 = (-vds - vgd_eff - BP(egisl)) / T0;
        if ((BP(agisl) <= 0.0) || (BP(bgisl) <= 0.0) || (T1 <= 0.0) || (BP(cgisl) <= 0.0) || (vbs > 0.0))
            Igisl = Ggisls = Ggislg = Ggislb = 0.0;
        else {
            dT1_dVd = 1.0 / T0;
            dT1_dVg = -dvgd_eff_dvg * dT1_dVd;
            T2 = BP(bgisl) / T1;
            if (T2 < 100.0) {
                Igisl = BP(agisl) * h->weffCJ * T1 * exp(-T2);
                T3 = Igisl * (1.0 + T2) / T1;
                Ggisls = T3 * dT1_dVd;
                Ggislg = T3 * dT1_dVg;
            } else {
                Igisl = BP(agisl) * h->weffCJ * 3.720075976e-44;
                Ggisls = Igisl * dT1_dVd;
                Ggislg = Igisl * dT1_dVg;
                Igisl *= T1;
            }
            T4 = vbs * vbs;
            T5 = -vbs * T4;
            T6 = BP(cgisl) + T5;
            T7 = T5 / T6;
            T8 = 3.0 * BP(cgisl) * T4 / T6 / T6;
            Ggisls = Ggisls * T7 + Igisl * T8;
            Ggislg = Ggislg * T7;
            Ggislb = -Igisl * T8;
            Igisl *= T7;
        }
    } else {
        T1 = (-vds - BP(rgisl) * vgd_eff - BP(egisl)) / T0;
        if ((BP(agisl) <= 0.0) || (BP(bgisl) <= 0.0) || (T1 <= 0.0) || (BP(cgisl) < 0.0))
            Igisl = Ggisls = Ggislg = Ggislb = 0.0;
        else {
            dT1_dVd = 1 / T0;
            dT1_dVg = -BP(rgisl) * dT1_dVd * dvgd_eff_dvg;
            T2 = BP(bgisl) / T1;
            if (T2 < EXPL_THRESHOLD) {
                Igisl = h->weffCJ * BP(agisl) * T1 * exp(-T2);
                T3 = Igisl / T1 * (T2 + 1);
                Ggisls = T3 * dT1_dVd;
                Ggislg = T3 * dT1_dVg;
            } else {
                T3 = h->weffCJ * BP(agisl) * MIN_EXPL;
                Igisl = T3 * T1;
                Ggisls = T3 * dT1_dVd;
                Ggislg = T3 * dT1_dVg;
            }
            T4 = vbs - BP(fgisl);
            if (T4 > MP(gidlclamp)) T4 = MP(gidlclamp);
            if (T4 == 0) T5 = EXPL_THRESHOLD;
            else T5 = BP(kgisl) / T4;
            if (T5 < EXPL_THRESHOLD) { T6 = exp(T5); Ggislb = -Igisl * T6 * T5 / T4; }
            else { T6 = MAX_EXPL; Ggislb = 0.0; }
            Ggisls *= T6; Ggislg *= T6; Igisl *= T6;
        }
        T1 = (vds - BP(rgidl) * vgs_eff - BP(egidl)) / T0;
        if ((BP(agidl) <= 0.0) || (BP(bgidl) <= 0.0) || (T1 <= 0.0) || (BP(cgidl) < 0.0))
            Igidl = Ggidld = Ggidlg = Ggidlb = 0.0;
        else {
            dT1_dVd = 1 / T0;
            dT1_dVg = -BP(rgidl) * dT1_dVd * dvgs_eff_dvg;
            T2 = BP(bgidl) / T1;
            if (T2 < EXPL_THRESHOLD) {
                Igidl = h->weffCJ * BP(agidl) * T1 * exp(-T2);
                T3 = Igidl / T1 * (T2 + 1);
                Ggidld = T3 * dT1_dVd;
                Ggidlg = T3 * dT1_dVg;
            } else {
                T3 = h->weffCJ * BP(agidl) * MIN_EXPL;
                Igidl = T3 * T1;
                Ggidld = T3 * dT1_dVd;
                Ggidlg = T3 * dT1_dVg;
            }
            T4 = vbd - BP(fgidl);
            if (T4 > MP(gidlclamp)) T4 = MP(gidlclamp);
            if (T4 == 0) T5 = EXPL_THRESHOLD;
            else T5 = BP(kgidl) / T4;
            if (T5 < EXPL_THRESHOLD) { T6 = exp(T5); Ggidlb = -Igidl * T6 * T5 / T4; }
            else { T6 = MAX_EXPL; Ggidlb = 0.0; }
            Ggidld *= T6; Ggidlg *= T6; Igidl *= T6;
        }
    }
    h->Igidl = Igidl; h->ggidld = Ggidld; h->ggidlg = Ggidlg; h->ggidlb = Ggidlb;
    h->Igisl = Igisl; h->ggisls = Ggisls; h->ggislg = Ggislg; h->ggislb = Ggislb;

    /* ---- gate tunnelling (b4ld.c:2527-2978) */
    if (((int)MP(igcMod) != 0) || ((int)MP(igbMod) != 0)) {
        Vfb = h->vfbzb;
        V3 = Vfb - Vgs_eff + Vbseff - DELTA_3;
        if (Vfb <= 0.0) T0 = sqrt(V3 * V3 - 4.0 * DELTA_3 * Vfb);
        else T0 = sqrt(V3 * V3 + 4.0 * DELTA_3 * Vfb);
        T1 = 0.5 * (1.0 + V3 / T0);
        Vfbeff = Vfb - 0.5 * (V3 + T0);
        dVfbeff_dVg = T1 * dVgs_eff_dVg;
        dVfbeff_dVb = -T1;
        Voxacc = Vfb - Vfbeff;
        dVoxacc_dVg = -dVfbeff_dVg;
        dVoxacc_dVb = -dVfbeff_dVb;
        if (Voxacc < 0.0) Voxacc = dVoxacc_dVg = dVoxacc_dVb = 0.0;
        T0 = 0.5 * h->k1ox;
        T3 = Vgs_eff - Vfbeff - Vbseff - Vgsteff;
        if (h->k1ox == 0.0)
            Voxdepinv = dVoxdepinv_dVg = dVoxdepinv_dVd = dVoxdepinv_dVb = 0.0;
        else if (T3 < 0.0) {
            Voxdepinv = -T3;
            dVoxdepinv_dVg = -dVgs_eff_dVg + dVfbeff_dVg + dVgsteff_dVg;
            dVoxdepinv_dVd = dVgsteff_dVd;
            dVoxdepinv_dVb = dVfbeff_dVb + 1.0 + dVgsteff_dVb;
        } else {
            T1 = sqrt(T0 * T0 + T3);
            T2 = T0 / T1;
            Voxdepinv = h->k1ox * (T1 - T0);
            dVoxdepinv_dVg = T2 * (dVgs_eff_dVg - dVfbeff_dVg - dVgsteff_dVg);
            dVoxdepinv_dVd = -T2 * dVgsteff_dVd;
            dVoxdepinv_dVb = -T2 * (dVfbeff_dVb + 1.0 + dVgsteff_dVb);
        }
        Voxdepinv += Vgsteff;
        dVoxdepinv_dVg += dVgsteff_dVg;
        dVoxdepinv_dVd += dVgsteff_dVd;
        dVoxdepinv_dVb += dVgsteff_dVb;
    }
    tmp = ((int)MP(tempMod) < 2) ? Vtm : Vtm0;
    if ((int)MP(igcMod)) {
        T0 = tmp * BP(nigc);
        if ((int)MP(igcMod) == 1) {
            VxNVt = (Vgs_eff - type * h->vth0) / T0;
            if (VxNVt > EXP_THRESHOLD) {
                Vaux = Vgs_eff - type * h->vth0;
                dVaux_dVg = dVgs_eff_dVg; dVaux_dVd = 0.0; dVaux_dVb = 0.0;
            }
        } else {
            VxNVt = (Vgs_eff - h->von) / T0;
            if (VxNVt > EXP_THRESHOLD) {
                Vaux = Vgs_eff - h->von;
                dVaux_dVg = dVgs_eff_dVg; dVaux_dVd = -dVth_dVd; dVaux_dVb = -dVth_dVb;
            }
        }
        if (VxNVt < -EXP_THRESHOLD) {
            Vaux = T0 * log(1.0 + MIN_EXP);
            dVaux_dVg = dVaux_dVd = dVaux_dVb = 0.0;
        } else if ((VxNVt >= -EXP_THRESHOLD) && (VxNVt <= EXP_THRESHOLD)) {
            ExpVxNVt = exp(VxNVt);
            Vaux = T0 * log(1.0 + ExpVxNVt);
            dVaux_dVg = ExpVxNVt / (1.0 + ExpVxNVt);
            if ((int)MP(igcMod) == 1) { dVaux_dVd = 0.0; dVaux_dVb = 0.0; }
            else { dVaux_dVd = -dVaux_dVg * dVth_dVd; dVaux_dVb = -dVaux_dVg * dVth_dVb; }
            dVaux_dVg *= dVgs_eff_dVg;
        }
        T2 = Vgs * Vaux;
        dT2_dVg = Vaux + Vgs * dVaux_dVg;
        dT2_dVd = Vgs * dVaux_dVd;
        dT2_dVb = Vgs * dVaux_dVb;
        T11 = h->Aechvb;
        T12 = h->Bechvb;
        T3 = BP(aigc) * BP(cigc) - BP(bigc);
        T4 = BP(bigc) * BP(cigc);
        T5 = T12 * (BP(aigc) + T3 * Voxdepinv - T4 * Voxdepinv * Voxdepinv);
        if (T5 > EXP_THRESHOLD) { T6 = MAX_EXP; dT6_dVg = dT6_dVd = dT6_dVb = 0.0; }
        else if (T5 < -EXP_THRESHOLD) { T6 = MIN_EXP; dT6_dVg = dT6_dVd = dT6_dVb = 0.0; }
        else {
            T6 = exp(T5);
            dT6_dVg = T6 * T12 * (T3 - 2.0 * T4 * Voxdepinv);
            dT6_dVd = dT6_dVg * dVoxdepinv_dVd;
            dT6_dVb = dT6_dVg * dVoxdepinv_dVb;
            dT6_dVg *= dVoxdepinv_dVg;
        }
        Igc = T11 * T2 * T6;
        dIgc_dVg = T11 * (T2 * dT6_dVg + T6 * dT2_dVg);
        dIgc_dVd = T11 * (T2 * dT6_dVd + T6 * dT2_dVd);
        dIgc_dVb = T11 * (T2 * dT6_dVb + T6 * dT2_dVb);
        if (MG(pigcd)) {
            Pigcd = BP(pigcd);
            dPigcd_dVg = dPigcd_dVd = dPigcd_dVb = 0.0;
        } else {
            T11 = -h->Bechvb;
            T12 = Vgsteff + 1.0e-20;
            T13 = T11 / T12 / T12;
            T14 = -T13 / T12;
            Pigcd = T13 * (1.0 - 0.5 * Vdseff / T12);
            dPigcd_dVg = T14 * (2.0 + 0.5 * (dVdseff_dVg - 3.0 * Vdseff / T12));
            dPigcd_dVd = 0.5 * T14 * dVdseff_dVd;
            dPigcd_dVb = 0.5 * T14 * dVdseff_dVb;
        }
        T7 = -Pigcd * Vdseff;
        dT7_dVg = -Vdseff * dPigcd_dVg - Pigcd * dVdseff_dVg;
        dT7_dVd = -Vdseff * dPigcd_dVd - Pigcd * dVdseff_dVd + dT7_dVg * dVgsteff_dVd;
        dT7_dVb = -Vdseff * dPigcd_dVb - Pigcd * dVdseff_dVb + dT7_dVg * dVgsteff_dVb;
        dT7_dVg *= dVgsteff_dVg;
        T8 = T7 * T7 + 2.0e-4;
        dT8_dVg = 2.0 * T7;
        dT8_dVd = dT8_dVg * dT7_dVd;
        dT8_dVb = dT8_dVg * dT7_dVb;
        dT8_dVg *= dT7_dVg;
        if (T7 > EXP_THRESHOLD) { T9 = MAX_EXP; dT9_dVg = dT9_dVd = dT9_dVb = 0.0; }
        else if (T7 < -EXP_THRESHOLD) { T9 = MIN_EXP; dT9_dVg = dT9_dVd = dT9_dVb = 0.0; }
        else { T9 = exp(T7); dT9_dVg = T9 * dT7_dVg; dT9_dVd = T9 * dT7_dVd; dT9_dVb = T9 * dT7_dVb; }
        T0 = T8 * T8;
        T1 = T9 - 1.0 + 1.0e-4;
        T10 = (T1 - T7) / T8;
        dT10_dVg = (dT9_dVg - dT7_dVg - T10 * dT8_dVg) / T8;
        dT10_dVd = (dT9_dVd - dT7_dVd - T10 * dT8_dVd) / T8;
        dT10_dVb = (dT9_dVb - dT7_dVb - T10 * dT8_dVb) / T8;
        Igcs = Igc * T10;
        dIgcs_dVg = dIgc_dVg * T10 + Igc * dT10_dVg;
        dIgcs_dVd = dIgc_dVd * T10 + Igc * dT10_dVd;
        dIgcs_dVb = dIgc_dVb * T10 + Igc * dT10_dVb;
        T1 = T9 - 1.0 - 1.0e-4;
        T10 = (T7 * T9 - T1) / T8;
        dT10_dVg = (dT7_dVg * T9 + (T7 - 1.0) * dT9_dVg - T10 * dT8_dVg) / T8;
        dT10_dVd = (dT7_dVd * T9 + (T7 - 1.0) * dT9_dVd - T10 * dT8_dVd) / T8;
        dT10_dVb = (dT7_dVb * T9 + (T7 - 1.0) * dT9_dVb - T10 * dT8_dVb) / T8;
        Igcd = Igc * T10;
        dIgcd_dVg = dIgc_dVg * T10 + Igc * dT10_dVg;
        dIgcd_dVd = dIgc_dVd * T10 + Igc * dT10_dVd;
        dIgcd_dVb = dIgc_dVb * T10 + Igc * dT10_dVb;
        h->Igcs = Igcs; h->gIgcsg = dIgcs_dVg; h->gIgcsd = dIgcs_dVd; h->gIgcsb = dIgcs_dVb * dVbseff_dVb;
        h->Igcd = Igcd; h->gIgcdg = dIgcd_dVg; h->gIgcdd = dIgcd_dVd; h->gIgcdb = dIgcd_dVb * dVbseff_dVb;

        T0 = vgs - (h->vfbsd + BP(vfbsdoff));
        vgs_eff = sqrt(T0 * T0 + 1.0e-4);
        dvgs_eff_dvg = T0 / vgs_eff;
        T2 = vgs * vgs_eff;
        dT2_dVg = vgs * dvgs_eff_dvg + vgs_eff;
        T11 = h->AechvbEdgeS;
        T12 = h->BechvbEdge;
        T3 = BP(aigs) * BP(cigs) - BP(bigs);
        T4 = BP(bigs) * BP(cigs);
        T5 = T12 * (BP(aigs) + T3 * vgs_eff - T4 * vgs_eff * vgs_eff);
        if (T5 > EXP_THRESHOLD) { T6 = MAX_EXP; dT6_dVg = 0.0; }
        else if (T5 < -EXP_THRESHOLD) { T6 = MIN_EXP; dT6_dVg = 0.0; }
        else { T6 = exp(T5); dT6_dVg = T6 * T12 * (T3 - 2.0 * T4 * vgs_eff) * dvgs_eff_dvg; }
        Igs = T11 * T2 * T6;
        dIgs_dVg = T11 * (T2 * dT6_dVg + T6 * dT2_dVg);
        dIgs_dVs = -dIgs_dVg;
        T0 = vgd - (h->vfbsd + BP(vfbsdoff));
        vgd_eff = sqrt(T0 * T0 + 1.0e-4);
        dvgd_eff_dvg = T0 / vgd_eff;
        T2 = vgd * vgd_eff;
        dT2_dVg = vgd * dvgd_eff_dvg + vgd_eff;
        T11 = h->AechvbEdgeD;
        T3 = BP(aigd) * BP(cigd) - BP(bigd);
        T4 = BP(bigd) * BP(cigd);
        T5 = T12 * (BP(aigd) + T3 * vgd_eff - T4 * vgd_eff * vgd_eff);
        if (T5 > EXP_THRESHOLD) { T6 = MAX_EXP; dT6_dVg = 0.0; }
        else if (T5 < -EXP_THRESHOLD) { T6 = MIN_EXP; dT6_dVg = 0.0; }
        else { T6 = exp(T5); dT6_dVg = T6 * T12 * (T3 - 2.0 * T4 * vgd_eff) * dvgd_eff_dvg; }
        Igd = T11 * T2 * T6;
        dIgd_dVg = T11 * (T2 * dT6_dVg + T6 * dT2_dVg);
        dIgd_dVd = -dIgd_dVg;
        h->Igs = Igs; h->gIgsg = dIgs_dVg; h->gIgss = dIgs_dVs;
        h->Igd = Igd; h->gIgdg = dIgd_dVg; h->gIgdd = dIgd_dVd;
    } else {
        h->Igcs = h->gIgcsg = h->gIgcsd = h->gIgcsb = 0.0;
        h->Igcd = h->gIgcdg = h->gIgcdd = h->gIgcdb = 0.0;
        h->Igs = h->gIgsg = h->gIgss = 0.0;
        h->Igd = h->gIgdg = h->gIgdd = 0.0;
    }
    if ((int)MP(igbMod)) {
        T0 = tmp * BP(nigbacc);
        T1 = -Vgs_eff + Vbseff + Vfb;
        VxNVt = T1 / T0;
        if (VxNVt > EXP_THRESHOLD) { Vaux = T1; dVaux_dVg = -dVgs_eff_dVg; dVaux_dVb = 1.0; }
        else if (VxNVt < -EXP_THRESHOLD) { Vaux = T0 * log(1.0 + MIN_EXP); dVaux_dVg = dVaux_dVb = 0.0; }
        else {
            ExpVxNVt = exp(VxNVt);
            Vaux = T0 * log(1.0 + ExpVxNVt);
            dVaux_dVb = ExpVxNVt / (1.0 + ExpVxNVt);
            dVaux_dVg = -dVaux_dVb * dVgs_eff_dVg;
        }
        T2 = (Vgs - Vbs) * Vaux;
        dT2_dVg = Vaux + (Vgs - Vbs) * dVaux_dVg;
        dT2_dVb = -Vaux + (Vgs - Vbs) * dVaux_dVb;
        T11 = 4.97232e-7 * h->weff * h->leff * h->ToxRatio;
        T12 = -7.45669e11 * toxe;
        T3 = BP(aigbacc) * BP(cigbacc) - BP(bigbacc);
        T4 = BP(bigbacc) * BP(cigbacc);
        T5 = T12 * (BP(aigbacc) + T3 * Voxacc - T4 * Voxacc * Voxacc);
        if (T5 > EXP_THRESHOLD) { T6 = MAX_EXP; dT6_dVg = dT6_dVb = 0.0; }
        else if (T5 < -EXP_THRESHOLD) { T6 = MIN_EXP; dT6_dVg = dT6_dVb = 0.0; }
        else {
            T6 = exp(T5);
            dT6_dVg = T6 * T12 * (T3 - 2.0 * T4 * Voxacc);
            dT6_dVb = dT6_dVg * dVoxacc_dVb;
            dT6_dVg *= dVoxacc_dVg;
        }
        Igbacc = T11 * T2 * T6;
        dIgbacc_dVg = T11 * (T2 * dT6_dVg + T6 * dT2_dVg);
        dIgbacc_dVb = T11 * (T2 * dT6_dVb + T6 * dT2_dVb);
        T0 = tmp * BP(nigbinv);
        T1 = Voxdepinv - BP(eigbinv);
        VxNVt = T1 / T0;
        if (VxNVt > EXP_THRESHOLD) {
            Vaux = T1; dVaux_dVg = dVoxdepinv_dVg; dVaux_dVd = dVoxdepinv_dVd; dVaux_dVb = dVoxdepinv_dVb;
        } else if (VxNVt < -EXP_THRESHOLD) {
            Vaux = T0 * log(1.0 + MIN_EXP); dVaux_dVg = dVaux_dVd = dVaux_dVb = 0.0;
        } else {
            ExpVxNVt = exp(VxNVt);
            Vaux = T0 * log(1.0 + ExpVxNVt);
            dVaux_dVg = ExpVxNVt / (1.0 + ExpVxNVt);
            dVaux_dVd = dVaux_dVg * dVoxdepinv_dVd;
            dVaux_dVb = dVaux_dVg * dVoxdepinv_dVb;
            dVaux_dVg *= dVoxdepinv_dVg;
        }
        T2 = (Vgs - Vbs) * Vaux;
        dT2_dVg = Vaux + (Vgs - Vbs) * dVaux_dVg;
        dT2_dVd = (Vgs - Vbs) * dVaux_dVd;
        dT2_dVb = -Vaux + (Vgs - Vbs) * dVaux_dVb;
        T11 *= 0.75610;
        T12 *= 1.31724;
        T3 = BP(aigbinv) * BP(cigbinv) - BP(bigbinv);
        T4 = BP(bigbinv) * BP(cigbinv);
        T5 = T12 * (BP(aigbinv) + T3 * Voxdepinv - T4 * Voxdepinv * Voxdepinv);
        if (T5 > EXP_THRESHOLD) { T6 = MAX_EXP; dT6_dVg = dT6_dVd = dT6_dVb = 0.0; }
        else if (T5 < -EXP_THRESHOLD) { T6 = MIN_EXP; dT6_dVg = dT6_dVd = dT6_dVb = 0.0; }
        else {
            T6 = exp(T5);
            dT6_dVg = T6 * T12 * (T3 - 2.0 * T4 * Voxdepinv);
            dT6_dVd = dT6_dVg * dVoxdepinv_dVd;
            dT6_dVb = dT6_dVg * dVoxdepinv_dVb;
            dT6_dVg *= dVoxdepinv_dVg;
        }
        Igbinv = T11 * T2 * T6;
        dIgbinv_dVg = T11 * (T2 * dT6_dVg + T6 * dT2_dVg);
        dIgbinv_dVd = T11 * (T2 * dT6_dVd + T6 * dT2_dVd);
        dIgbinv_dVb = T11 * (T2 * dT6_dVb + T6 * dT2_dVb);
        h->Igb = Igbinv + Igbacc;
        h->gIgbg = dIgbinv_dVg + dIgbacc_dVg;
        h->gIgbd = dIgbinv_dVd;
        h->gIgbb = (dIgbinv_dVb + dIgbacc_dVb) * dVbseff_dVb;
    } else {
        h->Igb = h->gIgbg = h->gIgbd = h->gIgbs = h->gIgbb = 0.0;
    }

    if (nf != 1.0) {   /* b4ld.c:2929-2978 */
        cdrain *= nf;
        h->gds *= nf; h->gm *= nf; h->gmbs *= nf; h->IdovVds *= nf;
        h->gbbs *= nf; h->gbgs *= nf; h->gbds *= nf; h->csub *= nf;
        h->Igidl *= nf; h->ggidld *= nf; h->ggidlg *= nf; h->ggidlb *= nf;
        h->Igisl *= nf; h->ggisls *= nf; h->ggislg *= nf; h->ggislb *= nf;
        h->Igcs *= nf; h->gIgcsg *= nf; h->gIgcsd *= nf; h->gIgcsb *= nf;
        h->Igcd *= nf; h->gIgcdg *= nf; h->gIgcdd *= nf; h->gIgcdb *= nf;
        h->Igs *= nf; h->gIgsg *= nf; h->gIgss *= nf;
        h->Igd *= nf; h->gIgdg *= nf; h->gIgdd *= nf;
        h->Igb *= nf; h->gIgbg *= nf; h->gIgbd *= nf; h->gIgbb *= nf;
    }
    h->ggidls = -(h->ggidld + h->ggidlg + h->ggidlb);
    h->ggisld = -(h->ggisls + h->ggislg + h->ggislb);
    h->gIgbs = -(h->gIgbg + h->gIgbd + h->gIgbb);
    h->gIgcss = -(h->gIgcsg + h->gIgcsd + h->gIgcsb);
    h->gIgcds = -(h->gIgcdg + h->gIgcdd + h->gIgcdb);
    h->cd = cdrain;

    if ((int)MP(tnoiMod) == 0) {   /* b4ld.c:2982-3004 (noise-only quantity qinv) */
        double Ab = Abulk0_Q * h->abulkCVfactor, Vds_ = Vgsteff / Ab, Vde;
        T0 = Vds_ - Vds - DELTA_4;
        T1 = sqrt(T0 * T0 + 4.0 * DELTA_4 * Vds_);
        if (T0 >= 0.0) Vde = Vds_ - 0.5 * (T0 + T1);
        else { T3 = (DELTA_4 + DELTA_4) / (T1 - T0); T4 = 1.0 - T3; Vde = Vds_ * T4; }
        if (Vds == 0.0) Vde = 0.0;
        T0 = Ab * Vde;
        T1 = 12.0 * (Vgsteff - 0.5 * T0 + 1.0e-20);
        T2 = Vde / T1;
        T3 = T0 * T2;
        h->qinv = Coxeff * h->weffCV * nf * h->leffCV * (Vgsteff - 0.5 * T0 + Ab * T3);
    }

    /* ---- C-V, capMod 2 (b4ld.c:3012-3914) */
    if ((MP(xpart) < 0) || (!ChargeComputationNeeded)) {
        qgate = qdrn = qsrc = qbulk = 0.0;
        h->cggb = h->cgsb = h->cgdb = 0.0;
        h->cdgb = h->cdsb = h->cddb = 0.0;
        h->cbgb = h->cbsb = h->cbdb = 0.0;
        h->csgb = h->cssb = h->csdb = 0.0;
        h->cgbb = h->csbb = h->cdbb = h->cbbb = 0.0;
        h->gtau = 0.0;
    } else {
        if (Vbseff < 0.0) { VbseffCV = Vbseff; dVbseffCV_dVb = 1.0; }
        else { VbseffCV = h->phi - Phis; dVbseffCV_dVb = -dPhis_dVb; }
        CoxWL = m->coxe * h->weffCV * h->leffCV * nf;
        if ((int)MP(cvchargeMod) == 0) {
            noff = n * BP(noff);
            dnoff_dVd = BP(noff) * dn_dVd;
            dnoff_dVb = BP(noff) * dn_dVb;
            T0 = Vtm * noff;
            voffcv = BP(voffcv);
            VgstNVt = (Vgst - voffcv) / T0;
            if (VgstNVt > EXP_THRESHOLD) {
                Vgsteff = Vgst - voffcv;
                dVgsteff_dVg = dVgs_eff_dVg;
                dVgsteff_dVd = -dVth_dVd;
                dVgsteff_dVb = -dVth_dVb;
            } else if (VgstNVt < -EXP_THRESHOLD) {
                Vgsteff = T0 * log(1.0 + MIN_EXP);
                dVgsteff_dVg = 0.0;
                dVgsteff_dVd = Vgsteff / noff;
                dVgsteff_dVb = dVgsteff_dVd * dnoff_dVb;
                dVgsteff_dVd *= dnoff_dVd;
            } else {
                ExpVgst = exp(VgstNVt);
                Vgsteff = T0 * log(1.0 + ExpVgst);
                dVgsteff_dVg = ExpVgst / (1.0 + ExpVgst);
                dVgsteff_dVd = -dVgsteff_dVg * (dVth_dVd + (Vgst - voffcv) / noff * dnoff_dVd)
                             + Vgsteff / noff * dnoff_dVd;
                dVgsteff_dVb = -dVgsteff_dVg * (dVth_dVb + (Vgst - voffcv) / noff * dnoff_dVb)
                             + Vgsteff / noff * dnoff_dVb;
                dVgsteff_dVg *= dVgs_eff_dVg;
            }
        } else {
            T0 = n * Vtm;
            T1 = h->mstarcv * Vgst;
            T2 = T1 / T0;
            if (T2 > EXP_THRESHOLD) {
                T10 = T1;
                dT10_dVg = h->mstarcv * dVgs_eff_dVg;
                dT10_dVd = -dVth_dVd * h->mstarcv;
                dT10_dVb = -dVth_dVb * h->mstarcv;
            } else if (T2 < -EXP_THRESHOLD) {
                T10 = Vtm * log(1.0 + MIN_EXP);
                dT10_dVg = 0.0;
                dT10_dVd = T10 * dn_dVd;
                dT10_dVb = T10 * dn_dVb;
                T10 *= n;
            } else {
                ExpVgst = exp(T2);
                T3 = Vtm * log(1.0 + ExpVgst);
                T10 = n * T3;
                dT10_dVg = h->mstarcv * ExpVgst / (1.0 + ExpVgst);
                dT10_dVb = T3 * dn_dVb - dT10_dVg * (dVth_dVb + Vgst * dn_dVb / n);
                dT10_dVd = T3 * dn_dVd - dT10_dVg * (dVth_dVd + Vgst * dn_dVd / n);
                dT10_dVg *= dVgs_eff_dVg;
            }
            T1 = h->voffcbncv - (1.0 - h->mstarcv) * Vgst;
            T2 = T1 / T0;
            if (T2 < -EXP_THRESHOLD) {
                T3 = m->coxe * MIN_EXP / h->cdep0;
                T9 = h->mstarcv + T3 * n;
                dT9_dVg = 0.0; dT9_dVd = dn_dVd * T3; dT9_dVb = dn_dVb * T3;
            } else if (T2 > EXP_THRESHOLD) {
                T3 = m->coxe * MAX_EXP / h->cdep0;
                T9 = h->mstarcv + T3 * n;
                dT9_dVg = 0.0; dT9_dVd = dn_dVd * T3; dT9_dVb = dn_dVb * T3;
            } else {
                ExpVgst = exp(T2);
                T3 = m->coxe / h->cdep0;
                T4 = T3 * ExpVgst;
                T5 = T1 * T4 / T0;
                T9 = h->mstarcv + n * T4;
                dT9_dVg = T3 * (h->mstarcv - 1.0) * ExpVgst / Vtm;
                dT9_dVb = T4 * dn_dVb - dT9_dVg * dVth_dVb - T5 * dn_dVb;
                dT9_dVd = T4 * dn_dVd - dT9_dVg * dVth_dVd - T5 * dn_dVd;
                dT9_dVg *= dVgs_eff_dVg;
            }
            Vgsteff = T10 / T9;
            T11 = T9 * T9;
            dVgsteff_dVg = (T9 * dT10_dVg - T10 * dT9_dVg) / T11;
            dVgsteff_dVd = (T9 * dT10_dVd - T10 * dT9_dVd) / T11;
            dVgsteff_dVb = (T9 * dT10_dVb - T10 * dT9_dVb) / T11;
        }

        V3 = h->vfbzb - Vgs_eff + VbseffCV - DELTA_3;
        if (h->vfbzb <= 0.0) T0 = sqrt(V3 * V3 - 4.0 * DELTA_3 * h->vfbzb);
        else T0 = sqrt(V3 * V3 + 4.0 * DELTA_3 * h->vfbzb);
        T1 = 0.5 * (1.0 + V3 / T0);
        Vfbeff = h->vfbzb - 0.5 * (V3 + T0);
        dVfbeff_dVg = T1 * dVgs_eff_dVg;
        dVfbeff_dVb = -T1 * dVbseffCV_dVb;
        Cox = h->coxp;
        Tox = 1.0e8 * h->toxp;
        T0 = (Vgs_eff - VbseffCV - h->vfbzb) / Tox;
        dT0_dVg = dVgs_eff_dVg / Tox;
        dT0_dVb = -dVbseffCV_dVb / Tox;
        tmp = T0 * BP(acde);
        if ((-EXP_THRESHOLD < tmp) && (tmp < EXP_THRESHOLD)) {
            Tcen = h->ldeb * exp(tmp);
            dTcen_dVg = BP(acde) * Tcen;
            dTcen_dVb = dTcen_dVg * dT0_dVb;
            dTcen_dVg *= dT0_dVg;
        } else if (tmp <= -EXP_THRESHOLD) {
            Tcen = h->ldeb * MIN_EXP;
            dTcen_dVg = dTcen_dVb = 0.0;
        } else {
            Tcen = h->ldeb * MAX_EXP;
            dTcen_dVg = dTcen_dVb = 0.0;
        }
        LINK = 1.0e-3 * h->toxp;
        V3 = h->ldeb - Tcen - LINK;
        V4 = sqrt(V3 * V3 + 4.0 * LINK * h->ldeb);
        Tcen = h->ldeb - 0.5 * (V3 + V4);
        T1 = 0.5 * (1.0 + V3 / V4);
        dTcen_dVg *= T1;
        dTcen_dVb *= T1;
        Ccen = epssub / Tcen;
        T2 = Cox / (Cox + Ccen);
        Coxeff = T2 * Ccen;
        T3 = -Ccen / Tcen;
        dCoxeff_dVg = T2 * T2 * T3;
        dCoxeff_dVb = dCoxeff_dVg * dTcen_dVb;
        dCoxeff_dVg *= dTcen_dVg;
        CoxWLcen = CoxWL * Coxeff / m->coxe;
        Qac0 = CoxWLcen * (Vfbeff - h->vfbzb);
        QovCox = Qac0 / Coxeff;
        dQac0_dVg = CoxWLcen * dVfbeff_dVg + QovCox * dCoxeff_dVg;
        dQac0_dVb = CoxWLcen * dVfbeff_dVb + QovCox * dCoxeff_dVb;
        T0 = 0.5 * h->k1ox;
        T3 = Vgs_eff - Vfbeff - VbseffCV - Vgsteff;
        if (h->k1ox == 0.0) { T1 = 0.0; T2 = 0.0; }
        else if (T3 < 0.0) { T1 = T0 + T3 / h->k1ox; T2 = CoxWLcen; }
        else { T1 = sqrt(T0 * T0 + T3); T2 = CoxWLcen * T0 / T1; }
        Qsub0 = CoxWLcen * h->k1ox * (T1 - T0);
        QovCox = Qsub0 / Coxeff;
        dQsub0_dVg = T2 * (dVgs_eff_dVg - dVfbeff_dVg - dVgsteff_dVg) + QovCox * dCoxeff_dVg;
        dQsub0_dVd = -T2 * dVgsteff_dVd;
        dQsub0_dVb = -T2 * (dVfbeff_dVb + dVbseffCV_dVb + dVgsteff_dVb) + QovCox * dCoxeff_dVb;
        if (h->k1ox <= 0.0) {
            Denomi = 0.25 * BP(moin) * Vtm;
            T0 = 0.5 * h->sqrtPhi;
        } else {
            Denomi = BP(moin) * Vtm * h->k1ox * h->k1ox;
            T0 = h->k1ox * h->sqrtPhi;
        }
        T1 = 2.0 * T0 + Vgsteff;
        DeltaPhi = Vtm * log(1.0 + T1 * Vgsteff / Denomi);
        dDeltaPhi_dVg = 2.0 * Vtm * (T1 - T0) / (Denomi + T1 * Vgsteff);
        T0 = Vgsteff - DeltaPhi - 0.001;
        dT0_dVg = 1.0 - dDeltaPhi_dVg;
        T1 = sqrt(T0 * T0 + Vgsteff * 0.004);
        VgDP = 0.5 * (T0 + T1);
        dVgDP_dVg = 0.5 * (dT0_dVg + (T0 * dT0_dVg + 0.002) / T1);
        Tox += Tox;
        T0 = (Vgsteff + h->vtfbphi2) / Tox;
        tmp = exp(MP(bdos) * 0.7 * log(T0));
        T1 = 1.0 + tmp;
        T2 = MP(bdos) * 0.7 * tmp / (T0 * Tox);
        Tcen = MP(ados) * 1.9e-9 / T1;
        dTcen_dVg = -Tcen * T2 / T1;
        dTcen_dVd = dTcen_dVg * dVgsteff_dVd;
        dTcen_dVb = dTcen_dVg * dVgsteff_dVb;
        dTcen_dVg *= dVgsteff_dVg;
        Ccen = epssub / Tcen;
        T0 = Cox / (Cox + Ccen);
        Coxeff = T0 * Ccen;
        T1 = -Ccen / Tcen;
        dCoxeff_dVg = T0 * T0 * T1;
        dCoxeff_dVd = dCoxeff_dVg * dTcen_dVd;
        dCoxeff_dVb = dCoxeff_dVg * dTcen_dVb;
        dCoxeff_dVg *= dTcen_dVg;
        CoxWLcen = CoxWL * Coxeff / m->coxe;
        AbulkCV = Abulk0_Q * h->abulkCVfactor;
        dAbulkCV_dVb = h->abulkCVfactor * dAbulk0_Q_dVb;
        VdsatCV = VgDP / AbulkCV;
        T0 = VdsatCV - Vds - DELTA_4;
        dT0_dVg = dVgDP_dVg / AbulkCV;
        dT0_dVb = -VdsatCV * dAbulkCV_dVb / AbulkCV;
        T1 = sqrt(T0 * T0 + 4.0 * DELTA_4 * VdsatCV);
        dT1_dVg = (T0 + DELTA_4 + DELTA_4) / T1;
        dT1_dVd = -T0 / T1;
        dT1_dVb = dT1_dVg * dT0_dVb;
        dT1_dVg *= dT0_dVg;
        if (T0 >= 0.0) {
            VdseffCV = VdsatCV - 0.5 * (T0 + T1);
            dVdseffCV_dVg = 0.5 * (dT0_dVg - dT1_dVg);
            dVdseffCV_dVd = 0.5 * (1.0 - dT1_dVd);
            dVdseffCV_dVb = 0.5 * (dT0_dVb - dT1_dVb);
        } else {
            T3 = (DELTA_4 + DELTA_4) / (T1 - T0);
            T4 = 1.0 - T3;
            T5 = VdsatCV * T3 / (T1 - T0);
            VdseffCV = VdsatCV * T4;
            dVdseffCV_dVg = dT0_dVg * T4 + T5 * (dT1_dVg - dT0_dVg);
            dVdseffCV_dVd = T5 * (dT1_dVd + 1.0);
            dVdseffCV_dVb = dT0_dVb * (T4 - T5) + T5 * dT1_dVb;
        }
        if (Vds == 0.0) { VdseffCV = 0.0; dVdseffCV_dVg = 0.0; dVdseffCV_dVb = 0.0; }
        T0 = AbulkCV * VdseffCV;
        T1 = VgDP;
        T2 = 12.0 * (T1 - 0.5 * T0 + 1.0e-20);
        T3 = T0 / T2;
        T4 = 1.0 - 12.0 * T3 * T3;
        T5 = AbulkCV * (6.0 * T0 * (4.0 * T1 - T0) / (T2 * T2) - 0.5);
        T6 = T5 * VdseffCV / AbulkCV;
        qgate = CoxWLcen * (T1 - T0 * (0.5 - T3));
        QovCox = qgate / Coxeff;
        Cgg1 = CoxWLcen * (T4 * dVgDP_dVg + T5 * dVdseffCV_dVg);
        Cgd1 = CoxWLcen * T5 * dVdseffCV_dVd + Cgg1 * dVgsteff_dVd + QovCox * dCoxeff_dVd;
        Cgb1 = CoxWLcen * (T5 * dVdseffCV_dVb + T6 * dAbulkCV_dVb) + Cgg1 * dVgsteff_dVb + QovCox * dCoxeff_dVb;
        Cgg1 = Cgg1 * dVgsteff_dVg + QovCox * dCoxeff_dVg;
        T7 = 1.0 - AbulkCV;
        T8 = T2 * T2;
        T9 = 12.0 * T7 * T0 * T0 / (T8 * AbulkCV);
        T10 = T9 * dVgDP_dVg;
        T11 = -T7 * T5 / AbulkCV;
        T12 = -(T9 * T1 / AbulkCV + VdseffCV * (0.5 - T0 / T2));
        qbulk = CoxWLcen * T7 * (0.5 * VdseffCV - T0 * VdseffCV / T2);
        QovCox = qbulk / Coxeff;
        Cbg1 = CoxWLcen * (T10 + T11 * dVdseffCV_dVg);
        Cbd1 = CoxWLcen * T11 * dVdseffCV_dVd + Cbg1 * dVgsteff_dVd + QovCox * dCoxeff_dVd;
        Cbb1 = CoxWLcen * (T11 * dVdseffCV_dVb + T12 * dAbulkCV_dVb) + Cbg1 * dVgsteff_dVb + QovCox * dCoxeff_dVb;
        Cbg1 = Cbg1 * dVgsteff_dVg + QovCox * dCoxeff_dVg;
        if (MP(xpart) > 0.5) {
            qsrc = -CoxWLcen * (T1 / 2.0 + T0 / 4.0 - 0.5 * T0 * T0 / T2);
            QovCox = qsrc / Coxeff;
            T2 += T2;
            T3 = T2 * T2;
            T7 = -(0.25 - 12.0 * T0 * (4.0 * T1 - T0) / T3);
            T4 = -(0.5 + 24.0 * T0 * T0 / T3) * dVgDP_dVg;
            T5 = T7 * AbulkCV;
            T6 = T7 * VdseffCV;
            Csg = CoxWLcen * (T4 + T5 * dVdseffCV_dVg);
            Csd = CoxWLcen * T5 * dVdseffCV_dVd + Csg * dVgsteff_dVd + QovCox * dCoxeff_dVd;
            Csb = CoxWLcen * (T5 * dVdseffCV_dVb + T6 * dAbulkCV_dVb) + Csg * dVgsteff_dVb + QovCox * dCoxeff_dVb;
            Csg = Csg * dVgsteff_dVg + QovCox * dCoxeff_dVg;
        } else if (MP(xpart) < 0.5) {
            T2 = T2 / 12.0;
            T3 = 0.5 * CoxWLcen / (T2 * T2);
            T4 = T1 * (2.0 * T0 * T0 / 3.0 + T1 * (T1 - 4.0 * T0 / 3.0)) - 2.0 * T0 * T0 * T0 / 15.0;
            qsrc = -T3 * T4;
            QovCox = qsrc / Coxeff;
            T8 = 4.0 / 3.0 * T1 * (T1 - T0) + 0.4 * T0 * T0;
            T5 = -2.0 * qsrc / T2 - T3 * (T1 * (3.0 * T1 - 8.0 * T0 / 3.0) + 2.0 * T0 * T0 / 3.0);
            T6 = AbulkCV * (qsrc / T2 + T3 * T8);
            T7 = T6 * VdseffCV / AbulkCV;
            Csg = T5 * dVgDP_dVg + T6 * dVdseffCV_dVg;
            Csd = Csg * dVgsteff_dVd + T6 * dVdseffCV_dVd + QovCox * dCoxeff_dVd;
            Csb = Csg * dVgsteff_dVb + T6 * dVdseffCV_dVb + T7 * dAbulkCV_dVb + QovCox * dCoxeff_dVb;
            Csg = Csg * dVgsteff_dVg + QovCox * dCoxeff_dVg;
        } else {
            qsrc = -0.5 * qgate;
            Csg = -0.5 * Cgg1;
            Csd = -0.5 * Cgd1;
            Csb = -0.5 * Cgb1;
        }
        qgate += Qac0 + Qsub0 - qbulk;
        qbulk -= (Qac0 + Qsub0);
        qdrn = -(qgate + qbulk + qsrc);
        Cbg = Cbg1 - dQac0_dVg - dQsub0_dVg;
        Cbd = Cbd1 - dQsub0_dVd;
        Cbb = Cbb1 - dQac0_dVb - dQsub0_dVb;
        Cgg = Cgg1 - Cbg;
        Cgd = Cgd1 - Cbd;
        Cgb = Cgb1 - Cbb;
        Cgb *= dVbseff_dVb;
        Cbb *= dVbseff_dVb;
        Csb *= dVbseff_dVb;
        h->cggb = Cgg;
        h->cgsb = -(Cgg + Cgd + Cgb);
        h->cgdb = Cgd;
        h->cdgb = -(Cgg + Cbg + Csg);
        h->cdsb = (Cgg + Cgd + Cgb + Cbg + Cbd + Cbb + Csg + Csd + Csb);
        h->cddb = -(Cgd + Cbd + Csd);
        h->cbgb = Cbg;
        h->cbsb = -(Cbg + Cbd + Cbb);
        h->cbdb = Cbd;
        h->csgb = -h->cggb - h->cdgb - h->cbgb;
        h->csdb = -h->cgdb - h->cddb - h->cbdb;
        h->cssb = -h->cgsb - h->cdsb - h->cbsb;
        h->cgbb = -h->cgdb - h->cggb - h->cgsb;
        h->cdbb = -h->cddb - h->cdgb - h->cdsb;
        h->cbbb = -h->cbgb - h->cbdb - h->cbsb;
        h->csbb = -h->cgbb - h->cdbb - h->cbbb;
        h->qgate = qgate; h->qbulk = qbulk; h->qdrn = qdrn;
        h->qsrc = -(qgate + qbulk + qdrn);
    }
    (void)Csd; (void)Cbd; (void)dbeta_dVg;

    /* ---- junction C-V (b4ld.c:3944-4063) */
    qbs = st0[S_qbs];
    qbd = st0[S_qbd];
    if (ChargeComputationNeeded) {
        czbd = m->DunitAreaTempJctCap * h->Adeff;
        czbs = m->SunitAreaTempJctCap * h->Aseff;
        czbdsw = m->DunitLengthSidewallTempJctCap * h->Pdeff;
        czbdswg = m->DunitLengthGateSidewallTempJctCap * h->weffCJ * nf;
        czbssw = m->SunitLengthSidewallTempJctCap * h->Pseff;
        czbsswg = m->SunitLengthGateSidewallTempJctCap * h->weffCJ * nf;
        junction_cv(vbs_jct, czbs, czbssw, czbsswg, MP(SbulkJctBotGradingCoeff),
                    MP(SbulkJctSideGradingCoeff), MP(SbulkJctGateSideGradingCoeff),
                    m->PhiBS, m->PhiBSWS, m->PhiBSWGS, &qbs, &h->capbs);
        junction_cv(vbd_jct, czbd, czbdsw, czbdswg, MP(DbulkJctBotGradingCoeff),
                    MP(DbulkJctSideGradingCoeff), MP(DbulkJctGateSideGradingCoeff),
                    m->PhiBD, m->PhiBSWD, m->PhiBSWGD, &qbd, &h->capbd);
        st0[S_qbs] = qbs;
        st0[S_qbd] = qbd;
    }

    /* ---- convergence check (b4ld.c:4066-4114, NEWCONV off) */
    if ((h->off == 0) || (!(ckt->mode & OR_MODEINITFIX))) {
        if (Check == 1) ckt->noncon++;
        else {
            double tol0, tol1, tol2, tol3, tol4, tol5, tol6;
            if (h->mode >= 0) Idtot = h->cd + h->csub + h->Igidl - h->cbd;
            else Idtot = h->cd + h->cbd - h->Igidl;
            tol0 = ckt->reltol * OR_MAX(fabs(cdhat), fabs(Idtot)) + ckt->abstol;
            tol1 = ckt->reltol * OR_MAX(fabs(cseshat), fabs(Isestot)) + ckt->abstol;
            tol2 = ckt->reltol * OR_MAX(fabs(cdedhat), fabs(Idedtot)) + ckt->abstol;
            tol3 = ckt->reltol * OR_MAX(fabs(cgshat), fabs(Igstot)) + ckt->abstol;
            tol4 = ckt->reltol * OR_MAX(fabs(cgdhat), fabs(Igdtot)) + ckt->abstol;
            tol5 = ckt->reltol * OR_MAX(fabs(cgbhat), fabs(Igbtot)) + ckt->abstol;
            if ((fabs(cdhat - Idtot) >= tol0) || (fabs(cseshat - Isestot) >= tol1)
                || (fabs(cdedhat - Idedtot) >= tol2)) ckt->noncon++;
            else if ((fabs(cgshat - Igstot) >= tol3) || (fabs(cgdhat - Igdtot) >= tol4)
                     || (fabs(cgbhat - Igbtot) >= tol5)) ckt->noncon++;
            else {
                Ibtot = h->cbs + h->cbd - h->Igidl - h->Igisl - h->csub;
                tol6 = ckt->reltol * OR_MAX(fabs(cbhat), fabs(Ibtot)) + ckt->abstol;
                if (fabs(cbhat - Ibtot) > tol6) ckt->noncon++;
            }
        }
    }
    st0[S_vds] = vds; st0[S_vgs] = vgs; st0[S_vbs] = vbs; st0[S_vbd] = vbd;
    st0[S_vges] = vges; st0[S_vgms] = vgms; st0[S_vdbs] = vdbs; st0[S_vdbd] = vdbd;
    st0[S_vsbs] = vsbs; st0[S_vses] = vses; st0[S_vdes] = vdes; st0[S_qdef] = qdef;

    mult_i = h->ip[BI_mult_i] * h->ip[BI_m];
    mult_q = h->ip[BI_mult_q] * h->ip[BI_m];

    if (!ChargeComputationNeeded) {
        /* line850 (b4ld.c:4650-4676) */
        ceqqg = ceqqb = ceqqd = 0.0;
        ceqqjd = ceqqjs = 0.0;
        gcdgb = gcddb = gcdsb = gcdbb = 0.0;
        gcsgb = gcsdb = gcssb = gcsbb = 0.0;
        gcggb = gcgdb = gcgsb = gcgbb = 0.0;
        gcbdb = gcbgb = gcbsb = gcbbb = 0.0;
        gcdbdb = gcsbsb = 0.0;
        dxpart = (h->mode > 0) ? 0.4 : 0.6;
        sxpart = 1.0 - dxpart;
        h->gtau = 0.0;
    } else {
        /* ---- overlap charges (b4ld.c:4132-4181), rgateMod 0/1 */
        T0 = vgd + DELTA_1;
        T1 = sqrt(T0 * T0 + 4.0 * DELTA_1);
        T2 = 0.5 * (T0 - T1);
        T3 = h->weffCV * BP(cgdl);
        T4 = sqrt(1.0 - 4.0 * T2 / BP(ckappad));
        cgdo = h->cgdoP + T3 - T3 * (1.0 - 1.0 / T4) * (0.5 - 0.5 * T0 / T1);
        qgdo = (h->cgdoP + T3) * vgd - T3 * (T2 + 0.5 * BP(ckappad) * (T4 - 1.0));
        T0 = vgs + DELTA_1;
        T1 = sqrt(T0 * T0 + 4.0 * DELTA_1);
        T2 = 0.5 * (T0 - T1);
        T3 = h->weffCV * BP(cgsl);
        T4 = sqrt(1.0 - 4.0 * T2 / BP(ckappas));
        cgso = h->cgsoP + T3 - T3 * (1.0 - 1.0 / T4) * (0.5 - 0.5 * T0 / T1);
        qgso = (h->cgsoP + T3) * vgs - T3 * (T2 + 0.5 * BP(ckappas) * (T4 - 1.0));
        if (nf != 1.0) { cgdo *= nf; cgso *= nf; qgdo *= nf; qgso *= nf; }
        h->cgdo_b = cgdo; h->qgdo = qgdo; h->cgso_b = cgso; h->qgso = qgso;

        /* ---- capacitive Jacobian (b4ld.c:4184-4567), trnqsMod 0 */
        ag0 = ckt->ag[0];
        if (h->mode > 0) {
            qdrn -= qgdo;
            gcggb = (h->cggb + cgdo + cgso + h->cgboP) * ag0;
            gcgdb = (h->cgdb - cgdo) * ag0;
            gcgsb = (h->cgsb - cgso) * ag0;
            gcgbb = -(gcggb + gcgdb + gcgsb);
            gcdgb = (h->cdgb - cgdo) * ag0;
            gcsgb = -(h->cggb + h->cbgb + h->cdgb + cgso) * ag0;
            gcbgb = (h->cbgb - h->cgboP) * ag0;
            qgb = h->cgboP * vgb;
            qgate += qgdo + qgso + qgb;
            qbulk -= qgb;
            qsrc = -(qgate + qbulk + qdrn);
            gcddb = (h->cddb + h->capbd + cgdo) * ag0;
            gcdsb = h->cdsb * ag0;
            gcsdb = -(h->cgdb + h->cbdb + h->cddb) * ag0;
            gcssb = (h->capbs + cgso - (h->cgsb + h->cbsb + h->cdsb)) * ag0;
            if (!h->rbodyMod) {
                gcdbb = -(gcdgb + gcddb + gcdsb);
                gcsbb = -(gcsgb + gcsdb + gcssb);
                gcbdb = (h->cbdb - h->capbd) * ag0;
                gcbsb = (h->cbsb - h->capbs) * ag0;
                gcdbdb = 0.0; gcsbsb = 0.0;
            } else {
                gcdbb = -(h->cddb + h->cdgb + h->cdsb) * ag0;
                gcsbb = -(gcsgb + gcsdb + gcssb) + h->capbs * ag0;
                gcbdb = h->cbdb * ag0;
                gcbsb = h->cbsb * ag0;
                gcdbdb = -h->capbd * ag0;
                gcsbsb = -h->capbs * ag0;
            }
            gcbbb = -(gcbdb + gcbgb + gcbsb);
            sxpart = 0.6;
            dxpart = 0.4;
        } else {
            qsrc = qdrn - qgso;
            gcggb = (h->cggb + cgdo + cgso + h->cgboP) * ag0;
            gcgdb = (h->cgsb - cgdo) * ag0;
            gcgsb = (h->cgdb - cgso) * ag0;
            gcgbb = -(gcggb + gcgdb + gcgsb);
            gcdgb = -(h->cggb + h->cbgb + h->cdgb + cgdo) * ag0;
            gcsgb = (h->cdgb - cgso) * ag0;
            gcbgb = (h->cbgb - h->cgboP) * ag0;
            qgb = h->cgboP * vgb;
            qgate += qgdo + qgso + qgb;
            qbulk -= qgb;
            qdrn = -(qgate + qbulk + qsrc);
            gcddb = (h->capbd + cgdo - (h->cgsb + h->cbsb + h->cdsb)) * ag0;
            gcdsb = -(h->cgdb + h->cbdb + h->cddb) * ag0;
            gcsdb = h->cdsb * ag0;
            gcssb = (h->cddb + h->capbs + cgso) * ag0;
            if (!h->rbodyMod) {
                gcdbb = -(gcdgb + gcddb + gcdsb);
                gcsbb = -(gcsgb + gcsdb + gcssb);
                gcbdb = (h->cbsb - h->capbd) * ag0;
                gcbsb = (h->cbdb - h->capbs) * ag0;
                gcdbdb = 0.0; gcsbsb = 0.0;
            } else {
                gcdbb = -(gcdgb + gcddb + gcdsb) + h->capbd * ag0;
                gcsbb = -(h->cddb + h->cdgb + h->cdsb) * ag0;
                gcbdb = h->cbsb * ag0;
                gcbsb = h->cbdb * ag0;
                gcdbdb = -h->capbd * ag0;
                gcsbsb = -h->capbs * ag0;
            }
            gcbbb = -(gcbgb + gcbdb + gcbsb);
            sxpart = 0.4;
            dxpart = 0.6;
        }

        /* ---- charge states and integration (b4ld.c:4581-4648) */
        st0[S_qg] = qgate;
        st0[S_qd] = qdrn - st0[S_qbd];
        st0[S_qs] = qsrc - st0[S_qbs];
        if (!h->rbodyMod) st0[S_qb] = qbulk + st0[S_qbd] + st0[S_qbs];
        else st0[S_qb] = qbulk;

        if (!(ckt->mode & OR_MODEINITSMSIG) && !(ckt->mode & OR_MODEDCTRANCURVE)) {
            if (ckt->mode & OR_MODEINITTRAN) {
                st1[S_qb] = st0[S_qb];
                st1[S_qg] = st0[S_qg];
                st1[S_qd] = st0[S_qd];
                if (h->rbodyMod) { st1[S_qbs] = st0[S_qbs]; st1[S_qbd] = st0[S_qbd]; }
            }
            or_integrate(ckt, &geq, &ceq, 0.0, st + S_qb);
            or_integrate(ckt, &geq, &ceq, 0.0, st + S_qg);
            or_integrate(ckt, &geq, &ceq, 0.0, st + S_qd);
            if (h->rbodyMod) {
                or_integrate(ckt, &geq, &ceq, 0.0, st + S_qbs);
                or_integrate(ckt, &geq, &ceq, 0.0, st + S_qbd);
            }
            /* line860 (b4ld.c:4689-4743) */
            cqgate = st0[S_cqg];
            cqbody = st0[S_cqb];
            cqdrn = st0[S_cqd];
            ceqqg = cqgate - gcggb * vgb + gcgdb * vbd + gcgsb * vbs;
            ceqqd = cqdrn - gcdgb * vgb + (gcddb + gcdbdb) * vbd - gcdbdb * vbd_jct + gcdsb * vbs;
            ceqqb = cqbody - gcbgb * vgb + gcbdb * vbd + gcbsb * vbs;
            if (h->rbodyMod) {
                ceqqjs = st0[S_cqbs] + gcsbsb * vbs_jct;
                ceqqjd = st0[S_cqbd] + gcdbdb * vbd_jct;
            }
            if (ckt->mode & OR_MODEINITTRAN) {
                st1[S_cqb] = st0[S_cqb];
                st1[S_cqg] = st0[S_cqg];
                st1[S_cqd] = st0[S_cqd];
                if (h->rbodyMod) { st1[S_cqbs] = st0[S_cqbs]; st1[S_cqbd] = st0[S_cqbd]; }
            }
        } else {
            /* small-signal / DC transfer: capacitances kept, no integration (line850) */
            ceqqg = ceqqb = ceqqd = 0.0;
            ceqqjd = ceqqjs = 0.0;
            gcdgb = gcddb = gcdsb = gcdbb = 0.0;
            gcsgb = gcsdb = gcssb = gcsbb = 0.0;
            gcggb = gcgdb = gcgsb = gcgbb = 0.0;
            gcbdb = gcbgb = gcbsb = gcbbb = 0.0;
            gcdbdb = gcsbsb = 0.0;
        }
        h->gtau = 0.0;
    }

    /* ---- equivalent current sources (b4ld.c:4750-4982), rdsMod 0, rgateMod 0/1 */
    if (h->mode >= 0) {
        Gm = h->gm; Gmbs = h->gmbs;
        FwdSum = Gm + Gmbs; RevSum = 0.0;
        ceqdrn = type * (cdrain - h->gds * vds - Gm * vgs - Gmbs * vbs);
        ceqbd = type * (h->csub + h->Igidl - (h->gbds + h->ggidld) * vds
                        - (h->gbgs + h->ggidlg) * vgs - (h->gbbs + h->ggidlb) * vbs);
        ceqbs = type * (h->Igisl + h->ggisls * vds - h->ggislg * vgd - h->ggislb * vbd);
        gbbdp = -(h->gbds);
        gbbsp = h->gbds + h->gbgs + h->gbbs;
        gbdpg = h->gbgs; gbdpdp = h->gbds; gbdpb = h->gbbs;
        gbdpsp = -(gbdpg + gbdpdp + gbdpb);
        gbspg = 0.0; gbspdp = 0.0; gbspb = 0.0; gbspsp = 0.0;
        if ((int)MP(igcMod)) {
            gIstotg = h->gIgsg + h->gIgcsg;
            gIstotd = h->gIgcsd;
            gIstots = h->gIgss + h->gIgcss;
            gIstotb = h->gIgcsb;
            Istoteq = type * (h->Igs + h->Igcs - gIstotg * vgs - h->gIgcsd * vds - h->gIgcsb * vbs);
            gIdtotg = h->gIgdg + h->gIgcdg;
            gIdtotd = h->gIgdd + h->gIgcdd;
            gIdtots = h->gIgcds;
            gIdtotb = h->gIgcdb;
            Idtoteq = type * (h->Igd + h->Igcd - h->gIgdg * vgd - h->gIgcdg * vgs
                              - h->gIgcdd * vds - h->gIgcdb * vbs);
        } else {
            gIstotg = gIstotd = gIstots = gIstotb = Istoteq = 0.0;
            gIdtotg = gIdtotd = gIdtots = gIdtotb = Idtoteq = 0.0;
        }
        if ((int)MP(igbMod)) {
            gIbtotg = h->gIgbg; gIbtotd = h->gIgbd; gIbtots = h->gIgbs; gIbtotb = h->gIgbb;
            Ibtoteq = type * (h->Igb - h->gIgbg * vgs - h->gIgbd * vds - h->gIgbb * vbs);
        } else gIbtotg = gIbtotd = gIbtots = gIbtotb = Ibtoteq = 0.0;
    } else {
        Gm = -h->gm; Gmbs = -h->gmbs;
        FwdSum = 0.0; RevSum = -(Gm + Gmbs);
        ceqdrn = -type * (cdrain + h->gds * vds + Gm * vgd + Gmbs * vbd);
        ceqbs = type * (h->csub + h->Igisl + (h->gbds + h->ggisls) * vds
                        - (h->gbgs + h->ggislg) * vgd - (h->gbbs + h->ggislb) * vbd);
        ceqbd = type * (h->Igidl - h->ggidld * vds - h->ggidlg * vgs - h->ggidlb * vbs);
        gbbsp = -(h->gbds);
        gbbdp = h->gbds + h->gbgs + h->gbbs;
        gbdpg = 0.0; gbdpsp = 0.0; gbdpb = 0.0; gbdpdp = 0.0;
        gbspg = h->gbgs; gbspsp = h->gbds; gbspb = h->gbbs;
        gbspdp = -(gbspg + gbspsp + gbspb);
        if ((int)MP(igcMod)) {
            gIstotg = h->gIgsg + h->gIgcdg;
            gIstotd = h->gIgcds;
            gIstots = h->gIgss + h->gIgcdd;
            gIstotb = h->gIgcdb;
            Istoteq = type * (h->Igs + h->Igcd - h->gIgsg * vgs - h->gIgcdg * vgd
                              + h->gIgcdd * vds - h->gIgcdb * vbd);
            gIdtotg = h->gIgdg + h->gIgcsg;
            gIdtotd = h->gIgdd + h->gIgcss;
            gIdtots = h->gIgcsd;
            gIdtotb = h->gIgcsb;
            Idtoteq = type * (h->Igd + h->Igcs - (h->gIgdg + h->gIgcsg) * vgd
                              + h->gIgcsd * vds - h->gIgcsb * vbd);
        } else {
            gIstotg = gIstotd = gIstots = gIstotb = Istoteq = 0.0;
            gIdtotg = gIdtotd = gIdtots = gIdtotb = Idtoteq = 0.0;
        }
        if ((int)MP(igbMod)) {
            gIbtotg = h->gIgbg; gIbtotd = h->gIgbs; gIbtots = h->gIgbd; gIbtotb = h->gIgbb;
            Ibtoteq = type * (h->Igb - h->gIgbg * vgd + h->gIgbd * vds - h->gIgbb * vbd);
        } else gIbtotg = gIbtotd = gIbtots = gIbtotb = Ibtoteq = 0.0;
    }
    if (((int)MP(igcMod) != 0) || ((int)MP(igbMod) != 0)) {
        gIgtotg = gIstotg + gIdtotg + gIbtotg;
        gIgtotd = gIstotd + gIdtotd + gIbtotd;
        gIgtots = gIstots + gIdtots + gIbtots;
        gIgtotb = gIstotb + gIdtotb + gIbtotb;
        Igtoteq = Istoteq + Idtoteq + Ibtoteq;
    } else gIgtotg = gIgtotd = gIgtots = gIgtotb = Igtoteq = 0.0;

    if (type > 0) {
        ceqjs = (h->cbs - h->gbs * vbs_jct);
        ceqjd = (h->cbd - h->gbd * vbd_jct);
    } else {
        ceqjs = -(h->cbs - h->gbs * vbs_jct);
        ceqjd = -(h->cbd - h->gbd * vbd_jct);
        ceqqg = -ceqqg;
        ceqqd = -ceqqd;
        ceqqb = -ceqqb;
        if (h->rbodyMod) { ceqqjs = -ceqqjs; ceqqjd = -ceqqjd; }
    }

    /* ---- RHS (b4ld.c:4992-5022 in the OpenMP form, applied as 5446-5472) */
    R[N_dp] += (mult_i * (ceqjd - ceqbd - ceqdrn + Idtoteq) - mult_q * ceqqd);
    R[N_gp] -= mult_q * ceqqg - mult_i * (-Igtoteq);
    if (!h->rbodyMod) {
        R[N_bp] += (mult_i * (ceqbd + ceqbs - ceqjd - ceqjs + Ibtoteq) - mult_q * ceqqb);
        R[N_sp] += (mult_i * (ceqdrn - ceqbs + ceqjs + Istoteq) + mult_q * (ceqqg + ceqqb + ceqqd));
    } else {
        R[N_db] -= mult_i * (ceqjd) + mult_q * ceqqjd;
        R[N_bp] += (mult_i * (ceqbd + ceqbs + Ibtoteq) - mult_q * ceqqb);
        R[N_sb] -= mult_i * (ceqjs) + mult_q * ceqqjs;
        R[N_sp] += (mult_i * (ceqdrn - ceqbs + ceqjs + Istoteq)
                    + mult_q * (ceqqd + ceqqg + ceqqb + ceqqjd + ceqqjs));
    }

    /* ---- Jacobian (b4ld.c:5060-5235 in the OpenMP form, applied as 5474-5621) */
    if (!h->rbodyMod) { gjbd = h->gbd; gjbs = h->gbs; }
    else gjbd = gjbs = 0.0;
    gdpr = h->drainConductance;
    gspr = h->sourceConductance;
    geltd = h->grgeltd;
    if (h->rgateMod == 1) {
        S[P_GEge] += mult_i * geltd;
        S[P_GPge] -= mult_i * geltd;
        S[P_GEgp] -= mult_i * geltd;
        S[P_GPgp] += mult_q * gcggb + mult_i * (geltd + gIgtotg);
        S[P_GPdp] += mult_q * gcgdb + mult_i * gIgtotd;
        S[P_GPsp] += mult_q * gcgsb + mult_i * gIgtots;
        S[P_GPbp] += mult_q * gcgbb + mult_i * gIgtotb;
    } else {
        S[P_GPgp] += mult_q * gcggb + mult_i * gIgtotg;
        S[P_GPdp] += mult_q * gcgdb + mult_i * gIgtotd;
        S[P_GPsp] += mult_q * gcgsb + mult_i * gIgtots;
        S[P_GPbp] += mult_q * gcgbb + mult_i * gIgtotb;
    }
    S[P_DPdp] += mult_i * (gdpr + h->gds + h->gbd + RevSum + gbdpdp - gIdtotd) + mult_q * gcddb;
    S[P_DPd] -= mult_i * (gdpr);
    S[P_DPgp] += mult_i * (Gm + gbdpg - gIdtotg) + mult_q * gcdgb;
    S[P_DPsp] -= mult_i * (h->gds + gIdtots + FwdSum - gbdpsp) - mult_q * gcdsb;
    S[P_DPbp] -= mult_i * (gjbd - Gmbs - gbdpb + gIdtotb) - mult_q * gcdbb;
    S[P_Ddp] -= mult_i * (gdpr);
    S[P_Dd] += mult_i * (gdpr);
    S[P_SPdp] -= mult_i * (h->gds + RevSum - gbspdp + gIstotd) - mult_q * gcsdb;
    S[P_SPgp] += mult_q * gcsgb + mult_i * (gbspg - Gm - gIstotg);
    S[P_SPsp] += mult_i * (gspr + h->gds + h->gbs - gIstots + FwdSum + gbspsp) + mult_q * gcssb;
    S[P_SPs] -= mult_i * (gspr);
    S[P_SPbp] -= mult_i * (gjbs + Gmbs - gbspb + gIstotb) - mult_q * gcsbb;
    S[P_Ssp] -= mult_i * (gspr);
    S[P_Ss] += mult_i * (gspr);
    S[P_BPdp] += mult_q * gcbdb - mult_i * (gjbd - gbbdp + gIbtotd);
    S[P_BPgp] += mult_q * gcbgb - mult_i * (h->gbgs + gIbtotg);
    S[P_BPsp] += mult_q * gcbsb - mult_i * (gjbs - gbbsp + gIbtots);
    S[P_BPbp] += mult_i * (gjbd + gjbs - h->gbbs - gIbtotb) + mult_q * gcbbb;
    /* GIDL / GISL stamps (b4ld.c:5561-5578) */
    S[P_DPdp] += mult_i * h->ggidld;
    S[P_DPgp] += mult_i * h->ggidlg;
    S[P_DPsp] -= mult_i * (h->ggidlg + h->ggidld + h->ggidlb);
    S[P_DPbp] += mult_i * h->ggidlb;
    S[P_BPdp] -= mult_i * h->ggidld;
    S[P_BPgp] -= mult_i * h->ggidlg;
    S[P_BPsp] += mult_i * (h->ggidlg + h->ggidld + h->ggidlb);
    S[P_BPbp] -= mult_i * h->ggidlb;
    S[P_SPdp] -= mult_i * (h->ggisls + h->ggislg + h->ggislb);
    S[P_SPgp] += mult_i * h->ggislg;
    S[P_SPsp] += mult_i * h->ggisls;
    S[P_SPbp] += mult_i * h->ggislb;
    S[P_BPdp] += mult_i * (h->ggislg + h->ggisls + h->ggislb);
    S[P_BPgp] -= mult_i * h->ggislg;
    S[P_BPsp] -= mult_i * h->ggisls;
    S[P_BPbp] -= mult_i * h->ggislb;
    if (h->rbodyMod) {   /* b4ld.c:5581-5606 */
        S[P_DPdb] += mult_q * gcdbdb - mult_i * h->gbd;
        S[P_SPsb] -= mult_i * h->gbs - mult_q * gcsbsb;
        S[P_DBdp] += mult_q * gcdbdb - mult_i * h->gbd;
        S[P_DBdb] += mult_i * (h->gbd + h->grbpd + h->grbdb) - mult_q * gcdbdb;
        S[P_DBbp] -= mult_i * h->grbpd;
        S[P_DBb] -= mult_i * h->grbdb;
        S[P_BPdb] -= mult_i * h->grbpd;
        S[P_BPb] -= mult_i * h->grbpb;
        S[P_BPsb] -= mult_i * h->grbps;
        S[P_BPbp] += mult_i * (h->grbpd + h->grbps + h->grbpb);
        S[P_SBsp] += mult_q * gcsbsb - mult_i * h->gbs;
        S[P_SBbp] -= mult_i * h->grbps;
        S[P_SBb] -= mult_i * h->grbsb;
        S[P_SBsb] += mult_i * (h->gbs + h->grbps + h->grbsb) - mult_q * gcsbsb;
        S[P_Bdb] -= mult_i * h->grbdb;
        S[P_Bbp] -= mult_i * h->grbpb;
        S[P_Bsb] -= mult_i * h->grbsb;
        S[P_Bb] += mult_i * (h->grbsb + h->grbdb + h->grbpb);
    }
    (void)sxpart; (void)dxpart; (void)Cgb; (void)dIgcs_dVg;
    return 0;
#undef NV
}

/* CKTterr (cktterr.c:13-78) */
void or_terr(int qcap, const ORCkt *ckt, double *timeStep, double trtol, double chgtol)
{
    static const double gearCoeff[] = { .5, .2222222222, .1363636364, .096, .07299270073, .05830903790 };
    static const double trapCoeff[] = { .5, .08333333333 };
    double volttol, chargetol, tol, del, diff[8], deltmp[8], factor = 0;
    int i, j, ccap = qcap + 1;
    volttol = ckt->abstol + ckt->reltol * OR_MAX(fabs(ckt->states[0][ccap]), fabs(ckt->states[1][ccap]));
    chargetol = OR_MAX(fabs(ckt->states[0][qcap]), fabs(ckt->states[1][qcap]));
    chargetol = ckt->reltol * OR_MAX(chargetol, chgtol) / ckt->delta;
    tol = OR_MAX(volttol, chargetol);
    for (i = ckt->order + 1; i >= 0; i--) diff[i] = ckt->states[i][qcap];
    for (i = 0; i <= ckt->order; i++) deltmp[i] = ckt->deltaOld[i];
    j = ckt->order;
    for (;;) {
        for (i = 0; i <= j; i++) diff[i] = (diff[i] - diff[i + 1]) / deltmp[i];
        if (--j < 0) break;
        for (i = 0; i <= j; i++) deltmp[i] = deltmp[i + 1] + ckt->deltaOld[i];
    }
    if (ckt->method == OR_GEAR) factor = gearCoeff[ckt->order - 1];
    else factor = trapCoeff[ckt->order - 1];
    del = trtol * tol / OR_MAX(ckt->abstol, factor * fabs(diff[0]));
    if (ckt->order == 2) del = sqrt(del);
    else if (ckt->order == 3) del = cbrt(del);
    else if (ckt->order > 3) del = exp(log(del) / ckt->order);
    *timeStep = OR_MIN(*timeStep, del);
}

void or_bsim4_trunc(const ORInst *h, const ORCkt *ckt, double *timeStep, double trtol, double chgtol)
{   /* b4trunc.c:32-71 */
    or_terr(h->state + S_qb, ckt, timeStep, trtol, chgtol);
    or_terr(h->state + S_qg, ckt, timeStep, trtol, chgtol);
    or_terr(h->state + S_qd, ckt, timeStep, trtol, chgtol);
    if (h->rbodyMod) {
        or_terr(h->state + S_qbs, ckt, timeStep, trtol, chgtol);
        or_terr(h->state + S_qbd, ckt, timeStep, trtol, chgtol);
    }
}
