/*
 * ORACLE (test infrastructure only).  Scalar C restatement of ngspice 46's
 * BSIM4 4.8.3 setup / temperature / load path, used by tests/ and by the
 * CPU baseline of bench.py to check the HIP engine.  Nothing in the engine
 * links or includes this code.
 *
 * Follows:
 *   src/spicelib/devices/bsim4/b4set.c   (defaults, internal nodes, TSTALLOC order)
 *   src/spicelib/devices/bsim4/b4temp.c  (model/size/instance temperature params)
 *   src/spicelib/devices/bsim4/b4geo.c   (S/D geometry)
 *   src/spicelib/devices/bsim4/b4ld.c    (I-V, C-V, limiting, convergence, stamps)
 *   src/spicelib/devices/bsim4/b4check.c (the parameter clamps that change values)
 *   src/spicelib/devices/devsup.c        (DEVfetlim / DEVlimvds / DEVpnjlim)
 *   src/maths/ni/niinteg.c               (NIintegrate)
 *
 * Supported selectors (others are rejected by or_bsim4_setup with an error):
 *   capMod 2, rdsMod 0, trnqsMod 0, acnqsMod 0, mtrlMod 0, rgateMod 0/1,
 *   rbodyMod 0/1, mobMod 0..6, dioMod 0..2, igcMod 0..2, igbMod 0/1,
 *   gidlMod 0/1, tempMod 0..3, cvchargeMod 0/1, version >= 4.8.1.
 */
#ifndef OR_BSIM4_REF_H
#define OR_BSIM4_REF_H

#include "bsim4_tables.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ckt->mode bits, same values as src/include/ngspice/cktdefs.h:171-199 */
#define OR_MODETRAN        0x1
#define OR_MODEAC          0x2
#define OR_MODEDCOP        0x10
#define OR_MODETRANOP      0x20
#define OR_MODEDCTRANCURVE 0x40
#define OR_MODEINITFLOAT   0x100
#define OR_MODEINITJCT     0x200
#define OR_MODEINITFIX     0x400
#define OR_MODEINITSMSIG   0x800
#define OR_MODEINITTRAN    0x1000
#define OR_MODEINITPRED    0x2000
#define OR_MODEUIC         0x10000

#define OR_TRAPEZOIDAL 1
#define OR_GEAR 2

enum {
#define X(f) BM_##f,
    BSIM4_MODEL_FIELDS(X)
#undef X
    BM__COUNT
};

enum {
#define X(f) BB_##f,
    BSIM4_BINNED(X)
#undef X
    BB__COUNT
};

/* instance parameter fields (unique) */
enum {
    BI_l, BI_w, BI_m, BI_nf, BI_sa, BI_sb, BI_sd, BI_sca, BI_scb, BI_scc, BI_sc,
    BI_min, BI_drainArea, BI_sourceArea, BI_drainPerimeter, BI_sourcePerimeter,
    BI_drainSquares, BI_sourceSquares, BI_off, BI_rbdb, BI_rbsb, BI_rbpb, BI_rbps,
    BI_rbpd, BI_delvto, BI_mulu0, BI_xgw, BI_ngcon, BI_wnflag, BI_trnqsMod,
    BI_acnqsMod, BI_rbodyMod, BI_rgateMod, BI_geoMod, BI_rgeoMod, BI_mult_i,
    BI_mult_q, BI_mult_fn, BI_icVGS, BI_icVDS, BI_icVBS, BI_dtemp,
    BI__COUNT
};

/* matrix pointer slots, in b4set.c:2587-2676 TSTALLOC order */
enum {
    P_DPbp, P_GPbp, P_SPbp, P_BPdp, P_BPgp, P_BPsp, P_BPbp, P_Dd, P_GPgp, P_Ss,
    P_DPdp, P_SPsp, P_Ddp, P_GPdp, P_GPsp, P_Ssp, P_DPsp, P_DPd, P_DPgp, P_SPgp,
    P_SPs, P_SPdp, P_Qq, P_Qbp, P_Qdp, P_Qsp, P_Qgp, P_DPq, P_SPq, P_GPq,
    /* rgateMod != 0 */
    P_GEge, P_GEgp, P_GPge, P_GEdp, P_GEsp, P_GEbp, P_GMdp, P_GMgp, P_GMgm,
    P_GMge, P_GMsp, P_GMbp, P_DPgm, P_GPgm, P_GEgm, P_SPgm, P_BPgm,
    /* rbodyMod 1/2 */
    P_DPdb, P_SPsb, P_DBdp, P_DBdb, P_DBbp, P_DBb, P_BPdb, P_BPb, P_BPsb,
    P_SBsp, P_SBbp, P_SBb, P_SBsb, P_Bdb, P_Bbp, P_Bsb, P_Bb,
    /* rdsMod */
    P_Dgp, P_Dsp, P_Dbp, P_Sdp, P_Sgp, P_Sbp,
    P__COUNT
};

/* node roles of one instance */
enum { N_d, N_g, N_s, N_b, N_dp, N_gp, N_gm, N_sp, N_db, N_bp, N_sb, N_q, N__COUNT };

/* state offsets, bsim4def.h:534-567 */
enum {
    S_vbd, S_vbs, S_vgs, S_vds, S_vdbs, S_vdbd, S_vsbs, S_vges, S_vgms, S_vses,
    S_vdes, S_qb, S_cqb, S_qg, S_cqg, S_qd, S_cqd, S_qgmid, S_cqgmid, S_qbs,
    S_cqbs, S_qbd, S_cqbd, S_qcheq, S_cqcheq, S_qcdump, S_cqcdump, S_qdef, S_qs,
    S__COUNT
};

typedef struct {
    double p[BM__COUNT];
    unsigned char given[BM__COUNT];
    /* model-level derived quantities (b4temp.c:90-417) */
    double coxe, coxp, vcrit, factor1, vtm0, vtm, Eg0, ni;
    double SjctTempSatCurDensity, SjctSidewallTempSatCurDensity,
           SjctGateSidewallTempSatCurDensity, DjctTempSatCurDensity,
           DjctSidewallTempSatCurDensity, DjctGateSidewallTempSatCurDensity;
    double SunitAreaTempJctCap, DunitAreaTempJctCap,
           SunitLengthSidewallTempJctCap, DunitLengthSidewallTempJctCap,
           SunitLengthGateSidewallTempJctCap, DunitLengthGateSidewallTempJctCap;
    double PhiBS, PhiBD, PhiBSWS, PhiBSWD, PhiBSWGS, PhiBSWGD;
    double njtsstemp, njtsswstemp, njtsswgstemp, njtsdtemp, njtsswdtemp, njtsswgdtemp;
    int type;          /* +1 NMOS, -1 PMOS */
    int checked;       /* b4check.c checks the first instance of a model only */
} ORModel;

typedef struct {
    double ip[BI__COUNT];
    unsigned char igiven[BI__COUNT];
    int node[N__COUNT];
    int state;                         /* offset into the state vectors */
    const ORModel *model;
    double bin[BB__COUNT];             /* size-dependent parameters */
    /* size-dependent derived parameters (pParam, b4temp.c:436-1692) */
    double leff, weff, leffCV, weffCV, weffCJ, dl, dw, dlc, dwc, dwj;
    double abulkCVfactor, vsattempP, rds0, rdswminP, rd0, rdwminP, rs0, rswminP;
    double u0tempP, tfactor, cgdoP, cgsoP, cgboP, phi, sqrtPhi, phis3, Xdep0,
           sqrtXdep0, litl, vbi, vfbsd, cdep0, ToxRatio, ToxRatioEdge, Aechvb,
           Bechvb, AechvbEdgeS, AechvbEdgeD, BechvbEdge, mstar, mstarcv,
           voffcbn, voffcbncv, ldeb, k1ox, theta0vb0, thetaRout, vfbzbfactor,
           ku0, kvth0, ku0temp, inv_od_ref, rho_ref, VgsteffVth, dvtp2factor, lc;
    /* instance derived (b4temp.c:1692-2400) */
    double u0temp, vsattemp, vth0, vfb, eta0, k2, vtfbphi1, vtfbphi2, vbsc,
           k2ox, vfbzb, cgso, cgdo, grbdb, grbpb, grbps, grbsb, grbpd,
           grgeltd, Pseff, Pdeff, Aseff, Adeff, sourceConductance,
           drainConductance, XExpBVS, XExpBVD, vjsmFwd, vjsmRev, vjdmFwd,
           vjdmRev, IVjsmFwd, IVjsmRev, IVjdmFwd, IVjdmRev, SslpFwd, SslpRev,
           DslpFwd, DslpRev, SjctTempRevSatCur, DjctTempRevSatCur,
           SswTempRevSatCur, DswTempRevSatCur, SswgTempRevSatCur,
           DswgTempRevSatCur, toxp, coxp, rbdb, rbsb, rbpb, rbps, rbpd;
    int rgateMod, rbodyMod, trnqsMod, acnqsMod, geoMod, rgeoMod, off, min;
    /* values cached by load for the next iteration (bypass/convergence) */
    double cd, csub, cbd, cbs, gbd, gbs, gm, gds, gmbs, gbbs, gbgs, gbds;
    double Igidl, Igisl, ggidld, ggidlg, ggidlb, ggidls, ggisls, ggislg, ggislb, ggisld;
    double Igs, Igd, Igcs, Igcd, Igb, gIgsg, gIgss, gIgdg, gIgdd, gIgcsg, gIgcsd,
           gIgcsb, gIgcss, gIgcdg, gIgcdd, gIgcdb, gIgcds, gIgbg, gIgbd, gIgbb, gIgbs;
    double gstot, gstotd, gstotg, gstots, gstotb, gdtot, gdtotd, gdtotg, gdtots, gdtotb;
    double von, vdsat, Vgsteff, Vdseff, ueff, thetavth, nstar, Abulk, EsatL,
           AbovVgst2Vtm, IdovVds, Coxeff, vgs_eff, vgd_eff, dvgs_eff_dvg,
           dvgd_eff_dvg, qinv, grdsw;
    double cggb, cgsb, cgdb, cdgb, cdsb, cddb, cbgb, cbsb, cbdb, csgb, cssb,
           csdb, cgbb, csbb, cdbb, cbbb, capbs, capbd;
    double qgate, qbulk, qdrn, qsrc, cgdo_b, qgdo, cgso_b, qgso;
    double gcrg, gcrgg, gcrgd, gcrgs, gcrgb, gtau;
    int mode;
    /* per-load outputs: matrix contributions per pointer slot, RHS per node */
    double stamp[P__COUNT];
    double rhs[N__COUNT];
} ORInst;

typedef struct {
    int mode;
    int order, method;
    double ag[7];
    double delta, deltaOld[7];
    double gmin, reltol, abstol, voltTol, temp;
    double *rhsOld;
    double *states[8];
    int noncon;
    int bypass;
} ORCkt;

/* netlist-style parameter setting: returns 0 or -1 for an unknown name */
int or_bsim4_model_init(ORModel *m);
int or_bsim4_model_set(ORModel *m, const char *name, double value);
int or_bsim4_inst_init(ORInst *h, const ORModel *m);
int or_bsim4_inst_set(ORInst *h, const char *name, double value);
/* defaults + selector checks (b4set.c); returns 0 or a negative error */
int or_bsim4_model_defaults(ORModel *m, double nomTemp);
int or_bsim4_inst_defaults(ORInst *h);
/* does this instance need internal nodes: bit0 dp, bit1 sp, bit2 gp, bit3 gm, bit4 body */
int or_bsim4_internal_nodes(const ORInst *h);
/* temperature / size dependent parameters (b4temp.c) */
int or_bsim4_model_temp(ORModel *m, double temp);
int or_bsim4_inst_temp(ORInst *h, double temp);
/* one instance evaluation (b4ld.c); updates ckt->noncon, states, h->stamp/rhs */
int or_bsim4_load(ORInst *h, ORCkt *ckt);
/* LTE (b4trunc.c + cktterr.c) */
void or_bsim4_trunc(const ORInst *h, const ORCkt *ckt, double *timeStep,
                    double trtol, double chgtol);

/* limiting helpers restated from devsup.c */
double or_fetlim(double vnew, double vold, double vto);
double or_limvds(double vnew, double vold);
double or_pnjlim(double vnew, double vold, double vt, double vcrit, int *icheck);
int or_integrate(ORCkt *ckt, double *geq, double *ceq, double cap, int qcap);
void or_terr(int qcap, const ORCkt *ckt, double *timeStep, double trtol, double chgtol);

#ifdef __cplusplus
}
#endif
#endif
