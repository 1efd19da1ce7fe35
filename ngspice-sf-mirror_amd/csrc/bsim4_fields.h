/*
 * Per-instance BSIM4 evaluation table: the values the device kernel reads for
 * one MOSFET, after model defaults, temperature and size dependence have been
 * resolved on the host.  Stored field-major ("struct of arrays"): field f of
 * instance i lives at table[f * stride + i], so the 64 lanes of a wavefront
 * read 64 consecutive doubles of each field.
 *
 * The names follow the BSIM4 4.8.3 quantities of the same name in
 * src/spicelib/devices/bsim4/b4temp.c / bsim4def.h.  Values that the
 * reference recomputes inside BSIM4load from instance constants only
 * (junction saturation currents, TAT emission voltages, zero-bias junction
 * capacitances, the gate-tunnelling prefactors) are precomputed here with the
 * same arithmetic.
 */
#ifndef NGMI_BSIM4_FIELDS_H
#define NGMI_BSIM4_FIELDS_H

#define NGMI_B4_FIELDS(X) \
  X(type) X(tvth0) X(icVDS) X(icVGS) X(icVBS) X(off) X(vcrit) X(vtm) X(vtm0) \
  X(coxe) X(factor1) X(toxe) X(TempRatio) \
  X(Nvtms) X(Nvtmd) X(IsatS) X(IsatD) X(xjbvs) X(bvs) X(xjbvd) X(bvd) \
  X(XExpBVS) X(XExpBVD) X(vjsmFwd) X(IVjsmFwd) X(vjsmRev) X(IVjsmRev) X(SslpFwd) X(SslpRev) \
  X(vjdmFwd) X(IVjdmFwd) X(vjdmRev) X(IVjdmRev) X(DslpFwd) X(DslpRev) \
  X(Nvtmrss) X(Nvtmrsd) X(Nvtmrssws) X(Nvtmrsswd) X(Nvtmrsswgs) X(Nvtmrsswgd) \
  X(vtss) X(vtsd) X(vtssws) X(vtsswd) X(vtsswgs) X(vtsswgd) \
  X(SjctTempRevSatCur) X(DjctTempRevSatCur) X(SswTempRevSatCur) X(DswTempRevSatCur) \
  X(SswgTempRevSatCur) X(DswgTempRevSatCur) \
  X(vbsc) X(phi) X(sqrtPhi) X(Xdep0) X(vbi) X(dvt2) X(dvt2w) X(dvt1) X(dvt0) X(dvt1w) \
  X(dvt0w) X(weff) X(leff) X(lpe0) X(lpeb) X(k1ox) X(k1) X(kt1) X(kt1l) X(kt2) X(w0) \
  X(eta0) X(etab) X(theta0vb0) X(k3) X(k3b) X(k2ox) X(cit) X(nfactor) X(cdsc) X(cdscb) \
  X(cdscd) X(dvtp0) X(dvtp1) X(dvtp4) X(dvtp2factor) \
  X(vfb) X(ngate) X(mstar) X(voffcbn) X(cdep0) \
  X(dwg) X(dwb) X(prwg) X(prwb) X(rds0) X(rdswmin) \
  X(xj) X(a0) X(ags) X(b0) X(b1) X(keta) X(ketac) \
  X(ua) X(ub) X(uc) X(ud) X(u0temp) X(vtfbphi1) X(eu) X(ucs) X(VgsteffVth) \
  X(vsattemp) X(a1) X(a2) X(delta) X(lambda) X(litl) \
  X(vtfbphi2) X(toxp) X(coxp) X(bdos) X(ados) \
  X(fprout) X(pvag) X(pclm) X(thetaRout) X(pdiblb) X(pditsd) X(pdits) X(pditsl) \
  X(pscbe1) X(pscbe2) X(alpha0) X(alpha1) X(beta0) X(vtl) X(tfactor) X(idovvdsc) \
  X(agidl) X(bgidl) X(cgidl) X(egidl) X(agisl) X(bgisl) X(cgisl) X(egisl) \
  X(rgidl) X(kgidl) X(fgidl) X(rgisl) X(kgisl) X(fgisl) X(gidlclamp) X(weffCJ) X(nf) \
  X(vfbzb) X(aigc) X(bigc) X(cigc) X(nigc) X(Aechvb) X(Bechvb) X(pigcd) \
  X(vfbsd) X(vfbsdoff) X(AechvbEdgeS) X(AechvbEdgeD) X(BechvbEdge) \
  X(aigs) X(bigs) X(cigs) X(aigd) X(bigd) X(cigd) \
  X(nigbacc) X(Tigb) X(aigbacc) X(bigbacc) X(cigbacc) X(nigbinv) X(eigbinv) \
  X(aigbinv) X(bigbinv) X(cigbinv) \
  X(abulkCVfactor) X(weffCV) X(leffCV) X(noff) X(voffcv) X(mstarcv) X(voffcbncv) \
  X(ldeb) X(acde) X(moin) X(xpart) \
  X(czbs) X(czbssw) X(czbsswg) X(czbd) X(czbdsw) X(czbdswg) \
  X(MJS) X(MJSWS) X(MJSWGS) X(MJD) X(MJSWD) X(MJSWGD) \
  X(PhiBS) X(PhiBSWS) X(PhiBSWGS) X(PhiBD) X(PhiBSWD) X(PhiBSWGD) \
  X(cgdl) X(cgsl) X(ckappad) X(ckappas) X(cgdoP) X(cgsoP) X(cgboP) \
  X(mult_i) X(mult_q) X(flags)

enum NgmiB4Field {
#define NGMI_X(f) B4F_##f,
    NGMI_B4_FIELDS(NGMI_X)
#undef NGMI_X
    B4F__COUNT
};

/* bits of the 'flags' field (selectors and given-flags the kernel branches on) */
#define B4FL_MOBMOD_MASK   0x7      /* mobMod 0..6 */
#define B4FL_DIOMOD_SHIFT  3        /* dioMod 0..2 (2 bits) */
#define B4FL_IGCMOD_SHIFT  5        /* igcMod 0..2 (2 bits) */
#define B4FL_IGBMOD        0x80
#define B4FL_GIDLMOD       0x100
#define B4FL_TEMPMOD_SHIFT 9        /* tempMod 0..3 (2 bits) */
#define B4FL_CVCHARGE      0x800
#define B4FL_PIGCD_GIVEN   0x1000
#define B4FL_LAMBDA        0x2000   /* lambda given and > 0 */
#define B4FL_VTL           0x4000   /* vtl given and > 0 */
/* the default card's selector word: dioMod 1, every other selector 0 */
#define B4FL_DEFAULT       (1 << B4FL_DIOMOD_SHIFT)

/* 16 Jacobian entries of one instance, row-major over (dp, gp, sp, bp) */
enum { B4S_DD, B4S_DG, B4S_DS, B4S_DB, B4S_GD, B4S_GG, B4S_GS, B4S_GB,
       B4S_SD, B4S_SG, B4S_SS, B4S_SB, B4S_BD, B4S_BG, B4S_BS, B4S_BB, B4S__COUNT };
/* 4 RHS entries (dp, gp, sp, bp) */
enum { B4R_D, B4R_G, B4R_S, B4R_B, B4R__COUNT };

/* state offsets of one instance (bsim4def.h:534-567) */
enum { B4X_vbd, B4X_vbs, B4X_vgs, B4X_vds, B4X_vdbs, B4X_vdbd, B4X_vsbs, B4X_vges,
       B4X_vgms, B4X_vses, B4X_vdes, B4X_qb, B4X_cqb, B4X_qg, B4X_cqg, B4X_qd, B4X_cqd,
       B4X_qgmid, B4X_cqgmid, B4X_qbs, B4X_cqbs, B4X_qbd, B4X_cqbd, B4X_qcheq,
       B4X_cqcheq, B4X_qcdump, B4X_cqcdump, B4X_qdef, B4X_qs, B4X__COUNT };

/* values cached across Newton iterations per instance (b4ld.c keeps them in
 * the instance struct for the next call's convergence check) */
#define NGMI_B4_CACHE(X) \
  X(cd) X(csub) X(cbd) X(cbs) X(gbd) X(gbs) X(gm) X(gds) X(gmbs) X(gbbs) X(gbgs) X(gbds) \
  X(Igidl) X(Igisl) X(ggidld) X(ggidlg) X(ggidlb) X(ggisls) X(ggislg) X(ggislb) \
  X(Igs) X(Igd) X(Igcs) X(Igcd) X(Igb) X(gIgsg) X(gIgdg) X(gIgcsg) X(gIgcsd) X(gIgcsb) \
  X(gIgcdg) X(gIgcdd) X(gIgcdb) X(gIgbg) X(gIgbd) X(gIgbb) X(von) X(mode)

enum NgmiB4Cache {
#define NGMI_X(f) B4C_##f,
    NGMI_B4_CACHE(NGMI_X)
#undef NGMI_X
    B4C__COUNT
};

#endif
