/*
 * Host-side BSIM4 model cards and instances (see bsim4_setup.cpp).
 */
#ifndef NGMI_BSIM4_SETUP_H
#define NGMI_BSIM4_SETUP_H

#include <cstdint>
#include "bsim4_tables.h"
#include "bsim4_fields.h"

namespace ngmi {

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

enum {
    BI_l, BI_w, BI_m, BI_nf, BI_sa, BI_sb, BI_sd, BI_sca, BI_scb, BI_scc, BI_sc, BI_min,
    BI_ad, BI_as, BI_pd, BI_ps, BI_nrd, BI_nrs, BI_off, BI_delvto, BI_mulu0, BI_xgw,
    BI_ngcon, BI_geoMod, BI_rgeoMod, BI_mult_i, BI_mult_q, BI_icVGS, BI_icVDS, BI_icVBS,
    BI_rgateMod, BI_rbodyMod, BI_trnqsMod, BI_acnqsMod, BI__COUNT
};

/* one .model card, after defaults and at one temperature */
struct B4Model {
    double p[BM__COUNT] = {};
    uint8_t given[BM__COUNT] = {};
    int type = 1;
    double temp = 300.15;
    double coxe = 0, coxp = 0, vcrit = 0, factor1 = 0, vtm0 = 0, vtm = 0, Eg0 = 0, ni = 0;
    double SjSat = 0, SjSwSat = 0, SjSwgSat = 0, DjSat = 0, DjSwSat = 0, DjSwgSat = 0;
    double SAreaCap = 0, DAreaCap = 0, SSwCap = 0, DSwCap = 0, SSwgCap = 0, DSwgCap = 0;
    double PhiBS = 0, PhiBD = 0, PhiBSWS = 0, PhiBSWD = 0, PhiBSWGS = 0, PhiBSWGD = 0;
    double njtss = 0, njtssws = 0, njtsswgs = 0, njtsd = 0, njtsswd = 0, njtsswgd = 0;
    bool checked = false;

    int set(const char *name, double value);        /* 0, or -1 for an unknown name */
    int finish(double nomTemp, double temp);        /* defaults + temperature; <0 if unsupported */
    const char *check() const;                      /* nullptr when supported */
};

struct B4Inst {
    B4Model *model = nullptr;
    double ip[BI__COUNT] = {};
    uint8_t igiven[BI__COUNT] = {};

    int set(const char *name, double value);
    void defaults();
    const char *check() const;
    /* writes B4F__COUNT values (field order of bsim4_fields.h) */
    int resolve(double *row);
};

}  // namespace ngmi

#endif
