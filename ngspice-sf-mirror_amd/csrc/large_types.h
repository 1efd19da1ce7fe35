/*
 * Device view of one large circuit's transient (large_host.cpp /
 * large_kernels.hip).
 *
 * Unknowns 1..n (0 = ground, always 0): nodes 1..nnodes, then one branch
 * current per voltage source in source order.  Iterates are [n+1] device
 * vectors.  Per-element values the gathers read:
 *   gsrc: [ncap geq | 16 nm MOSFET stamps (row k of instance i at k nm + i)]
 *   rsrc: [ncap ceq | 4 nm MOSFET right-hand sides | nvs source values]
 * State levels (CKTstate0.. rotated by pointer on the host):
 *   capacitors [2][ncap] (charge, current) per level,
 *   MOSFETs    [29][nm] per level (bsim4def.h:534-567 order).
 */
#ifndef NGMI_LARGE_TYPES_H
#define NGMI_LARGE_TYPES_H

#include <cstdint>
#include "bsim4_fields.h"

namespace ngmi {

/* levels kept: trapezoidal or Gear up to order 2 reads CKTstate0..3 */
constexpr int LG_NLEV = 4;

struct LgPulse {
    int pulse;                 /* 0: DC */
    double dc;
    double v1, v2, td, tr, tf, pw, per;   /* vsrcload.c:96-125 defaults applied */
};

struct LgDev {
    int64_t n, nnodes, nm, ncap, nvs;
    const int *mos_nodes;      /* [4][nm] d g s b */
    double *vterm;             /* [4][nm] */
    const int *cap_p, *cap_n;
    const double *cap_c;
    const LgPulse *pulse;      /* [nvs] */
    double *gsrc, *rsrc;
    double reltol, abstol, vntol, chgtol, trtol;
};

/* the scalars of one Newton iteration or LTE evaluation */
struct LgStep {
    int mode, order, method;
    double ag[7];
    double time, delta, deltaOld[7], srcFact;
    double *cs[LG_NLEV];       /* capacitor states by level */
    double *ms[LG_NLEV];       /* MOSFET states by level */
};

}  // namespace ngmi

#endif
