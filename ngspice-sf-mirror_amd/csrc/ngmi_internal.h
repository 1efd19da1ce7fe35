/*
 * Internal helpers shared by the engine's translation units: error text,
 * HIP status checks, and the device-side view of a BSIM4 instance table.
 */
#ifndef NGMI_INTERNAL_H
#define NGMI_INTERNAL_H

#include <hip/hip_runtime.h>
#include <cstdint>
#include "../../include/ngspice_mi355x.h"
#include "bsim4_fields.h"

namespace ngmi {

/* thread-local last error */
void set_error(const char *fmt, ...);

/* launch-shape knobs (ngmi_tune); defaults are the measured best */
enum {
    TUNE_EVAL_WAVES = 0,
    TUNE_MOS_WAVES = 1,
    TUNE_STREAMS = 2,
    TUNE_SPECIALISE = 3,
    TUNE_ORDER_SLACK = 4,
    TUNE_GRAPH = 5,
    TUNE_GUESS_FUSED = 6,
    TUNE_GUESS_GRID = 7,
    TUNE__COUNT
};
int tune_get(int knob);
/* a BSIM4 kernel's occupancy variant for a launch of this many instances:
 * the knob's value, or for 0 (automatic) 2, or 1 when one wave per SIMD
 * holds every instance (its occupancy then buys nothing and the spill-free
 * 1-wave variant has the shorter latency) */
int waves_for(int knob, int64_t instances);

#define NGMI_HIP(call)                                                          \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess) {                                                 \
            ::ngmi::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                              __FILE__, __LINE__);                              \
            return NGMI_E_DEVICE;                                               \
        }                                                                       \
    } while (0)

/*
 * Device view of a BSIM4 evaluation table.  Each field is either a
 * per-instance column (itab[col * n + i]) or, when every instance of a
 * model card has the same value, a per-model column (mtab[m * mcols +
 * col], model-major so that one card's values share cache lines).
 * code[f] = col for a per-instance field and ~col for a per-model one; the
 * code array lives in device memory and is read with a scalar load next to
 * each use (a kernel-argument copy would pin ~190 SGPRs for the whole
 * kernel).
 */
struct B4Tab {
    const double *itab;
    const double *mtab;
    const int *model_of;
    int64_t n;       /* instances (row stride of itab) */
    int nm;          /* models */
    int mcols;       /* per-model columns (row stride of mtab) */
    const int *code;
    int flags;       /* the selector word when every instance shares it, else -1 */
};

}  // namespace ngmi

#endif
