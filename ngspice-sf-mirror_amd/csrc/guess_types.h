/*
 * The learned Newton guess's output rule, shared by the engine's apply
 * kernel (tran_kernels.hip) and the fused network kernel (mlp_kernels.hip).
 *
 * A feature row carries (lin, act): the node's linear extrapolation and its
 * recent motion.  The node starts its Newton run at lin + act * clamp(y, +-4)
 * for the network output y, so the network predicts the extrapolation's
 * error in units of the node's own motion.
 */
#ifndef NGMI_GUESS_TYPES_H
#define NGMI_GUESS_TYPES_H

#include <hip/hip_runtime.h>

namespace ngmi {

constexpr double GUESS_CLAMP = 4.0;   /* |network output| limit, in units of activity */

/* an explicit FMA: left to the compiler, the contraction of b.x + b.y * g
 * differs between call sites, and so would the last bit */
__device__ __forceinline__ double guess_value(double2 b, float y)
{
    const double g = fmin(fmax((double)y, -GUESS_CLAMP), GUESS_CLAMP);
    return fma(b.y, g, b.x);
}

}  // namespace ngmi

#endif
