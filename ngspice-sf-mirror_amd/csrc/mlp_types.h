/*
 * Shapes, parameter layout and element-wise rules of the Newton initial-guess
 * network (kernels in mlp_kernels.hip, host side in mlp_host.hip).  Every
 * kernel that rounds to BF16 or applies the activation uses these helpers, so
 * the layer-by-layer path and the fused inference kernel agree bit for bit.
 */
#ifndef NGMI_MLP_TYPES_H
#define NGMI_MLP_TYPES_H

#include <cstdint>
#include <hip/hip_runtime.h>

namespace ngmi {
constexpr int MLP_IN = 16;       /* features per node */
constexpr int MLP_HID = 256;     /* hidden width, three hidden layers */
constexpr int MLP_IN_PAD = 32;   /* device input width: 16 features + 16 zero columns */

/* parameter blocks, in this order in one FP32 array (and its BF16 copy):
 * W1 [256][32], b1, W2 [256][256], b2, W3 [256][256], b3, w4 [256], b4 */
constexpr int64_t MLP_SZ_W1 = (int64_t)MLP_HID * MLP_IN_PAD, MLP_SZ_W2 = (int64_t)MLP_HID * MLP_HID;
constexpr int64_t MLP_OFF_W1 = 0, MLP_OFF_B1 = MLP_OFF_W1 + MLP_SZ_W1, MLP_OFF_W2 = MLP_OFF_B1 + MLP_HID,
                  MLP_OFF_B2 = MLP_OFF_W2 + MLP_SZ_W2, MLP_OFF_W3 = MLP_OFF_B2 + MLP_HID,
                  MLP_OFF_B3 = MLP_OFF_W3 + MLP_SZ_W2, MLP_OFF_W4 = MLP_OFF_B3 + MLP_HID,
                  MLP_OFF_B4 = MLP_OFF_W4 + MLP_HID, MLP_NPARAM = MLP_OFF_B4 + 1;

/* the head's dot product runs in blocks of 32 hidden units, each summed in
 * order, and the block sums are added to the bias in order */
constexpr int MLP_HEAD_BLOCK = 32;

/* FP32 -> BF16, round to nearest even (one v_cvt_pk_bf16_f32 on gfx950) */
__device__ __forceinline__ unsigned short f2bf(float f) { return __builtin_bit_cast(unsigned short, (__bf16)f); }

__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned int)h << 16); }

/* SiLU z * sigmoid(z), with the hardware reciprocal (no IEEE division) */
__device__ __forceinline__ float silu(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

/* one hidden unit: the accumulator starts at the unit's bias, so it ends
 * at the pre-activation z; output BF16 */
__device__ __forceinline__ unsigned short hidden_out(float z) { return f2bf(silu(z)); }

}  // namespace ngmi

#endif
