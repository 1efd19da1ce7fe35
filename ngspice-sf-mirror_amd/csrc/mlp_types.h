/*
 * Shapes of the Newton initial-guess network (see mlp_kernels.hip).
 */
#ifndef NGMI_MLP_TYPES_H
#define NGMI_MLP_TYPES_H

namespace ngmi {
constexpr int MLP_IN = 16;       /* features per node */
constexpr int MLP_HID = 256;     /* hidden width, three hidden layers */
}

#endif
