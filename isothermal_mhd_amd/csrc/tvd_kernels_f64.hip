// FP64 instantiation of the K98 TVD + CTU + CT kernels: no FMA contraction, IEEE division
// and square root, so results reproduce the reference's FP64 arithmetic bit for bit.
#include "tvd_kernels_impl.hpp"

namespace imhd {
IMHD_INST_TVD(double)
}  // namespace imhd
