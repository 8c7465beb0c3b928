// FP64 instantiation of the VL2 + CT kernels: no FMA contraction, IEEE division and square
// root, so results reproduce the reference's FP64 arithmetic bit for bit.
#include "vl_kernels_impl.hpp"

namespace imhd {
IMHD_INST_VL(double)
}  // namespace imhd
