// FP32 instantiation of the K98 TVD + CTU + CT kernels (FMA contraction; fdiv/fsqrt use the
// hardware reciprocal and square root).  Face B, corner and centre EMFs and the CFL speed
// stay FP64.
#include "tvd_kernels_impl.hpp"

namespace imhd {
IMHD_INST_TVD(float)
}  // namespace imhd
