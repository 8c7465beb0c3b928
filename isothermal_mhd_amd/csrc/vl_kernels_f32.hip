// FP32 instantiation of the VL2 + CT kernels (compiled with FMA contraction; fdiv/fsqrt use
// the hardware reciprocal and square root).  The corner-EMF sum, face B and the CFL speed
// stay FP64.
#include "vl_kernels_impl.hpp"

namespace imhd {
IMHD_INST_VL(float)
}  // namespace imhd
