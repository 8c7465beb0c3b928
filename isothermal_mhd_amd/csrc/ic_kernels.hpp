// Argument block and launch entry points of the seeded random-field start
// (imhd_init_turbulence).
#pragma once
#include "kernels.hpp"

namespace imhd {

constexpr int ICSEG = 16;   // z-planes per thread of k_ic_cells (velocity sums held in registers)

template <class real>
struct IcArgs {
    real* q[7];                 // state set written (interior cells)
    double* b[3];               // face field of that set
    real* qp[MAXPASS];          // passive scalars of that set
    const double* Bv;           // y-pass output [3][K+1][ny][nx] complex, null: v = 0
    const double* twz;          // [nz][K+1] (cos, sin) of 2 pi kz k / nz
    double rho0, sigma;
    double b0[3];
    unsigned seed_lo, seed_hi;
    int K, npass, passive, _pad;
};

void launch_ic_ypass(hipStream_t st, int nx, int ny, int K, const double* A, const double* twy, double* Bout);
template <class real> void launch_ic_cells(hipStream_t st, const BoxDesc& B, const IcArgs<real>& a);

}  // namespace imhd
