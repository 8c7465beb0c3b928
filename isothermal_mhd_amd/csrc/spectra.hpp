// Device diagnostics of the reference's analysis program (3D_MPI/Analysis/src/
// analysis_main.f90): shell-binned power spectra (compute_pk) and the density PDF
// (compute_density_pdf), formed in situ from the device state instead of from snapshot
// files.  The analysis program works on the REAL*4 snapshot; so do these kernels.
#pragma once
#include "kernels.hpp"

namespace imhd {

enum { SPEC_DENSITY = 0, SPEC_VELOCITY = 1, SPEC_BFIELD = 2, SPEC_VPAR = 3, SPEC_VPERP = 4 };

template <class real>
struct SpecPackArgs {
    const real* q[7];
    int which, comp;   // field kind, component 0..2
    float* out;        // n0 x n1 x n2 interior, x fastest
};

template <class real>
void launch_spec_pack(hipStream_t st, const BoxDesc& B, const SpecPackArgs<real>& a);
// |F|^2 of the octant 0 <= kx, ky, kz < n/2 of an R2C transform (layout [z][y][x/2+1]),
// added in REAL*4 to the running sum s (first: start from 0)
void launch_spec_accum(hipStream_t st, const float* F, int n, float* s, int first);
// bins of compute_pk: doubled sums (zero mode single), k = k0 sqrt(kx^2+ky^2+kz^2) in
// REAL*4, bin INT(k / dk); pk accumulates in FP64
void launch_spec_bin(hipStream_t st, const float* s, int n, float k0, float dk, double* pk);
// Mean cell-centred field of the interior, summed in FP64 into sum[3] (zeroed by the caller)
template <class real>
void launch_mean_b(hipStream_t st, const BoxDesc& B, const real* bx, const real* by, const real* bz, double* sum);
// mode_decomposition: per octant mode k != 0, the v_perp transforms F[3] projected onto the
// Alfven / fast / slow displacement vectors about B0 (b0sum / n^3); P holds 8 x n shells
void launch_mode_bin(hipStream_t st, const float* F0, const float* F1, const float* F2, int n, const double* b0sum,
                     float k0, float dk, double* P);
template <class real>
void launch_density_hist(hipStream_t st, const BoxDesc& B, const real* rho, float drho, int nbins,
                         unsigned long long* counts);

}  // namespace imhd
