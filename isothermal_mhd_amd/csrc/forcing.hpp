// Turbulence forcing kernels: application of a driver-supplied force field as the
// reference's turbulence driver does it (turb_driver_mod.f90:130-192, 395-457).
#pragma once
#include "kernels.hpp"

namespace imhd {

constexpr int FORCE_NSUM = 5;   // 0.5 sum rho dv^2 (x2), sum dv.m, sum rho dv (3)
constexpr int FORCE_NPAR = 4;   // per box: constE, <rho dv> dx^3 (3)

template <class real>
struct ForceSumArgs {
    const real* q[4];          // rho, m at t^n
    const double* delv[3];     // force field
    double* part;              // FORCE_NSUM per 256-item block of the plan
};

struct ForceFinArgs {
    const double* part;
    double* par;               // FORCE_NPAR per box
    double edot, dt_turb;
    unsigned int* err;         // set when the energy normalisation is undefined
};

template <class real>
struct ForceApplyArgs {
    const real* rho0;          // rho at t^n
    real* m[3];                // momentum of the new state
    const double* delv[3];
    const double* par;
    unsigned long long* dvmax; // striped per-box max |delta m| / rho0
};

// ---- Ornstein-Uhlenbeck forcing (native loop) ----
// Mode m of box b: integer wave vector n_m (one of each +/- pair), complex amplitude
// a_m (3 components).  Per step: a <- f a + sigma sqrt(1 - f^2) P(k) xi, f = exp(-dt/T),
// xi complex Gaussian from Philox4x32-10 keyed by (seed, box), counter (step, m, d, 0).
// Field a(x) = sum_m 2 Re(a_m e^{i k_m.x}) from per-axis twiddle tables; applied as
// dm = dt (rho a - <rho a>) on interior cells (no net momentum injection).
constexpr int OU_MAXMODE = 128;

struct OuDev {
    const int* n;            // 3 per mode
    double* are;             // [box][mode][3] real parts
    double* aim;             // imaginary parts
    unsigned int* step;      // per box OU step counter
    const double* tw;        // twiddles: per box, axis-major blocks of (2K+1) x N_axis complex
    const long long* twoff;  // per box offset into tw
    int nmode, K;
    double t_corr, sigma;
    unsigned int seed_lo, seed_hi;
    int solenoidal;
    const double* dt;        // per box
    double* col;             // per box, per (j, k) column: sum over the modes with the same
                             // n_x of a_m e_y e_z, [k-1][j-1][n_x + K][d] (re, im)
    const long long* coloff; // per box offset into col (doubles)
};

template <class real>
struct OuApplyArgs {
    OuDev ou;
    real* q[7];              // new state (rho, m read; m written; all read for the CFL)
    double* part;            // ou_part_rows partial rows of 3 sums
    double* mean;            // 3 per box
    unsigned long long* cmax;// non-null: fused CFL speed of the forced cells
    double cs;
};
// row groups per x-block of the OU kick: 256 x-blocks of 256 cells at 1024^3, four waves
// per SIMD; more (1024) ran the 1024^3 kick 25 % slower.  The <rho a> pass takes cpx
// cells per thread and OU_RY cpx row groups.
constexpr int OU_RY = 256;

void launch_ou_update(hipStream_t st, int nbox, const BoxDesc* boxes, const OuDev& ou, int init);
void launch_ou_columns(hipStream_t st, int nbox, int maxcols, const BoxDesc* boxes, const OuDev& ou);
template <class real>
void launch_ou_sums(hipStream_t st, int nbox, int maxnx, const BoxDesc* boxes, const OuApplyArgs<real>& a);
template <class real>
void launch_ou_kick(hipStream_t st, int nbox, int maxnx, const BoxDesc* boxes, const OuApplyArgs<real>& a);
void launch_ou_finalize(hipStream_t st, const BoxDesc* boxes, int nbox, int maxnx, int K, const double* part,
                        double* mean);
size_t ou_part_rows(int nbox, int maxnx);   // partial rows of <rho a> (3 doubles each)

// ---- the reference's random force (generate_rand_force) ----
// Inverse DFT passes along y and z of the direct transform (turb_driver_mod.f90:461-486):
// out(i) = sum_{j=-(K-1)}^{K-1} in(j) e^{+i theta0 i j}, summed in that order, twiddles
// (cos, sin) precomputed per (i, j).  Slots s = k + K, k = -K .. K-1.
struct RandForceArgs {
    const double* xr;   // x-pass output [d][kz slot][ky slot][ix]
    const double* xi;
    const double* twy;  // [iy][j + K - 1] (cos, sin)
    const double* twz;  // [iz][j + K - 1]
    double* yr;         // y-pass output [d][kz slot][iy][ix]
    double* yi;
    double* delv[3];    // the box's padded force field
    int K, P[3];
    double idft_pre;    // 1 / (nx ny nz)
};
void launch_rand_force_y(hipStream_t st, const RandForceArgs& a);
void launch_rand_force_z(hipStream_t st, const RandForceArgs& a);

template <class real>
void launch_force_sums(const Launch& L, const ForceSumArgs<real>& a);
void launch_force_finalize(hipStream_t st, const BoxDesc* boxes, const PlanBox* plan, int nbox, const ForceFinArgs& a);
template <class real>
void launch_force_apply(const Launch& L, const ForceApplyArgs<real>& a);

}  // namespace imhd
