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

template <class real>
void launch_force_sums(const Launch& L, const ForceSumArgs<real>& a);
void launch_force_finalize(hipStream_t st, const BoxDesc* boxes, const PlanBox* plan, int nbox, const ForceFinArgs& a);
template <class real>
void launch_force_apply(const Launch& L, const ForceApplyArgs<real>& a);

}  // namespace imhd
