// Forcing application kernels (FP64 arithmetic, no FMA contraction).
//
// get_rho_turb's sums are reduced deterministically in two passes: each 256-item block of
// the interior plan writes its FORCE_NSUM partial sums (fixed tree order), then one
// workgroup per box adds its blocks' partials in a fixed order and solves for the
// constant-energy normalisation.  Results are therefore identical run to run and between
// eager launches and graph replays; they differ from the reference's sequential sums only
// by summation order (about 1e-16 relative).
#include "forcing.hpp"

namespace imhd {

__device__ __forceinline__ void block_sum5(double (*sh)[256], double* v) {
    const int t = threadIdx.x;
#pragma unroll
    for (int s = 0; s < FORCE_NSUM; s++) sh[s][t] = v[s];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
#pragma unroll
            for (int s = 0; s < FORCE_NSUM; s++) sh[s][t] = sh[s][t] + sh[s][t + w];
        }
        __syncthreads();
    }
}

// per cell (turb_driver_mod.f90:159-176): rho (dv.dv), dv.m, rho dv
template <class real>
__global__ __launch_bounds__(256) void k_force_sums(const BoxDesc* __restrict__ boxes, Plan plan, ForceSumArgs<real> a) {
    __shared__ double sh[FORCE_NSUM][256];
    int b, i, j, k;
    double v[FORCE_NSUM] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (plan_item(plan, b, i, j, k)) {
        const long long c = cell_index(boxes[b], i, j, k);
        const double rho = (double)a.q[0][c];
        const double d1 = a.delv[0][c], d2 = a.delv[1][c], d3 = a.delv[2][c];
        v[0] = rho * (d1 * d1 + d2 * d2 + d3 * d3);
        v[1] = d1 * (double)a.q[1][c] + d2 * (double)a.q[2][c] + d3 * (double)a.q[3][c];
        v[2] = rho * d1;
        v[3] = rho * d2;
        v[4] = rho * d3;
    }
    block_sum5(sh, v);
    if (threadIdx.x == 0) {
        const unsigned blk = xcd_block(blockIdx.x, gridDim.x);   // the plan block this workgroup served
#pragma unroll
        for (int s = 0; s < FORCE_NSUM; s++) a.part[(size_t)blk * FORCE_NSUM + s] = sh[s][0];
    }
}

// one workgroup per box: add the box's block partials, then add_random_force's
// normalisation (turb_driver_mod.f90:428-442)
__global__ __launch_bounds__(256) void k_force_finalize(const BoxDesc* __restrict__ boxes, const PlanBox* __restrict__ plan,
                                                        ForceFinArgs a) {
    __shared__ double sh[FORCE_NSUM][256];
    const int b = blockIdx.x;
    const PlanBox& pb = plan[b];
    const long long first = pb.start / 256;
    const long long nblk = ((long long)pb.r[0] * pb.r[1] * pb.r[2] + 255) / 256;
    double v[FORCE_NSUM] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (long long q = threadIdx.x; q < nblk; q += 256)
#pragma unroll
        for (int s = 0; s < FORCE_NSUM; s++) v[s] = v[s] + a.part[(size_t)(first + q) * FORCE_NSUM + s];
    block_sum5(sh, v);
    if (threadIdx.x != 0) return;
    const double quad1 = 0.5 * sh[0][0], quad2 = sh[1][0];
    const double dx = boxes[b].dx, dx3 = dx * dx * dx;
    double* par = a.par + (size_t)b * FORCE_NPAR;
    if (quad1 <= 1.0e-20) {          // the reference stops here (turb_driver_mod.f90:420-425)
        atomicOr(a.err, 1u);
        par[0] = par[1] = par[2] = par[3] = 0.0;
        return;
    }
    const double tmp1 = a.edot * a.dt_turb / dx3;
    const double tmp2 = sqrt(quad2 * quad2 + (4.0 * tmp1 * quad1));
    par[0] = 0.5 * (-quad2 + tmp2) / quad1;
    par[1] = sh[2][0] * dx3;
    par[2] = sh[3][0] * dx3;
    par[3] = sh[4][0] * dx3;
}

// delta m = constE (rho0 dv - <rho dv>) on interior cells, max |delta m| / rho0
// (turb_driver_mod.f90:444-455)
template <class real>
__global__ __launch_bounds__(256) void k_force_apply(const BoxDesc* __restrict__ boxes, Plan plan, ForceApplyArgs<real> a) {
    int b, i, j, k;
    double m = 0.0;
    if (plan_item_box(plan, b, i, j, k)) {
        const long long c = cell_index(boxes[b], i, j, k);
        const double* par = a.par + (size_t)b * FORCE_NPAR;
        const double r0 = (double)a.rho0[c];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double dm = par[0] * (r0 * a.delv[d][c] - par[1 + d]);
            a.m[d][c] = (real)((double)a.m[d][c] + dm);
            m = fmax(m, fabs(dm) / r0);
        }
    }
    atomic_max_nonneg(a.dvmax, b, m);
}

template <class real>
void launch_force_sums(const Launch& L, const ForceSumArgs<real>& a) {
    if (L.plan.total > 0) hipLaunchKernelGGL((k_force_sums<real>), dim3(L.blocks()), dim3(256), 0, L.stream, L.boxes, L.plan, a);
}
void launch_force_finalize(hipStream_t st, const BoxDesc* boxes, const PlanBox* plan, int nbox, const ForceFinArgs& a) {
    hipLaunchKernelGGL(k_force_finalize, dim3(nbox), dim3(256), 0, st, boxes, plan, a);
}
template <class real>
void launch_force_apply(const Launch& L, const ForceApplyArgs<real>& a) {
    if (L.plan.total > 0) hipLaunchKernelGGL((k_force_apply<real>), dim3(L.blocks()), dim3(256), 0, L.stream, L.boxes, L.plan, a);
}

template void launch_force_sums<double>(const Launch&, const ForceSumArgs<double>&);
template void launch_force_sums<float>(const Launch&, const ForceSumArgs<float>&);
template void launch_force_apply<double>(const Launch&, const ForceApplyArgs<double>&);
template void launch_force_apply<float>(const Launch&, const ForceApplyArgs<float>&);

}  // namespace imhd
