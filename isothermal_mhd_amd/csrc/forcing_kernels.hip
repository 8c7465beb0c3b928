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

// ---------------------------------------------------------------------------------
// Ornstein-Uhlenbeck forcing

// one thread per (box, mode): OU update of the 3 complex components (init: draw from
// the stationary distribution, counter word 3 = 1).  The box's step counter is advanced
// afterwards by k_ou_bump.
__global__ void k_ou_update(int nbox, const BoxDesc* __restrict__ boxes, OuDev ou, int init) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nbox * ou.nmode) return;
    const int b = t / ou.nmode, m = t - b * ou.nmode;
    const unsigned step = ou.step[b];
    double xr[3], xi[3];
    for (int d = 0; d < 3; d++) {
        unsigned o[4];
        philox4x32(step, (unsigned)m, (unsigned)d, init ? 1u : 0u, ou.seed_lo, ou.seed_hi ^ (unsigned)b, o);
        xr[d] = box_muller(o[0], o[1]);
        xi[d] = box_muller(o[2], o[3]);
    }
    if (ou.solenoidal) {   // project onto the plane normal to the physical k = 2 pi n_d / L_d
        const int* n = ou.n + 3 * m;
        const BoxDesc& B = boxes[b];
        double k[3];
        for (int d = 0; d < 3; d++) k[d] = (double)n[d] / ((double)B.n[d] * B.dx);
        double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
        double dr = k[0] * xr[0] + k[1] * xr[1] + k[2] * xr[2];
        double di = k[0] * xi[0] + k[1] * xi[1] + k[2] * xi[2];
        for (int d = 0; d < 3; d++) {
            xr[d] = xr[d] - k[d] * dr / k2;
            xi[d] = xi[d] - k[d] * di / k2;
        }
    }
    double* ar = ou.are + ((size_t)b * ou.nmode + m) * 3;
    double* ai = ou.aim + ((size_t)b * ou.nmode + m) * 3;
    if (init) {
        for (int d = 0; d < 3; d++) { ar[d] = ou.sigma * xr[d]; ai[d] = ou.sigma * xi[d]; }
    } else {
        const double f = exp(-ou.dt[b] / ou.t_corr);
        const double g = ou.sigma * sqrt(1.0 - f * f);
        for (int d = 0; d < 3; d++) {
            ar[d] = f * ar[d] + g * xr[d];
            ai[d] = f * ai[d] + g * xi[d];
        }
    }
}

__global__ void k_ou_bump(int nbox, unsigned int* step) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nbox) step[b] = step[b] + 1u;
}

// a(x) of cell (i, j, k) of box b (interior indices from 1)
// Field of the OU modes at interior cell (i, j, k): a(x) = sum_m 2 Re(a_m e^{i k_m.x}),
// evaluated separably.  k_ou_columns forms, per (j, k) column and per n_x, the sum over
// the modes with that n_x of a_m e_y(n_y, j) e_z(n_z, k); each cell then sums over the
// 2K+1 values of n_x only: 2 Re(e_x(n_x, i) D(n_x; j, k)).
__global__ __launch_bounds__(256) void k_ou_columns(int nbox, const BoxDesc* __restrict__ boxes, OuDev ou) {
    const int b = (int)blockIdx.y;
    const BoxDesc& B = boxes[b];
    const int col = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (col >= B.n[1] * B.n[2]) return;
    const int j = col % B.n[1] + 1, k = col / B.n[1] + 1;
    const int W = 2 * ou.K + 1;
    const double* tx = ou.tw + ou.twoff[b];
    const double* ty = tx + (size_t)2 * W * B.n[0];
    const double* tz = ty + (size_t)2 * W * B.n[1];
    const double* ar = ou.are + (size_t)b * ou.nmode * 3;
    const double* ai = ou.aim + (size_t)b * ou.nmode * 3;
    double* D = ou.col + ou.coloff[b] + (size_t)col * W * 6;
    for (int slot = 0; slot < W; slot++) {
        double dr[3] = {0.0, 0.0, 0.0}, di[3] = {0.0, 0.0, 0.0};
        for (int m = 0; m < ou.nmode; m++) {
            const int* n = ou.n + 3 * m;
            if (n[0] + ou.K != slot) continue;
            const double* ey = ty + 2 * ((size_t)(n[1] + ou.K) * B.n[1] + (j - 1));
            const double* ez = tz + 2 * ((size_t)(n[2] + ou.K) * B.n[2] + (k - 1));
            const double er = ey[0] * ez[0] - ey[1] * ez[1], ei = ey[0] * ez[1] + ey[1] * ez[0];
            for (int d = 0; d < 3; d++) {
                dr[d] = dr[d] + (ar[3 * m + d] * er - ai[3 * m + d] * ei);
                di[d] = di[d] + (ar[3 * m + d] * ei + ai[3 * m + d] * er);
            }
        }
        for (int d = 0; d < 3; d++) {
            D[slot * 6 + 2 * d] = dr[d];
            D[slot * 6 + 2 * d + 1] = di[d];
        }
    }
}

// The OU passes walk interior rows: a workgroup holds 256 consecutive x-cells (x-block xb)
// and takes the rows (j, k) r = ry, ry + OU_RY, ...; the x-twiddles of its cell stay in
// registers (KC = K <= 4; KC = 0: any K, loaded per row), and the row's mode sums D are
// uniform over the workgroup.  Per cell the field a(x) = sum_slot 2 Re(e_x D) is summed in
// slot order, as ou_field does (FAST = false, the FP64 state's arithmetic).  FAST (an FP32
// state) forms each slot's term as two FMAs with the twiddles prescaled by 2 and -2: the
// same field to FP64 rounding, in 30 instead of 75 operations per cell.
template <int KC, bool FAST>
struct OuRowField {
    static constexpr int WC = KC > 0 ? 2 * KC + 1 : 1;
    double xr[WC], xi[WC];
    const double* tx;
    int W, nx, i;
    __device__ __forceinline__ OuRowField() {}
    __device__ __forceinline__ OuRowField(const OuDev& ou, const BoxDesc& B, int b, int i_) : i(i_) {
        W = 2 * ou.K + 1;
        nx = B.n[0];
        tx = ou.tw + ou.twoff[b];
        if constexpr (KC > 0) {
#pragma unroll
            for (int s = 0; s < WC; s++) {
                const double* ex = tx + 2 * ((size_t)s * nx + (i - 1));
                xr[s] = FAST ? 2.0 * ex[0] : ex[0];
                xi[s] = FAST ? -2.0 * ex[1] : ex[1];
            }
        }
    }
    __device__ __forceinline__ void eval(const double* D, double* a) const {
        a[0] = a[1] = a[2] = 0.0;
        if constexpr (KC > 0) {
#pragma unroll
            for (int s = 0; s < WC; s++)
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    if constexpr (FAST) a[d] = fma(xr[s], D[s * 6 + 2 * d], fma(xi[s], D[s * 6 + 2 * d + 1], a[d]));
                    else a[d] = a[d] + 2.0 * (xr[s] * D[s * 6 + 2 * d] - xi[s] * D[s * 6 + 2 * d + 1]);
                }
        } else {
            for (int s = 0; s < W; s++) {
                const double* ex = tx + 2 * ((size_t)s * nx + (i - 1));
                const double er = ex[0], ei = ex[1];
                for (int d = 0; d < 3; d++) {
                    if constexpr (FAST) a[d] = fma(2.0 * er, D[s * 6 + 2 * d], fma(-2.0 * ei, D[s * 6 + 2 * d + 1], a[d]));
                    else a[d] = a[d] + 2.0 * (er * D[s * 6 + 2 * d] - ei * D[s * 6 + 2 * d + 1]);
                }
            }
        }
    }
};

// <rho a> partial sums: one partial row of 3 sums per workgroup (box, x-block, row group),
// reduced in a fixed order by k_ou_finalize.  A thread takes CPX cells of each of its rows
// (x, x + 256, ..), so one fetch of a row's mode sums D (scalar loads, a dependent round
// trip) serves CPX independent density loads, and a workgroup of 256 CPX cells fetches D
// once where 256-cell x-blocks fetched it CPX times; the row groups are CPX times as many,
// so the grid keeps its size.  Per thread the sums run row by row, cells in x order.
template <class real, int KC, int CPX>
__global__ __launch_bounds__(256) void k_ou_sums(const BoxDesc* __restrict__ boxes, OuApplyArgs<real> a, int nxb,
                                                 int nry) {
    __shared__ double sh[FORCE_NSUM][256];
    const int b = blockIdx.z;
    const BoxDesc& B = boxes[b];
    const int i0 = (int)blockIdx.x * 256 * CPX + (int)threadIdx.x + 1;
    const int ny = B.n[1], nrow = B.n[1] * B.n[2];
    const int W = 2 * a.ou.K + 1;
    const double* col = a.ou.col + a.ou.coloff[b];
    double v[FORCE_NSUM] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (i0 <= B.n[0]) {
        bool act[CPX];
        int iu[CPX];   // a cell past the box end stands in for the first (its weight is 0)
#pragma unroll
        for (int u = 0; u < CPX; u++) {
            act[u] = i0 + 256 * u <= B.n[0];
            iu[u] = act[u] ? i0 + 256 * u : i0;
        }
        OuRowField<KC, sizeof(real) == 4> F[CPX];
#pragma unroll
        for (int u = 0; u < CPX; u++) F[u] = OuRowField<KC, sizeof(real) == 4>(a.ou, B, b, iu[u]);
#pragma unroll 2
        for (int r = (int)blockIdx.y; r < nrow; r += nry) {
            const int k = r / ny, j = r - k * ny;
            double rho[CPX];
#pragma unroll
            for (int u = 0; u < CPX; u++) {
                const double q0 = (double)a.q[0][cell_index(B, iu[u], j + 1, k + 1)];
                rho[u] = act[u] ? q0 : 0.0;
            }
#pragma unroll
            for (int u = 0; u < CPX; u++) {
                double f[3];
                F[u].eval(col + (size_t)r * W * 6, f);
                v[0] = v[0] + rho[u] * f[0];
                v[1] = v[1] + rho[u] * f[1];
                v[2] = v[2] + rho[u] * f[2];
            }
        }
    }
    block_sum5(sh, v);
    if (threadIdx.x < 3)
        a.part[(((size_t)b * nxb + blockIdx.x) * nry + blockIdx.y) * 3 + threadIdx.x] = sh[threadIdx.x][0];
}

// kick: m += dt (rho a - <rho a>) on every interior cell, and (cmax non-null) the CFL
// speed of the kicked cell: FP64 from the FP64 state, FP32 (hardware rcp/rsqrt/sqrt) from
// an FP32 state, maximum in FP64
template <class real, int KC>
__global__ __launch_bounds__(256) void k_ou_kick(const BoxDesc* __restrict__ boxes, OuApplyArgs<real> a) {
    const int b = blockIdx.z;
    const BoxDesc& B = boxes[b];
    const int i = (int)blockIdx.x * 256 + (int)threadIdx.x + 1;
    const int ny = B.n[1], nrow = B.n[1] * B.n[2];
    const int W = 2 * a.ou.K + 1;
    const double* col = a.ou.col + a.ou.coloff[b];
    const double dt = a.ou.dt[b];
    const double* mean = a.mean + 3 * b;
    double cfl = 0.0;
    if (i <= B.n[0]) {
        const OuRowField<KC, sizeof(real) == 4> F(a.ou, B, b, i);
        const double m0 = mean[0], m1 = mean[1], m2 = mean[2];
#pragma unroll 2
        for (int r = (int)blockIdx.y; r < nrow; r += OU_RY) {
            const int k = r / ny, j = r - k * ny;
            const long long c = cell_index(B, i, j + 1, k + 1);
            double f[3];
            F.eval(col + (size_t)r * W * 6, f);
            real q7[7];
            q7[0] = a.q[0][c];
            const double rho = (double)q7[0];
            q7[1] = (real)((double)a.q[1][c] + dt * (rho * f[0] - m0));
            q7[2] = (real)((double)a.q[2][c] + dt * (rho * f[1] - m1));
            q7[3] = (real)((double)a.q[3][c] + dt * (rho * f[2] - m2));
            a.q[1][c] = q7[1];
            a.q[2][c] = q7[2];
            a.q[3][c] = q7[3];
            if (a.cmax != nullptr) {
                q7[4] = a.q[4][c];
                q7[5] = a.q[5][c];
                q7[6] = a.q[6][c];
                if constexpr (sizeof(real) == 4) cfl = fmax(cfl, (double)cell_fast_speed_f(q7, (float)a.cs));
                else cfl = fmax(cfl, cell_fast_speed(q7, a.cs));
            }
        }
    }
    if (a.cmax != nullptr) atomic_max_nonneg(a.cmax, b, cfl);
}

// per box: <rho a> = sum / cells, its nrows partial rows (x-blocks x row groups of the
// <rho a> pass) in a fixed order
__global__ __launch_bounds__(256) void k_ou_finalize(const BoxDesc* __restrict__ boxes, const double* part,
                                                     double* mean, int nrows) {
    __shared__ double sh[FORCE_NSUM][256];
    const int b = blockIdx.x;
    const BoxDesc& B = boxes[b];
    const double ncell = (double)B.n[0] * (double)B.n[1] * (double)B.n[2];
    double v[FORCE_NSUM] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int w = threadIdx.x; w < nrows; w += 256)
        for (int s = 0; s < 3; s++) v[s] = v[s] + part[((size_t)b * nrows + w) * 3 + s];
    block_sum5(sh, v);
    if (threadIdx.x == 0)
        for (int s = 0; s < 3; s++) mean[3 * b + s] = sh[s][0] / ncell;
}

void launch_ou_update(hipStream_t st, int nbox, const BoxDesc* boxes, const OuDev& ou, int init) {
    const int n = nbox * ou.nmode;
    hipLaunchKernelGGL(k_ou_update, dim3((n + 255) / 256), dim3(256), 0, st, nbox, boxes, ou, init);
    if (!init) hipLaunchKernelGGL(k_ou_bump, dim3((nbox + 255) / 256), dim3(256), 0, st, nbox, ou.step);
}
void launch_ou_columns(hipStream_t st, int nbox, int maxcols, const BoxDesc* boxes, const OuDev& ou) {
    hipLaunchKernelGGL(k_ou_columns, dim3((unsigned)((maxcols + 255) / 256), (unsigned)nbox), dim3(256), 0, st, nbox,
                       boxes, ou);
}
// x-blocks of 256 cells covering the widest box
static int ou_nxb(int maxnx) { return (maxnx + 255) / 256; }
// the <rho a> pass: cells per thread (K <= 4: 4 for boxes 1024 cells wide or more, 2 from
// 512), its x-blocks of 256 cpx cells and its row groups
static int ou_cpx(int maxnx, int K) { return K < 1 || K > 4 ? 1 : (maxnx >= 1024 ? 4 : (maxnx >= 512 ? 2 : 1)); }
static int ou_sum_nxb(int maxnx, int cpx) { return (maxnx + 256 * cpx - 1) / (256 * cpx); }

template <class real, int KC>
static void launch_ou_sums_k(hipStream_t st, dim3 g, int cpx, const BoxDesc* boxes, const OuApplyArgs<real>& a,
                             int nxb, int nry) {
    if (KC == 0 || cpx == 1)
        hipLaunchKernelGGL((k_ou_sums<real, KC, 1>), g, dim3(256), 0, st, boxes, a, nxb, nry);
    else if (cpx == 2)
        hipLaunchKernelGGL((k_ou_sums<real, KC, 2>), g, dim3(256), 0, st, boxes, a, nxb, nry);
    else
        hipLaunchKernelGGL((k_ou_sums<real, KC, 4>), g, dim3(256), 0, st, boxes, a, nxb, nry);
}
template <class real>
void launch_ou_sums(hipStream_t st, int nbox, int maxnx, const BoxDesc* boxes, const OuApplyArgs<real>& a) {
    const int cpx = ou_cpx(maxnx, a.ou.K), nxb = ou_sum_nxb(maxnx, cpx), nry = OU_RY * cpx;
    const dim3 g((unsigned)nxb, (unsigned)nry, (unsigned)nbox);
    switch (a.ou.K) {
        case 1: launch_ou_sums_k<real, 1>(st, g, cpx, boxes, a, nxb, nry); break;
        case 2: launch_ou_sums_k<real, 2>(st, g, cpx, boxes, a, nxb, nry); break;
        case 3: launch_ou_sums_k<real, 3>(st, g, cpx, boxes, a, nxb, nry); break;
        case 4: launch_ou_sums_k<real, 4>(st, g, cpx, boxes, a, nxb, nry); break;
        default: launch_ou_sums_k<real, 0>(st, g, 1, boxes, a, nxb, nry); break;
    }
}
template <class real>
void launch_ou_kick(hipStream_t st, int nbox, int maxnx, const BoxDesc* boxes, const OuApplyArgs<real>& a) {
    const dim3 g((unsigned)ou_nxb(maxnx), OU_RY, (unsigned)nbox);
    switch (a.ou.K) {
        case 1: hipLaunchKernelGGL((k_ou_kick<real, 1>), g, dim3(256), 0, st, boxes, a); break;
        case 2: hipLaunchKernelGGL((k_ou_kick<real, 2>), g, dim3(256), 0, st, boxes, a); break;
        case 3: hipLaunchKernelGGL((k_ou_kick<real, 3>), g, dim3(256), 0, st, boxes, a); break;
        case 4: hipLaunchKernelGGL((k_ou_kick<real, 4>), g, dim3(256), 0, st, boxes, a); break;
        default: hipLaunchKernelGGL((k_ou_kick<real, 0>), g, dim3(256), 0, st, boxes, a); break;
    }
}
void launch_ou_finalize(hipStream_t st, const BoxDesc* boxes, int nbox, int maxnx, int K, const double* part,
                        double* mean) {
    const int cpx = ou_cpx(maxnx, K);
    const int nrows = ou_sum_nxb(maxnx, cpx) * OU_RY * cpx;
    hipLaunchKernelGGL(k_ou_finalize, dim3(nbox), dim3(256), 0, st, boxes, part, mean, nrows);
}
size_t ou_part_rows(int nbox, int maxnx) {
    size_t r = 0;
    for (int cpx = 1; cpx <= 4; cpx *= 2) r = std::max(r, (size_t)ou_sum_nxb(maxnx, cpx) * OU_RY * cpx);
    return (size_t)nbox * r;
}
template void launch_ou_sums<double>(hipStream_t, int, int, const BoxDesc*, const OuApplyArgs<double>&);
template void launch_ou_sums<float>(hipStream_t, int, int, const BoxDesc*, const OuApplyArgs<float>&);
template void launch_ou_kick<double>(hipStream_t, int, int, const BoxDesc*, const OuApplyArgs<double>&);
template void launch_ou_kick<float>(hipStream_t, int, int, const BoxDesc*, const OuApplyArgs<float>&);

template void launch_force_sums<double>(const Launch&, const ForceSumArgs<double>&);
template void launch_force_sums<float>(const Launch&, const ForceSumArgs<float>&);
template void launch_force_apply<double>(const Launch&, const ForceApplyArgs<double>&);
template void launch_force_apply<float>(const Launch&, const ForceApplyArgs<float>&);

// ---------------------------------------------------------------------------------
// generate_rand_force: y and z passes of the direct inverse DFT (turb_driver_mod.f90:
// 310-385), per output value the sum over j = -(K-1) .. K-1 in the reference's order and
// arithmetic form (in_re cos + (-in_im) sin; in_im cos - (-in_re) sin).  Compiled without
// FMA contraction, with the host's twiddles, so the field is the host restatement's bit
// for bit.
// grid: x over the plane (ix fastest), y over the remaining index (d * S + kz slot for
// the y pass, iz for the z pass): 32-bit index arithmetic, no 64-bit division
__global__ __launch_bounds__(256) void k_rand_force_y(RandForceArgs a) {
    const int S = 2 * a.K, J = 2 * a.K - 1;
    const int plane = a.P[0] * a.P[1];
    const int xy = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (xy >= plane) return;
    const int ix = xy % a.P[0], iy = xy / a.P[0];
    const int dz = (int)blockIdx.y;   // d * S + kz slot
    const double sg = -1.0;
    double re = 0.0, im = 0.0;
    for (int j = -(a.K - 1); j <= a.K - 1; j++) {
        const long long xo = ((long long)dz * S + (j + a.K)) * a.P[0] + ix;
        const double xr = a.xr[xo], xi = a.xi[xo];
        const double c = a.twy[2 * (iy * J + j + a.K - 1)], s = a.twy[2 * (iy * J + j + a.K - 1) + 1];
        re = re + (xr * c + sg * xi * s);
        im = im + (xi * c - sg * xr * s);
    }
    const long long o = (long long)dz * plane + xy;
    a.yr[o] = re;
    a.yi[o] = im;
}

// z pass, z-marching: each thread keeps its column's y-pass values (3 components x
// (2K-1) wave numbers, complex) in registers and walks RF_ZSEG planes, so the y-pass
// buffer is read once per segment instead of once per plane.  KC > 0: K <= KC, values
// cached; KC == 0: any K, values re-read per plane.  The per-value arithmetic and the
// order of the j sum are those of the plain pass.
constexpr int RF_ZSEG = 32;

template <int KC>
__global__ __launch_bounds__(256) void k_rand_force_z(RandForceArgs a) {
    const int S = 2 * a.K, J = 2 * a.K - 1;
    const int plane = a.P[0] * a.P[1];
    const int xy = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (xy >= plane) return;
    const int z0 = (int)blockIdx.y * RF_ZSEG;
    const int z1 = min(z0 + RF_ZSEG, a.P[2]);
    const double sg = -1.0;
    constexpr int JC = KC > 0 ? 2 * KC - 1 : 1;
    double cr[3][JC], ci[3][JC];
    if (KC > 0) {
        for (int d = 0; d < 3; d++)
            for (int t = 0; t < JC; t++) {
                cr[d][t] = 0.0;
                ci[d][t] = 0.0;
                if (t < J) {
                    const long long yo = ((long long)d * S + (t + 1)) * plane + xy;
                    cr[d][t] = a.yr[yo];
                    ci[d][t] = a.yi[yo];
                }
            }
    }
    for (int iz = z0; iz < z1; iz++) {
        const long long g = (long long)iz * plane + xy;
        for (int d = 0; d < 3; d++) {
            double re = 0.0;
            if (KC > 0) {
#pragma unroll
                for (int t = 0; t < JC; t++) {
                    if (t < J) {
                        const double c = a.twz[2 * (iz * J + t)], s = a.twz[2 * (iz * J + t) + 1];
                        re = re + (cr[d][t] * c + sg * ci[d][t] * s);
                    }
                }
            } else {
                for (int t = 0; t < J; t++) {
                    const long long yo = ((long long)d * S + (t + 1)) * plane + xy;
                    const double c = a.twz[2 * (iz * J + t)], s = a.twz[2 * (iz * J + t) + 1];
                    re = re + (a.yr[yo] * c + sg * a.yi[yo] * s);
                }
            }
            a.delv[d][g] = re * a.idft_pre;
        }
    }
}

void launch_rand_force_y(hipStream_t st, const RandForceArgs& a) {
    const int plane = a.P[0] * a.P[1];
    hipLaunchKernelGGL(k_rand_force_y, dim3((unsigned)((plane + 255) / 256), (unsigned)(3 * 2 * a.K)), dim3(256), 0,
                       st, a);
}

void launch_rand_force_z(hipStream_t st, const RandForceArgs& a) {
    const int plane = a.P[0] * a.P[1];
    const dim3 grid((unsigned)((plane + 255) / 256), (unsigned)((a.P[2] + RF_ZSEG - 1) / RF_ZSEG));
    if (a.K <= 4)
        hipLaunchKernelGGL(k_rand_force_z<4>, grid, dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL(k_rand_force_z<0>, grid, dim3(256), 0, st, a);
}

}  // namespace imhd
