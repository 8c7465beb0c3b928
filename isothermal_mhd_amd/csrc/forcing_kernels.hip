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

// Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11): 10 rounds of two 32x32->64
// multiplies with the Weyl key schedule.
__device__ __forceinline__ void philox4x32(unsigned c0, unsigned c1, unsigned c2, unsigned c3, unsigned k0,
                                           unsigned k1, unsigned* o) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        unsigned long long p0 = (unsigned long long)0xD2511F53u * c0;
        unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c2;
        unsigned y0 = (unsigned)(p1 >> 32) ^ c1 ^ k0, y1 = (unsigned)p1;
        unsigned y2 = (unsigned)(p0 >> 32) ^ c3 ^ k1, y3 = (unsigned)p0;
        c0 = y0; c1 = y1; c2 = y2; c3 = y3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    o[0] = c0; o[1] = c1; o[2] = c2; o[3] = c3;
}

// Box-Muller (cosine branch) from two 32-bit words mapped to (0, 1)
__device__ __forceinline__ double box_muller(unsigned a, unsigned b) {
    const double s = 2.3283064365386963e-10;   // 2^-32
    double u1 = ((double)a + 0.5) * s, u2 = ((double)b + 0.5) * s;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

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

__device__ __forceinline__ void ou_field(const OuDev& ou, const BoxDesc& B, int b, int i, int j, int k, double* a) {
    const int W = 2 * ou.K + 1;
    const double* tx = ou.tw + ou.twoff[b];
    const double* D = ou.col + ou.coloff[b] + ((size_t)(k - 1) * B.n[1] + (j - 1)) * W * 6;
    a[0] = a[1] = a[2] = 0.0;
    for (int slot = 0; slot < W; slot++) {
        const double* ex = tx + 2 * ((size_t)slot * B.n[0] + (i - 1));
        const double xr = ex[0], xi = ex[1];
        for (int d = 0; d < 3; d++) a[d] = a[d] + 2.0 * (xr * D[slot * 6 + 2 * d] - xi * D[slot * 6 + 2 * d + 1]);
    }
}

template <class real>
__global__ __launch_bounds__(256) void k_ou_pass(const BoxDesc* __restrict__ boxes, Plan plan, OuApplyArgs<real> a) {
    __shared__ double sh[FORCE_NSUM][256];
    int b, i, j, k;
    const bool act = plan_item_box(plan, b, i, j, k);
    double v[FORCE_NSUM] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double cfl = 0.0;
    if (act) {
        const BoxDesc& B = boxes[b];
        const long long c = cell_index(B, i, j, k);
        double f[3];
        ou_field(a.ou, B, b, i, j, k, f);
        const double rho = (double)a.q[0][c];
        if (a.pass == 0) {
            v[0] = rho * f[0];
            v[1] = rho * f[1];
            v[2] = rho * f[2];
        } else {
            const double dt = a.ou.dt[b];
            const double* mean = a.mean + 3 * b;
            for (int d = 0; d < 3; d++) a.q[1 + d][c] = (real)((double)a.q[1 + d][c] + dt * (rho * f[d] - mean[d]));
            if (a.cmax != nullptr) {
                double q7[7];
                for (int g = 0; g < 7; g++) q7[g] = (double)a.q[g][c];
                cfl = cell_fast_speed(q7, a.cs);
            }
        }
    }
    if (a.pass == 0) {
        block_sum5(sh, v);
        if (threadIdx.x == 0) {
            const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
            for (int s = 0; s < 3; s++) a.part[(size_t)blk * 3 + s] = sh[s][0];
        }
    } else if (a.cmax != nullptr) {
        atomic_max_nonneg(a.cmax, b, cfl);
    }
}

// per box: <rho a> = sum / cells, fixed summation order
__global__ __launch_bounds__(256) void k_ou_finalize(const BoxDesc* __restrict__ boxes, const PlanBox* __restrict__ plan,
                                                     const double* part, double* mean) {
    __shared__ double sh[FORCE_NSUM][256];
    const int b = blockIdx.x;
    const PlanBox& pb = plan[b];
    const long long first = pb.start / 256;
    const long long ncell = (long long)pb.r[0] * pb.r[1] * pb.r[2];
    const long long nblk = (ncell + 255) / 256;
    double v[FORCE_NSUM] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (long long q = threadIdx.x; q < nblk; q += 256)
        for (int s = 0; s < 3; s++) v[s] = v[s] + part[(size_t)(first + q) * 3 + s];
    block_sum5(sh, v);
    if (threadIdx.x == 0)
        for (int s = 0; s < 3; s++) mean[3 * b + s] = sh[s][0] / (double)ncell;
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
template <class real>
void launch_ou_pass(const Launch& L, const OuApplyArgs<real>& a) {
    if (L.plan.total > 0) hipLaunchKernelGGL((k_ou_pass<real>), dim3(L.blocks()), dim3(256), 0, L.stream, L.boxes, L.plan, a);
}
void launch_ou_finalize(hipStream_t st, const BoxDesc* boxes, const PlanBox* plan, int nbox, const double* part,
                        double* mean) {
    hipLaunchKernelGGL(k_ou_finalize, dim3(nbox), dim3(256), 0, st, boxes, plan, part, mean);
}
template void launch_ou_pass<double>(const Launch&, const OuApplyArgs<double>&);
template void launch_ou_pass<float>(const Launch&, const OuApplyArgs<float>&);

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
