// Ghost fill, CFL reduction, diagnostics, passive-scalar advection and layout/precision
// conversion kernels.
#include "aux_kernels.hpp"

#include <mutex>

namespace imhd {

// ---------------------------------------------------------------------------------
// Ghost fill (boundary_conditions_mod.f90:20-234, one rank).  The reference fills x,
// then y, then z over whole planes; the composition equals reading every ghost cell
// from (map_x(i), map_y(j), map_z(k)) with an independent per-axis source map, which is
// what this kernel does in one pass (sources are always interior cells).
__device__ __forceinline__ int bc_map(int i, int n, int bc) {
    if (i >= 1 && i <= n) return i;
    if (bc == 2) return i < 1 ? i + n : i - n;   // periodic: q(-g) = q(n-g), q(n+1+g) = q(1+g)
    return i < 1 ? 1 : n;                         // open: copy the edge cell
}

template <class real>
__global__ __launch_bounds__(256) void k_fill_ghosts(const BoxDesc* __restrict__ boxes, Plan plan, GhostArgs<real> a) {
    int b, i, j, k;
    if (!plan_item(plan, b, i, j, k)) return;
    const BoxDesc& B = boxes[b];
    const int w = a.depth;   // nb for q / bface, 3 for passives (passive_mod.f90:52)
    if (i < 1 - w || j < 1 - w || k < 1 - w || i > B.n[0] + w || j > B.n[1] + w || k > B.n[2] + w) return;
    if (i >= 1 && i <= B.n[0] && j >= 1 && j <= B.n[1] && k >= 1 && k <= B.n[2]) return;
    const int si = bc_map(i, B.n[0], B.bc), sj = bc_map(j, B.n[1], B.bc), sk = bc_map(k, B.n[2], B.bc);
    const long long d = cell_index(B, i, j, k), s = cell_index(B, si, sj, sk);
    for (int f = 0; f < a.nreal; f++) a.r[f][d] = a.r[f][s];
    for (int f = 0; f < a.ndbl; f++) a.d[f][d] = a.d[f][s];
}

// Same fill, enumerating only the ghost band of depth w around the interior: the plan gives
// each box r[0] = number of band cells; the linear index is decoded as z-slabs, then
// y-slabs (interior k), then x-slabs (interior j, k).
template <class real>
__global__ __launch_bounds__(256) void k_fill_ghost_band(const BoxDesc* __restrict__ boxes, Plan plan, GhostArgs<real> a) {
    int b, g, dum1, dum2;
    if (!plan_item(plan, b, g, dum1, dum2)) return;
    const BoxDesc& B = boxes[b];
    const int w = a.depth;
    const long long E0 = B.n[0] + 2 * w, E1 = B.n[1] + 2 * w;
    const long long nA = E0 * E1 * 2 * w, nB = E0 * 2 * w * B.n[2];
    long long t = g;
    int i, j, k;
    if (t < nA) {
        long long kk = t / (E0 * E1), rem = t - kk * E0 * E1;
        k = kk < w ? 1 - w + (int)kk : B.n[2] + 1 + (int)(kk - w);
        j = 1 - w + (int)(rem / E0);
        i = 1 - w + (int)(rem % E0);
    } else if (t < nA + nB) {
        t -= nA;
        long long kk = t / (E0 * 2 * w), rem = t - kk * E0 * 2 * w;
        k = 1 + (int)kk;
        long long jj = rem / E0;
        j = jj < w ? 1 - w + (int)jj : B.n[1] + 1 + (int)(jj - w);
        i = 1 - w + (int)(rem % E0);
    } else {
        t -= nA + nB;
        long long u = t / (2 * w), ii = t - u * 2 * w;
        i = ii < w ? 1 - w + (int)ii : B.n[0] + 1 + (int)(ii - w);
        j = 1 + (int)(u % B.n[1]);
        k = 1 + (int)(u / B.n[1]);
    }
    const int si = bc_map(i, B.n[0], B.bc), sj = bc_map(j, B.n[1], B.bc), sk = bc_map(k, B.n[2], B.bc);
    const long long d = cell_index(B, i, j, k), s = cell_index(B, si, sj, sk);
    for (int f = 0; f < a.nreal; f++) a.r[f][d] = a.r[f][s];
    for (int f = 0; f < a.ndbl; f++) a.d[f][d] = a.d[f][s];
}

// Cells outside [lo, n + hi_off] (any axis) take their t^n values: the state the
// reference's restore (vanleer_mod.f90:404-405) leaves outside the final update range.
template <class real>
__global__ __launch_bounds__(256) void k_restore_shell(const BoxDesc* __restrict__ boxes, Plan plan, ShellArgs<real> a) {
    int b, i, j, k;
    if (!plan_item(plan, b, i, j, k)) return;
    const BoxDesc& B = boxes[b];
    const bool in = i >= a.lo && j >= a.lo && k >= a.lo && i <= B.n[0] + a.hi_off && j <= B.n[1] + a.hi_off &&
                    k <= B.n[2] + a.hi_off;
    if (in) return;
    const long long c = cell_index(B, i, j, k);
    for (int f = 0; f < 7; f++) a.dst[f][c] = a.src[f][c];
    for (int f = 0; f < 3; f++) a.dstb[f][c] = a.srcb[f][c];
}

// ---------------------------------------------------------------------------------
// CFL: max over interior cells of |v_d| + c_f,d (vanleer_mod.f90:3266-3334), FP64
// accumulation; wavefront shuffle reduction and one integer atomicMax per wave.
template <class real>
__global__ __launch_bounds__(256) void k_max_speed(const BoxDesc* __restrict__ boxes, Plan plan, QPtrs<real> Q,
                                                   double cs, unsigned long long* cmax) {
    const real* const* q = Q.q;
    int b = 0, i, j, k;
    double m = 0.0;
    bool act = plan_item(plan, b, i, j, k);
    if (act) {
        const BoxDesc& B = boxes[b];
        long long c = cell_index(B, i, j, k);
        double asqr = cs * cs;
        double rho = q[0][c];
        double isrho = 1.0 / sqrt(rho);
        double bx = (double)q[4][c] * isrho, by = (double)q[5][c] * isrho, bz = (double)q[6][c] * isrho;
        double btsqr = by * by + bz * bz, bsqr = bx * bx + btsqr, x = asqr - bsqr;
        double tmp = sqrt(x * x + 4.0 * asqr * btsqr);
        double cfx = sqrt(0.5 * (asqr + bsqr + tmp));
        btsqr = bx * bx + bz * bz; bsqr = by * by + btsqr; x = asqr - bsqr;
        tmp = sqrt(x * x + 4.0 * asqr * btsqr);
        double cfy = sqrt(0.5 * (asqr + bsqr + tmp));
        btsqr = bx * bx + by * by; bsqr = bz * bz + btsqr; x = asqr - bsqr;
        tmp = sqrt(x * x + 4.0 * asqr * btsqr);
        double cfz = sqrt(0.5 * (asqr + bsqr + tmp));
        double vx = (double)q[1][c] / rho, vy = (double)q[2][c] / rho, vz = (double)q[3][c] / rho;
        m = max3(fabs(vx) + cfx, fabs(vy) + cfy, fabs(vz) + cfz);
    }
    // interior plans are 64-aligned per box: a wave never spans two boxes
    atomic_max_nonneg(cmax, b, m);
}

// dt = COUR dx / max(cmax, smax + delv_max), then MIN(t_end - t, dt) (driver.f90:432-448, 475)
__global__ void k_new_dt(int nbox, const BoxDesc* __restrict__ boxes, const unsigned long long* cmax,
                         const unsigned long long* smax, const unsigned long long* dvmax, const double* cour,
                         double t_end, double* t, double* dt, double* dt_prev, int advance_time) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    double cm = striped_max(cmax, b);
    double sm = striped_max(smax, b) + striped_max(dvmax, b);   // smax + delv_max (vanleer_mod.f90:3323)
    double c = fmax(cm, sm);
    if (advance_time) t[b] = t[b] + dt[b];
    dt_prev[b] = dt[b];
    double dn = cour[b] * boxes[b].dx / c;
    double rem = t_end - t[b];
    dt[b] = rem < dn ? rem : dn;
}

__global__ void k_reset_u64(unsigned long long* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// ---------------------------------------------------------------------------------
// Diagnostics (driver.f90:492-553) in FP64 plus total mass and max|div B|.
//
// Deterministic two-pass reduction: DIAG_WG workgroups per box each walk a fixed set of
// the box's interior cells (cell loc = w*256 + t + s*DIAG_WG*256, x fastest), sum serially
// per thread, then in a fixed tree per workgroup, and write one partial row of DIAG_STRIDE
// doubles (7 sums + the maximum); one workgroup per box then adds its DIAG_WG rows in a
// fixed order.  The result bits do not depend on scheduling, so repeated calls agree
// exactly; no memory has to start at zero (every partial and result is written).
__device__ __forceinline__ void block_reduce8(double (*sh)[256], double* v) {
    const int t = threadIdx.x;
#pragma unroll
    for (int s = 0; s < DIAG_STRIDE; s++) sh[s][t] = v[s];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
#pragma unroll
            for (int s = 0; s < DIAG_STRIDE - 1; s++) sh[s][t] = sh[s][t] + sh[s][t + w];
            sh[DIAG_STRIDE - 1][t] = fmax(sh[DIAG_STRIDE - 1][t], sh[DIAG_STRIDE - 1][t + w]);
        }
        __syncthreads();
    }
}

template <class real>
__global__ __launch_bounds__(256) void k_diag_part(const BoxDesc* __restrict__ boxes, DiagArgs<real> a) {
    __shared__ double sh[DIAG_STRIDE][256];
    const int b = blockIdx.y;
    const BoxDesc& B = boxes[b];
    const unsigned long long nx = B.n[0], nxy = nx * (unsigned long long)B.n[1], ncell = nxy * B.n[2];
    const long long sy = B.P[0], sz = (long long)B.P[0] * B.P[1];
    double v[DIAG_STRIDE] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (unsigned long long loc = (unsigned long long)blockIdx.x * 256 + threadIdx.x; loc < ncell;
         loc += (unsigned long long)DIAG_WG * 256) {
        unsigned long long k, r, j, i;
        if (ncell <= 0xffffffffull) {   // 32-bit index arithmetic whenever the box allows it
            const unsigned l32 = (unsigned)loc, k32 = l32 / (unsigned)nxy, r32 = l32 - k32 * (unsigned)nxy;
            const unsigned j32 = r32 / (unsigned)nx;
            k = k32; j = j32; i = r32 - j32 * (unsigned)nx;
        } else {
            k = loc / nxy; r = loc - k * nxy; j = r / nx; i = r - j * nx;
        }
        const long long c = cell_index(B, (int)i + 1, (int)j + 1, (int)k + 1);
        const double r0 = a.q[0][c];
        const double rho = fmin(fmax(r0, 1e-15), 1e10);
        const double m1 = a.q[1][c], m2 = a.q[2][c], m3 = a.q[3][c];
        const double vsqr = fmin((m1 * m1 + m2 * m2 + m3 * m3) / (rho * rho), 1e5);
        const double db = (a.b[0][c] - a.b[0][c - 1]) + (a.b[1][c] - a.b[1][c - sy]) + (a.b[2][c] - a.b[2][c - sz]);
        v[0] = v[0] + rho * vsqr;
        v[1] = v[1] + vsqr;
        v[2] = v[2] + m1;
        v[3] = v[3] + m2;
        v[4] = v[4] + m3;
        v[5] = v[5] + db;
        v[6] = v[6] + r0;
        v[7] = fmax(v[7], fabs(db));
    }
    block_reduce8(sh, v);
    if (threadIdx.x < DIAG_STRIDE) a.part[((size_t)b * DIAG_WG + blockIdx.x) * DIAG_STRIDE + threadIdx.x] = sh[threadIdx.x][0];
}

// one workgroup per box: its DIAG_WG partial rows in a fixed order -> out[b * DIAG_STRIDE + s]
__global__ __launch_bounds__(256) void k_diag_final(const double* __restrict__ part, double* __restrict__ out) {
    __shared__ double sh[DIAG_STRIDE][256];
    const int b = blockIdx.x;
    double v[DIAG_STRIDE] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int w = threadIdx.x; w < DIAG_WG; w += 256) {
        const double* p = part + ((size_t)b * DIAG_WG + w) * DIAG_STRIDE;
#pragma unroll
        for (int s = 0; s < DIAG_STRIDE - 1; s++) v[s] = v[s] + p[s];
        v[DIAG_STRIDE - 1] = fmax(v[DIAG_STRIDE - 1], p[DIAG_STRIDE - 1]);
    }
    block_reduce8(sh, v);
    if (threadIdx.x < DIAG_STRIDE) out[(size_t)b * DIAG_STRIDE + threadIdx.x] = sh[threadIdx.x][0];
}

// ---------------------------------------------------------------------------------
// Passive scalars (passive_mod.f90): velocities, upwind face values, advective update.

// copy_velocity (passive_mod.f90:204-244)
template <class real>
__global__ __launch_bounds__(256) void k_copy_velocity(const BoxDesc* __restrict__ boxes, Plan plan, VelArgs<real> a) {
    int b, i, j, k;
    if (!plan_item(plan, b, i, j, k)) return;
    long long c = cell_index(boxes[b], i, j, k);
    real irho = real(1) / a.q[0][c];
    if (!a.stage) {
        for (int d = 0; d < 3; d++) a.vel[d][c] = a.q[1 + d][c] * irho;
    } else {
        for (int d = 0; d < 3; d++) a.vta[d][c] = real(0.5) * (a.vel[d][c] + a.q[1 + d][c] * irho);
    }
}

// face value by upwinding on the mean adjacent cell velocity (passive_mod.f90:594-627);
// corrector stage reconstructs with minmod PLM (passive_mod.f90:510-557)
template <class real>
__global__ __launch_bounds__(256) void k_pass_flux(const BoxDesc* __restrict__ boxes, Plan plan, PassFluxArgs<real> a) {
    int b, i, j, k;
    if (!plan_item(plan, b, i, j, k)) return;
    const BoxDesc& B = boxes[b];
    long long c = cell_index(B, i, j, k);
    long long s = a.dir == 0 ? 1LL : (a.dir == 1 ? (long long)B.P[0] : (long long)B.P[0] * B.P[1]);
    real vavg = real(0.5) * (a.v[c] + a.v[c + s]);
    real theta = vavg >= real(0) ? real(1) : real(-1);
    for (int p = 0; p < a.npass; p++) {
        const real* w = a.s[p];
        real uL, uR;
        if (!a.stage) {
            uL = w[c];
            uR = w[c + s];
        } else {
            real wm = w[c - s], w0 = w[c], w1 = w[c + s], w2 = w[c + 2 * s];
            real dL = w0 - wm, dR = w1 - w0;
            real d0 = real(0.5) * (sgn1(dL) + sgn1(dR)) * min2(fabs(dL), fabs(dR));
            dL = w1 - w0; dR = w2 - w1;
            real d1 = real(0.5) * (sgn1(dL) + sgn1(dR)) * min2(fabs(dL), fabs(dR));
            uL = w0 + real(0.5) * d0;
            uR = w1 - real(0.5) * d1;
        }
        a.F[p][c] = real(0.5) * ((real(1) + theta) * uL + (real(1) - theta) * uR);
    }
}

// update_passive_state_3d (passive_mod.f90:562-590): advective, not conservative
template <class real>
__global__ __launch_bounds__(256) void k_pass_update(const BoxDesc* __restrict__ boxes, Plan plan, PassUpdArgs<real> a) {
    int b, i, j, k;
    if (!plan_item(plan, b, i, j, k)) return;
    const BoxDesc& B = boxes[b];
    long long c = cell_index(B, i, j, k);
    const long long sx = 1, sy = B.P[0], sz = (long long)B.P[0] * B.P[1];
    real dtdx = (real)(a.factor * a.dt[b] / B.dx);
    real vx = a.v[0][c], vy = a.v[1][c], vz = a.v[2][c];
    for (int p = 0; p < a.npass; p++) {
        const real* F0 = a.F[0][p];
        const real* F1 = a.F[1][p];
        const real* F2 = a.F[2][p];
        a.dst[p][c] = a.src[p][c] - dtdx * (vx * (F0[c] - F0[c - sx]) + vy * (F1[c] - F1[c - sy]) + vz * (F2[c] - F2[c - sz]));
    }
}

// 1/rho of the velocity reads: IEEE division in FP64 (bitwise with the reference), the
// hardware reciprocal in the FP32 path
template <class real>
__device__ __forceinline__ real pass_inv(real r) {
    if constexpr (sizeof(real) == 4) return __builtin_amdgcn_rcpf(r);
    else return real(1) / r;
}

template <class real>
__device__ __forceinline__ real pass_vel(const PassStageArgs<real>& a, int d, long long p) {
    real irho = pass_inv(a.qo[0][p]);
    real v = a.qo[1 + d][p] * irho;   // copy_velocity(.FALSE.)
    if (a.stage) {
        real irn = pass_inv(a.qn[0][p]);
        v = real(0.5) * (v + a.qn[1 + d][p] * irn);  // copy_velocity(.TRUE.)
    }
    return v;
}

template <class real>
__device__ __forceinline__ real minmod_slope(real wm, real w0, real wp) {
    real dL = w0 - wm, dR = wp - w0;
    return real(0.5) * (sgn1(dL) + sgn1(dR)) * min2(fabs(dL), fabs(dR));
}

// face value at the face between p and p + s (passive_mod.f90:510-557, 594-627)
template <class real>
__device__ __forceinline__ real pass_face(const real* w, long long p, long long s, real vavg, int stage) {
    real uL, uR;
    if (!stage) {
        uL = w[p];
        uR = w[p + s];
    } else {
        real wm = w[p - s], w0 = w[p], w1 = w[p + s], w2 = w[p + 2 * s];
        uL = w0 + real(0.5) * minmod_slope(wm, w0, w1);
        uR = w1 - real(0.5) * minmod_slope(w0, w1, w2);
    }
    real theta = vavg >= real(0) ? real(1) : real(-1);
    return real(0.5) * ((real(1) + theta) * uL + (real(1) - theta) * uR);
}

template <class real>
__global__ __launch_bounds__(256) void k_pass_stage(const BoxDesc* __restrict__ boxes, Plan plan, PassStageArgs<real> a) {
    int b, i, j, k;
    if (!plan_item(plan, b, i, j, k)) return;
    const BoxDesc& B = boxes[b];
    const long long c = cell_index(B, i, j, k);
    const long long st[3] = {1, (long long)B.P[0], (long long)B.P[0] * B.P[1]};
    real dtdx = (real)(a.factor * a.dt[b] / B.dx);
    real vc[3], va_hi[3], va_lo[3];
    for (int d = 0; d < 3; d++) {
        vc[d] = pass_vel(a, d, c);
        va_hi[d] = real(0.5) * (vc[d] + pass_vel(a, d, c + st[d]));
        va_lo[d] = real(0.5) * (pass_vel(a, d, c - st[d]) + vc[d]);
    }
    for (int p = 0; p < a.npass; p++) {
        const real* w = a.rec[p];
        real dF[3];
        for (int d = 0; d < 3; d++)
            dF[d] = pass_face(w, c, st[d], va_hi[d], a.stage) - pass_face(w, c - st[d], st[d], va_lo[d], a.stage);
        a.dst[p][c] = a.src[p][c] - dtdx * (vc[0] * dF[0] + vc[1] * dF[1] + vc[2] * dF[2]);
    }
}

// z-marching form of k_pass_stage: one thread walks a z-segment of an (i, j) column.  At
// the z-neighbours only v_z is needed, and the lower z-face value of cell k is the upper
// one of cell k-1 (same operands in the same order, so the same bits); both are carried
// in registers, with the scalar's z-window, so each plane of rho, m and the scalar is
// loaded once per segment instead of once per neighbour.
template <class real, int STAGE>
__device__ __forceinline__ real pass_face_v(real wm, real w0, real w1, real w2, real vavg) {
    real uL, uR;
    if (!STAGE) {
        uL = w0;
        uR = w1;
    } else {
        uL = w0 + real(0.5) * minmod_slope(wm, w0, w1);
        uR = w1 - real(0.5) * minmod_slope(w0, w1, w2);
    }
    real theta = vavg >= real(0) ? real(1) : real(-1);
    return real(0.5) * ((real(1) + theta) * uL + (real(1) - theta) * uR);
}

template <class real, int STAGE, bool VARR>
__global__ __launch_bounds__(256) void k_pass_stage_z(const BoxDesc* __restrict__ boxes, Plan plan, PassStageArgs<real> a) {
    // velocities: read from a.vel (formed by the fused MHD stages) or formed here from rho, m
    auto V = [&](int d, long long p) -> real { return VARR ? a.vel[d][p] : pass_vel(a, d, p); };
    int b, i, j, sg;
    if (!plan_item(plan, b, i, j, sg)) return;
    const BoxDesc& B = boxes[b];
    const PlanBox& pb = plan.boxes[b];
    const int kb = pb.zlo + sg * pb.seg, ke = min(kb + pb.seg, pb.zlo + pb.zn);
    const long long sy = B.P[0], sz = (long long)B.P[0] * B.P[1];
    const real dtdx = (real)(a.factor * a.dt[b] / B.dx);
    const long long c0 = cell_index(B, i, j, kb);
    if constexpr (VARR) {
        // velocities are loads: one passive at a time keeps the fewest live registers
        for (int p = 0; p < a.npass; p++) {
            const real* w = a.rec[p];
            long long c = c0;
            // carried: v_z at k, the scalar at k-1, k, k+1 (k, k+1 for the predictor), and the
            // value at the face between k-1 and k
            real vz1 = V(2, c);
            real flo, wa, wb, wc;
            {
                const real vz0 = V(2, c - sz);
                const real va = real(0.5) * (vz0 + vz1);
                if (STAGE) {
                    const real wmm = w[c - 2 * sz];
                    wa = w[c - sz]; wb = w[c]; wc = w[c + sz];
                    flo = pass_face_v<real, 1>(wmm, wa, wb, wc, va);
                } else {
                    wa = w[c - sz]; wb = w[c]; wc = wb;
                    flo = pass_face_v<real, 0>(wa, wa, wb, wb, va);
                }
            }
    #pragma unroll 1
            for (int k = kb; k < ke; k++, c += sz) {
                // one live cell index instead of a strength-reduced pointer per array and offset
                __asm__ volatile("" : "+v"(c));
                const real vx = V(0, c), vy = V(1, c);
                const real vz2 = V(2, c + sz);
                const real vxh = real(0.5) * (vx + V(0, c + 1));
                const real vxl = real(0.5) * (V(0, c - 1) + vx);
                const real vyh = real(0.5) * (vy + V(1, c + sy));
                const real vyl = real(0.5) * (V(1, c - sy) + vy);
                const real vzh = real(0.5) * (vz1 + vz2);
                real fhi;
                if (STAGE) {
                    const real wd = w[c + 2 * sz];
                    fhi = pass_face_v<real, 1>(wa, wb, wc, wd, vzh);
                    wa = wb; wb = wc; wc = wd;
                } else {
                    const real wn = w[c + sz];
                    fhi = pass_face_v<real, 0>(wb, wb, wn, wn, vzh);
                    wb = wn;
                }
                const real dF0 = pass_face(w, c, (long long)1, vxh, STAGE) - pass_face(w, c - 1, (long long)1, vxl, STAGE);
                const real dF1 = pass_face(w, c, sy, vyh, STAGE) - pass_face(w, c - sy, sy, vyl, STAGE);
                const real dF2 = fhi - flo;
                a.dst[p][c] = a.src[p][c] - dtdx * (vx * dF0 + vy * dF1 + vz1 * dF2);
                flo = fhi;
                vz1 = vz2;
            }
        }
    } else {
        // The velocities of a cell are formed once for all passives.  Carried per passive: the
        // scalar at k-1, k, k+1 (k, k+1 for the predictor) and the value at the face between
        // k-1 and k; v_z at k.
        real vz1 = V(2, c0);
        real flo[MAXPASS], wa[MAXPASS], wb[MAXPASS], wc[MAXPASS];
        {
            const real vz0 = V(2, c0 - sz);
            const real va = real(0.5) * (vz0 + vz1);
    #pragma unroll
            for (int p = 0; p < MAXPASS; p++) {
                if (p >= a.npass) break;
                const real* w = a.rec[p];
                if (STAGE) {
                    const real wmm = w[c0 - 2 * sz];
                    wa[p] = w[c0 - sz]; wb[p] = w[c0]; wc[p] = w[c0 + sz];
                    flo[p] = pass_face_v<real, 1>(wmm, wa[p], wb[p], wc[p], va);
                } else {
                    wa[p] = w[c0 - sz]; wb[p] = w[c0]; wc[p] = wb[p];
                    flo[p] = pass_face_v<real, 0>(wa[p], wa[p], wb[p], wb[p], va);
                }
            }
        }
        long long c = c0;
    #pragma unroll 1
        for (int k = kb; k < ke; k++, c += sz) {
            // one live cell index instead of a strength-reduced pointer per array and offset
            __asm__ volatile("" : "+v"(c));
            const real vx = V(0, c), vy = V(1, c);
            const real vz2 = V(2, c + sz);
            const real vxh = real(0.5) * (vx + V(0, c + 1));
            const real vxl = real(0.5) * (V(0, c - 1) + vx);
            const real vyh = real(0.5) * (vy + V(1, c + sy));
            const real vyl = real(0.5) * (V(1, c - sy) + vy);
            const real vzh = real(0.5) * (vz1 + vz2);
    #pragma unroll
            for (int p = 0; p < MAXPASS; p++) {
                if (p >= a.npass) break;
                const real* w = a.rec[p];
                real fhi;
                if (STAGE) {
                    const real wd = w[c + 2 * sz];
                    fhi = pass_face_v<real, 1>(wa[p], wb[p], wc[p], wd, vzh);
                    wa[p] = wb[p]; wb[p] = wc[p]; wc[p] = wd;
                } else {
                    const real wn = w[c + sz];
                    fhi = pass_face_v<real, 0>(wb[p], wb[p], wn, wn, vzh);
                    wb[p] = wn;
                }
                const real dF0 = pass_face(w, c, (long long)1, vxh, STAGE) - pass_face(w, c - 1, (long long)1, vxl, STAGE);
                const real dF1 = pass_face(w, c, sy, vyh, STAGE) - pass_face(w, c - sy, sy, vyl, STAGE);
                const real dF2 = fhi - flo[p];
                a.dst[p][c] = a.src[p][c] - dtdx * (vx * dF0 + vy * dF1 + vz1 * dF2);
                flo[p] = fhi;
            }
            vz1 = vz2;
        }
    }
}

// ---------------------------------------------------------------------------------
// host <-> device layout and precision conversion (reference FP64 arrays)
template <class real>
__global__ void k_convert_in(const double* __restrict__ src, real* __restrict__ dst, long long n) {
    long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) dst[g] = (real)src[g];
}
template <class real>
__global__ void k_convert_out(const real* __restrict__ src, double* __restrict__ dst, long long n) {
    long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) dst[g] = (double)src[g];
}
// passive arrays: the reference's bounds are [3-nb, n+nb-2] per axis; ours are the full
// ghost-padded box.  dir 0: host-layout src -> device; dir 1: device -> host layout.
// one chunk [g0, g0 + m) of the host-layout index; hsrc / hdst hold that chunk only
template <class real>
__global__ void k_pass_layout(BoxDesc B, const double* hsrc, double* hdst, real* dev, int to_device, long long g0,
                              long long m) {
    const long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= m) return;
    const long long g = g0 + l;
    int PX = B.n[0] + 2 * B.nb - 4, PY = B.n[1] + 2 * B.nb - 4;
    int ii = (int)(g % PX);
    long long t = g / PX;
    int jj = (int)(t % PY);
    int kk = (int)(t / PY);
    long long c = cell_index(B, ii + 3 - B.nb, jj + 3 - B.nb, kk + 3 - B.nb);
    if (to_device) dev[c] = (real)hsrc[l];
    else hdst[l] = (double)dev[c];
}

// Snapshot packing: each workgroup reads 256 consecutive interior cells field by field
// (coalesced), transposes them through LDS and writes their 256 x (7 + npo) values as one
// contiguous, coalesced run.  double -> float is the round-to-nearest conversion of SNGL.
template <class real, class out_t>
__global__ __launch_bounds__(256) void k_pack_cube(BoxDesc B, CubeArgs<real> a, out_t* __restrict__ out) {
    __shared__ out_t s[(7 + MAXPASS) * 256];
    const int V = 7 + a.npo;
    const long long ncell = (long long)B.n[0] * B.n[1] * a.nk;
    const long long base = (long long)blockIdx.x * 256;
    const long long g = base + threadIdx.x;
    if (g < ncell) {
        const int i = (int)(g % B.n[0]);
        const long long t = g / B.n[0];
        const int j = (int)(t % B.n[1]);
        const int k = a.k0 + (int)(t / B.n[1]);
        const long long c = cell_index(B, i + 1, j + 1, k + 1);
        out_t* row = s + threadIdx.x * V;
        for (int v = 0; v < 7; v++) row[v] = (out_t)a.q[v][c];
        for (int p = 0; p < a.npo; p++) row[7 + p] = (out_t)a.qp[p][c];
    }
    __syncthreads();
    const long long left = ncell - base;
    const int nv = (int)(left < 256 ? left : 256) * V;
    for (int e = threadIdx.x; e < nv; e += 256) out[base * V + e] = s[e];
}

// Flux protection (vanleer_mod.f90:2084-2200), whole padded box.  The first-pass update is
// update_cell's expression without the floor (final_check_unphysical runs after the
// recompute), in the context's precision (FP32: the FP32 update's dt/dx and sums).
template <class real>
__global__ __launch_bounds__(256) void k_prot_flag(const BoxDesc* __restrict__ boxes, Plan plan, ProtArgs<real> a) {
    int b, i, j, k;
    if (!plan_item(plan, b, i, j, k)) return;
    const BoxDesc& B = boxes[b];
    const long long c = cell_index(B, i, j, k);
    const int lo = a.lo;
    const bool in = i >= lo && j >= lo && k >= lo && i <= B.n[0] + a.hi_off && j <= B.n[1] + a.hi_off &&
                    k <= B.n[2] + a.hi_off;
    unsigned char f = 0;
    if (in) {
        const long long sx = 1, sy = B.P[0], sz = (long long)B.P[0] * B.P[1];
        const real dtdx = (real)(a.factor * a.dt[b] / B.dx);
        const real div = (a.F[0][0][c] - a.F[0][0][c - sx]) + (a.F[1][0][c] - a.F[1][0][c - sy]) +
                         (a.F[2][0][c] - a.F[2][0][c - sz]);
        f = (a.src[0][c] - dtdx * div) < (real)a.thr ? 1 : 0;
    }
    a.flag[c] = f;
}

template <class real>
__global__ __launch_bounds__(256) void k_prot_fix(const BoxDesc* __restrict__ boxes, Plan plan, ProtArgs<real> a) {
    int b, i, j, k;
    if (!plan_item(plan, b, i, j, k)) return;
    const BoxDesc& B = boxes[b];
    if (i >= B.n[0] + B.nb || j >= B.n[1] + B.nb || k >= B.n[2] + B.nb) return;   // m + 1 stays in the box
    const long long c = cell_index(B, i, j, k);
    const long long s[3] = {1, B.P[0], (long long)B.P[0] * B.P[1]};
    const bool f0 = a.flag[c] != 0;
    bool any = f0;
    for (int d = 0; d < 3; d++) {
        const bool fn = a.flag[c + s[d]] != 0;
        any = any || fn;
        if (f0 || fn)   // face m + 1/2 along d borders a marked cell
            for (int q = 0; q < 4; q++) a.F[d][q][c] = a.Fp[d][q][c];
    }
    if (any)
        for (int d = 0; d < 3; d++) a.Ec[d][c] = a.Ecp[d][c];
}

template <class real>
void launch_prot_flag(const Launch& L, const ProtArgs<real>& a) {
    if (L.plan.total > 0)
        hipLaunchKernelGGL(k_prot_flag<real>, dim3(L.blocks()), dim3(256), 0, L.stream, L.boxes, L.plan, a);
}
template <class real>
void launch_prot_fix(const Launch& L, const ProtArgs<real>& a) {
    if (L.plan.total > 0)
        hipLaunchKernelGGL(k_prot_fix<real>, dim3(L.blocks()), dim3(256), 0, L.stream, L.boxes, L.plan, a);
}
template void launch_prot_flag<double>(const Launch&, const ProtArgs<double>&);
template void launch_prot_flag<float>(const Launch&, const ProtArgs<float>&);
template void launch_prot_fix<double>(const Launch&, const ProtArgs<double>&);
template void launch_prot_fix<float>(const Launch&, const ProtArgs<float>&);

// ---------------------------------------------------------------------------------
#define L256(k, ...) hipLaunchKernelGGL(k, dim3(L.blocks()), dim3(256), 0, L.stream, __VA_ARGS__)

template <class real>
void launch_fill_ghosts(const Launch& L, const GhostArgs<real>& a) {
    if (L.plan.total > 0) L256((k_fill_ghosts<real>), L.boxes, L.plan, a);
}
template <class real>
void launch_fill_ghost_band(const Launch& L, const GhostArgs<real>& a) {
    if (L.plan.total > 0) L256((k_fill_ghost_band<real>), L.boxes, L.plan, a);
}
template <class real>
void launch_restore_shell(const Launch& L, const ShellArgs<real>& a) {
    if (L.plan.total > 0) L256((k_restore_shell<real>), L.boxes, L.plan, a);
}
template <class real>
void launch_pass_stage(const Launch& L, const PassStageArgs<real>& a) {
    if (L.plan.total <= 0) return;
    // the per-cell form forms its velocities itself (a.vel only caches them)
    if (!L.column) L256((k_pass_stage<real>), L.boxes, L.plan, a);
    else if (a.vel[0] != nullptr) {
        if (a.stage) L256((k_pass_stage_z<real, 1, true>), L.boxes, L.plan, a);
        else L256((k_pass_stage_z<real, 0, true>), L.boxes, L.plan, a);
    } else if (a.stage) L256((k_pass_stage_z<real, 1, false>), L.boxes, L.plan, a);
    else L256((k_pass_stage_z<real, 0, false>), L.boxes, L.plan, a);
}
template <class real>
void launch_max_speed(const Launch& L, const QPtrs<real>& q, double cs, unsigned long long* cmax) {
    if (L.plan.total > 0) L256((k_max_speed<real>), L.boxes, L.plan, q, cs, cmax);
}
template <class real>
void launch_diag(hipStream_t st, const BoxDesc* boxes, int nbox, const DiagArgs<real>& a) {
    hipLaunchKernelGGL((k_diag_part<real>), dim3(DIAG_WG, nbox), dim3(256), 0, st, boxes, a);
    hipLaunchKernelGGL(k_diag_final, dim3(nbox), dim3(256), 0, st, (const double*)a.part, a.out);
}
// Device scratch of imhd_diagnostics_static, the entry point that allocates nothing: the
// partial rows (DIAG_WG x DIAG_STRIDE doubles per box) and the results (DIAG_STRIDE per
// box) of up to DIAG_STATIC_MAXBOX boxes, 4.2 MB.  Every device holds its own copy, for
// the life of the process; calls on one device share it.  Host code never passes the
// array's name to a kernel: it is a host-side stand-in, not a device address.
static __device__ __align__(256) double scratch[(size_t)DIAG_STATIC_MAXBOX * (DIAG_WG + 1) * DIAG_STRIDE];

// The device address of the current device's copy of `scratch`, looked up once per
// device (with that device current) and kept.
hipError_t diag_static_scratch(double** out) {
    static std::mutex mu;
    static double* addr[DIAG_STATIC_MAXDEV] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= DIAG_STATIC_MAXDEV) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[dev] = static_cast<double*>(p);
    }
    *out = addr[dev];
    return hipSuccess;
}

template <class real>
void launch_copy_velocity(const Launch& L, const VelArgs<real>& a) {
    if (L.plan.total > 0) L256((k_copy_velocity<real>), L.boxes, L.plan, a);
}
template <class real>
void launch_pass_flux(const Launch& L, const PassFluxArgs<real>& a) {
    if (L.plan.total > 0) L256((k_pass_flux<real>), L.boxes, L.plan, a);
}
template <class real>
void launch_pass_update(const Launch& L, const PassUpdArgs<real>& a) {
    if (L.plan.total > 0) L256((k_pass_update<real>), L.boxes, L.plan, a);
}
void launch_new_dt(hipStream_t st, int nbox, const BoxDesc* boxes, const unsigned long long* cmax,
                   const unsigned long long* smax, const unsigned long long* dvmax, const double* cour, double t_end,
                   double* t, double* dt, double* dt_prev, int advance_time) {
    hipLaunchKernelGGL(k_new_dt, dim3((nbox + 63) / 64), dim3(64), 0, st, nbox, boxes, cmax, smax, dvmax, cour, t_end,
                       t, dt, dt_prev, advance_time);
}
void launch_reset_u64(hipStream_t st, unsigned long long* p, int n) {
    hipLaunchKernelGGL(k_reset_u64, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
}
template <class real>
void launch_convert_in(hipStream_t st, const double* src, real* dst, long long n) {
    hipLaunchKernelGGL((k_convert_in<real>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, dst, n);
}
template <class real>
void launch_convert_out(hipStream_t st, const real* src, double* dst, long long n) {
    hipLaunchKernelGGL((k_convert_out<real>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, dst, n);
}
template <class real>
void launch_pass_layout(hipStream_t st, const BoxDesc& B, const double* hsrc, double* hdst, real* dev, int to_device,
                        long long g0, long long m) {
    if (m <= 0) return;
    hipLaunchKernelGGL((k_pass_layout<real>), dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, B, hsrc, hdst, dev,
                       to_device, g0, m);
}

template <class real, class out_t>
void launch_pack_cube(hipStream_t st, const BoxDesc& B, const CubeArgs<real>& a, out_t* out) {
    const long long ncell = (long long)B.n[0] * B.n[1] * a.nk;
    if (ncell <= 0) return;
    hipLaunchKernelGGL((k_pack_cube<real, out_t>), dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, st, B, a,
                       out);
}

#define INST_AUX(T)                                                                                           \
    template void launch_fill_ghosts<T>(const Launch&, const GhostArgs<T>&);                                  \
    template void launch_fill_ghost_band<T>(const Launch&, const GhostArgs<T>&);                              \
    template void launch_restore_shell<T>(const Launch&, const ShellArgs<T>&);                                \
    template void launch_pass_stage<T>(const Launch&, const PassStageArgs<T>&);                               \
    template void launch_max_speed<T>(const Launch&, const QPtrs<T>&, double, unsigned long long*);                 \
    template void launch_diag<T>(hipStream_t, const BoxDesc*, int, const DiagArgs<T>&);                      \
    template void launch_copy_velocity<T>(const Launch&, const VelArgs<T>&);                                  \
    template void launch_pass_flux<T>(const Launch&, const PassFluxArgs<T>&);                                 \
    template void launch_pass_update<T>(const Launch&, const PassUpdArgs<T>&);                                \
    template void launch_convert_in<T>(hipStream_t, const double*, T*, long long);                            \
    template void launch_convert_out<T>(hipStream_t, const T*, double*, long long);                           \
    template void launch_pass_layout<T>(hipStream_t, const BoxDesc&, const double*, double*, T*, int, long long, long long); \
    template void launch_pack_cube<T, float>(hipStream_t, const BoxDesc&, const CubeArgs<T>&, float*);         \
    template void launch_pack_cube<T, double>(hipStream_t, const BoxDesc&, const CubeArgs<T>&, double*);
INST_AUX(double)
INST_AUX(float)

}  // namespace imhd
