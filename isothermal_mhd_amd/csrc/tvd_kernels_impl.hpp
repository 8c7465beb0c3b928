// K98 Roe-TVD / HLLE + CTU + CT cell-update kernels (Kim et al. 1998 TVD fluxes with
// corner-transport-upwind half updates and constrained transport), following
// isotvd_mod.f90 of the reference.  Instantiated per precision by tvd_kernels_f64.hip (no
// FMA contraction: bitwise with the reference) and tvd_kernels_f32.hip.
//
// Per step:   3 x (eig -> flux) predictor sweeps -> corner EMF -> CT half step and the
//             corrector centre EMF -> 3 x (CTU half update -> eig -> flux) corrector
//             sweeps -> corner EMF -> final update (hydro, CT faces, cell-centred B).
// A Roe-TVD interface flux needs the flux limiters of its two neighbouring interfaces:
// k_tvd_eig writes g~_k and alpha_k per face, k_tvd_flux combines three faces' values and
// recomputes the right eigenvectors of its own face column by column (no 6x6 matrix is
// held in registers).
#pragma once
#include "riemann.hpp"
#include "tvd_kernels.hpp"

namespace imhd {

__device__ __forceinline__ long long tvd_stride(const BoxDesc& B, int d) {
    return d == 0 ? 1LL : (d == 1 ? (long long)B.P[0] : (long long)B.P[0] * (long long)B.P[1]);
}

// eps(k) of the entropy fix: single-precision literals widened (isotvd_mod.f90:108)
__device__ __forceinline__ double tvd_eps(int k) { return (k == 1 || k == 4) ? 0.0 : (double)0.2f; }

// Roe (arithmetic) average, wave speeds, renormalisation factors (isotvd_mod.f90:1856-2020)
template <class real>
struct TvdWave {
    real qa[7];
    real cf, ca, cslow, af, as, by, bz, sbx, signbt, isrho, amc;
    real ev[6];
};

template <class real>
__device__ __forceinline__ void tvd_wave(const real* a, const real* b, real cs, TvdWave<real>& W) {
    const real asqr = cs * cs;
    real rho2 = real(0.5) * (a[0] + b[0]);
    real isrho = fdiv(real(1), fsqrt(rho2));
    W.qa[0] = rho2;
    W.qa[1] = real(0.5) * (fdiv(a[1], a[0]) + fdiv(b[1], b[0]));
    W.qa[2] = real(0.5) * (fdiv(a[2], a[0]) + fdiv(b[2], b[0]));
    W.qa[3] = real(0.5) * (fdiv(a[3], a[0]) + fdiv(b[3], b[0]));
    W.qa[4] = a[4] * isrho;
    W.qa[5] = real(0.5) * (a[5] + b[5]) * isrho;
    W.qa[6] = real(0.5) * (a[6] + b[6]) * isrho;
    W.isrho = isrho;
    const real vx2 = W.qa[1], bx = W.qa[4];
    real bsqr = W.qa[4] * W.qa[4] + W.qa[5] * W.qa[5] + W.qa[6] * W.qa[6];
    real btsqr = W.qa[5] * W.qa[5] + W.qa[6] * W.qa[6];
    real tmp1 = asqr + bsqr, x = asqr - bsqr;
    real tmp2 = fsqrt(x * x + real(4) * asqr * btsqr);
    W.cslow = fsqrt(max2(real(0), real(0.5) * (tmp1 - tmp2)));
    W.ca = fabs(bx);
    W.cf = fsqrt(real(0.5) * (tmp1 + tmp2));
    W.ev[0] = vx2 - W.cf;
    W.ev[1] = vx2 - W.ca;
    W.ev[2] = vx2 - W.cslow;
    W.ev[3] = vx2 + W.cslow;
    W.ev[4] = vx2 + W.ca;
    W.ev[5] = vx2 + W.cf;
    real bt = fsqrt(W.qa[5] * W.qa[5] + W.qa[6] * W.qa[6]);
    const real cf = W.cf, ca = W.ca, csl = W.cslow;
    // degeneracy rule of the TVD module ("MODIFIED V2", isotvd_mod.f90:1999-2009)
    if ((bt < (real)ZERO && fabs(ca * ca - asqr) < (real)ZERO) || (fabs(cf * cf - csl * csl) < (real)ZERO)) {
        W.af = real(1);
        W.as = real(1);
    } else {
        real t = fdiv(real(1), fsqrt(max2(real(0), cf * cf - csl * csl)));
        W.af = fsqrt(max2(real(0), cf * cf - ca * ca)) * t;
        W.as = fsqrt(max2(real(0), cf * cf - asqr)) * t;
    }
    if (bt < (real)ZERO) {
        W.by = fdiv(real(1), fsqrt(real(2)));
        W.bz = W.by;
    } else {
        real t = fdiv(real(1), bt);
        W.by = W.qa[5] * t;
        W.bz = W.qa[6] * t;
    }
    W.signbt = (fabs(W.qa[5]) >= (real)ZERO) ? sgn1(W.qa[5]) : sgn1(W.qa[6]);
    W.amc = cs * cs - ca * ca;
    W.sbx = sgn1(W.qa[4]);
}

// characteristic amplitudes alpha_k = L_k . (q_R - q_L) (isotvd_mod.f90:2163-2333)
template <class real>
__device__ __forceinline__ void tvd_ck(const real* a, const real* b, real cs, const TvdWave<real>& W, real* ck) {
    const real* qa = W.qa;
    const real af = W.af, as = W.as, cf = W.cf, ca = W.ca, csl = W.cslow, by = W.by, bz = W.bz, sbx = W.sbx;
    real dq[6] = {b[0] - a[0], b[1] - a[1], b[2] - a[2], b[3] - a[3], b[5] - a[5], b[6] - a[6]};
    real ith1 = fdiv(real(1), real(2) * (af * af * (cs * cs) + as * as * (cf * cf)));
    real ith2 = fdiv(real(1), real(2) * (af * af * cs * cf + as * as * csl * ca));
    real sq = fsqrt(qa[0]);
    real Lm[6], Lp[6];
    auto dot = [&](const real* L) {
        return dq[0] * L[0] + dq[1] * L[1] + dq[2] * L[2] + dq[3] * L[3] + dq[4] * L[4] + dq[5] * L[5];
    };
    {
        real t1 = ith1 * af * (cs * cs);
        real t2 = ith2 * (-(af * cs * qa[1]) + as * csl * sbx * (by * qa[2] + bz * qa[3]));
        real t3 = ith2 * as * csl * sbx;
        real t4 = ith1 * as * cf * sq;
        Lm[0] = t1 - t2; Lp[0] = t1 + t2;
        Lm[1] = -(ith2 * af * cs); Lp[1] = -Lm[1];
        Lm[2] = t3 * by; Lp[2] = -Lm[2];
        Lm[3] = t3 * bz; Lp[3] = -Lm[3];
        Lm[4] = t4 * by; Lp[4] = Lm[4];
        Lm[5] = t4 * bz; Lp[5] = Lm[5];
        if (W.amc < (real)ZERO)
            for (int m = 0; m < 6; m++) { Lm[m] = W.signbt * Lm[m]; Lp[m] = W.signbt * Lp[m]; }
        ck[0] = dot(Lm);
        ck[5] = dot(Lp);
    }
    {
        real t1 = real(0.5) * (bz * qa[2] - by * qa[3]) * sbx, t2 = real(0.5) * sbx, t3 = real(0.5) * sq;
        Lm[0] = -t1; Lp[0] = t1;
        Lm[1] = real(0); Lp[1] = real(0);
        Lm[2] = t2 * bz; Lp[2] = -Lm[2];
        Lm[3] = -(t2 * by); Lp[3] = -Lm[3];
        Lm[4] = t3 * bz; Lp[4] = Lm[4];
        Lm[5] = -(t3 * by); Lp[5] = Lm[5];
        ck[1] = dot(Lm);
        ck[4] = dot(Lp);
    }
    {
        real t1 = ith1 * as * (cf * cf);
        real t2 = ith2 * (as * ca * qa[1] + af * cf * sbx * (by * qa[2] + bz * qa[3]));
        real t3 = ith2 * af * cf * sbx;
        real t4 = ith1 * af * cf * sq;
        Lm[0] = t1 + t2; Lp[0] = t1 - t2;
        Lm[1] = -(ith2 * as * ca); Lp[1] = -Lm[1];
        Lm[2] = -(t3 * by); Lp[2] = -Lm[2];
        Lm[3] = -(t3 * bz); Lp[3] = -Lm[3];
        Lm[4] = -(t4 * by); Lp[4] = Lm[4];
        Lm[5] = -(t4 * bz); Lp[5] = Lm[5];
        if (W.amc > (real)ZERO)
            for (int m = 0; m < 6; m++) { Lm[m] = W.signbt * Lm[m]; Lp[m] = W.signbt * Lp[m]; }
        ck[2] = dot(Lm);
        ck[3] = dot(Lp);
    }
}

// column k of the right eigenvector matrix (isotvd_mod.f90:2026-2141)
template <class real>
__device__ __forceinline__ void tvd_rcol(int k, real cs, const TvdWave<real>& W, real* r) {
    const real* qa = W.qa;
    if (k == 0 || k == 5) {
        real t1 = W.af * qa[2], t2 = W.as * W.by * qa[4], t3 = W.af * qa[3], t4 = W.as * W.bz * qa[4];
        real t5 = W.as * W.cf * W.isrho;
        r[0] = W.af;
        r[1] = W.af * W.ev[k];
        r[2] = k == 0 ? t1 + t2 : t1 - t2;
        r[3] = k == 0 ? t3 + t4 : t3 - t4;
        r[4] = t5 * W.by;
        r[5] = t5 * W.bz;
        if (W.amc < (real)ZERO)
            for (int m = 0; m < 6; m++) r[m] = W.signbt * r[m];
    } else if (k == 1 || k == 4) {
        real s = k == 1 ? W.sbx : -W.sbx;
        r[0] = real(0);
        r[1] = real(0);
        r[2] = s * W.bz;
        r[3] = k == 1 ? -(W.sbx * W.by) : W.sbx * W.by;
        r[4] = W.isrho * W.bz;
        r[5] = -(W.isrho * W.by);
    } else {
        real s1 = W.as * qa[2], s2 = W.af * W.by * cs * W.sbx, s3 = W.as * qa[3], s4 = W.af * W.bz * cs * W.sbx;
        real s5 = fdiv(W.af * (cs * cs) * W.isrho, W.cf);
        r[0] = W.as;
        r[1] = W.as * W.ev[k];
        r[2] = k == 2 ? s1 - s2 : s1 + s2;
        r[3] = k == 2 ? s3 - s4 : s3 + s4;
        r[4] = -(s5 * W.by);
        r[5] = -(s5 * W.bz);
        if (W.amc > (real)ZERO)
            for (int m = 0; m < 6; m++) r[m] = W.signbt * r[m];
    }
}

// K98 eqn. 2.31 entropy function
template <class real>
__device__ __forceinline__ real tvd_Q(real chi, real e) {
    if (fabs(chi) < real(2) * e) return real(0.25) * fdiv(chi * chi, e) + e;
    return fabs(chi);
}

// minmod limiter of the TVD module (isotvd_mod.f90:2426-2429)
template <class real>
__device__ __forceinline__ real tvd_lim(real g1, real g2) {
    real s = sgn1(g2);
    return s * max2(real(0), min2(fabs(g2), g1 * s));
}

// cell flux with the cell's own face field in slot 4 (isotvd_mod.f90:2364-2382)
template <class real>
__device__ __forceinline__ void tvd_cell_flux(const real* u, real cs, real* f) {
    real rho = u[0], vx = fdiv(u[1], rho), vy = fdiv(u[2], rho), vz = fdiv(u[3], rho);
    real Bx = u[4], By = u[5], Bz = u[6];
    f[0] = rho * vx;
    f[1] = rho * (vx * vx + cs * cs) + real(0.5) * (By * By + Bz * Bz - Bx * Bx);
    f[2] = rho * vx * vy - Bx * By;
    f[3] = rho * vx * vz - Bx * Bz;
    f[4] = vx * By - vy * Bx;
    f[5] = vx * Bz - vz * Bx;
}

template <class real>
__device__ __forceinline__ void tvd_load(const TvdSweepArgs<real>& a, long long c, real* u) {
#pragma unroll
    for (int r = 0; r < 7; r++)
        if (r != 4) u[r] = a.w[r][c];
    u[4] = (real)a.bn[c];
}

template <class real>
__device__ __forceinline__ void tvd_store(const TvdSweepArgs<real>& a, long long c, const real* F) {
    a.F[0][c] = F[0];
    a.F[1][c] = F[1];
    a.F[2][c] = F[2];
    a.F[3][c] = F[3];
    a.Et1[c] = -F[5];
    a.Et2[c] = F[6];
}

// per face: g~_k and alpha_k (isotvd_mod.f90:1856-2333, 2400-2424)
template <class real>
__device__ __forceinline__ void tvd_eig_face(const real* uL, const real* uR, real cs, real dtdx, real* gt, real* ck) {
    TvdWave<real> W;
    tvd_wave(uL, uR, cs, W);
    tvd_ck(uL, uR, cs, W, ck);
#pragma unroll
    for (int w = 0; w < 6; w++) {
        real chi = dtdx * W.ev[w];
        real Qk = tvd_Q(chi, (real)tvd_eps(w));
        gt[w] = real(0.5) * (Qk - chi * chi) * ck[w];
    }
}

// Roe-TVD flux of one face from its states, its alpha_k and the g~_k of the face below
// (gm), itself (g0) and above (gp) (isotvd_mod.f90:2400-2466); F[4] = 0.
template <class real>
__device__ __forceinline__ void tvd_flux_face(const real* uL, const real* uR, real cs, real dtdx, real dxdt,
                                              const real* gm, const real* g0, const real* gp, const real* ck, real* F) {
    TvdWave<real> W;
    tvd_wave(uL, uR, cs, W);
    real FT[6] = {real(0), real(0), real(0), real(0), real(0), real(0)};
#pragma unroll
    for (int w = 0; w < 6; w++) {
        const real ckw = ck[w];
        const real gk0 = tvd_lim(gm[w], g0[w]), gk1 = tvd_lim(g0[w], gp[w]);
        real gam = (fabs(ckw) > (real)ZERO) ? fdiv(gk1 - gk0, ckw) : real(0);
        real chi = dtdx * W.ev[w] + gam;
        real Qk = tvd_Q(chi, (real)tvd_eps(w));
        real betak = Qk * ckw - (gk0 + gk1);
        real r[6];
        tvd_rcol(w, cs, W, r);
#pragma unroll
        for (int m = 0; m < 6; m++) FT[m] = FT[m] - betak * r[m];
    }
    real fa[6], fb[6];
    tvd_cell_flux(uL, cs, fa);
    tvd_cell_flux(uR, cs, fb);
    F[0] = real(0.5) * (dxdt * FT[0] + fa[0] + fb[0]);
    F[1] = real(0.5) * (dxdt * FT[1] + fa[1] + fb[1]);
    F[2] = real(0.5) * (dxdt * FT[2] + fa[2] + fb[2]);
    F[3] = real(0.5) * (dxdt * FT[3] + fa[3] + fb[3]);
    F[4] = real(0);
    F[5] = real(0.5) * (dxdt * FT[4] + fa[4] + fb[4]);
    F[6] = real(0.5) * (dxdt * FT[5] + fa[5] + fb[5]);
}

template <class real>
__device__ __forceinline__ void tvd_dts(const TvdSweepArgs<real>& a, const BoxDesc& B, int b, real& dtdx, real& dxdt) {
    const double dtd = a.dt[b] / B.dx;
    dtdx = (real)(a.factor == 1.0 ? dtd : a.factor * dtd);
    dxdt = (real)(B.dx / a.dt[b]);
}

template <class real>
__global__ __launch_bounds__(256) void k_tvd_eig(const BoxDesc* __restrict__ boxes, Plan plan, TvdSweepArgs<real> a) {
    int b, i, j, k;
    if (!plan_item(plan, b, i, j, k)) return;
    const BoxDesc& B = boxes[b];
    const long long c = cell_index(B, i, j, k), s = tvd_stride(B, a.dir);
    real uL[7], uR[7], gt[6], ck[6], dtdx, dxdt;
    tvd_load(a, c, uL);
    tvd_load(a, c + s, uR);
    tvd_dts(a, B, b, dtdx, dxdt);
    tvd_eig_face(uL, uR, (real)a.cs, dtdx, gt, ck);
#pragma unroll
    for (int w = 0; w < 6; w++) {
        a.gt[w][c] = gt[w];
        a.ck[w][c] = ck[w];
    }
}

template <class real>
__global__ __launch_bounds__(256) void k_tvd_flux(const BoxDesc* __restrict__ boxes, Plan plan, TvdSweepArgs<real> a) {
    int b, i, j, k;
    if (!plan_item(plan, b, i, j, k)) return;
    const BoxDesc& B = boxes[b];
    const long long c = cell_index(B, i, j, k), s = tvd_stride(B, a.dir);
    real uL[7], uR[7], gm[6], g0[6], gp[6], ck[6], F[7], dtdx, dxdt;
    tvd_load(a, c, uL);
    tvd_load(a, c + s, uR);
    tvd_dts(a, B, b, dtdx, dxdt);
#pragma unroll
    for (int w = 0; w < 6; w++) {
        gm[w] = a.gt[w][c - s];
        g0[w] = a.gt[w][c];
        gp[w] = a.gt[w][c + s];
        ck[w] = a.ck[w][c];
    }
    tvd_flux_face(uL, uR, (real)a.cs, dtdx, dxdt, gm, g0, gp, ck, F);
    tvd_store(a, c, F);
}

// Fused Roe-TVD sweep, eigen-system and flux in one launch (g~_k and alpha_k never leave
// the chip).  The sweep's flux faces [lo, lo + r) per axis come from a face plan (build_plan
// cases 17 / 19); each face's flux needs g~_k of its two neighbours along the sweep.
//
// x sweep: one wave per (row, segment of TVD_XSEG = 62 flux faces).  Lane l forms g~_k
// and alpha_k of face x0 - 1 + l; lanes 1 .. 62 then take their neighbours' g~_k by wave
// shuffles and form the flux of their own face from the states they already hold.  No
// LDS and no barrier; every lane of the wave reaches the shuffles.
template <class real>
__global__ __launch_bounds__(256) void k_tvd_sweep_x(const BoxDesc* __restrict__ boxes, Plan plan, TvdSweepArgs<real> a) {
    const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
    const long long g = (long long)blk * 256 + threadIdx.x;
    const long long gw = g - (threadIdx.x & 63);   // first item of this wave
    if (gw >= plan.total) return;                  // whole wave
    int b = 0;
    if (plan.nbox > 1) {
        int lo = 0, hi = plan.nbox - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (plan.boxes[mid].start <= gw) lo = mid; else hi = mid - 1;
        }
        b = lo;
    }
    const PlanBox& pb = plan.boxes[b];
    const long long wseg = (gw - pb.start) >> 6;   // (row, segment) of this wave
    const int nseg = pb.zn;
    if (wseg >= (long long)pb.r[1] * pb.r[2] * nseg) return;   // whole wave
    const int lane = (int)(threadIdx.x & 63);
    const int sgi = (int)(wseg % nseg);
    const long long row = wseg / nseg;
    const int j = pb.lo[1] + (int)(row % pb.r[1]);
    const int k = pb.lo[2] + (int)(row / pb.r[1]);
    const int x0 = pb.lo[0] + sgi * TVD_XSEG;
    const int x1 = min(x0 + TVD_XSEG - 1, pb.lo[0] + pb.r[0] - 1);   // last flux face of the segment
    const int x = x0 - 1 + lane;                                     // this lane's face
    const bool eig = x <= x1 + 1;
    const BoxDesc& B = boxes[b];
    real uL[7], uR[7], ck[6], gt[6], dtdx, dxdt;
    tvd_dts(a, B, b, dtdx, dxdt);
    const long long c = cell_index(B, eig ? x : x0, j, k);
    if (eig) {
        tvd_load(a, c, uL);
        tvd_load(a, c + 1, uR);
        tvd_eig_face(uL, uR, (real)a.cs, dtdx, gt, ck);
    } else {
#pragma unroll
        for (int w = 0; w < 6; w++) gt[w] = ck[w] = real(0);
    }
    real gm[6], gp[6];
#pragma unroll
    for (int w = 0; w < 6; w++) {
        gm[w] = __shfl_up(gt[w], 1, 64);
        gp[w] = __shfl_down(gt[w], 1, 64);
    }
    if (lane >= 1 && lane <= TVD_XSEG && x <= x1) {
        real F[7];
        tvd_flux_face(uL, uR, (real)a.cs, dtdx, dxdt, gm, gt, gp, ck, F);
        tvd_store(a, c, F);
    }
}

// y / z sweep: threads along x (coalesced), each marching a segment of pb.seg flux faces
// along the sweep axis with a three-face window of g~_k in registers.  Items enumerate
// (x, transverse index, segment), x fastest.
// FP64: at most 256 registers, two waves per SIMD (the unbounded allocation took 268 and
// left one: 1371 us per call at 256^3; bounded, 6 registers spill).
template <class real>
__global__ __launch_bounds__(256, sizeof(real) == 8 ? 2 : 1) void k_tvd_sweep_m(const BoxDesc* __restrict__ boxes, Plan plan,
                                                                              TvdSweepArgs<real> a) {
    const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
    const long long g = (long long)blk * blockDim.x + threadIdx.x;
    if (g >= plan.total) return;
    int b = 0;
    if (plan.nbox > 1) {
        int lo = 0, hi = plan.nbox - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (plan.boxes[mid].start <= g) lo = mid; else hi = mid - 1;
        }
        b = lo;
    }
    const PlanBox& pb = plan.boxes[b];
    const int d = a.dir, ta = d == 1 ? 2 : 1;   // sweep axis, other transverse axis
    const long long loc = g - pb.start;
    const long long cols = (long long)pb.r[0] * pb.r[ta];
    if (loc >= cols * pb.zn) return;
    const int i = pb.lo[0] + (int)(loc % pb.r[0]);
    const long long t = loc / pb.r[0];
    const int it = pb.lo[ta] + (int)(t % pb.r[ta]);
    const int sgi = (int)(t / pb.r[ta]);
    const int f0 = pb.lo[d] + sgi * pb.seg;
    const int f1 = min(f0 + pb.seg - 1, pb.lo[d] + pb.r[d] - 1);
    const BoxDesc& B = boxes[b];
    const long long s = tvd_stride(B, d);
    const long long c0 = d == 1 ? cell_index(B, i, f0, it) : cell_index(B, i, it, f0);
    const real cs = (real)a.cs;
    real dtdx, dxdt;
    tvd_dts(a, B, b, dtdx, dxdt);
    // One loop body forms every face's g~_k / alpha_k and every flux, so each face gets the
    // same instruction sequence wherever the segment boundaries fall (FP32 contraction
    // included): the fluxes stay translation-invariant, as periodic images require.
    // Iteration f: eigen-system of face f (cells f, f+1); flux of face f-1 once its upper
    // neighbour's g~_k is known.
    real uA[7], uB[7], uC[7], gm[6], g0[6], gp[6], ck0[6], ckp[6];
    tvd_load(a, c0 - s, uB);
#pragma unroll
    for (int w = 0; w < 6; w++) {
        gm[w] = g0[w] = gp[w] = real(0);
        ck0[w] = ckp[w] = real(0);
    }
#pragma unroll
    for (int r = 0; r < 7; r++) uA[r] = uB[r];
    long long c = c0 - s;   // face f
#pragma unroll 1
    for (int f = f0 - 1; f <= f1 + 1; f++, c += s) {
        tvd_load(a, c + s, uC);
        tvd_eig_face(uB, uC, cs, dtdx, gp, ckp);
        if (f > f0) {
            real F[7];
            tvd_flux_face(uA, uB, cs, dtdx, dxdt, gm, g0, gp, ck0, F);
            tvd_store(a, c - s, F);
        }
#pragma unroll
        for (int w = 0; w < 6; w++) {
            gm[w] = g0[w];
            g0[w] = gp[w];
            ck0[w] = ckp[w];
        }
#pragma unroll
        for (int r = 0; r < 7; r++) {
            uA[r] = uB[r];
            uB[r] = uC[r];
        }
    }
}

// HLLE of the TVD module (isotvd_mod.f90:2473-2578): adjacent cell states, each with its
// own face field; same arithmetic as the VL module's HLLE without the smax bookkeeping.
template <class real>
__global__ __launch_bounds__(256) void k_tvd_hlle(const BoxDesc* __restrict__ boxes, Plan plan, TvdSweepArgs<real> a) {
    int b, i, j, k;
    if (!plan_item(plan, b, i, j, k)) return;
    const BoxDesc& B = boxes[b];
    const long long c = cell_index(B, i, j, k), s = tvd_stride(B, a.dir);
    real uL[7], uR[7], F[7];
    tvd_load(a, c, uL);
    tvd_load(a, c + s, uR);
    real sm = real(0);
    hlle(uL, uR, (real)(a.cs * a.cs), F, sm);
    tvd_store(a, c, F);
}

__device__ __forceinline__ void tvd_upw(double v, double L14, double R14, double L34, double R34, double& o14,
                                        double& o34) {
    if (v > ZERO) { o14 = L14; o34 = L34; }
    else if (v < -ZERO) { o14 = R14; o34 = R34; }
    else { o14 = 0.5 * (L14 + R14); o34 = 0.5 * (L34 + R34); }
}

// Corner EMFs of the TVD module (isotvd_mod.f90:993-1174): Gardiner-Stone upwinding with
// gradients 2/dx (E_face - E_centre) and the correction scaled by dx/8.  Centre EMF: the
// predictor forms (B x m)/rho from the t^n state (isotvd_mod.f90:947-953); the corrector
// reads the half-step centre EMF.
template <class real>
__global__ __launch_bounds__(256) void k_tvd_corner(const BoxDesc* __restrict__ boxes, Plan plan, TvdCornerArgs<real> a) {
    int b, i, j, k;
    if (!plan_item(plan, b, i, j, k)) return;
    const BoxDesc& B = boxes[b];
    const long long c = cell_index(B, i, j, k);
    const long long sx = 1, sy = B.P[0], sz = (long long)B.P[0] * B.P[1];
    const double dx = B.dx, idx2 = 2.0 / dx;
    // x / rho: FP64 divides (the reference's arithmetic); an FP32 state multiplies by its
    // hardware reciprocal
    auto divr = [](double x, real rho) -> double {
        if constexpr (sizeof(real) == 4) return x * (double)__builtin_amdgcn_rcpf(rho);
        else return x / (double)rho;
    };
    // Every load comes before the arithmetic (the only stores go to Ec), so the loads of a
    // corner triple are in flight together: face EMFs, rho and m of the cells the upwinding
    // velocities take, and the centre EMFs, read (corrector) or formed from the t^n state of
    // the seven cells c, +x, +y, +z, +x+y, +x+z, +y+z (predictor, isotvd_mod.f90:947-953).
    const double ez_y0 = (double)a.E[5][c], ez_y1 = (double)a.E[5][c + sx];
    const double ez_x0 = (double)a.E[4][c], ez_x1 = (double)a.E[4][c + sy];
    const double ey_z0 = (double)a.E[3][c], ey_z1 = (double)a.E[3][c + sx];
    const double ey_x0 = (double)a.E[2][c], ey_x1 = (double)a.E[2][c + sz];
    const double ex_z0 = (double)a.E[1][c], ex_z1 = (double)a.E[1][c + sy];
    const double ex_y0 = (double)a.E[0][c], ex_y1 = (double)a.E[0][c + sz];
    const real r0 = a.q[0][c], r1 = a.q[0][c + sx], r2 = a.q[0][c + sy], r3 = a.q[0][c + sz];
    const real mx0 = a.q[1][c], mx1 = a.q[1][c + sx];
    const real my0 = a.q[2][c], my2 = a.q[2][c + sy];
    const real mz0 = a.q[3][c], mz3 = a.q[3][c + sz];
    // centre EMFs: Ez at c, +x, +y, +x+y; Ey at c, +x, +z, +x+z; Ex at c, +y, +z, +y+z
    double ez[4], ey[4], ex[4];
    if (a.Ecc[0] != nullptr) {
        const long long pz[4] = {c, c + sx, c + sy, c + sx + sy};
        const long long py[4] = {c, c + sx, c + sz, c + sx + sz};
        const long long px[4] = {c, c + sy, c + sz, c + sy + sz};
#pragma unroll
        for (int n = 0; n < 4; n++) {
            ez[n] = a.Ecc[2][pz[n]];
            ey[n] = a.Ecc[1][py[n]];
            ex[n] = a.Ecc[0][px[n]];
        }
    } else {
        const long long p7[7] = {c, c + sx, c + sy, c + sz, c + sx + sy, c + sx + sz, c + sy + sz};
        real w[7][7];   // rho, m, B of each cell (the values no centre EMF takes are not loaded)
#pragma unroll
        for (int n = 0; n < 7; n++) {
#pragma unroll
            for (int f = 0; f < 4; f++) w[n][f] = a.q[f][p7[n]];
#pragma unroll
            for (int f = 0; f < 3; f++) w[n][4 + f] = a.qb[f][p7[n]];
        }
        // (B x m) / rho, component comp, of cell n
        auto cen = [&](int comp, int n) -> double {
            const real* u = w[n];
            if (comp == 0) return divr((double)u[5] * (double)u[3] - (double)u[6] * (double)u[2], u[0]);
            if (comp == 1) return divr((double)u[6] * (double)u[1] - (double)u[4] * (double)u[3], u[0]);
            return divr((double)u[4] * (double)u[2] - (double)u[5] * (double)u[1], u[0]);
        };
        const int nz[4] = {0, 1, 2, 4}, ny[4] = {0, 1, 3, 5}, nx[4] = {0, 2, 3, 6};
#pragma unroll
        for (int n = 0; n < 4; n++) {
            ez[n] = cen(2, nz[n]);
            ey[n] = cen(1, ny[n]);
            ex[n] = cen(0, nx[n]);
        }
    }
    const double vx = 0.5 * (divr((double)mx0, r0) + divr((double)mx1, r1));
    const double vy = 0.5 * (divr((double)my0, r0) + divr((double)my2, r2));
    const double vz = 0.5 * (divr((double)mz0, r0) + divr((double)mz3, r3));
    double a14, a34, b14, b34;
    {   // Ez at (i+1/2, j+1/2, k)
        const double e00 = ez[0], e10 = ez[1], e01 = ez[2], e11 = ez[3];
        const double y0 = ez_y0, y1 = ez_y1, x0 = ez_x0, x1 = ez_x1;
        tvd_upw(vx, idx2 * (y0 - e00), idx2 * (y1 - e10), idx2 * (e01 - y0), idx2 * (e11 - y1), a14, a34);
        tvd_upw(vy, idx2 * (x0 - e00), idx2 * (x1 - e01), idx2 * (e10 - x0), idx2 * (e11 - x1), b14, b34);
        a.Ec[2][c] = 0.25 * (x0 + x1 + y0 + y1) + 0.125 * dx * ((a14 - a34) + (b14 - b34));
    }
    {   // Ey at (i+1/2, j, k+1/2)
        const double e00 = ey[0], e10 = ey[1], e01 = ey[2], e11 = ey[3];
        const double z0 = ey_z0, z1 = ey_z1, x0 = ey_x0, x1 = ey_x1;
        tvd_upw(vx, idx2 * (z0 - e00), idx2 * (z1 - e10), idx2 * (e01 - z0), idx2 * (e11 - z1), a14, a34);
        tvd_upw(vz, idx2 * (x0 - e00), idx2 * (x1 - e01), idx2 * (e10 - x0), idx2 * (e11 - x1), b14, b34);
        a.Ec[1][c] = 0.25 * (x0 + x1 + z0 + z1) + 0.125 * dx * ((a14 - a34) + (b14 - b34));
    }
    {   // Ex at (i, j+1/2, k+1/2)
        const double e00 = ex[0], e10 = ex[1], e01 = ex[2], e11 = ex[3];
        const double z0 = ex_z0, z1 = ex_z1, y0 = ex_y0, y1 = ex_y1;
        tvd_upw(vy, idx2 * (z0 - e00), idx2 * (z1 - e10), idx2 * (e01 - z0), idx2 * (e11 - z1), a14, a34);
        tvd_upw(vz, idx2 * (y0 - e00), idx2 * (y1 - e01), idx2 * (e10 - y0), idx2 * (e11 - y1), b14, b34);
        a.Ec[0][c] = 0.25 * (y0 + y1 + z0 + z1) + 0.125 * dx * ((a14 - a34) + (b14 - b34));
    }
}

// Corrector centre EMF from the unsplit half update of rho, m with the predictor fluxes
// and the face average of the half-step face fields (isotvd_mod.f90:876-906)
template <class real>
__device__ __forceinline__ void tvd_center_calc(const real* q, const real* const (*Fb)[4], const double* const* bh,
                                                double* const* Ecc, long long c, long long sx, long long sy,
                                                long long sz, real dtdx) {
    real h[4];
#pragma unroll
    for (int f = 0; f < 4; f++)
        h[f] = q[f] - dtdx * (Fb[0][f][c] - Fb[0][f][c - sx] + Fb[1][f][c] - Fb[1][f][c - sy] + Fb[2][f][c] -
                              Fb[2][f][c - sz]);
    double vx = (double)h[1] / (double)h[0], vy = (double)h[2] / (double)h[0], vz = (double)h[3] / (double)h[0];
    double Bx = 0.5 * (bh[0][c] + bh[0][c - sx]);
    double By = 0.5 * (bh[1][c] + bh[1][c - sy]);
    double Bz = 0.5 * (bh[2][c] + bh[2][c - sz]);
    Ecc[0][c] = By * vz - Bz * vy;
    Ecc[1][c] = Bz * vx - Bx * vz;
    Ecc[2][c] = Bx * vy - By * vx;
}
template <class real>
__global__ __launch_bounds__(256) void k_tvd_center(const BoxDesc* __restrict__ boxes, Plan plan, TvdCenterArgs<real> a) {
    int b, i, j, k;
    if (!plan_item(plan, b, i, j, k)) return;
    const BoxDesc& B = boxes[b];
    const long long c = cell_index(B, i, j, k);
    const long long sx = 1, sy = B.P[0], sz = (long long)B.P[0] * B.P[1];
    const real dtdx = (real)(0.5 * a.dt[b] / B.dx);
    const real q[4] = {a.q[0][c], a.q[1][c], a.q[2][c], a.q[3][c]};
    tvd_center_calc(q, a.Fb, a.bh, a.Ecc, c, sx, sy, sz, dtdx);
}

// minmod-like multidimensional source of the half updates (isotvd_mod.f90:1208-1212)
__device__ __forceinline__ double tvd_src(double v, double da, double db) {
    if ((-da * db) > ZERO) return -v * sgn1(da) * fmin(fabs(da), fabs(db));
    return 0.0;
}

// halfupdate_{x,y,z}_3d (isotvd_mod.f90:1179-1650) over [3-nb, n+nb-2]: 1/2 dt update of
// rho, m with the two transverse predictor fluxes and the B_n-gradient momentum source,
// of the two transverse cell-centred fields with the predictor corner EMFs and the
// minmod source; then the density floor.  The normal cell-centred field is not needed by
// the following sweep (it reads the face field) and is not written.
// One cell's half update along direction d, without stores: hq = rho, m of the half state,
// bt = its two transverse cell-centred fields (in the order of the global field index),
// floored = the density floor was applied.
template <class real>
__device__ __forceinline__ void tvd_half_calc(const TvdHalfArgs<real>& a, int d, long long c, const long long* st,
                                              const real* q, const double* dB, double dtd, real* hq, real* bt,
                                              bool& floored) {
    const long long sx = st[0], sy = st[1], sz = st[2];
    const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
    const real dtdx = (real)dtd;
    const real dn = (real)dB[d];
    const real sm[3] = {q[4] * dn, q[5] * dn, q[6] * dn};
#pragma unroll
    for (int f = 0; f < 4; f++) {
        real sum = a.Fb[t1][f][c] - a.Fb[t1][f][c - st[t1]] + a.Fb[t2][f][c] - a.Fb[t2][f][c - st[t2]];
        hq[f] = f == 0 ? q[0] - dtdx * sum : q[f] - dtdx * sum + dtdx * sm[f - 1];
    }
    floored = hq[0] < (real)MIN_DENS;
    if (floored) hq[0] = (real)MIN_DENS;
    const double* Ec = a.Ec[d];
    const double dBx = dB[0], dBy = dB[1], dBz = dB[2];
    if (d == 0) {
        double vy = (double)(q[2] / q[0]), vz = (double)(q[3] / q[0]);
        double sBy = tvd_src(vy, dBz, dBx), sBz = tvd_src(vz, dBy, dBx);
        double by = (double)q[5] - 0.5 * dtd * ((Ec[c] - Ec[c - sz]) + (Ec[c - sy] - Ec[c - sy - sz])) + dtd * sBy;
        double bz = (double)q[6] + 0.5 * dtd * ((Ec[c] - Ec[c - sy]) + (Ec[c - sz] - Ec[c - sy - sz])) + dtd * sBz;
        bt[0] = (real)by;
        bt[1] = (real)bz;
    } else if (d == 1) {
        double vx = (double)(q[1] / q[0]), vz = (double)(q[3] / q[0]);
        double sBz = tvd_src(vz, dBx, dBy), sBx = tvd_src(vx, dBz, dBy);
        double bx = (double)q[4] + 0.5 * dtd * ((Ec[c] - Ec[c - sz]) + (Ec[c - sx] - Ec[c - sx - sz])) + dtd * sBx;
        double bz = (double)q[6] - 0.5 * dtd * ((Ec[c] - Ec[c - sx]) + (Ec[c - sz] - Ec[c - sx - sz])) + dtd * sBz;
        bt[0] = (real)bx;
        bt[1] = (real)bz;
    } else {
        double vx = (double)(q[1] / q[0]), vy = (double)(q[2] / q[0]);
        double sBx = tvd_src(vx, dBy, dBz), sBy = tvd_src(vy, dBx, dBz);
        double bx = (double)q[4] - 0.5 * dtd * ((Ec[c] - Ec[c - sy]) + (Ec[c - sx] - Ec[c - sx - sy])) + dtd * sBx;
        double by = (double)q[5] + 0.5 * dtd * ((Ec[c] - Ec[c - sx]) + (Ec[c - sy] - Ec[c - sx - sy])) + dtd * sBy;
        bt[0] = (real)bx;
        bt[1] = (real)by;
    }
}
__device__ __forceinline__ void tvd_half_bt_fields(int d, int& f1, int& f2) {
    f1 = d == 0 ? 5 : 4;
    f2 = d == 2 ? 5 : 6;
}

template <class real>
__global__ __launch_bounds__(256) void k_tvd_half(const BoxDesc* __restrict__ boxes, Plan plan, TvdHalfArgs<real> a) {
    int b, i, j, k;
    if (!plan_item(plan, b, i, j, k)) return;
    const BoxDesc& B = boxes[b];
    const long long c = cell_index(B, i, j, k);
    const long long st[3] = {1, B.P[0], (long long)B.P[0] * B.P[1]};
    const int d = a.dir;
    const double dtd = 0.5 * a.dt[b] / B.dx;
    const double dB[3] = {a.b0[0][c] - a.b0[0][c - st[0]], a.b0[1][c] - a.b0[1][c - st[1]],
                          a.b0[2][c] - a.b0[2][c - st[2]]};
    real q[7];
#pragma unroll
    for (int f = 0; f < 7; f++) q[f] = a.q[f][c];
    real hq[4], bt[2];
    bool floored;
    tvd_half_calc(a, d, c, st, q, dB, dtd, hq, bt, floored);
    if (floored) atomicAdd(a.floor_count, 1u);
#pragma unroll
    for (int f = 0; f < 4; f++) a.dst[f][c] = hq[f];
    int f1, f2;
    tvd_half_bt_fields(d, f1, f2);
    a.dst[f1][c] = bt[0];
    a.dst[f2][c] = bt[1];
}

// The three half updates of a cell in one pass, each into the set its sweep reads
// (dst3[d]); all values are formed before any store, so the loads the directions share
// (the t^n state and face field, each direction's predictor-flux differences, used by the
// two others) are issued once.  Same arithmetic and floor count as three k_tvd_half.
template <class real>
__global__ __launch_bounds__(256) void k_tvd_half3(const BoxDesc* __restrict__ boxes, Plan plan, TvdHalfArgs<real> a,
                                                   TvdHalfDst<real> o) {
    int b, i, j, k;
    if (!plan_item(plan, b, i, j, k)) return;
    const BoxDesc& B = boxes[b];
    const long long c = cell_index(B, i, j, k);
    const long long st[3] = {1, B.P[0], (long long)B.P[0] * B.P[1]};
    const double dtd = 0.5 * a.dt[b] / B.dx;
    const double dB[3] = {a.b0[0][c] - a.b0[0][c - st[0]], a.b0[1][c] - a.b0[1][c - st[1]],
                          a.b0[2][c] - a.b0[2][c - st[2]]};
    real q[7];
#pragma unroll
    for (int f = 0; f < 7; f++) q[f] = a.q[f][c];
    real hq[3][4], bt[3][2];
    bool fl[3];
#pragma unroll
    for (int d = 0; d < 3; d++) tvd_half_calc(a, d, c, st, q, dB, dtd, hq[d], bt[d], fl[d]);
    const unsigned nf = (unsigned)fl[0] + (unsigned)fl[1] + (unsigned)fl[2];
    if (nf) atomicAdd(a.floor_count, nf);
#pragma unroll
    for (int d = 0; d < 3; d++) {
#pragma unroll
        for (int f = 0; f < 4; f++) o.dst[d][f][c] = hq[d][f];
        int f1, f2;
        tvd_half_bt_fields(d, f1, f2);
        o.dst[d][f1][c] = bt[d][0];
        o.dst[d][f2][c] = bt[d][1];
    }
    if (o.Ecc[0] != nullptr) {   // the corrector centre EMF from the same loads (case 22's range)
        const int nb = B.nb;
        auto in = [&](int v, int a) { return v >= 4 - nb && v <= B.n[a] + nb - 3; };
        if (in(i, 0) && in(j, 1) && in(k, 2))
            tvd_center_calc(q, a.Fb, o.bh, o.Ecc, c, st[0], st[1], st[2], (real)dtd);
    }
}

// update_CT_3d(.FALSE.) (isotvd_mod.f90:1794-1831) on [3-nb, n+nb-3], t^n faces elsewhere
template <class real>
__global__ __launch_bounds__(256) void k_tvd_ct_pred(const BoxDesc* __restrict__ boxes, Plan plan, TvdCtArgs<real> a) {
    int b, i, j, k;
    if (!plan_item(plan, b, i, j, k)) return;
    const BoxDesc& B = boxes[b];
    const long long c = cell_index(B, i, j, k);
    const long long sx = 1, sy = B.P[0], sz = (long long)B.P[0] * B.P[1];
    const int nb = B.nb;
    const bool in = i >= 3 - nb && j >= 3 - nb && k >= 3 - nb && i <= B.n[0] + nb - 3 && j <= B.n[1] + nb - 3 &&
                    k <= B.n[2] + nb - 3;
    // all loads before the first store (b1 and b0 are different sets)
    const double b0x = a.b0[0][c], b0y = a.b0[1][c], b0z = a.b0[2][c];
    if (!in) {
        a.b1[0][c] = b0x;
        a.b1[1][c] = b0y;
        a.b1[2][c] = b0z;
        return;
    }
    const double dtd = 0.5 * (a.dt[b] / B.dx);
    const double* Ex = a.Ec[0];
    const double* Ey = a.Ec[1];
    const double* Ez = a.Ec[2];
    const double ez0 = Ez[c], ez_y = Ez[c - sy], ez_x = Ez[c - sx];
    const double ey0 = Ey[c], ey_z = Ey[c - sz], ey_x = Ey[c - sx];
    const double ex0 = Ex[c], ex_z = Ex[c - sz], ex_y = Ex[c - sy];
    a.b1[0][c] = b0x - dtd * ((ez0 - ez_y) - (ey0 - ey_z));
    a.b1[1][c] = b0y - dtd * ((ex0 - ex_z) - (ez0 - ez_x));
    a.b1[2][c] = b0z - dtd * ((ey0 - ey_x) - (ex0 - ex_y));
}

// update_state_vector_3d + update_CT_3d(.TRUE.) (isotvd_mod.f90:1655-1851): rho, m on the
// interior with the corrector fluxes and the floor; faces on [5-nb, n+nb-5] from the t^n
// faces; cell-centred B on the interior from the new faces (lower faces recomputed, so the
// kernel never reads a face it writes).  Everything else keeps its t^n value, as the
// reference's restore of q_3d and bface_3d leaves it.
template <class real>
__global__ __launch_bounds__(256) void k_tvd_final(const BoxDesc* __restrict__ boxes, Plan plan, TvdFinalArgs<real> a) {
    int b, i, j, k;
    if (!plan_item(plan, b, i, j, k)) return;
    const BoxDesc& B = boxes[b];
    const long long c = cell_index(B, i, j, k);
    const long long sx = 1, sy = B.P[0], sz = (long long)B.P[0] * B.P[1];
    const int nb = B.nb;
    const double dtd = a.dt[b] / B.dx;
    const real dtdx = (real)dtd;
    const bool interior = i >= 1 && j >= 1 && k >= 1 && i <= B.n[0] && j <= B.n[1] && k <= B.n[2];
    const bool face = i >= 5 - nb && j >= 5 - nb && k >= 5 - nb && i <= B.n[0] + nb - 5 && j <= B.n[1] + nb - 5 &&
                      k <= B.n[2] + nb - 5;
    const bool comp = a.comp[0] != nullptr;
    // Every load comes before the first store: the kernel reads the t^n set, the fluxes and
    // the corner EMFs and writes the other set (the Kahan terms only at its own cell), so the
    // loads of a cell are in flight together.
    double bs[3];   // t^n faces at c
#pragma unroll
    for (int d = 0; d < 3; d++) bs[d] = a.src.b[d][c];
    real qs[7];
#pragma unroll
    for (int f = 0; f < 7; f++) qs[f] = a.src.q[f][c];
    // corner EMFs: Ez at c, c-y, c-x, c-x-y; Ey at c, c-z, c-x, c-x-z; Ex at c, c-z, c-y, c-y-z
    double ez[4] = {0.0, 0.0, 0.0, 0.0}, ey[4] = {0.0, 0.0, 0.0, 0.0}, ex[4] = {0.0, 0.0, 0.0, 0.0};
    if (face || interior) {
        const double* Ex = a.Ec[0];
        const double* Ey = a.Ec[1];
        const double* Ez = a.Ec[2];
        ez[0] = Ez[c]; ez[1] = Ez[c - sy]; ez[2] = Ez[c - sx]; ez[3] = Ez[c - sx - sy];
        ey[0] = Ey[c]; ey[1] = Ey[c - sz]; ey[2] = Ey[c - sx]; ey[3] = Ey[c - sx - sz];
        ex[0] = Ex[c]; ex[1] = Ex[c - sz]; ex[2] = Ex[c - sy]; ex[3] = Ex[c - sy - sz];
    }
    double bl[3] = {0.0, 0.0, 0.0};   // t^n lower faces c-x (x), c-y (y), c-z (z)
    real fc[4][3], fm[4][3], cp[4];    // fluxes at c and at c minus the unit step; Kahan terms
    if (interior) {
        bl[0] = a.src.b[0][c - sx];
        bl[1] = a.src.b[1][c - sy];
        bl[2] = a.src.b[2][c - sz];
#pragma unroll
        for (int f = 0; f < 4; f++) {
            fc[f][0] = a.F[0][f][c]; fm[f][0] = a.F[0][f][c - sx];
            fc[f][1] = a.F[1][f][c]; fm[f][1] = a.F[1][f][c - sy];
            fc[f][2] = a.F[2][f][c]; fm[f][2] = a.F[2][f][c - sz];
            cp[f] = comp ? a.comp[f][c] : real(0);
        }
    }
    // update_CT_3d: b - dt/dx ((E1 - E1') - (E2 - E2'))
    auto ct = [&](double b0, double e1, double e1m, double e2, double e2m) { return b0 - dtd * ((e1 - e1m) - (e2 - e2m)); };
    double bx, by, bz;
    if (face) {
        bx = ct(bs[0], ez[0], ez[1], ey[0], ey[1]);
        by = ct(bs[1], ex[0], ex[1], ez[0], ez[2]);
        bz = ct(bs[2], ey[0], ey[2], ex[0], ex[2]);
    } else {
        bx = bs[0]; by = bs[1]; bz = bs[2];
    }
    a.dst.b[0][c] = bx;
    a.dst.b[1][c] = by;
    a.dst.b[2][c] = bz;
    if (!interior) {
#pragma unroll
        for (int f = 0; f < 7; f++) a.dst.q[f][c] = qs[f];
        return;
    }
#pragma unroll
    for (int f = 0; f < 4; f++) {
        real div = (fc[f][0] - fm[f][0]) + (fc[f][1] - fm[f][1]) + (fc[f][2] - fm[f][2]);
        real q0 = qs[f];
        real v;
        if (comp) {  // compensated (Kahan) accumulation of the FP32 state
            real y = -(dtdx * div) - cp[f];
            v = q0 + y;
            cp[f] = (v - q0) - y;
        } else {
            v = q0 - dtdx * div;
        }
        if (f == 0 && v < (real)MIN_DENS) {
            v = (real)MIN_DENS;
            cp[0] = real(0);
            atomicAdd(a.floor_count, 1u);
        }
        a.dst.q[f][c] = v;
        if (comp) a.comp[f][c] = cp[f];
    }
    // cell-centred B from the new faces; the lower faces recomputed (never read back)
    a.dst.q[4][c] = (real)(0.5 * (bx + ct(bl[0], ez[2], ez[3], ey[2], ey[3])));
    a.dst.q[5][c] = (real)(0.5 * (by + ct(bl[1], ex[2], ex[3], ez[1], ez[3])));
    a.dst.q[6][c] = (real)(0.5 * (bz + ct(bl[2], ey[1], ey[3], ex[1], ex[3])));
}

// get_max_speed_tvd (isotvd_mod.f90:2607-2673) over [0, n]: fast speed of the
// interface-mean state per direction
template <class real>
__global__ __launch_bounds__(256) void k_tvd_max_speed(const BoxDesc* __restrict__ boxes, Plan plan, QPtrs<real> Q,
                                                       double cs, unsigned long long* cmax) {
    int b = 0, i, j, k;
    double m = 0.0;
    bool act = plan_item_box(plan, b, i, j, k);
    if (act) {
        const BoxDesc& B = boxes[b];
        const long long c = cell_index(B, i, j, k);
        const long long st[3] = {1, B.P[0], (long long)B.P[0] * B.P[1]};
        const double asqr = cs * cs;
        auto q = [&](int f, long long p) -> double { return (double)Q.q[f][p]; };
        double cf[3], v[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const long long n = c + st[d];
            double rho2 = 0.5 * (q(0, c) + q(0, n));
            double isrho = 1.0 / sqrt(rho2);
            double bb[3];
            for (int e = 0; e < 3; e++) bb[e] = 0.5 * (q(4 + e, c) + q(4 + e, n)) * isrho;
            const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
            double btsqr = bb[t1] * bb[t1] + bb[t2] * bb[t2];
            double bsqr = bb[d] * bb[d] + btsqr, x = asqr - bsqr;
            double tmp = sqrt(x * x + 4.0 * asqr * btsqr);
            cf[d] = sqrt(0.5 * (asqr + bsqr + tmp));
            v[d] = 0.5 * ((q(1 + d, c) / q(0, c)) + (q(1 + d, n) / q(0, n)));
        }
        m = max3(fabs(v[0]) + cf[0], fabs(v[1]) + cf[1], fabs(v[2]) + cf[2]);
    }
    atomic_max_nonneg(cmax, b, m);
}

template <class real>
void launch_tvd_eig(const Launch& L, const TvdSweepArgs<real>& a) {
    if (L.plan.total > 0) hipLaunchKernelGGL((k_tvd_eig<real>), dim3(L.blocks()), dim3(256), 0, L.stream, L.boxes, L.plan, a);
}
template <class real>
void launch_tvd_flux(const Launch& L, const TvdSweepArgs<real>& a) {
    if (L.plan.total > 0) hipLaunchKernelGGL((k_tvd_flux<real>), dim3(L.blocks()), dim3(256), 0, L.stream, L.boxes, L.plan, a);
}
template <class real>
void launch_tvd_sweep(const Launch& L, const TvdSweepArgs<real>& a) {
    if (L.plan.total <= 0) return;
    if (a.dir == 0)
        hipLaunchKernelGGL((k_tvd_sweep_x<real>), dim3(L.blocks()), dim3(256), 0, L.stream, L.boxes, L.plan, a);
    else
        hipLaunchKernelGGL((k_tvd_sweep_m<real>), dim3(L.blocks()), dim3(256), 0, L.stream, L.boxes, L.plan, a);
}
template <class real>
void launch_tvd_hlle(const Launch& L, const TvdSweepArgs<real>& a) {
    if (L.plan.total > 0) hipLaunchKernelGGL((k_tvd_hlle<real>), dim3(L.blocks()), dim3(256), 0, L.stream, L.boxes, L.plan, a);
}
template <class real>
void launch_tvd_corner(const Launch& L, const TvdCornerArgs<real>& a) {
    if (L.plan.total > 0) hipLaunchKernelGGL((k_tvd_corner<real>), dim3(L.blocks()), dim3(256), 0, L.stream, L.boxes, L.plan, a);
}
template <class real>
void launch_tvd_center(const Launch& L, const TvdCenterArgs<real>& a) {
    if (L.plan.total > 0) hipLaunchKernelGGL((k_tvd_center<real>), dim3(L.blocks()), dim3(256), 0, L.stream, L.boxes, L.plan, a);
}
template <class real>
void launch_tvd_half(const Launch& L, const TvdHalfArgs<real>& a) {
    if (L.plan.total > 0) hipLaunchKernelGGL((k_tvd_half<real>), dim3(L.blocks()), dim3(256), 0, L.stream, L.boxes, L.plan, a);
}
template <class real>
void launch_tvd_half3(const Launch& L, const TvdHalfArgs<real>& a, const TvdHalfDst<real>& o) {
    if (L.plan.total > 0)
        hipLaunchKernelGGL((k_tvd_half3<real>), dim3(L.blocks()), dim3(256), 0, L.stream, L.boxes, L.plan, a, o);
}
template <class real>
void launch_tvd_ct_pred(const Launch& L, const TvdCtArgs<real>& a) {
    if (L.plan.total > 0) hipLaunchKernelGGL((k_tvd_ct_pred<real>), dim3(L.blocks()), dim3(256), 0, L.stream, L.boxes, L.plan, a);
}
template <class real>
void launch_tvd_final(const Launch& L, const TvdFinalArgs<real>& a) {
    if (L.plan.total > 0) hipLaunchKernelGGL((k_tvd_final<real>), dim3(L.blocks()), dim3(256), 0, L.stream, L.boxes, L.plan, a);
}
template <class real>
void launch_tvd_max_speed(const Launch& L, const QPtrs<real>& q, double cs, unsigned long long* cmax) {
    if (L.plan.total > 0)
        hipLaunchKernelGGL((k_tvd_max_speed<real>), dim3(L.blocks()), dim3(256), 0, L.stream, L.boxes, L.plan, q, cs, cmax);
}

#define IMHD_INST_TVD(T)                                                            \
    template void launch_tvd_eig<T>(const Launch&, const TvdSweepArgs<T>&);         \
    template void launch_tvd_flux<T>(const Launch&, const TvdSweepArgs<T>&);        \
    template void launch_tvd_sweep<T>(const Launch&, const TvdSweepArgs<T>&);       \
    template void launch_tvd_hlle<T>(const Launch&, const TvdSweepArgs<T>&);        \
    template void launch_tvd_corner<T>(const Launch&, const TvdCornerArgs<T>&);     \
    template void launch_tvd_center<T>(const Launch&, const TvdCenterArgs<T>&);     \
    template void launch_tvd_half<T>(const Launch&, const TvdHalfArgs<T>&);         \
    template void launch_tvd_half3<T>(const Launch&, const TvdHalfArgs<T>&, const TvdHalfDst<T>&); \
    template void launch_tvd_ct_pred<T>(const Launch&, const TvdCtArgs<T>&);        \
    template void launch_tvd_final<T>(const Launch&, const TvdFinalArgs<T>&);       \
    template void launch_tvd_max_speed<T>(const Launch&, const QPtrs<T>&, double, unsigned long long*);
}  // namespace imhd
