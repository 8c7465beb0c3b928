// VL2 + CT cell-update kernels (Stone & Gardiner 2009 unsplit predictor-corrector with
// upwind constrained transport), following vanleer_mod.f90 of the reference.
// Instantiated per precision by vl_kernels_f64.hip (no FMA contraction: bitwise with the
// reference) and vl_kernels_f32.hip (contraction and hardware rcp/sqrt).
//
// Per stage:  k_flux (x3 directions)  ->  k_corner  ->  k_update
// Fluxes and face EMFs go to HBM once per face; the corner EMF is summed in FP64; the
// update writes a second state buffer (no copy/restore of the t^n state).
#pragma once
#include "kernels.hpp"
#include "riemann.hpp"
#include "ppm.hpp"

namespace imhd {

__device__ __forceinline__ long long dir_stride(const BoxDesc& B, int d) {
    return d == 0 ? 1LL : (d == 1 ? (long long)B.P[0] : (long long)B.P[0] * (long long)B.P[1]);
}

template <class real>
__device__ __forceinline__ void store_flux(const SweepArgs<real>& a, long long c, const real* F) {
    a.F[0][c] = F[0];
    a.F[1][c] = F[1];
    a.F[2][c] = F[2];
    a.F[3][c] = F[3];
    a.Et1[c] = -F[5];
    a.Et2[c] = F[6];
}

// First-order predictor fluxes: interface states are the adjacent cell values with the
// face-centred normal field (vanleer_mod.f90:443-467).
template <class real, int RS>
__device__ __forceinline__ real flux_pred_calc(const SweepArgs<real>& a, long long c, long long s, real* F) {
    real uL[7], uR[7];
#pragma unroll
    for (int r = 0; r < 7; r++) {
        if (r == 4) continue;
        uL[r] = a.w[r][c];
        uR[r] = a.w[r][c + s];
    }
    uL[4] = bn_at(a, c);
    uR[4] = uL[4];
    real smax = real(0);
    riemann<RS>(uL, uR, (real)a.cs, F, smax);
    return smax;
}
template <class real, int RS>
__device__ __forceinline__ double flux_pred_face(const SweepArgs<real>& a, long long c, long long s) {
    real F[7];
    const real smax = flux_pred_calc<real, RS>(a, c, s, F);
    store_flux(a, c, F);
    return (double)smax;
}

template <class real, int RS>
__global__ __launch_bounds__(256) void k_flux_pred(const BoxDesc* __restrict__ boxes, Plan plan, SweepArgs<real> a) {
    int b, i, j, k;
    bool act = plan_item_box(plan, b, i, j, k);
    double sm = 0.0;
    if (act) {
        const BoxDesc& B = boxes[b];
        sm = flux_pred_face<real, RS>(a, cell_index(B, i, j, k), dir_stride(B, a.dir));
    }
    atomic_max_nonneg(a.smax, b, sm);
}

// Primitive cell values of the rotated state r at t: rho, v_n, v_t1, v_t2, B_t1, B_t2.
template <class real>
__device__ __forceinline__ void prim_rot(real rho, real m1, real m2, real m3, real bt1, real bt2, real* W) {
    real irho = fdiv(real(1), rho);
    W[0] = rho;
    W[1] = m1 * irho;
    W[2] = m2 * irho;
    W[3] = m3 * irho;
    W[4] = bt1;
    W[5] = bt2;
}
template <class real>
__device__ __forceinline__ void prim_load(const SweepArgs<real>& a, long long ct, real* W) {
    prim_rot(a.w[0][ct], a.w[1][ct], a.w[2][ct], a.w[3][ct], a.w[5][ct], a.w[6][ct], W);
}

// PLM interface states from the primitive values of cells c-1 .. c+2 and the Riemann flux.
template <class real, int RS>
__device__ __forceinline__ real plm_flux(const SweepArgs<real>& a, const real (*W)[6], real wb, real* F) {
    const real beta = (real)a.mc_beta;
    real d1[6], d2[6];
#pragma unroll
    for (int m = 0; m < 6; m++) {
        real dL = W[1][m] - W[0][m], dR = W[2][m] - W[1][m], dC = real(0.5) * (dL + dR);
        d1[m] = sgn1(dC) * min3(beta * fabs(dL), beta * fabs(dR), fabs(dC));
        dL = W[2][m] - W[1][m]; dR = W[3][m] - W[2][m]; dC = real(0.5) * (dL + dR);
        d2[m] = sgn1(dC) * min3(beta * fabs(dL), beta * fabs(dR), fabs(dC));
    }
    real uL[7], uR[7];
    uL[0] = W[1][0] + real(0.5) * d1[0];
    uL[1] = W[1][1] + real(0.5) * d1[1];
    uL[2] = W[1][2] + real(0.5) * d1[2];
    uL[3] = W[1][3] + real(0.5) * d1[3];
    uL[4] = wb;
    uL[5] = W[1][4] + real(0.5) * d1[4];
    uL[6] = W[1][5] + real(0.5) * d1[5];
    uR[0] = W[2][0] - real(0.5) * d2[0];
    uR[1] = W[2][1] - real(0.5) * d2[1];
    uR[2] = W[2][2] - real(0.5) * d2[2];
    uR[3] = W[2][3] - real(0.5) * d2[3];
    uR[4] = wb;
    uR[5] = W[2][4] - real(0.5) * d2[4];
    uR[6] = W[2][5] - real(0.5) * d2[5];
    uL[1] = uL[1] * uL[0]; uL[2] = uL[2] * uL[0]; uL[3] = uL[3] * uL[0];
    uR[1] = uR[1] * uR[0]; uR[2] = uR[2] * uR[0]; uR[3] = uR[3] * uR[0];
    real smax = real(0);
    riemann<RS>(uL, uR, (real)a.cs, F, smax);
    return smax;
}

// Corrector fluxes: primitive PLM reconstruction with the MC-beta slope
// sign(dC) min(beta|dL|, beta|dR|, |dC|)  -- no zeroing at extrema, as in the
// reference (vanleer_mod.f90:1160-1215, 1467-1481) -- then the Riemann solver.
template <class real, int RS>
__device__ __forceinline__ real flux_corr_calc(const SweepArgs<real>& a, long long c, long long s, real* F) {
    real W[4][6];
#pragma unroll
    for (int t = 0; t < 4; t++) prim_load(a, c + (long long)(t - 1) * s, W[t]);
    return plm_flux<real, RS>(a, W, bn_at(a, c), F);
}
template <class real, int RS>
__device__ __forceinline__ double flux_corr_face(const SweepArgs<real>& a, long long c, long long s) {
    real F[7];
    const real smax = flux_corr_calc<real, RS>(a, c, s, F);
    store_flux(a, c, F);
    return (double)smax;
}

template <class real, int RS>
__global__ __launch_bounds__(256) void k_flux_corr(const BoxDesc* __restrict__ boxes, Plan plan, SweepArgs<real> a) {
    int b, i, j, k;
    bool act = plan_item_box(plan, b, i, j, k);
    double sm = 0.0;
    if (act) {
        const BoxDesc& B = boxes[b];
        sm = flux_corr_face<real, RS>(a, cell_index(B, i, j, k), dir_stride(B, a.dir));
    }
    atomic_max_nonneg(a.smax, b, sm);
}

// Corrector fluxes with PPM reconstruction (characteristic: the reference default;
// primitive: the compiled branch), six cells c-2 .. c+3 per interface.
template <class real, int RS, bool CHAR>
__device__ __forceinline__ real ppm_flux_calc(const SweepArgs<real>& a, long long c, long long s, real* F) {
    real W[6][6], wbc[6];
#pragma unroll
    for (int t = 0; t < 6; t++) {
        long long ct = c + (long long)(t - 2) * s;
        real rho = a.w[0][ct];
        real irho = fdiv(real(1), rho);
        W[t][0] = rho;
        W[t][1] = a.w[1][ct] * irho;
        W[t][2] = a.w[2][ct] * irho;
        W[t][3] = a.w[3][ct] * irho;
        W[t][4] = a.w[5][ct];
        W[t][5] = a.w[6][ct];
        wbc[t] = bn_at(a, ct);
    }
    real uL[7], uR[7];
    ppm_states<real, CHAR>(W, wbc, (real)a.cs, (real)a.mc_beta, uL, uR);
    real smax = real(0);
    riemann<RS>(uL, uR, (real)a.cs, F, smax);
    return smax;
}
template <class real, int RS, bool CHAR>
__device__ __forceinline__ double flux_corr_ppm_face(const SweepArgs<real>& a, long long c, long long s) {
    real F[7];
    const real smax = ppm_flux_calc<real, RS, CHAR>(a, c, s, F);
    store_flux(a, c, F);
    return (double)smax;
}

template <class real, int RS, bool CHAR>
__global__ __launch_bounds__(256) void k_flux_corr_ppm(const BoxDesc* __restrict__ boxes, Plan plan, SweepArgs<real> a) {
    int b, i, j, k;
    bool act = plan_item_box(plan, b, i, j, k);
    double sm = 0.0;
    if (act) {
        const BoxDesc& B = boxes[b];
        sm = flux_corr_ppm_face<real, RS, CHAR>(a, cell_index(B, i, j, k), dir_stride(B, a.dir));
    }
    atomic_max_nonneg(a.smax, b, sm);
}

// All three directional sweeps of a stage in one launch: the plan is the union of the
// three directions' ranges (their common transverse range), and a cell computes its
// d-face only when its index along d is inside direction d's normal range
// [lo_off, n_d + hi_off].  Same per-face arithmetic as the single-direction kernels; the
// cell's own state is fetched from HBM once for all three faces.
template <class real, int RS, int KIND>   // KIND 0 predictor, 1 PLM, 2 PPM primitive, 3 PPM characteristic
__global__ __launch_bounds__(256) void k_flux3(const BoxDesc* __restrict__ boxes, Plan plan, Sweep3Args<real> a) {
    int b, i, j, k;
    bool act = plan_item_box(plan, b, i, j, k);
    double sm = 0.0;
    if (act) {
        const BoxDesc& B = boxes[b];
        const long long c = cell_index(B, i, j, k);
        const int idx[3] = {i, j, k};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            if (idx[d] < a.lo_off || idx[d] > B.n[d] + a.hi_off) continue;
            const long long s = dir_stride(B, d);
            double f;
            if (KIND == 0) f = flux_pred_face<real, RS>(a.d[d], c, s);
            else if (KIND == 1) f = flux_corr_face<real, RS>(a.d[d], c, s);
            else if (KIND == 2) f = flux_corr_ppm_face<real, RS, false>(a.d[d], c, s);
            else f = flux_corr_ppm_face<real, RS, true>(a.d[d], c, s);
            sm = fmax(sm, f);
        }
    }
    atomic_max_nonneg(a.d[0].smax, b, sm);
}

template <class T>
__device__ __forceinline__ void upw(T v, T L14, T R14, T L34, T R34, T& o14, T& o34) {
    if (v > T(ZERO)) { o14 = L14; o34 = L34; }
    else if (v < -T(ZERO)) { o14 = R14; o34 = R34; }
    else { o14 = T(0.5) * (L14 + R14); o34 = T(0.5) * (L34 + R34); }
}

// Gardiner-Stone upwind corner EMFs (vanleer_mod.f90:1884-2076), summed in FP64.  The
// cell-centre reference EMF E = -v x B (vanleer_mod.f90:1823-1862) is formed on the fly.
template <class real>
__global__ __launch_bounds__(256) void k_corner(const BoxDesc* __restrict__ boxes, Plan plan, CornerArgs<real> a) {
    int b, i, j, k;
    if (!plan_item(plan, b, i, j, k)) return;
    const BoxDesc& B = boxes[b];
    const long long c = cell_index(B, i, j, k);
    const long long sx = 1, sy = B.P[0], sz = (long long)B.P[0] * B.P[1];
    auto Q = [&](int f, long long p) -> double { return (double)a.q[f][p]; };
    auto Exc = [&](long long p) -> double { double ir = 1.0 / Q(0, p); return (Q(5, p) * Q(3, p) - Q(6, p) * Q(2, p)) * ir; };
    auto Eyc = [&](long long p) -> double { double ir = 1.0 / Q(0, p); return (Q(6, p) * Q(1, p) - Q(4, p) * Q(3, p)) * ir; };
    auto Ezc = [&](long long p) -> double { double ir = 1.0 / Q(0, p); return (Q(4, p) * Q(2, p) - Q(5, p) * Q(1, p)) * ir; };
    auto E = [&](int comp, int slot, long long p) -> double { return (double)a.E[2 * comp + slot][p]; };
    const double vx = 0.5 * (Q(1, c) / Q(0, c) + Q(1, c + sx) / Q(0, c + sx));
    const double vy = 0.5 * (Q(2, c) / Q(0, c) + Q(2, c + sy) / Q(0, c + sy));
    const double vz = 0.5 * (Q(3, c) / Q(0, c) + Q(3, c + sz) / Q(0, c + sz));
    double a14, a34, b14, b34;
    {   // Ez at (i+1/2, j+1/2, k)
        double e00 = Ezc(c), e10 = Ezc(c + sx), e01 = Ezc(c + sy), e11 = Ezc(c + sx + sy);
        double y0 = E(2, 1, c), y1 = E(2, 1, c + sx), x0 = E(2, 0, c), x1 = E(2, 0, c + sy);
        upw(vx, 2.0 * (y0 - e00), 2.0 * (y1 - e10), 2.0 * (e01 - y0), 2.0 * (e11 - y1), a14, a34);
        upw(vy, 2.0 * (x0 - e00), 2.0 * (x1 - e01), 2.0 * (e10 - x0), 2.0 * (e11 - x1), b14, b34);
        a.Ec[2][c] = 0.25 * (x0 + x1 + y0 + y1) + 0.125 * ((a14 - a34) + (b14 - b34));
    }
    {   // Ey at (i+1/2, j, k+1/2)
        double e00 = Eyc(c), e10 = Eyc(c + sx), e01 = Eyc(c + sz), e11 = Eyc(c + sx + sz);
        double z0 = E(1, 1, c), z1 = E(1, 1, c + sx), x0 = E(1, 0, c), x1 = E(1, 0, c + sz);
        upw(vx, 2.0 * (z0 - e00), 2.0 * (z1 - e10), 2.0 * (e01 - z0), 2.0 * (e11 - z1), a14, a34);
        upw(vz, 2.0 * (x0 - e00), 2.0 * (x1 - e01), 2.0 * (e10 - x0), 2.0 * (e11 - x1), b14, b34);
        a.Ec[1][c] = 0.25 * (x0 + x1 + z0 + z1) + 0.125 * ((a14 - a34) + (b14 - b34));
    }
    {   // Ex at (i, j+1/2, k+1/2)
        double e00 = Exc(c), e10 = Exc(c + sy), e01 = Exc(c + sz), e11 = Exc(c + sy + sz);
        double z0 = E(0, 1, c), z1 = E(0, 1, c + sy), y0 = E(0, 0, c), y1 = E(0, 0, c + sz);
        upw(vy, 2.0 * (z0 - e00), 2.0 * (z1 - e10), 2.0 * (e01 - z0), 2.0 * (e11 - z1), a14, a34);
        upw(vz, 2.0 * (y0 - e00), 2.0 * (y1 - e01), 2.0 * (e10 - y0), 2.0 * (e11 - y1), b14, b34);
        a.Ec[0][c] = 0.25 * (y0 + y1 + z0 + z1) + 0.125 * ((a14 - a34) + (b14 - b34));
    }
}

// z-marching form of k_corner: one thread walks a z-segment of an (i, j) column.  What
// cell k reads at plane k+1 (the centre EMFs Ex, Ey and v_z there, and the slot-0 face
// EMFs of Ey and Ex) is what cell k+1 reads at its own plane, so those values are carried
// in registers; every value is formed by the same expression as in k_corner.
template <class real>
__global__ __launch_bounds__(256) void k_corner_z(const BoxDesc* __restrict__ boxes, Plan plan, CornerArgs<real> a) {
    int b, i, j, sg;
    if (!plan_item(plan, b, i, j, sg)) return;
    const BoxDesc& B = boxes[b];
    const PlanBox& pb = plan.boxes[b];
    const int kb = pb.zlo + sg * pb.seg, ke = min(kb + pb.seg, pb.zlo + pb.zn);
    const long long sx = 1, sy = B.P[0], sz = (long long)B.P[0] * B.P[1];
    long long c = cell_index(B, i, j, kb);
    auto Q = [&](int f, long long p) -> double { return (double)a.q[f][p]; };
    auto Exc = [&](long long p) -> double { double ir = 1.0 / Q(0, p); return (Q(5, p) * Q(3, p) - Q(6, p) * Q(2, p)) * ir; };
    auto Eyc = [&](long long p) -> double { double ir = 1.0 / Q(0, p); return (Q(6, p) * Q(1, p) - Q(4, p) * Q(3, p)) * ir; };
    auto Ezc = [&](long long p) -> double { double ir = 1.0 / Q(0, p); return (Q(4, p) * Q(2, p) - Q(5, p) * Q(1, p)) * ir; };
    auto E = [&](int comp, int slot, long long p) -> double { return (double)a.E[2 * comp + slot][p]; };
    // carried plane-k values
    double wz = Q(3, c) / Q(0, c);
    double ey0 = Eyc(c), ey1 = Eyc(c + sx), ex0 = Exc(c), ex1 = Exc(c + sy);
    double fy = E(1, 0, c), fx = E(0, 0, c);
#pragma unroll 1
    for (int k = kb; k < ke; k++, c += sz) {
        // one live cell index instead of a strength-reduced pointer per array and offset
        __asm__ volatile("" : "+v"(c));
        const double wzn = Q(3, c + sz) / Q(0, c + sz);
        const double vx = 0.5 * (Q(1, c) / Q(0, c) + Q(1, c + sx) / Q(0, c + sx));
        const double vy = 0.5 * (Q(2, c) / Q(0, c) + Q(2, c + sy) / Q(0, c + sy));
        const double vz = 0.5 * (wz + wzn);
        double a14, a34, b14, b34;
        {   // Ez at (i+1/2, j+1/2, k)
            double e00 = Ezc(c), e10 = Ezc(c + sx), e01 = Ezc(c + sy), e11 = Ezc(c + sx + sy);
            double y0 = E(2, 1, c), y1 = E(2, 1, c + sx), x0 = E(2, 0, c), x1 = E(2, 0, c + sy);
            upw(vx, 2.0 * (y0 - e00), 2.0 * (y1 - e10), 2.0 * (e01 - y0), 2.0 * (e11 - y1), a14, a34);
            upw(vy, 2.0 * (x0 - e00), 2.0 * (x1 - e01), 2.0 * (e10 - x0), 2.0 * (e11 - x1), b14, b34);
            a.Ec[2][c] = 0.25 * (x0 + x1 + y0 + y1) + 0.125 * ((a14 - a34) + (b14 - b34));
        }
        {   // Ey at (i+1/2, j, k+1/2)
            double e00 = ey0, e10 = ey1, e01 = Eyc(c + sz), e11 = Eyc(c + sx + sz);
            double z0 = E(1, 1, c), z1 = E(1, 1, c + sx), x0 = fy, x1 = E(1, 0, c + sz);
            upw(vx, 2.0 * (z0 - e00), 2.0 * (z1 - e10), 2.0 * (e01 - z0), 2.0 * (e11 - z1), a14, a34);
            upw(vz, 2.0 * (x0 - e00), 2.0 * (x1 - e01), 2.0 * (e10 - x0), 2.0 * (e11 - x1), b14, b34);
            a.Ec[1][c] = 0.25 * (x0 + x1 + z0 + z1) + 0.125 * ((a14 - a34) + (b14 - b34));
            ey0 = e01; ey1 = e11; fy = x1;
        }
        {   // Ex at (i, j+1/2, k+1/2)
            double e00 = ex0, e10 = ex1, e01 = Exc(c + sz), e11 = Exc(c + sy + sz);
            double z0 = E(0, 1, c), z1 = E(0, 1, c + sy), y0 = fx, y1 = E(0, 0, c + sz);
            upw(vy, 2.0 * (z0 - e00), 2.0 * (z1 - e10), 2.0 * (e01 - z0), 2.0 * (e11 - z1), a14, a34);
            upw(vz, 2.0 * (y0 - e00), 2.0 * (y1 - e01), 2.0 * (e10 - y0), 2.0 * (e11 - y1), b14, b34);
            a.Ec[0][c] = 0.25 * (y0 + y1 + z0 + z1) + 0.125 * ((a14 - a34) + (b14 - b34));
            ex0 = e01; ex1 = e11; fx = y1;
        }
        wz = wzn;
    }
}

// Unsplit conservative update of rho and rho v (vanleer_mod.f90:2080-2116) with the
// density floor (:3243-3262), CT face update (:2218-2236) and face-average cell-centred B
// (:2241-2256).  The lower faces needed by the face average are recomputed from the
// corner EMFs, so the kernel never reads a face it writes.
template <class real>
__device__ __forceinline__ void update_cell(const BoxDesc& B, const PlanBox& pb, int b, int i, int j, int k,
                                            const UpdateArgs<real>& a, double& cfl) {
    const long long c = cell_index(B, i, j, k);
    const long long sx = 1, sy = B.P[0], sz = (long long)B.P[0] * B.P[1];
    const double dtd = a.factor * a.dt[b] / B.dx;
    const real dtdx = (real)dtd;
#pragma unroll
    for (int f = 0; f < 4; f++) {
        real div = (a.F[0][f][c] - a.F[0][f][c - sx]) + (a.F[1][f][c] - a.F[1][f][c - sy]) +
                   (a.F[2][f][c] - a.F[2][f][c - sz]);
        real q0 = a.src.q[f][c];
        real v;
        if (a.comp[0] != nullptr) {  // compensated (Kahan) accumulation of the FP32 state
            real y = -(dtdx * div) - a.comp[f][c];
            v = q0 + y;
            a.comp[f][c] = (v - q0) - y;
        } else {
            v = q0 - dtdx * div;
        }
        if (f == 0 && v < (real)MIN_DENS) {
            v = (real)MIN_DENS;
            if (a.comp[0] != nullptr) a.comp[0][c] = real(0);
            atomicAdd(a.floor_count, 1u);
        }
        a.dst.q[f][c] = v;
    }
    const double* Ex = a.Ec[0];
    const double* Ey = a.Ec[1];
    const double* Ez = a.Ec[2];
    double bx = a.src.b[0][c] - dtd * ((Ez[c] - Ez[c - sy]) - (Ey[c] - Ey[c - sz]));
    double by = a.src.b[1][c] - dtd * ((Ex[c] - Ex[c - sz]) - (Ez[c] - Ez[c - sx]));
    double bz = a.src.b[2][c] - dtd * ((Ey[c] - Ey[c - sx]) - (Ex[c] - Ex[c - sy]));
    a.dst.b[0][c] = bx;
    a.dst.b[1][c] = by;
    a.dst.b[2][c] = bz;
    const int sh = a.ccb_lo_shift;
    if (i >= pb.lo[0] + sh && j >= pb.lo[1] + sh && k >= pb.lo[2] + sh) {
        long long m = c - sx;
        double bxm = a.src.b[0][m] - dtd * ((Ez[m] - Ez[m - sy]) - (Ey[m] - Ey[m - sz]));
        m = c - sy;
        double bym = a.src.b[1][m] - dtd * ((Ex[m] - Ex[m - sz]) - (Ez[m] - Ez[m - sx]));
        m = c - sz;
        double bzm = a.src.b[2][m] - dtd * ((Ey[m] - Ey[m - sx]) - (Ex[m] - Ex[m - sy]));
        a.dst.q[4][c] = (real)(0.5 * (bx + bxm));
        a.dst.q[5][c] = (real)(0.5 * (by + bym));
        a.dst.q[6][c] = (real)(0.5 * (bz + bzm));
    } else {
        a.dst.q[4][c] = a.src.q[4][c];
        a.dst.q[5][c] = a.src.q[5][c];
        a.dst.q[6][c] = a.src.q[6][c];
    }
    // CFL speed of the new state on interior cells (the driver's compute_new_time_step
    // reads the same interior values after the passive step and ghost fill)
    if (a.cmax != nullptr && i >= 1 && j >= 1 && k >= 1 && i <= B.n[0] && j <= B.n[1] && k <= B.n[2]) {
        double q7[7];
        for (int f = 0; f < 7; f++) q7[f] = (double)a.dst.q[f][c];
        cfl = cell_fast_speed(q7, a.cs);
    }
}

template <class real>
__global__ __launch_bounds__(256) void k_update(const BoxDesc* __restrict__ boxes, Plan plan, UpdateArgs<real> a) {
    int b, i, j, k;
    bool act = plan_item_box(plan, b, i, j, k);
    double cfl = 0.0;
    if (act) update_cell(boxes[b], plan.boxes[b], b, i, j, k, a, cfl);
    if (a.cmax != nullptr) atomic_max_nonneg(a.cmax, b, cfl);
}

// Column (z-marching) form of k_update: one thread walks seg cells of an (i, j) column and
// carries the k-1 values it needs -- the z-fluxes F_z(k-1), the corner EMFs Ex, Ey at k-1
// and at the (i-1)/(j-1) neighbours, and the freshly updated B_z face of k-1, which is the
// lower face of the cell-centred average -- instead of re-reading the k-1 plane.  The
// arithmetic per cell is update_cell's.
template <class real>
__global__ __launch_bounds__(256) void k_update_z(const BoxDesc* __restrict__ boxes, Plan plan, UpdateArgs<real> a) {
    int b, i, j, sg;
    bool act = plan_item_box(plan, b, i, j, sg);
    double cfl = 0.0;
    if (act) {
        const BoxDesc& B = boxes[b];
        const PlanBox& pb = plan.boxes[b];
        const int kb = pb.zlo + sg * pb.seg, ke = min(kb + pb.seg, pb.zlo + pb.zn);
        const long long sx = 1, sy = B.P[0], sz = (long long)B.P[0] * B.P[1];
        const double dtd = a.factor * a.dt[b] / B.dx;
        const real dtdx = (real)dtd;
        const double* Ex = a.Ec[0];
        const double* Ey = a.Ec[1];
        const double* Ez = a.Ec[2];
        const int sh = a.ccb_lo_shift;
        const bool ccb_ij = i >= pb.lo[0] + sh && j >= pb.lo[1] + sh;
        const bool int_ij = i >= 1 && j >= 1 && i <= B.n[0] && j <= B.n[1];
        long long c = cell_index(B, i, j, kb);
        // carried k-1 values, loaded once per segment
        real fz[4];
#pragma unroll
        for (int f = 0; f < 4; f++) fz[f] = a.F[2][f][c - sz];
        double ey_m = Ey[c - sz], ex_m = Ex[c - sz], eyx_m = Ey[c - sx - sz], exy_m = Ex[c - sy - sz];
        double bz_m = a.src.b[2][c - sz] - dtd * ((ey_m - eyx_m) - (ex_m - exy_m));
#pragma unroll 1
        for (int k = kb; k < ke; k++, c += sz) {
            // keep one live cell index instead of a strength-reduced pointer per array and
            // neighbour offset (register pressure): addresses are formed from c each step
            __asm__ volatile("" : "+v"(c));
#pragma unroll
            for (int f = 0; f < 4; f++) {
                const real fzc = a.F[2][f][c];
                real div = (a.F[0][f][c] - a.F[0][f][c - sx]) + (a.F[1][f][c] - a.F[1][f][c - sy]) + (fzc - fz[f]);
                fz[f] = fzc;
                real q0 = a.src.q[f][c];
                real v;
                if (a.comp[0] != nullptr) {
                    real y = -(dtdx * div) - a.comp[f][c];
                    v = q0 + y;
                    a.comp[f][c] = (v - q0) - y;
                } else {
                    v = q0 - dtdx * div;
                }
                if (f == 0 && v < (real)MIN_DENS) {
                    v = (real)MIN_DENS;
                    if (a.comp[0] != nullptr) a.comp[0][c] = real(0);
                    atomicAdd(a.floor_count, 1u);
                }
                a.dst.q[f][c] = v;
            }
            const double ex = Ex[c], ey = Ey[c], ez = Ez[c];
            const double ez_y = Ez[c - sy], ez_x = Ez[c - sx], ey_x = Ey[c - sx], ex_y = Ex[c - sy];
            double bx = a.src.b[0][c] - dtd * ((ez - ez_y) - (ey - ey_m));
            double by = a.src.b[1][c] - dtd * ((ex - ex_m) - (ez - ez_x));
            double bz = a.src.b[2][c] - dtd * ((ey - ey_x) - (ex - ex_y));
            a.dst.b[0][c] = bx;
            a.dst.b[1][c] = by;
            a.dst.b[2][c] = bz;
            if (ccb_ij && k >= pb.zlo + sh) {
                long long m = c - sx;
                double bxm = a.src.b[0][m] - dtd * ((ez_x - Ez[m - sy]) - (ey_x - eyx_m));
                m = c - sy;
                double bym = a.src.b[1][m] - dtd * ((ex_y - exy_m) - (ez_y - Ez[m - sx]));
                a.dst.q[4][c] = (real)(0.5 * (bx + bxm));
                a.dst.q[5][c] = (real)(0.5 * (by + bym));
                a.dst.q[6][c] = (real)(0.5 * (bz + bz_m));
            } else {
                a.dst.q[4][c] = a.src.q[4][c];
                a.dst.q[5][c] = a.src.q[5][c];
                a.dst.q[6][c] = a.src.q[6][c];
            }
            ey_m = ey;
            ex_m = ex;
            eyx_m = ey_x;
            exy_m = ex_y;
            bz_m = bz;
            if (a.cmax != nullptr && int_ij && k >= 1 && k <= B.n[2]) {
                double q7[7];
                for (int f = 0; f < 7; f++) q7[f] = (double)a.dst.q[f][c];
                cfl = fmax(cfl, cell_fast_speed(q7, a.cs));
            }
        }
    }
    if (a.cmax != nullptr) atomic_max_nonneg(a.cmax, b, cfl);
}

template <class real, int RS>
void launch_flux_predictor(const Launch& L, const SweepArgs<real>& a) {
    if (L.plan.total <= 0) return;
    hipLaunchKernelGGL((k_flux_pred<real, RS>), dim3(L.blocks()), dim3(256), 0, L.stream, L.boxes, L.plan, a);
}
template <class real, int RS>
void launch_flux_corrector(const Launch& L, const SweepArgs<real>& a) {
    if (L.plan.total <= 0) return;
    if (a.recon == 2)
        hipLaunchKernelGGL((k_flux_corr_ppm<real, RS, true>), dim3(L.blocks()), dim3(256), 0, L.stream, L.boxes, L.plan, a);
    else if (a.recon == 1)
        hipLaunchKernelGGL((k_flux_corr_ppm<real, RS, false>), dim3(L.blocks()), dim3(256), 0, L.stream, L.boxes, L.plan, a);
    else
        hipLaunchKernelGGL((k_flux_corr<real, RS>), dim3(L.blocks()), dim3(256), 0, L.stream, L.boxes, L.plan, a);
}
template <class real, int RS>
void launch_flux3(const Launch& L, const Sweep3Args<real>& a, int kind) {
    if (L.plan.total <= 0) return;
    const dim3 g(L.blocks()), t(256);
    if (kind == 0) hipLaunchKernelGGL((k_flux3<real, RS, 0>), g, t, 0, L.stream, L.boxes, L.plan, a);
    else if (kind == 1) hipLaunchKernelGGL((k_flux3<real, RS, 1>), g, t, 0, L.stream, L.boxes, L.plan, a);
    else if (kind == 2) hipLaunchKernelGGL((k_flux3<real, RS, 2>), g, t, 0, L.stream, L.boxes, L.plan, a);
    else hipLaunchKernelGGL((k_flux3<real, RS, 3>), g, t, 0, L.stream, L.boxes, L.plan, a);
}
template <class real>
void launch_corner(const Launch& L, const CornerArgs<real>& a) {
    if (L.plan.total <= 0) return;
    if (L.column)
        hipLaunchKernelGGL((k_corner_z<real>), dim3(L.blocks()), dim3(256), 0, L.stream, L.boxes, L.plan, a);
    else
        hipLaunchKernelGGL((k_corner<real>), dim3(L.blocks()), dim3(256), 0, L.stream, L.boxes, L.plan, a);
}
template <class real>
void launch_update(const Launch& L, const UpdateArgs<real>& a) {
    if (L.plan.total <= 0) return;
    if (L.column)
        hipLaunchKernelGGL((k_update_z<real>), dim3(L.blocks()), dim3(256), 0, L.stream, L.boxes, L.plan, a);
    else
        hipLaunchKernelGGL((k_update<real>), dim3(L.blocks()), dim3(256), 0, L.stream, L.boxes, L.plan, a);
}

}  // namespace imhd
#include "vl_stage_impl.hpp"
namespace imhd {

#define IMHD_INST_RS(T, RS)                                                              \
    template void launch_flux_predictor<T, RS>(const Launch&, const SweepArgs<T>&); \
    template void launch_flux_corrector<T, RS>(const Launch&, const SweepArgs<T>&); \
    template void launch_flux3<T, RS>(const Launch&, const Sweep3Args<T>&, int);     \
    template void launch_vl_stage<T, RS>(hipStream_t, const BoxDesc*, const TilePlan&, const StageArgs<T>&, int, int); \
    template int vl_stage_blocks_per_cu<T, RS>(int, int);
#define IMHD_INST_VL(T)                                                      \
    IMHD_INST_RS(T, RS_ROE)                                                  \
    IMHD_INST_RS(T, RS_HLLD)                                                 \
    IMHD_INST_RS(T, RS_HLLE)                                                 \
    template void launch_corner<T>(const Launch&, const CornerArgs<T>&);     \
    template void launch_update<T>(const Launch&, const UpdateArgs<T>&);
}  // namespace imhd
