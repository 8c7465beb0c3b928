// One VL2 + CT stage in one launch: directional fluxes (vanleer_mod.f90:429-938), upwind
// corner EMFs (:1884-2076) and the conservative + CT update (:2080-2259), with the per-face
// and per-corner arithmetic of flux_*_calc, k_corner and k_update_z, so every value has
// the same bits as on the three-kernel path.  Fluxes, face EMFs and corner EMFs never go
// to HBM: a workgroup holds a TX x TY tile of (i, j) columns (one-cell halo) and
// marches a z-segment; per plane it
//   A  solves the x- and y-faces of plane k+1 and the z-face k+1/2, and forms the centre
//      EMFs and velocities of plane k+1 (face EMFs and centre values double-buffered by
//      plane parity in LDS);
//   D  forms the corner EMFs Ez(k), Ey(k+1/2), Ex(k+1/2), and the x/y flux differences of
//      plane k+1 while its fluxes are in LDS;
//   G  updates the cells of plane k, keeping the index-k corner EMFs it reads for plane k+1.
// Two barriers per plane; FP32 LDS: 40 KB per 16 x 16 tile, 80 KB per 32 x 16 tile.
// Every face, corner and cell is masked to the reference's own index range for that
// quantity, so the reads and the signal-speed maximum are those of the separate kernels.
//
// The FP32 PLM corrector (SH below) reconstructs from primitive values kept in LDS: each
// plane's primitives are formed once per cell (from the cell's own load of plane k+2, one
// march step ahead) into a plane-parity buffer; the x-slopes are formed once per cell and
// passed to the x-neighbour by a DPP lane shift, the y-faces read their four-cell stencil
// from the buffer, and the z-slope of a plane is formed once and carried to the next
// plane's face.  Only a tile's edge lanes load neighbour cells from HBM.
#pragma once
#include <type_traits>
#include "kernels.hpp"

namespace imhd {

// waves per SIMD the register allocation must allow: in FP32 the LDS of a 16 x 16 tile
// (40 KB) fits 4 workgroups per CU and that of a 32 x 16 tile (80 KB) 2, 16 waves per CU
// either way; FP64 needs more of both and gets 2 waves per SIMD; the characteristic PPM
// corrector (KIND 3: six-cell stencils and 6 x 6 eigenvector matrices, about 210
// registers) gets 2 in FP32 too, rather than spill at 128; the shared-slope PLM corrector
// (KIND 4) gets 3 on 16 x 16 tiles (three workgroups per CU) and 4 otherwise
template <class real, int KIND, int N>
constexpr int vs_waves() {
    if constexpr (sizeof(real) == 8) return 1;
    if constexpr (KIND == 4) return N == 256 ? 3 : 4;   // 16 x 16: 168 registers, three workgroups per CU
    if constexpr (KIND == 5) return 4;
    return KIND == 3 ? 2 : 4;
}
#define VS_WAVES (vs_waves<real, KIND, TX * TY>())

// x or y face (direction D) of the fused stage.  Predictor: the cell c on the low side is
// the thread's own, held in registers in the global field order, and only the neighbour is
// loaded (same arithmetic as flux_pred_calc).  PLM corrector (KIND 1): flux_corr_calc's
// arithmetic with the own cell from registers and the other three of its four-cell
// stencil loaded.  PPM correctors (KIND 2 primitive, 3 characteristic): ppm_flux_calc,
// which loads its six-cell stencil.
template <class real, int RS, int KIND, int D>
__device__ __forceinline__ real stage_face_own(const SweepArgs<real>& a, const real* U, long long c, long long s,
                                               real* F) {
    if constexpr (KIND == 0) {
        real uL[7], uR[7];
#pragma unroll
        for (int r = 0; r < 7; r++) {
            if (r == 4) continue;
            uL[r] = U[perm(D, r)];
            uR[r] = a.w[r][c + s];
        }
        uL[4] = bn_at(a, c);
        uR[4] = uL[4];
        real smax = real(0);
        riemann<RS>(uL, uR, (real)a.cs, F, smax);
        return smax;
    } else if constexpr (KIND == 1) {
        real W[4][6];
        prim_load(a, c - s, W[0]);
        prim_rot(U[0], U[perm(D, 1)], U[perm(D, 2)], U[perm(D, 3)], U[perm(D, 5)], U[perm(D, 6)], W[1]);
        prim_load(a, c + s, W[2]);
        prim_load(a, c + 2 * s, W[3]);
        return plm_flux<real, RS>(a, W, bn_at(a, c), F);
    } else {
        return ppm_flux_calc<real, RS, KIND == 3>(a, c, s, F);
    }
}
// z face k+1/2: predictor between the thread's own cells of planes k (UL) and k+1 (UR),
// both in registers; FP64 corrector with those two of its four stencil cells from
// registers; FP32 corrector as above (holding them spills at 128 VGPRs).
template <class real, int RS, int KIND>
__device__ __forceinline__ real stage_face_z(const SweepArgs<real>& a, const real* UL, const real* UR, long long c,
                                             long long s, real* F) {
    if constexpr (KIND == 0) {
        real uL[7], uR[7];
#pragma unroll
        for (int r = 0; r < 7; r++) {
            if (r == 4) continue;
            uL[r] = UL[perm(2, r)];
            uR[r] = UR[perm(2, r)];
        }
        uL[4] = bn_at(a, c);
        uR[4] = uL[4];
        real smax = real(0);
        riemann<RS>(uL, uR, (real)a.cs, F, smax);
        return smax;
    } else if constexpr (KIND >= 2) {
        return ppm_flux_calc<real, RS, KIND == 3>(a, c, s, F);
    } else if constexpr (sizeof(real) == 8) {
        // FP64 (2 waves per SIMD, registers to spare): planes k and k+1 from registers
        real W[4][6];
        prim_load(a, c - s, W[0]);
        prim_rot(UL[0], UL[perm(2, 1)], UL[perm(2, 2)], UL[perm(2, 3)], UL[perm(2, 5)], UL[perm(2, 6)], W[1]);
        prim_rot(UR[0], UR[perm(2, 1)], UR[perm(2, 2)], UR[perm(2, 3)], UR[perm(2, 5)], UR[perm(2, 6)], W[2]);
        prim_load(a, c + 2 * s, W[3]);
        return plm_flux<real, RS>(a, W, bn_at(a, c), F);
    } else {
        return flux_corr_calc<real, RS>(a, c, s, F);
    }
}

// Values of the lanes l - 1 and l + 1 of the wavefront (DPP wave_shr:1 / wave_shl:1, all
// lanes active).  A tile row is TX lanes (16 or 32, dividing 64), so within a row these are
// the x-neighbours; at a row's first and last lane they are not, and those lanes use loaded
// values instead.
// (bound_ctrl: the lane without a source lane reads 0; no old value to initialise)
__device__ __forceinline__ float lane_prev(float x) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float lane_next(float x) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ double lane_prev_d(double x) {
    const long long v = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)v, 0x138, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(v >> 32), 0x138, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// MC-beta limited slope sign(dC) min(beta|dL|, beta|dR|, |dC|) of one primitive from three
// cell values (plm_flux's arithmetic)
template <class real>
__device__ __forceinline__ real mc_slope(real wm, real w0, real wp, real beta) {
    const real dL = w0 - wm, dR = wp - w0, dC = real(0.5) * (dL + dR);
    // sign(dC) m for m >= 0 is copysign(m, dC), bit for bit (a multiply by +-1 is exact)
    return copysign(min3(beta * fabs(dL), beta * fabs(dR), fabs(dC)), dC);
}
// Primitive slot r (rho, v_n, v_t1, v_t2, B_t1, B_t2) of direction D in a primitive vector
// in global order (rho, v_x, v_y, v_z, B_x, B_y, B_z)
template <int D>
__host__ __device__ constexpr int rslot(int r) { return perm(D, r < 4 ? r : r + 1); }
// Riemann flux from PLM interface primitives (rotated, 6 each) and the face's normal field:
// plm_flux's interface states and call
template <class real, int RS>
__device__ __forceinline__ real plm_solve(const real* sl, const real* sr, real wb, real cs, real* F) {
    real uL[7], uR[7];
    uL[0] = sl[0]; uL[4] = wb; uL[5] = sl[4]; uL[6] = sl[5];
    uR[0] = sr[0]; uR[4] = wb; uR[5] = sr[4]; uR[6] = sr[5];
    uL[1] = sl[1] * uL[0]; uL[2] = sl[2] * uL[0]; uL[3] = sl[3] * uL[0];
    uR[1] = sr[1] * uR[0]; uR[2] = sr[2] * uR[0]; uR[3] = sr[3] * uR[0];
    real smax = real(0);
    if constexpr (RS == RS_HLLD && sizeof(real) == 4) hlld<real, true>(uL, uR, cs * cs, F, smax, sl + 1, sr + 1);
    else riemann<RS>(uL, uR, cs, F, smax);
    return smax;
}

// Direction d's sweep view of the flux state: the rotated slots are the global fields
// permuted at compile time, so the three views share the seven field pointers (and their
// scalar registers) instead of passing 21 copies.
// The predictor's (kinds 0 and 5) flux state is the state at t^n, whose faces are never the
// FP32 half-step ones, so its view reads the FP64 faces without testing.
template <class real, int D, int KIND>
__device__ __forceinline__ SweepArgs<real> stage_sweep(const StageArgs<real>& a) {
    SweepArgs<real> s{};
#pragma unroll
    for (int r = 0; r < 7; r++) s.w[r] = r == 4 ? nullptr : a.qf[perm(D, r)];
    s.bn = a.bf[D];
    s.bnr = KIND == 0 || KIND == 5 ? nullptr : a.bfr[D];
    s.dir = D;
    s.cs = a.cs;
    s.mc_beta = a.mc_beta;
    return s;
}

template <class real, int RS, int KIND, int TX, int TY>
__global__ __launch_bounds__(TX * TY, VS_WAVES) void k_vl_stage(const BoxDesc* __restrict__ boxes, TilePlan plan,
                                                      StageArgs<real> a) {
    constexpr int N = TX * TY;
    // FP32 keeps one buffer of the x/y fluxes and of the corner Ex, Ey (forming the flux
    // differences one plane ahead and carrying the index k-1 corners in registers): 40 KB of
    // LDS per 16 x 16 tile.  FP64, which registers limit to 2 waves per SIMD anyway, keeps
    // two (by plane / index parity) and fewer live registers.
    constexpr bool ONE = sizeof(real) == 4;
    constexpr int NB = ONE ? 1 : 2;
    // FP32 PLM corrector: primitives of two planes in LDS (28 KB per 32 x 16 tile), x-flux
    // differences by lane shift (no x slot in sF), centre values in FP32: 80 KB in all.  Its
    // flux state is the half step, whose normal faces are the FP32 ones in a.bfr.
    constexpr bool SH = ONE && KIND == 4;
    // FP32 predictor (KIND 5) the same way: the own cell of plane k+2 to a plane-parity LDS
    // buffer, the x-neighbour by DPP lane shift, the y-neighbour from the buffer
    constexpr bool SHP = ONE && KIND == 5;
    constexpr bool SHX = SH || SHP;
    constexpr bool PRED = KIND == 0 || KIND == 5;
    constexpr int FK = KIND == 4 ? 1 : (KIND == 5 ? 0 : KIND);   // per-face arithmetic of the other paths
    constexpr int FY = SHX ? 0 : 1;   // sF slot of the y face
    using ctype = typename std::conditional<SHX, float, double>::type;
    __shared__ real sF[NB][SHX ? 1 : 2][4][N];   // [buffer][x, y face][flux, global order]
    __shared__ real sEf[2][2][N];      // [plane parity][x face: E(2,0); y face: E(2,1)]
    __shared__ real sEz3[2][N];        // z face k+1/2: E(1,1), E(0,1)
    __shared__ ctype sC[2][5][N];      // [plane parity][centre Ex, Ey, Ez, m_x/rho, m_y/rho]
    // corner EMFs, summed in FP64; FP32 stores them in FP32 (their inputs, the FP32 face
    // EMFs, carry FP32 rounding already) and keeps two buffers of Ex, Ey, so that the update
    // reads the index k-1 corners instead of carrying them in registers.  The LDS-stencil
    // stages, whose centre values are FP32 as well, also sum them in FP32 (mtype): an FP64
    // sum of FP32 inputs that is rounded to FP32 for storage gains a few FP32 ulps at most
    // and costs a conversion per input.  The face-field update stays FP64.
    using etype = typename std::conditional<ONE, float, double>::type;
    using mtype = ctype;
    constexpr int NBE = ONE ? 2 : NB;
    __shared__ etype sEz[N];           // corner Ez of plane k
    __shared__ etype sEc[NBE][2][N];   // [buffer][Ex, Ey] corner EMFs of index k (at k+1/2)
    // SH: [plane parity][primitive, global order]; SHP: [plane parity][conserved, global order]
    __shared__ real sW[SHX ? 2 : 1][SHX ? 7 : 1][SHX ? N : 1];
    // SH: the x-edge lanes' outer neighbours (rotated primitives, direction 0), loaded with
    // the row loads of plane k+2, whose cache lines they share, and read back one plane
    // later by the same lane: [plane parity][slot][2 ty + (lane at the high edge)]
    __shared__ real sGx[SH ? 2 : 1][SH ? 6 : 1][SH ? 2 * TY : 1];

    const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
    int b = 0;
    if (plan.nbox > 1) {
        int lo = 0, hi = plan.nbox - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (plan.boxes[mid].start <= (long long)blk) lo = mid; else hi = mid - 1;
        }
        b = lo;
    }
    const TileBox& tb = plan.boxes[b];
    const long long loc = (long long)blk - tb.start;
    const long long ntile = (long long)tb.nt[0] * tb.nt[1];
    long long tile;
    int sg, nsg;
    if (loc < (long long)tb.s * ntile) {
        sg = (int)(loc / ntile);
        tile = loc - (long long)sg * ntile;
        nsg = tile < tb.k1 ? tb.s + 1 : tb.s;
    } else {
        sg = tb.s;
        tile = loc - (long long)tb.s * ntile;
        nsg = tb.s + 1;
    }
    const int tX = (int)(tile % tb.nt[0]), tY = (int)(tile / tb.nt[0]);
    const BoxDesc& B = boxes[b];
    const int t = threadIdx.x, tx = t % TX, ty = t / TX;
    const int i = tb.x0 + tX * (TX - 2) + tx, j = tb.y0 + tY * (TY - 2) + ty;
    const long long L = tb.zhi - tb.zlo + 1;
    const int ks = tb.zlo + (int)(sg * L / nsg), ke = tb.zlo + (int)((sg + 1) * L / nsg);
    const int nb = B.nb, n0 = B.n[0], n1 = B.n[1], n2 = B.n[2];
    const long long sx = 1, sy = B.P[0], sz = (long long)B.P[0] * B.P[1];
    auto inr = [](int v, int lo, int hi) { return v >= lo && v <= hi; };
    const bool cell_ij = inr(i, 1 - nb, n0 + nb) && inr(j, 1 - nb, n1 + nb);
    const bool fx_ij = inr(i, a.fn_lo, n0 + a.fn_hi) && inr(j, a.ft_lo, n1 + a.ft_hi);
    const bool fy_ij = inr(j, a.fn_lo, n1 + a.fn_hi) && inr(i, a.ft_lo, n0 + a.ft_hi);
    const bool fz_ij = inr(i, a.ft_lo, n0 + a.ft_hi) && inr(j, a.ft_lo, n1 + a.ft_hi);
    const bool c_ij = tx < TX - 1 && ty < TY - 1 && inr(i, a.c_lo, n0 + a.c_hi) && inr(j, a.c_lo, n1 + a.c_hi);
    const bool u_ij = tx >= 1 && tx < TX - 1 && ty >= 1 && ty < TY - 1 && inr(i, a.u_lo, n0 + a.u_hi) &&
                      inr(j, a.u_lo, n1 + a.u_hi);
    const int sh = a.ccb_lo_shift;
    const bool ccb_ij = i >= a.u_lo + sh && j >= a.u_lo + sh;
    const bool int_ij = i >= 1 && j >= 1 && i <= n0 && j <= n1;
    // The per-lane conditions as bits of one vector register.  FP32 re-reads it every plane
    // (an empty asm marks it as rewritten), so that the conditions are formed where used
    // instead of being held as 64-bit lane masks in scalar registers across the march, which
    // the allocator would spill.
    enum { LB_CELL, LB_FX, LB_FY, LB_FZ, LB_C, LB_U, LB_CCB, LB_INT, LB_X0, LB_XE, LB_Y0, LB_YE, LB_XF, LB_YF };
    unsigned lane_bits = (unsigned)cell_ij << LB_CELL | (unsigned)fx_ij << LB_FX | (unsigned)fy_ij << LB_FY |
                         (unsigned)fz_ij << LB_FZ | (unsigned)c_ij << LB_C | (unsigned)u_ij << LB_U |
                         (unsigned)ccb_ij << LB_CCB | (unsigned)int_ij << LB_INT | (unsigned)(tx == 0) << LB_X0 |
                         (unsigned)(tx == TX - 1) << LB_XE | (unsigned)(ty == 0) << LB_Y0 |
                         (unsigned)(ty == TY - 2) << LB_YE | (unsigned)(tx < TX - 1) << LB_XF |
                         (unsigned)(ty < TY - 1) << LB_YF;
    auto lb = [&](int bit) -> bool { return (lane_bits >> bit) & 1u; };
    const double dtd = a.factor * a.dt[b] / B.dx;
    const real dtdx = (real)dtd;
    const SweepArgs<real> sw0 = stage_sweep<real, 0, KIND>(a), sw1 = stage_sweep<real, 1, KIND>(a),
                          sw2 = stage_sweep<real, 2, KIND>(a);
    // update base face field: the predictor's base is its flux state
    const double* const sb[3] = {a.src.b[0], a.src.b[1], a.src.b[2]};

    // the thread's own cell of plane p of the flux state, global field order: loaded once
    // per plane for the centre values, the three faces and (predictor) the update
    auto own = [&](int p, real* U) {
        if (lb(LB_CELL) && inr(p, 1 - nb, n2 + nb)) {
            const long long c = cell_index(B, i, j, p);
#pragma unroll
            for (int f = 0; f < 7; f++) U[f] = a.qf[f][c];
        } else {
#pragma unroll
            for (int f = 0; f < 7; f++) U[f] = real(0);
        }
    };
    // centre EMFs and velocities of plane p (k_corner's Exc, Eyc, Ezc, m/rho); m_z/rho returned
    auto cells = [&](int p, const real* U) -> double {
        double vz = 0.0;
        const int q = p & 1;
        if (lb(LB_CELL) && inr(p, 1 - nb, n2 + nb)) {
            const double r = (double)U[0], m1 = (double)U[1], m2 = (double)U[2], m3 = (double)U[3],
                         b4 = (double)U[4], b5 = (double)U[5], b6 = (double)U[6];
            const double ir = 1.0 / r;
            sC[q][0][t] = (b5 * m3 - b6 * m2) * ir;
            sC[q][1][t] = (b6 * m1 - b4 * m3) * ir;
            sC[q][2][t] = (b4 * m2 - b5 * m1) * ir;
            if (ONE) {   // FP32 path: m (1/rho) for the velocities (FP64 keeps the division)
                sC[q][3][t] = m1 * ir;
                sC[q][4][t] = m2 * ir;
                vz = m3 * ir;
            } else {
                sC[q][3][t] = m1 / r;
                sC[q][4][t] = m2 / r;
                vz = m3 / r;
            }
        }
        return vz;
    };
    // x- and y-faces of plane p: fluxes and Ez slots to LDS; the Ey (x-face) and Ex (y-face)
    // slots stay with the thread (only its own column reads them)
    auto xyfaces = [&](int p, const real* U, real& ey_x, real& ex_y, real& sm) {
        const int q = p & 1, fq = ONE ? 0 : q;
        const long long c = cell_index(B, i, j, p);
        const bool pt = inr(p, a.ft_lo, n2 + a.ft_hi);
        real F[7];
        if (lb(LB_FX) && pt) {
            sm = fmax(sm, stage_face_own<real, RS, FK, 0>(sw0, U, c, sx, F));
            sF[fq][0][0][t] = F[0]; sF[fq][0][1][t] = F[1]; sF[fq][0][2][t] = F[2]; sF[fq][0][3][t] = F[3];
            sEf[q][0][t] = -F[5];
            ey_x = F[6];
        } else {
            sF[fq][0][0][t] = sF[fq][0][1][t] = sF[fq][0][2][t] = sF[fq][0][3][t] = real(0);
            sEf[q][0][t] = real(0);
            ey_x = real(0);
        }
        if (lb(LB_FY) && pt) {
            sm = fmax(sm, stage_face_own<real, RS, FK, 1>(sw1, U, c, sy, F));
            sF[fq][FY][0][t] = F[0]; sF[fq][FY][2][t] = F[1]; sF[fq][FY][3][t] = F[2]; sF[fq][FY][1][t] = F[3];
            ex_y = -F[5];
            sEf[q][1][t] = F[6];
        } else {
            sF[fq][FY][0][t] = sF[fq][FY][1][t] = sF[fq][FY][2][t] = sF[fq][FY][3][t] = real(0);
            sEf[q][1][t] = real(0);
            ex_y = real(0);
        }
    };
    // z-face between planes p and p+1
    auto zface = [&](int p, bool live, const real* UL, const real* UR, real* fz, real& sm) {
        if (live && lb(LB_FZ) && inr(p, a.fn_lo, n2 + a.fn_hi)) {
            real F[7];
            sm = fmax(sm, stage_face_z<real, RS, FK>(sw2, UL, UR, cell_index(B, i, j, p), sz, F));
            fz[0] = F[0]; fz[3] = F[1]; fz[1] = F[2]; fz[2] = F[3];
            sEz3[0][t] = -F[5];
            sEz3[1][t] = F[6];
        } else {
            fz[0] = fz[1] = fz[2] = fz[3] = real(0);
            sEz3[0][t] = sEz3[1][t] = real(0);
        }
    };

    // SH: conserved values of cell (ii, jj) of plane p of the flux state, zero outside the
    // box's cells: all seven in global order (load_cell), or the six of direction D's
    // rotated primitive slots (load_rot: rho, m_n, m_t1, m_t2, B_t1, B_t2); to_prims turns
    // either into primitives with prim_rot's arithmetic (zero stays zero)
    auto cell_ok = [&](int ii, int jj, int p) {
        return inr(ii, 1 - nb, n0 + nb) && inr(jj, 1 - nb, n1 + nb) && inr(p, 1 - nb, n2 + nb);
    };
    auto load_cell = [&](int ii, int jj, int p, real* u) {
        if (cell_ok(ii, jj, p)) {
            const long long c = cell_index(B, ii, jj, p);
#pragma unroll
            for (int f = 0; f < 7; f++) u[f] = a.qf[f][c];
        } else {
#pragma unroll
            for (int f = 0; f < 7; f++) u[f] = real(0);
        }
    };
    auto load_rot = [&](auto dir, int ii, int jj, int p, real* g) {
        constexpr int D = decltype(dir)::value;
        if (cell_ok(ii, jj, p)) {
            const long long c = cell_index(B, ii, jj, p);
            g[0] = a.qf[0][c];
            g[1] = a.qf[perm(D, 1)][c];
            g[2] = a.qf[perm(D, 2)][c];
            g[3] = a.qf[perm(D, 3)][c];
            g[4] = a.qf[perm(D, 5)][c];
            g[5] = a.qf[perm(D, 6)][c];
        }
    };
    auto to_prims = [&](real* u) {
        const real ir = u[0] != real(0) ? fdiv(real(1), u[0]) : real(0);
        u[1] = u[1] * ir;
        u[2] = u[2] * ir;
        u[3] = u[3] * ir;
    };

    real sm = real(0), cfl = real(0);   // running maxima (the working precision holds them exactly)
    real fz_lo[4], fz_hi[4];
    real eyx_lo, eyx_hi, exy_lo, exy_hi;   // own x-face Ey slot and y-face Ex slot, planes k, k+1
    real U0[7], U1[7];   // own cell of planes k and k+1
    ctype vz_lo = ctype(0), vz_hi;
    eyx_lo = exy_lo = real(0);
#pragma unroll
    for (int f = 0; f < 4; f++) fz_lo[f] = fz_hi[f] = real(0);
    double bz_m = 0.0;
    bool have_bzm = false;
    // carried from the update phase of plane k-1: corner EMFs of index k-1 (own and the
    // i-1 / j-1 neighbours), and the x/y flux differences of plane k (formed one plane ahead,
    // while the fluxes of that plane are in LDS)
    double ey_m = 0.0, ex_m = 0.0, eyx_m = 0.0, exy_m = 0.0;
    real dxy[4] = {real(0), real(0), real(0), real(0)}, dxy_n[4];
    // SH: x-flux differences of plane k+1 (by lane shift), and the upper PLM state of plane k
    // at the z-face k+1/2 (rotated primitives), formed with plane k's slope one step earlier
    real dxn[4] = {real(0), real(0), real(0), real(0)}, uLz[6];
    // SHP: the predictor's flux state is its update base, so the face field its x- and
    // y-faces of plane k+1 read is the update's of the next plane: carried (plane k)
    double cbx = 0.0, cby = 0.0;
    const int gslot = SH ? 2 * ty + (lb(LB_XE) ? 1 : 0) : 0;
    if constexpr (SH) {
        real W[7];
        load_cell(i, j, ks - 2, W);
        to_prims(W);
#pragma unroll
        for (int f = 0; f < 7; f++) sW[(ks - 2) & 1][f][t] = W[f];
        load_cell(i, j, ks - 1, W);
        to_prims(W);
#pragma unroll
        for (int f = 0; f < 7; f++) sW[(ks - 1) & 1][f][t] = W[f];
        if (lb(LB_X0) || lb(LB_XE)) {
            real g[6] = {real(0), real(0), real(0), real(0), real(0), real(0)};
            load_rot(std::integral_constant<int, 0>{}, lb(LB_X0) ? i - 1 : i + 1, j, ks - 1, g);
            to_prims(g);
#pragma unroll
            for (int r = 0; r < 6; r++) sGx[(ks - 1) & 1][r][gslot] = g[r];
        }
#pragma unroll
        for (int r = 0; r < 6; r++) uLz[r] = real(0);
        __syncthreads();
    }
    if constexpr (SHP) {
        real W[7];
        load_cell(i, j, ks - 1, W);
#pragma unroll
        for (int f = 0; f < 7; f++) sW[(ks - 1) & 1][f][t] = W[f];
        __syncthreads();
    }
    // The first pass (k = ks - 2) only forms plane ks - 1.  Every plane's faces and centre
    // values come from this one copy of the loop body, so a plane shared by two z-segments
    // (and the corner EMFs formed from it) gets the same bits in both, also where FMA
    // contraction is on.
#pragma unroll 1
    for (int k = ks - 2; k < ke; k++) {
        if constexpr (ONE) asm volatile("" : "+v"(lane_bits));
        const bool live = k >= ks - 1;   // uniform over the workgroup
        double fbx = 0.0, fby = 0.0, fbz = 0.0;   // SHP: face field at (i, j, k+1) x, y and (i, j, k) z
        // A: plane k+1 and the z-face k+1/2
        if constexpr (SH) {
            const int q1 = (k + 1) & 1, q2 = k & 1;   // buffers of planes k+1 and k (refilled with k+2)
            const int p = k + 1;
            const bool pt = inr(p, a.ft_lo, n2 + a.ft_hi);
            const real beta = (real)a.mc_beta, cs = (real)a.cs;
            const long long c = cell_index(B, i, j, p);
            const bool zlive = live && lb(LB_FZ) && inr(k, a.fn_lo, n2 + a.fn_hi);
            const bool ylive = lb(LB_FY) && pt && lb(LB_YF);   // row TY-1's y-faces are formed by the next tile row
            const bool xlive = lb(LB_FX) && pt && lb(LB_XF);   // column TX-1's x-faces by the next tile
            // the phase's HBM loads first, all in flight together: the own cell of plane k+2,
            // the edge lanes' outer neighbours and the faces' normal field (index 0 for idle lanes)
            real u2[7], gy[6], gx2[6];
            load_cell(i, j, k + 2, u2);
            gy[0] = gy[1] = gy[2] = gy[3] = gy[4] = gy[5] = real(0);
            gx2[0] = gx2[1] = gx2[2] = gx2[3] = gx2[4] = gx2[5] = real(0);
            const bool xedge = lb(LB_X0) || lb(LB_XE);
            if (xedge) load_rot(std::integral_constant<int, 0>{}, lb(LB_X0) ? i - 1 : i + 1, j, k + 2, gx2);
            if (lb(LB_Y0) || lb(LB_YE)) load_rot(std::integral_constant<int, 1>{}, i, lb(LB_Y0) ? j - 1 : j + 2, p, gy);
            const real wbz = a.bfr[2][zlive ? cell_index(B, i, j, k) : 0];
            const real wby = a.bfr[1][ylive ? c : 0];
            const real wbx = a.bfr[0][xlive ? c : 0];
            real F[7];
            {   // z: plane k+1's slope from planes k, k+1, k+2 (own cells); face k+1/2
                real sr[6], uLn[6];
                to_prims(u2);
#pragma unroll
                for (int r = 0; r < 6; r++) {
                    const int f = rslot<2>(r);
                    const real w1 = sW[q1][f][t];
                    const real d = mc_slope(sW[q2][f][t], w1, u2[f], beta);
                    sr[r] = w1 - real(0.5) * d;
                    uLn[r] = w1 + real(0.5) * d;
                }
#pragma unroll
                for (int f = 0; f < 7; f++) sW[q2][f][t] = u2[f];
                if (xedge) {
                    to_prims(gx2);
#pragma unroll
                    for (int r = 0; r < 6; r++) sGx[q2][r][gslot] = gx2[r];
                }
                if (zlive) {
                    sm = fmax(sm, plm_solve<real, RS>(uLz, sr, wbz, cs, F));
                    fz_hi[0] = F[0]; fz_hi[3] = F[1]; fz_hi[1] = F[2]; fz_hi[2] = F[3];
                    sEz3[0][t] = -F[5];
                    sEz3[1][t] = F[6];
                } else {
                    fz_hi[0] = fz_hi[1] = fz_hi[2] = fz_hi[3] = real(0);
                    sEz3[0][t] = sEz3[1][t] = real(0);
                }
#pragma unroll
                for (int r = 0; r < 6; r++) uLz[r] = uLn[r];
            }
            vz_hi = ctype(0);
            if (lb(LB_CELL) && inr(p, 1 - nb, n2 + nb)) {   // centre EMFs and velocities
                real W[7];   // own primitives of plane k+1
#pragma unroll
                for (int f = 0; f < 7; f++) W[f] = sW[q1][f][t];
                sC[q1][0][t] = W[5] * W[3] - W[6] * W[2];
                sC[q1][1][t] = W[6] * W[1] - W[4] * W[3];
                sC[q1][2][t] = W[4] * W[2] - W[5] * W[1];
                sC[q1][3][t] = W[1];
                sC[q1][4][t] = W[2];
                vz_hi = W[3];
            }
            {   // y-face j+1/2: four-cell stencil from the buffer (edge rows load the cell beyond)
                const int tm = ty > 0 ? t - TX : t, tp1 = ty < TY - 1 ? t + TX : t, tp2 = ty < TY - 2 ? t + 2 * TX : t;
                real sl[6], sr[6];
                to_prims(gy);
#pragma unroll
                for (int r = 0; r < 6; r++) {
                    const int f = rslot<1>(r);
                    const real w0 = sW[q1][f][t], w1 = sW[q1][f][tp1], lm = sW[q1][f][tm], l2 = sW[q1][f][tp2];
                    const real wm = lb(LB_Y0) ? gy[r] : lm;
                    const real w2 = lb(LB_YE) ? gy[r] : l2;
                    sl[r] = w0 + real(0.5) * mc_slope(wm, w0, w1, beta);
                    sr[r] = w1 - real(0.5) * mc_slope(w0, w1, w2, beta);
                }
                if (ylive) {
                    sm = fmax(sm, plm_solve<real, RS>(sl, sr, wby, cs, F));
                    sF[0][FY][0][t] = F[0]; sF[0][FY][2][t] = F[1]; sF[0][FY][3][t] = F[2]; sF[0][FY][1][t] = F[3];
                    exy_hi = -F[5];
                    sEf[q1][1][t] = F[6];
                } else {
                    sF[0][FY][0][t] = sF[0][FY][1][t] = sF[0][FY][2][t] = sF[0][FY][3][t] = real(0);
                    sEf[q1][1][t] = real(0);
                    exy_hi = real(0);
                }
            }
            {   // x-face i+1/2: each lane forms its cell's slope; the right state comes from lane + 1
                real sl[6], rr[6], sr[6], fx[4], gx[6];
#pragma unroll
                for (int r = 0; r < 6; r++) gx[r] = sGx[q1][r][gslot];   // used by the edge lanes only
#pragma unroll
                for (int r = 0; r < 6; r++) {
                    const int f = rslot<0>(r);
                    const real w0 = sW[q1][f][t], lp = lane_prev(w0), ln = lane_next(w0);
                    const real wm = lb(LB_X0) ? gx[r] : lp;
                    const real wp = lb(LB_XE) ? gx[r] : ln;
                    const real d = mc_slope(wm, w0, wp, beta);
                    sl[r] = w0 + real(0.5) * d;
                    rr[r] = w0 - real(0.5) * d;
                }
#pragma unroll
                for (int r = 0; r < 6; r++) sr[r] = lane_next(rr[r]);
                if (xlive) {
                    sm = fmax(sm, plm_solve<real, RS>(sl, sr, wbx, cs, F));
                    fx[0] = F[0]; fx[1] = F[1]; fx[2] = F[2]; fx[3] = F[3];
                    sEf[q1][0][t] = -F[5];
                    eyx_hi = F[6];
                } else {
                    fx[0] = fx[1] = fx[2] = fx[3] = real(0);
                    sEf[q1][0][t] = real(0);
                    eyx_hi = real(0);
                }
#pragma unroll
                for (int f = 0; f < 4; f++) dxn[f] = fx[f] - lane_prev(fx[f]);
            }
        } else if constexpr (SHP) {
            const int q1 = (k + 1) & 1, q2 = k & 1;   // buffers of planes k+1 and k (refilled with k+2)
            const int p = k + 1;
            const bool pt = inr(p, a.ft_lo, n2 + a.ft_hi);
            const real cs = (real)a.cs;
            const long long c = cell_index(B, i, j, p);
            {   // own cell of plane k+2 into the buffer plane k used
                real u2[7];
                load_cell(i, j, k + 2, u2);
#pragma unroll
                for (int f = 0; f < 7; f++) sW[q2][f][t] = u2[f];
            }
            fbz = a.bf[2][cell_ok(i, j, k) ? cell_index(B, i, j, k) : 0];
            fbx = a.bf[0][cell_ok(i, j, p) ? c : 0];
            fby = a.bf[1][cell_ok(i, j, p) ? c : 0];
#pragma unroll
            for (int f = 0; f < 7; f++) U1[f] = sW[q1][f][t];   // own cell of plane k+1
            real F[7];
            if (live && lb(LB_FZ) && inr(k, a.fn_lo, n2 + a.fn_hi)) {   // z-face k+1/2 from planes k and k+1
                real uL[7], uR[7];
#pragma unroll
                for (int r = 0; r < 7; r++) {
                    if (r == 4) continue;
                    uL[r] = U0[perm(2, r)];
                    uR[r] = U1[perm(2, r)];
                }
                uL[4] = uR[4] = (real)fbz;
                real smx = real(0);
                riemann<RS>(uL, uR, cs, F, smx);
                sm = fmax(sm, smx);
                fz_hi[0] = F[0]; fz_hi[3] = F[1]; fz_hi[1] = F[2]; fz_hi[2] = F[3];
                sEz3[0][t] = -F[5];
                sEz3[1][t] = F[6];
            } else {
                fz_hi[0] = fz_hi[1] = fz_hi[2] = fz_hi[3] = real(0);
                sEz3[0][t] = sEz3[1][t] = real(0);
            }
            vz_hi = ctype(0);
            if (lb(LB_CELL) && inr(p, 1 - nb, n2 + nb)) {   // centre EMFs and velocities
                const real ir = fdiv(real(1), U1[0]);
                const real vx = U1[1] * ir, vy = U1[2] * ir, vz = U1[3] * ir;
                sC[q1][0][t] = U1[5] * vz - U1[6] * vy;
                sC[q1][1][t] = U1[6] * vx - U1[4] * vz;
                sC[q1][2][t] = U1[4] * vy - U1[5] * vx;
                sC[q1][3][t] = vx;
                sC[q1][4][t] = vy;
                vz_hi = vz;
            }
            if (lb(LB_FY) && pt && lb(LB_YF)) {   // y-face j+1/2, neighbour from the buffer
                const int tp1 = t + TX;
                real uL[7], uR[7];
#pragma unroll
                for (int r = 0; r < 7; r++) {
                    if (r == 4) continue;
                    uL[r] = U1[perm(1, r)];
                    uR[r] = sW[q1][perm(1, r)][tp1];
                }
                uL[4] = uR[4] = (real)fby;
                real smx = real(0);
                riemann<RS>(uL, uR, cs, F, smx);
                sm = fmax(sm, smx);
                sF[0][FY][0][t] = F[0]; sF[0][FY][2][t] = F[1]; sF[0][FY][3][t] = F[2]; sF[0][FY][1][t] = F[3];
                exy_hi = -F[5];
                sEf[q1][1][t] = F[6];
            } else {
                sF[0][FY][0][t] = sF[0][FY][1][t] = sF[0][FY][2][t] = sF[0][FY][3][t] = real(0);
                sEf[q1][1][t] = real(0);
                exy_hi = real(0);
            }
            {   // x-face i+1/2, neighbour from lane + 1
                real un[7], fx[4];
#pragma unroll
                for (int f = 0; f < 7; f++) un[f] = lane_next(U1[f]);
                if (lb(LB_FX) && pt && lb(LB_XF)) {
                    real uL[7], uR[7];
#pragma unroll
                    for (int r = 0; r < 7; r++) {
                        if (r == 4) continue;
                        uL[r] = U1[r];
                        uR[r] = un[r];
                    }
                    uL[4] = uR[4] = (real)fbx;
                    real smx = real(0);
                    riemann<RS>(uL, uR, cs, F, smx);
                    sm = fmax(sm, smx);
                    fx[0] = F[0]; fx[1] = F[1]; fx[2] = F[2]; fx[3] = F[3];
                    sEf[q1][0][t] = -F[5];
                    eyx_hi = F[6];
                } else {
                    fx[0] = fx[1] = fx[2] = fx[3] = real(0);
                    sEf[q1][0][t] = real(0);
                    eyx_hi = real(0);
                }
#pragma unroll
                for (int f = 0; f < 4; f++) dxn[f] = fx[f] - lane_prev(fx[f]);
            }
        } else {
            own(k + 1, U1);
            vz_hi = cells(k + 1, U1);
            xyfaces(k + 1, U1, eyx_hi, exy_hi, sm);
            zface(k, live, U0, U1, fz_hi, sm);
        }
        __syncthreads();
        // D: corner EMFs of index k, x/y flux differences.  The update's inputs at t^n are
        // loaded first, so that their latency overlaps the corner work and the barrier.
        const int q0 = k & 1, q1 = (k + 1) & 1, eq = NBE == 1 ? 0 : q0;
        const bool upd = k >= ks && lb(LB_U) && inr(k, a.u_lo, n2 + a.u_hi);
        const bool ccb = lb(LB_CCB) && k >= a.u_lo + sh;
        real pq[4] = {real(0), real(0), real(0), real(0)}, pc[4] = {real(0), real(0), real(0), real(0)};
        double pb[3] = {0.0, 0.0, 0.0}, pbm[3] = {0.0, 0.0, 0.0};
        double cbx_left = 0.0;   // SHP: lane - 1's carried x face field (all lanes take part)
        if constexpr (SHP) cbx_left = lane_prev_d(cbx);
        if (upd) {
            const long long c = cell_index(B, i, j, k);
#pragma unroll
            for (int f = 0; f < 4; f++) pq[f] = PRED ? U0[f] : a.src.q[f][c];   // predictor: the flux state is the base
            if (!PRED && a.comp[0] != nullptr) {
#pragma unroll
                for (int f = 0; f < 4; f++) pc[f] = a.comp[f][c];
            }
            if constexpr (SHP) {
                pb[0] = cbx;
                pb[1] = cby;
                pb[2] = fbz;
            } else {
#pragma unroll
                for (int d = 0; d < 3; d++) pb[d] = sb[d][c];
            }
            if (ccb) {
                if constexpr (SHP) {
                    if (tx == 0) pbm[0] = sb[0][c - sx];   // a row's first lane: lane - 1 is another row
                    else pbm[0] = cbx_left;
                } else {
                    pbm[0] = sb[0][c - sx];
                }
                pbm[1] = sb[1][c - sy];
                if (!have_bzm) pbm[2] = sb[2][c - sz];
            }
        }
        if (live && lb(LB_C) && inr(k, a.c_lo, n2 + a.c_hi)) {
            const mtype vx = mtype(0.5) * ((mtype)sC[q0][3][t] + (mtype)sC[q0][3][t + 1]);
            const mtype vy = mtype(0.5) * ((mtype)sC[q0][4][t] + (mtype)sC[q0][4][t + TX]);
            const mtype vz = mtype(0.5) * ((mtype)vz_lo + (mtype)vz_hi);
            // k_corner's corner sums with the factor 2 of its upwinded differences taken out:
            // upw commutes with the exact scaling by 2, so 0.25 (a - b) here has the bits of
            // 0.125 (2a - 2b) there
            constexpr mtype qtr = mtype(0.25);
            mtype a14, a34, b14, b34;
            {   // Ez at (i+1/2, j+1/2, k)
                mtype e00 = sC[q0][2][t], e10 = sC[q0][2][t + 1], e01 = sC[q0][2][t + TX], e11 = sC[q0][2][t + TX + 1];
                mtype y0 = sEf[q0][1][t], y1 = sEf[q0][1][t + 1], x0 = sEf[q0][0][t], x1 = sEf[q0][0][t + TX];
                upw(vx, y0 - e00, y1 - e10, e01 - y0, e11 - y1, a14, a34);
                upw(vy, x0 - e00, x1 - e01, e10 - x0, e11 - x1, b14, b34);
                sEz[t] = qtr * (x0 + x1 + y0 + y1) + qtr * ((a14 - a34) + (b14 - b34));
            }
            {   // Ey at (i+1/2, j, k+1/2)
                mtype e00 = sC[q0][1][t], e10 = sC[q0][1][t + 1], e01 = sC[q1][1][t], e11 = sC[q1][1][t + 1];
                mtype z0 = sEz3[0][t], z1 = sEz3[0][t + 1], x0 = eyx_lo, x1 = eyx_hi;
                upw(vx, z0 - e00, z1 - e10, e01 - z0, e11 - z1, a14, a34);
                upw(vz, x0 - e00, x1 - e01, e10 - x0, e11 - x1, b14, b34);
                sEc[eq][1][t] = qtr * (x0 + x1 + z0 + z1) + qtr * ((a14 - a34) + (b14 - b34));
            }
            {   // Ex at (i, j+1/2, k+1/2)
                mtype e00 = sC[q0][0][t], e10 = sC[q0][0][t + TX], e01 = sC[q1][0][t], e11 = sC[q1][0][t + TX];
                mtype z0 = sEz3[1][t], z1 = sEz3[1][t + TX], y0 = exy_lo, y1 = exy_hi;
                upw(vy, z0 - e00, z1 - e10, e01 - z0, e11 - z1, a14, a34);
                upw(vz, y0 - e00, y1 - e01, e10 - y0, e11 - y1, b14, b34);
                sEc[eq][0][t] = qtr * (y0 + y1 + z0 + z1) + qtr * ((a14 - a34) + (b14 - b34));
            }
        }
        if (ONE) {   // plane k+1, while its fluxes are in the single buffer
            if (lb(LB_U)) {
#pragma unroll
                for (int f = 0; f < 4; f++) {
                    const real dx = SHX ? dxn[f] : sF[0][0][f][t] - sF[0][0][f][t - 1];
                    dxy_n[f] = dx + (sF[0][FY][f][t] - sF[0][FY][f][t - TX]);
                }
            }
        } else if (upd) {   // plane k, from its parity buffer
#pragma unroll
            for (int f = 0; f < 4; f++)
                dxy[f] = (sF[q0][0][f][t] - sF[q0][0][f][t - 1]) + (sF[q0][FY][f][t] - sF[q0][FY][f][t - TX]);
        }
        __syncthreads();
        // G: update of plane k (k_update_z's arithmetic)
        if (upd) {
            const long long c = cell_index(B, i, j, k);
            real nq[7];   // the new state of the cell, for its CFL speed
            real oq[4];   // rho, m at t^n
#pragma unroll
            for (int f = 0; f < 4; f++) {
                real div = dxy[f] + (fz_hi[f] - fz_lo[f]);
                const real q0v = pq[f];
                oq[f] = q0v;
                real v;
                if (!PRED && a.comp[0] != nullptr) {   // Kahan terms: corrector only
                    real y = -(dtdx * div) - pc[f];
                    v = q0v + y;
                    a.comp[f][c] = (v - q0v) - y;
                } else {
                    v = q0v - dtdx * div;
                }
                if (f == 0 && v < (real)MIN_DENS) {
                    v = (real)MIN_DENS;
                    if (!PRED && a.comp[0] != nullptr) a.comp[0][c] = real(0);
                    atomicAdd(a.floor_count, 1u);
                }
                a.dst.q[f][c] = v;
                nq[f] = v;
            }
            const double ex = sEc[eq][0][t], ey = sEc[eq][1][t], ez = sEz[t];
            const double ez_y = sEz[t - TX], ez_x = sEz[t - 1], ey_x = sEc[eq][1][t - 1], ex_y = sEc[eq][0][t - TX];
            if (NBE == 2) {
                const int qm = (k - 1) & 1;
                ey_m = sEc[qm][1][t]; ex_m = sEc[qm][0][t]; eyx_m = sEc[qm][1][t - 1]; exy_m = sEc[qm][0][t - TX];
            }
            double bx = pb[0] - dtd * ((ez - ez_y) - (ey - ey_m));
            double by = pb[1] - dtd * ((ex - ex_m) - (ez - ez_x));
            double bz = pb[2] - dtd * ((ey - ey_x) - (ex - ex_y));
            if (PRED && a.bhw[0] != nullptr) {   // FP32 half step: the corrector reads these as FP32
                a.bhw[0][c] = (real)bx;
                a.bhw[1][c] = (real)by;
                a.bhw[2][c] = (real)bz;
            } else {
                a.dst.b[0][c] = bx;
                a.dst.b[1][c] = by;
                a.dst.b[2][c] = bz;
            }
            if (ccb) {
                const double ez_xy = sEz[t - 1 - TX];
                double bxm = pbm[0] - dtd * ((ez_x - ez_xy) - (ey_x - eyx_m));
                double bym = pbm[1] - dtd * ((ex_y - exy_m) - (ez_y - ez_xy));
                if (!have_bzm) bz_m = pbm[2] - dtd * ((ey_m - eyx_m) - (ex_m - exy_m));
                nq[4] = (real)(0.5 * (bx + bxm));
                nq[5] = (real)(0.5 * (by + bym));
                nq[6] = (real)(0.5 * (bz + bz_m));
            } else {
                nq[4] = PRED ? U0[4] : a.src.q[4][c];
                nq[5] = PRED ? U0[5] : a.src.q[5][c];
                nq[6] = PRED ? U0[6] : a.src.q[6][c];
            }
            a.dst.q[4][c] = nq[4];
            a.dst.q[5][c] = nq[5];
            a.dst.q[6][c] = nq[6];
            bz_m = bz;
            have_bzm = true;
            ey_m = ey; ex_m = ex; eyx_m = ey_x; exy_m = ex_y;
            if (a.pvel[0] != nullptr) {   // passive velocities (copy_velocity's arithmetic)
                const real io = ONE ? (real)__builtin_amdgcn_rcpf((float)oq[0]) : real(1) / oq[0];
                if (PRED) {
#pragma unroll
                    for (int d = 0; d < 3; d++) a.pvel[d][c] = oq[1 + d] * io;
                } else {
                    const real in = ONE ? (real)__builtin_amdgcn_rcpf((float)nq[0]) : real(1) / nq[0];
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        const real vo = oq[1 + d] * io;
                        a.pvel[d][c] = real(0.5) * (vo + nq[1 + d] * in);
                    }
                }
            }
            if (!PRED && a.cmax != nullptr && lb(LB_INT) && k >= 1 && k <= n2) {   // corrector only
                if constexpr (ONE) {
                    cfl = fmax(cfl, cell_fast_speed_f(nq, (float)a.cs));
                } else {
                    double q7[7];
                    for (int f = 0; f < 7; f++) q7[f] = (double)nq[f];
                    cfl = fmax(cfl, cell_fast_speed(q7, a.cs));
                }
            }
        } else {
            have_bzm = false;
            if (NBE == 1 && lb(LB_U)) { ey_m = sEc[0][1][t]; ex_m = sEc[0][0][t]; eyx_m = sEc[0][1][t - 1]; exy_m = sEc[0][0][t - TX]; }
        }
#pragma unroll
        for (int f = 0; f < 4; f++) {
            fz_lo[f] = fz_hi[f];
            if (ONE) dxy[f] = dxy_n[f];
        }
#pragma unroll
        for (int f = 0; f < 7; f++) U0[f] = U1[f];
        if constexpr (SHP) { cbx = fbx; cby = fby; }
        eyx_lo = eyx_hi;
        exy_lo = exy_hi;
        vz_lo = vz_hi;
    }
    atomic_max_nonneg(a.smax, b, (double)sm);
    if (!PRED && a.cmax != nullptr) atomic_max_nonneg(a.cmax, b, (double)cfl);
}

// kind: 0 predictor, 1 PLM corrector, 2 PPM-primitive corrector, 3 PPM-characteristic
// corrector; shape 2 exists in FP32 only
template <class real, int RS, int KIND>
void launch_vl_stage_k(hipStream_t st, const BoxDesc* boxes, const TilePlan& plan, const StageArgs<real>& a,
                       int shape) {
    if constexpr (sizeof(real) == 4) {
        if (shape == 2) {
            hipLaunchKernelGGL((k_vl_stage<real, RS, KIND, vs_tx(2), vs_ty(2)>), dim3((unsigned)plan.total),
                               dim3(vs_tx(2) * vs_ty(2)), 0, st, boxes, plan, a);
            return;
        }
    }
    if (shape == 1)
        hipLaunchKernelGGL((k_vl_stage<real, RS, KIND, vs_tx(1), vs_ty(1)>), dim3((unsigned)plan.total),
                           dim3(vs_tx(1) * vs_ty(1)), 0, st, boxes, plan, a);
    else
        hipLaunchKernelGGL((k_vl_stage<real, RS, KIND, vs_tx(0), vs_ty(0)>), dim3((unsigned)plan.total),
                           dim3(vs_tx(0) * vs_ty(0)), 0, st, boxes, plan, a);
}
template <class real, int RS, int KIND>
int vl_stage_blocks_per_cu_k(int shape) {
    int n = 0;
    hipError_t e;
    if constexpr (sizeof(real) == 4) {
        if (shape == 2) {
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_vl_stage<real, RS, KIND, vs_tx(2), vs_ty(2)>,
                                                             vs_tx(2) * vs_ty(2), 0);
            return e == hipSuccess && n > 0 ? n : 1;
        }
    }
    if (shape == 1)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_vl_stage<real, RS, KIND, vs_tx(1), vs_ty(1)>,
                                                         vs_tx(1) * vs_ty(1), 0);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_vl_stage<real, RS, KIND, vs_tx(0), vs_ty(0)>,
                                                         vs_tx(0) * vs_ty(0), 0);
    return e == hipSuccess && n > 0 ? n : 1;
}
template <class real, int RS>
int vl_stage_blocks_per_cu(int kind, int shape) {
    switch (kind) {
        case 0: return vl_stage_blocks_per_cu_k<real, RS, 0>(shape);
        case 1: return vl_stage_blocks_per_cu_k<real, RS, 1>(shape);
        case 2: return vl_stage_blocks_per_cu_k<real, RS, 2>(shape);
        case 4:
            if constexpr (sizeof(real) == 4) return vl_stage_blocks_per_cu_k<real, RS, 4>(shape);
            return vl_stage_blocks_per_cu_k<real, RS, 1>(shape);
        case 5:
            if constexpr (sizeof(real) == 4) return vl_stage_blocks_per_cu_k<real, RS, 5>(shape);
            return vl_stage_blocks_per_cu_k<real, RS, 0>(shape);
        default: return vl_stage_blocks_per_cu_k<real, RS, 3>(shape);
    }
}

template <class real, int RS>
void launch_vl_stage(hipStream_t st, const BoxDesc* boxes, const TilePlan& plan, const StageArgs<real>& a, int kind,
                     int shape) {
    if (plan.total <= 0) return;
    switch (kind) {
        case 0: launch_vl_stage_k<real, RS, 0>(st, boxes, plan, a, shape); break;
        case 1: launch_vl_stage_k<real, RS, 1>(st, boxes, plan, a, shape); break;
        case 2: launch_vl_stage_k<real, RS, 2>(st, boxes, plan, a, shape); break;
        case 4:   // FP32 only; an FP64 context runs the PLM corrector as kind 1
            if constexpr (sizeof(real) == 4) launch_vl_stage_k<real, RS, 4>(st, boxes, plan, a, shape);
            else launch_vl_stage_k<real, RS, 1>(st, boxes, plan, a, shape);
            break;
        case 5:   // FP32 only; an FP64 context runs the predictor as kind 0
            if constexpr (sizeof(real) == 4) launch_vl_stage_k<real, RS, 5>(st, boxes, plan, a, shape);
            else launch_vl_stage_k<real, RS, 0>(st, boxes, plan, a, shape);
            break;
        default: launch_vl_stage_k<real, RS, 3>(st, boxes, plan, a, shape); break;
    }
}

}  // namespace imhd
