// Third-order (PPM) corrector-stage reconstruction of the VL integrator: the reference's
// default characteristic variant (vanleer_mod.f90:941-1141, 1595-1818) and the
// primitive variant as compiled (vanleer_mod.f90:1217-1334; the block after GO TO 1000 is
// dead).  One thread builds the two states of one interface from the six cells
// c-2 .. c+3 along the sweep.  Arithmetic follows the reference's order (see riemann.hpp).
#pragma once
#include "imhd_common.hpp"

namespace imhd {

// Right (columns) and left (rows) eigenvectors of the isothermal primitive system of one
// cell, S08 eqs. A12/A22 as stated in vanleer_mod.f90:941-1141.  wb is the normal face
// field of the cell's upper face (the reference passes wb(i,j) there).
template <class real>
__device__ __forceinline__ void prim_eigvecs(const real* w, real wb, real cs, real R[6][6], real L[6][6]) {
    const real asqr = cs * cs;
    const real rho = w[0], irho = fdiv(real(1), rho), rhos = fsqrt(rho), irhos = fdiv(real(1), rhos);
    const real bx = irhos * wb, by = irhos * w[4], bz = irhos * w[5];
    const real bsqr = bx * bx + by * by + bz * bz;
    const real btsqr = by * by + bz * bz;
    const real tmp1 = asqr + bsqr, x = asqr - bsqr;
    const real tmp2 = fsqrt(x * x + real(4) * asqr * btsqr);
    const real cslow = fsqrt(max2(real(0), real(0.5) * (tmp1 - tmp2)));
    const real cf = fsqrt(real(0.5) * (tmp1 + tmp2));
    real af, as;
    if (fabs(cf * cf - cslow * cslow) < real(ZERO)) { af = real(1); as = real(0); }
    else {
        af = fsqrt(max2(fdiv(asqr - cslow * cslow, cf * cf - cslow * cslow), real(0)));
        as = fsqrt(max2(fdiv(cf * cf - asqr, cf * cf - cslow * cslow), real(0)));
    }
    real by_, bz_;
    if (btsqr < real(ZERO)) { by_ = fdiv(real(1), fsqrt(real(2))); bz_ = by_; }
    else { by_ = fdiv(by, fsqrt(btsqr)); bz_ = fdiv(bz, fsqrt(btsqr)); }
    const real S = sgn1(bx);
    const real Cff = cf * af, Css = cslow * as, Qf = cf * af * S, Qs = cslow * as * S;
    const real Af = cs * af * rhos, As = cs * as * rhos;
    R[0][0] = rho * af; R[1][0] = -Cff; R[2][0] = Qs * by_; R[3][0] = Qs * bz_; R[4][0] = As * by_; R[5][0] = As * bz_;
    R[0][1] = real(0); R[1][1] = real(0); R[2][1] = -bz_; R[3][1] = by_; R[4][1] = -bz_ * S * rhos; R[5][1] = by_ * S * rhos;
    R[0][2] = rho * as; R[1][2] = -Css; R[2][2] = -Qf * by_; R[3][2] = -Qf * bz_; R[4][2] = -Af * by_; R[5][2] = -Af * bz_;
    R[0][3] = rho * as; R[1][3] = Css; R[2][3] = Qf * by_; R[3][3] = Qf * bz_; R[4][3] = -Af * by_; R[5][3] = -Af * bz_;
    R[0][4] = real(0); R[1][4] = real(0); R[2][4] = bz_; R[3][4] = -by_; R[4][4] = -bz_ * S * rhos; R[5][4] = by_ * S * rhos;
    R[0][5] = rho * af; R[1][5] = Cff; R[2][5] = -Qs * by_; R[3][5] = -Qs * bz_; R[4][5] = As * by_; R[5][5] = As * bz_;
    const real Nf = fdiv(real(0.5), asqr), Ns = Nf;
    L[0][0] = Nf * af * asqr * irho; L[0][1] = -(Nf * Cff); L[0][2] = Nf * Qs * by_; L[0][3] = Nf * Qs * bz_;
    L[0][4] = Nf * As * by_ * irho; L[0][5] = Nf * As * bz_ * irho;
    L[1][0] = real(0); L[1][1] = real(0); L[1][2] = -(real(0.5) * bz_); L[1][3] = real(0.5) * by_;
    L[1][4] = -(real(0.5) * bz_ * S * irhos); L[1][5] = real(0.5) * by_ * S * irhos;
    L[2][0] = Ns * as * asqr * irho; L[2][1] = -(Ns * Css); L[2][2] = -(Ns * Qf * by_); L[2][3] = -(Ns * Qf * bz_);
    L[2][4] = -(Ns * Af * by_ * irho); L[2][5] = -(Ns * Af * bz_ * irho);
    L[3][0] = Ns * as * asqr * irho; L[3][1] = Ns * Css; L[3][2] = Ns * Qf * by_; L[3][3] = Ns * Qf * bz_;
    L[3][4] = -(Ns * Af * by_ * irho); L[3][5] = -(Ns * Af * bz_ * irho);
    L[4][0] = real(0); L[4][1] = real(0); L[4][2] = real(0.5) * bz_; L[4][3] = -(real(0.5) * by_);
    L[4][4] = -(real(0.5) * bz_ * S * irhos); L[4][5] = real(0.5) * by_ * S * irhos;
    L[5][0] = Nf * af * asqr * irho; L[5][1] = Nf * Cff; L[5][2] = -(Nf * Qs * by_); L[5][3] = -(Nf * Qs * bz_);
    L[5][4] = Nf * As * by_ * irho; L[5][5] = Nf * As * bz_ * irho;
}

template <class real>
__device__ __forceinline__ real dot6(const real* a, const real* b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3] + a[4] * b[4] + a[5] * b[5];
}

template <class real>
__device__ __forceinline__ real mc_lim(real dL, real dR, real dC, real beta) {
    if (dL * dR > real(0)) return sgn1(dC) * min3(beta * fabs(dL), beta * fabs(dR), fabs(dC));
    return real(0);
}

// CW84 monotonicity (vanleer_mod.f90:1313-1332 primitive, :1724-1743 characteristic)
template <class real>
__device__ __forceinline__ void ppm_monotone(real w, real dw, real& wL, real& wR, bool flat_to_linear) {
    real tmp1 = (wR - w) * (w - wL);
    real tmp2 = real(6) * (wR - wL) * (w - real(0.5) * (wL + wR));
    real d = wR - wL;
    real tmp3 = d * d;
    if (tmp1 <= real(0)) {
        if (flat_to_linear) { wL = w - real(0.5) * dw; wR = w + real(0.5) * dw; }
        else { wL = w; wR = w; }
    } else if (tmp2 > tmp3) {
        wL = real(3) * w - real(2) * wR;
    } else if (tmp2 < -tmp3) {
        wR = real(3) * w - real(2) * wL;
    }
}

// characteristic slope difference of cell n: R (da_{n+1} - da_n) with both limited slopes
// projected on cell n's left eigenvectors (vanleer_mod.f90:1595-1657).  W holds cells
// n-1 .. n+2 at W[0..3].
template <class real>
__device__ __forceinline__ void char_delw(const real (*W)[6], real wb, real cs, real beta, real* delw) {
    real R[6][6], L[6][6];
    prim_eigvecs(W[1], wb, cs, R, L);
    real dm[6];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        real aL1 = dot6(L[k], W[0]), aC = dot6(L[k], W[1]), aR1 = dot6(L[k], W[2]), aR2 = dot6(L[k], W[3]);
        real dL1 = aC - aL1, dR1 = aR1 - aC, dC1 = real(0.5) * (dL1 + dR1);
        real dL2 = aR1 - aC, dR2 = aR2 - aR1, dC2 = real(0.5) * (dL2 + dR2);
        dm[k] = mc_lim(dL2, dR2, dC2, beta) - mc_lim(dL1, dR1, dC1, beta);
    }
#pragma unroll
    for (int k = 0; k < 6; k++)
        delw[k] = dm[0] * R[k][0] + dm[1] * R[k][1] + dm[2] * R[k][2] + dm[3] * R[k][3] + dm[4] * R[k][4] + dm[5] * R[k][5];
}

// Interface states between cells c and c+1 from W[0..5] = cells c-2 .. c+3 and the normal
// field; wbc[t] is the normal face field of cell c-2+t.  uL / uR are conserved, rotated.
template <class real, bool CHAR>
__device__ __forceinline__ void ppm_states(const real (*W)[6], const real* wbc, real cs, real beta, real* uL, real* uR) {
    const real sixth = fdiv(real(1), real(6));
    real wLc[6], wRc[6], wLn[6], wRn[6];
    if constexpr (CHAR) {
        real dm1[6], d0[6], dp1[6];   // delw of cells c-1, c, c+1
        char_delw<real>(W + 0, wbc[1], cs, beta, dm1);
        char_delw<real>(W + 1, wbc[2], cs, beta, d0);
        char_delw<real>(W + 2, wbc[3], cs, beta, dp1);
#pragma unroll
        for (int k = 0; k < 6; k++) {
            wLc[k] = real(0.5) * (W[1][k] + W[2][k]) - sixth * dm1[k];
            wRc[k] = real(0.5) * (W[2][k] + W[3][k]) - sixth * d0[k];
            wLn[k] = real(0.5) * (W[2][k] + W[3][k]) - sixth * d0[k];
            wRn[k] = real(0.5) * (W[3][k] + W[4][k]) - sixth * dp1[k];
            ppm_monotone(W[2][k], real(0), wLc[k], wRc[k], false);
            ppm_monotone(W[3][k], real(0), wLn[k], wRn[k], false);
        }
    } else {
#pragma unroll
        for (int k = 0; k < 6; k++) {
            real d[4];   // limited slopes of cells c-1 .. c+2
#pragma unroll
            for (int t = 0; t < 4; t++) {
                real dL = W[t + 1][k] - W[t][k], dR = W[t + 2][k] - W[t + 1][k], dC = real(0.5) * (dL + dR);
                real mabs = min2(fabs(dL), fabs(dR));
                d[t] = (dL * dR > real(0)) ? sgn1(dC) * min2(beta * mabs, fabs(dC)) : real(0);
            }
            wLc[k] = real(0.5) * (W[1][k] + W[2][k]) + sixth * (d[0] - d[1]);
            wRc[k] = real(0.5) * (W[2][k] + W[3][k]) + sixth * (d[1] - d[2]);
            wLn[k] = real(0.5) * (W[2][k] + W[3][k]) + sixth * (d[1] - d[2]);
            wRn[k] = real(0.5) * (W[3][k] + W[4][k]) + sixth * (d[2] - d[3]);
            ppm_monotone(W[2][k], d[1], wLc[k], wRc[k], true);
            ppm_monotone(W[3][k], d[2], wLn[k], wRn[k], true);
        }
    }
    const real wb = wbc[2];
    uL[0] = wRc[0]; uL[1] = wRc[1]; uL[2] = wRc[2]; uL[3] = wRc[3]; uL[4] = wb; uL[5] = wRc[4]; uL[6] = wRc[5];
    uR[0] = wLn[0]; uR[1] = wLn[1]; uR[2] = wLn[2]; uR[3] = wLn[3]; uR[4] = wb; uR[5] = wLn[4]; uR[6] = wLn[5];
    uL[1] = uL[1] * uL[0]; uL[2] = uL[2] * uL[0]; uL[3] = uL[3] * uL[0];
    uR[1] = uR[1] * uR[0]; uR[2] = uR[2] * uR[0]; uR[3] = uR[3] * uR[0];
}

}  // namespace imhd
