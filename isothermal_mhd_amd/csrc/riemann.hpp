// Interface Riemann solvers of the isothermal MHD system (one interface per thread).
// States are rotated so the face normal is "x": u = (rho, m_n, m_t1, m_t2, B_n, B_t1, B_t2).
// Flux component 4 (the normal field) is identically zero.
//
// Each solver keeps the reference's operation order so that the FP64 instance agrees
// bit for bit with the reference build (no FMA contraction in this library).
#pragma once
#include "imhd_common.hpp"

namespace imhd {

template <class real>
__device__ __forceinline__ real fast_speed(const real* u, real asqr) {  // vanleer_mod.f90:3048-3053
    real bsqr = fdiv(u[4] * u[4] + u[5] * u[5] + u[6] * u[6], u[0]);
    real btsqr = fdiv(u[5] * u[5] + u[6] * u[6], u[0]);
    real tmp1 = asqr + bsqr;
    real x = asqr - bsqr;
    real tmp2 = fsqrt(x * x + real(4) * asqr * btsqr);
    return fsqrt(real(0.5) * (tmp1 + tmp2));
}

template <class real>
__device__ __forceinline__ void phys_flux(const real* u, real asqr, real vx, real vy, real vz, real* F) {
    F[0] = u[1];
    F[1] = u[0] * (vx * vx + asqr) + real(0.5) * (u[5] * u[5] + u[6] * u[6] - u[4] * u[4]);
    F[2] = u[1] * vy - u[4] * u[5];
    F[3] = u[1] * vz - u[4] * u[6];
    F[4] = real(0);
    F[5] = vx * u[5] - vy * u[4];
    F[6] = vx * u[6] - vz * u[4];
}

// HLLE with Einfeldt-type speeds clamped with 0 (vanleer_mod.f90:2867-2981)
template <class real>
__device__ __forceinline__ void hlle(const real* uL, const real* uR, real asqr, real* F, real& smax) {
    real uavg[7], FL[7], FR[7];
#pragma unroll
    for (int m = 0; m < 7; m++) uavg[m] = real(0.5) * (uL[m] + uR[m]);
    real vxL = fdiv(uL[1], uL[0]), vyL = fdiv(uL[2], uL[0]), vzL = fdiv(uL[3], uL[0]);
    real vxR = fdiv(uR[1], uR[0]), vyR = fdiv(uR[2], uR[0]), vzR = fdiv(uR[3], uR[0]);
    phys_flux(uL, asqr, vxL, vyL, vzL, FL);
    phys_flux(uR, asqr, vxR, vyR, vzR, FR);
    real cfL = fast_speed(uL, asqr), cfR = fast_speed(uR, asqr), cfa = fast_speed(uavg, asqr);
    real vxa = fdiv(uavg[1], uavg[0]);
    real sL = min3(vxL - cfL, vxa - cfa, real(0));
    real sR = max3(vxR + cfR, vxa + cfa, real(0));
    smax = max2(smax, max3(fabs(vxL) + cfL, fabs(vxR) + cfR, fabs(vxa) + cfa));
    real isRL = fdiv(real(1), sR - sL);
#pragma unroll
    for (int m = 0; m < 7; m++) F[m] = (sR * FL[m] - sL * FR[m] + sR * sL * (uR[m] - uL[m])) * isRL;
}

// Isothermal HLLD of Mignone (2007) as the reference states it (vanleer_mod.f90:2986-3239).
// VGIVEN (FP32 PLM states): the caller passes the interface velocities it formed the
// momenta from (vl, vr), instead of dividing the momenta by rho again.
template <class real, bool VGIVEN = false>
__device__ __forceinline__ void hlld(const real* uL, const real* uR, real asqr, real* F, real& smax,
                                     const real* vl = nullptr, const real* vr = nullptr) {
    real uavg[7], FL[7], FR[7];
#pragma unroll
    for (int m = 0; m < 7; m++) uavg[m] = real(0.5) * (uL[m] + uR[m]);
    const real Bx = uL[4];
    if constexpr (sizeof(real) == 4) uavg[4] = Bx;   // both sides carry the face's B_n: exact
    real vxL, vyL, vzL, vxR, vyR, vzR;
    if constexpr (VGIVEN) {
        vxL = vl[0]; vyL = vl[1]; vzL = vl[2];
        vxR = vr[0]; vyR = vr[1]; vzR = vr[2];
    } else {
        vxL = fdiv(uL[1], uL[0]); vyL = fdiv(uL[2], uL[0]); vzL = fdiv(uL[3], uL[0]);
        vxR = fdiv(uR[1], uR[0]); vyR = fdiv(uR[2], uR[0]); vzR = fdiv(uR[3], uR[0]);
    }
    phys_flux(uL, asqr, vxL, vyL, vzL, FL);
    phys_flux(uR, asqr, vxR, vyR, vzR, FR);
    real cfL = fast_speed(uL, asqr), cfR = fast_speed(uR, asqr), cfa = fast_speed(uavg, asqr);
    real vxa = fdiv(uavg[1], uavg[0]);
    real sL = min3(vxL - cfL, vxa - cfa, real(0));
    real sR = max3(vxR + cfR, vxa + cfa, real(0));
    smax = max2(smax, max3(fabs(vxL) + cfL, fabs(vxR) + cfR, fabs(vxa) + cfa));
    if (sL >= real(0)) {
#pragma unroll
        for (int m = 0; m < 7; m++) F[m] = FL[m];
        return;
    }
    if (sR <= real(0)) {
#pragma unroll
        for (int m = 0; m < 7; m++) F[m] = FR[m];
        return;
    }
    real isRL = fdiv(real(1), sR - sL);
    real uHLL0 = (sR * uR[0] - sL * uL[0] - FR[0] + FL[0]) * isRL;
    real uHLL1 = (sR * uR[1] - sL * uL[1] - FR[1] + FL[1]) * isRL;
    real FHLL0 = (sR * FL[0] - sL * FR[0] + sR * sL * (uR[0] - uL[0])) * isRL;
    real FHLL1 = (sR * FL[1] - sL * FR[1] + sR * sL * (uR[1] - uL[1])) * isRL;
    const real rho_s = uHLL0, momx_s = uHLL1;
    // 1/rho*, 1/sqrt(rho*) and sqrt(rho*): FP32 forms all three from one hardware rsqrt
    real irho_s, isrho_s, srho_s;
    if constexpr (sizeof(real) == 4) {
        isrho_s = __builtin_amdgcn_rsqf(rho_s);
        irho_s = isrho_s * isrho_s;
        srho_s = rho_s * isrho_s;
    } else {
        irho_s = real(1) / rho_s;
        isrho_s = sqrt(irho_s);
        srho_s = sqrt(rho_s);
    }
    const real vx_s = sizeof(real) == 4 ? FHLL0 * irho_s : FHLL0 / rho_s;
    const real sLs = vx_s - fabs(Bx) * isrho_s;
    const real sRs = vx_s + fabs(Bx) * isrho_s;
    const real zero = real(ZERO);
    real uLs[7], uRs[7];
    uLs[0] = rho_s; uLs[1] = momx_s;
    if (fabs(sL - sLs) < zero || fabs(sL - sRs) < zero) {
#pragma unroll
        for (int m = 2; m < 7; m++) uLs[m] = uL[m];
    } else {
        real is = fdiv(real(1), (sL - sLs) * (sL - sRs));
        real t1 = Bx * (vx_s - vxL) * is;
        real d = sL - vxL;
        real t2 = (uL[0] * (d * d) - Bx * Bx) * irho_s * is;
        uLs[2] = rho_s * vyL - t1 * uL[5];
        uLs[3] = rho_s * vzL - t1 * uL[6];
        uLs[4] = uL[4];
        uLs[5] = t2 * uL[5];
        uLs[6] = t2 * uL[6];
    }
    uRs[0] = rho_s; uRs[1] = momx_s;
    if (fabs(sR - sLs) < zero || fabs(sR - sRs) < zero) {
#pragma unroll
        for (int m = 2; m < 7; m++) uRs[m] = uR[m];
    } else {
        real is = fdiv(real(1), (sR - sLs) * (sR - sRs));
        real t1 = Bx * (vx_s - vxR) * is;
        real d = sR - vxR;
        real t2 = (uR[0] * (d * d) - Bx * Bx) * irho_s * is;
        uRs[2] = rho_s * vyR - t1 * uR[5];
        uRs[3] = rho_s * vzR - t1 * uR[6];
        uRs[4] = uR[4];
        uRs[5] = t2 * uR[5];
        uRs[6] = t2 * uR[6];
    }
    if (sLs >= real(0)) {
#pragma unroll
        for (int m = 0; m < 7; m++) F[m] = FL[m] + sL * (uLs[m] - uL[m]);
    } else if (sRs <= real(0)) {
#pragma unroll
        for (int m = 0; m < 7; m++) F[m] = FR[m] + sR * (uRs[m] - uR[m]);
    } else {
        const real X = srho_s * sgn1(Bx);
        const real iX = sizeof(real) == 4 ? isrho_s * sgn1(Bx) : real(1) / X;
        real uC2 = real(0.5) * (uLs[2] + uRs[2] + X * (uRs[5] - uLs[5]));
        real uC3 = real(0.5) * (uLs[3] + uRs[3] + X * (uRs[6] - uLs[6]));
        real uC5 = real(0.5) * (uLs[5] + uRs[5] + iX * (uRs[2] - uLs[2]));
        real uC6 = real(0.5) * (uLs[6] + uRs[6] + iX * (uRs[3] - uLs[3]));
        F[0] = FHLL0;
        F[1] = FHLL1;
        F[2] = uC2 * vx_s - Bx * uC5;
        F[3] = uC3 * vx_s - Bx * uC6;
        F[4] = real(0);
        F[5] = vx_s * uC5 - (sizeof(real) == 4 ? (uC2 * Bx) * irho_s : (uC2 * Bx) / rho_s);
        F[6] = vx_s * uC6 - (sizeof(real) == 4 ? (uC3 * Bx) * irho_s : (uC3 * Bx) / rho_s);
    }
}

// Roe flux with arithmetic-mean state, K98 eigenvectors ("MODIFIED V1" degeneracy rules)
// and the reference's Harten-type entropy fix, vanleer_mod.f90:2264-2863.
template <class real>
__device__ __forceinline__ void roe(const real* qL, const real* qR, real cs, real* F, real& smax) {
    const real asqr = cs * cs;
    const real zero = real(ZERO);
    real qa[7];
    real rho2 = real(0.5) * (qL[0] + qR[0]);
    real isrho = fdiv(real(1), fsqrt(rho2));
    qa[0] = rho2;
    qa[1] = real(0.5) * (fdiv(qL[1], qL[0]) + fdiv(qR[1], qR[0]));
    qa[2] = real(0.5) * (fdiv(qL[2], qL[0]) + fdiv(qR[2], qR[0]));
    qa[3] = real(0.5) * (fdiv(qL[3], qL[0]) + fdiv(qR[3], qR[0]));
    qa[4] = qL[4] * isrho;
    qa[5] = real(0.5) * (qL[5] + qR[5]) * isrho;
    qa[6] = real(0.5) * (qL[6] + qR[6]) * isrho;
    real vx2 = qa[1], bx = qa[4];
    real bsqr = qa[4] * qa[4] + qa[5] * qa[5] + qa[6] * qa[6];
    real btsqr = qa[5] * qa[5] + qa[6] * qa[6];
    real tmp1 = asqr + bsqr, x = asqr - bsqr;
    real tmp2 = fsqrt(x * x + real(4) * asqr * btsqr);
    real cslow = fsqrt(max2(real(0), real(0.5) * (tmp1 - tmp2)));
    real ca = fabs(bx);
    real cf = fsqrt(real(0.5) * (tmp1 + tmp2));
    real ev[6] = {vx2 - cf, vx2 - ca, vx2 - cslow, vx2 + cslow, vx2 + ca, vx2 + cf};
    smax = max2(smax, fabs(vx2) + cf);
    real bt = fsqrt(qa[5] * qa[5] + qa[6] * qa[6]);
    real af, as;
    if (fabs(cf * cf - cslow * cslow) < zero) { af = real(1); as = real(1); }
    else if (fabs(cf * cf - ca * ca) < zero) { af = real(0); as = real(1); }
    else if (fabs(cf * cf - asqr) < zero) { af = real(1); as = real(0); }
    else {
        real t = fdiv(real(1), fsqrt(max2(real(0), cf * cf - cslow * cslow)));
        af = fsqrt(max2(real(0), cf * cf - ca * ca)) * t;
        as = fsqrt(max2(real(0), cf * cf - asqr)) * t;
    }
    real by_, bz_;
    if (bt < zero) { by_ = fdiv(real(1), fsqrt(real(2))); bz_ = fdiv(real(1), fsqrt(real(2))); }
    else { real t = fdiv(real(1), bt); by_ = qa[5] * t; bz_ = qa[6] * t; }
    real signbt = (fabs(qa[5]) >= zero) ? sgn1(qa[5]) : sgn1(qa[6]);
    real amc = cs * cs - ca * ca;
    real sbx = sgn1(qa[4]);
    real R[6][6];
    {
        real t1 = af * qa[2], t2 = as * by_ * qa[4], t3 = af * qa[3], t4 = as * bz_ * qa[4];
        real t5 = as * cf * isrho;
        R[0][0] = af; R[1][0] = af * ev[0]; R[2][0] = t1 + t2; R[3][0] = t3 + t4; R[4][0] = t5 * by_; R[5][0] = t5 * bz_;
        R[0][5] = af; R[1][5] = af * ev[5]; R[2][5] = t1 - t2; R[3][5] = t3 - t4; R[4][5] = t5 * by_; R[5][5] = t5 * bz_;
        if (amc < zero) {
#pragma unroll
            for (int m = 0; m < 6; m++) { R[m][0] = signbt * R[m][0]; R[m][5] = signbt * R[m][5]; }
        }
        R[0][1] = real(0); R[1][1] = real(0); R[2][1] = sbx * bz_; R[3][1] = -sbx * by_;
        R[4][1] = isrho * bz_; R[5][1] = -isrho * by_;
        R[0][4] = real(0); R[1][4] = real(0); R[2][4] = -sbx * bz_; R[3][4] = sbx * by_;
        R[4][4] = isrho * bz_; R[5][4] = -isrho * by_;
        real s1 = as * qa[2], s2 = af * by_ * cs * sbx, s3 = as * qa[3], s4 = af * bz_ * cs * sbx;
        real s5 = fdiv(af * (cs * cs) * isrho, cf);
        R[0][2] = as; R[1][2] = as * ev[2]; R[2][2] = s1 - s2; R[3][2] = s3 - s4; R[4][2] = -s5 * by_; R[5][2] = -s5 * bz_;
        R[0][3] = as; R[1][3] = as * ev[3]; R[2][3] = s1 + s2; R[3][3] = s3 + s4; R[4][3] = -s5 * by_; R[5][3] = -s5 * bz_;
        if (amc > zero) {
#pragma unroll
            for (int m = 0; m < 6; m++) { R[m][2] = signbt * R[m][2]; R[m][3] = signbt * R[m][3]; }
        }
    }
    real dq[6] = {qR[0] - qL[0], qR[1] - qL[1], qR[2] - qL[2], qR[3] - qL[3], qR[5] - qL[5], qR[6] - qL[6]};
    real ith1 = fdiv(real(1), real(2) * (af * af * (cs * cs) + as * as * (cf * cf)));
    real ith2 = fdiv(real(1), real(2) * (af * af * cs * cf + as * as * cslow * ca));
    real ck[6];
    {
        real t1 = ith1 * af * (cs * cs);
        real t2 = ith2 * (-(af * cs * qa[1]) + as * cslow * sbx * (by_ * qa[2] + bz_ * qa[3]));
        real t3 = ith2 * as * cslow * sbx;
        real t4 = ith1 * as * cf * fsqrt(qa[0]);
        real Lm[6] = {t1 - t2, -(ith2 * af * cs), t3 * by_, t3 * bz_, t4 * by_, t4 * bz_};
        real Lp[6] = {t1 + t2, -Lm[1], -Lm[2], -Lm[3], Lm[4], Lm[5]};
        if (amc < zero) {
#pragma unroll
            for (int m = 0; m < 6; m++) { Lm[m] = signbt * Lm[m]; Lp[m] = signbt * Lp[m]; }
        }
        ck[0] = dq[0] * Lm[0] + dq[1] * Lm[1] + dq[2] * Lm[2] + dq[3] * Lm[3] + dq[4] * Lm[4] + dq[5] * Lm[5];
        ck[5] = dq[0] * Lp[0] + dq[1] * Lp[1] + dq[2] * Lp[2] + dq[3] * Lp[3] + dq[4] * Lp[4] + dq[5] * Lp[5];
    }
    {
        real t1 = real(0.5) * (bz_ * qa[2] - by_ * qa[3]) * sbx;
        real t2 = real(0.5) * sbx;
        real t3 = real(0.5) * fsqrt(qa[0]);
        real Lm[6] = {-t1, real(0), t2 * bz_, -(t2 * by_), t3 * bz_, -(t3 * by_)};
        real Lp[6] = {t1, real(0), -Lm[2], -Lm[3], Lm[4], Lm[5]};
        ck[1] = dq[0] * Lm[0] + dq[1] * Lm[1] + dq[2] * Lm[2] + dq[3] * Lm[3] + dq[4] * Lm[4] + dq[5] * Lm[5];
        ck[4] = dq[0] * Lp[0] + dq[1] * Lp[1] + dq[2] * Lp[2] + dq[3] * Lp[3] + dq[4] * Lp[4] + dq[5] * Lp[5];
    }
    {
        real t1 = ith1 * as * (cf * cf);
        real t2 = ith2 * (as * ca * qa[1] + af * cf * sbx * (by_ * qa[2] + bz_ * qa[3]));
        real t3 = ith2 * af * cf * sbx;
        real t4 = ith1 * af * cf * fsqrt(qa[0]);
        real Lm[6] = {t1 + t2, -(ith2 * as * ca), -(t3 * by_), -(t3 * bz_), -(t4 * by_), -(t4 * bz_)};
        real Lp[6] = {t1 - t2, -Lm[1], -Lm[2], -Lm[3], Lm[4], Lm[5]};
        if (amc > zero) {
#pragma unroll
            for (int m = 0; m < 6; m++) { Lm[m] = signbt * Lm[m]; Lp[m] = signbt * Lp[m]; }
        }
        ck[2] = dq[0] * Lm[0] + dq[1] * Lm[1] + dq[2] * Lm[2] + dq[3] * Lm[3] + dq[4] * Lm[4] + dq[5] * Lm[5];
        ck[3] = dq[0] * Lp[0] + dq[1] * Lp[1] + dq[2] * Lp[2] + dq[3] * Lp[3] + dq[4] * Lp[4] + dq[5] * Lp[5];
    }
    real FLc[6], FRc[6];
    {
        real rho = qL[0], vx = fdiv(qL[1], rho), vy = fdiv(qL[2], rho), vz = fdiv(qL[3], rho), Bxx = qL[4], By = qL[5], Bz = qL[6];
        FLc[0] = rho * vx;
        FLc[1] = rho * (vx * vx + cs * cs) + real(0.5) * (By * By + Bz * Bz - Bxx * Bxx);
        FLc[2] = rho * vx * vy - Bxx * By;
        FLc[3] = rho * vx * vz - Bxx * Bz;
        FLc[4] = vx * By - vy * Bxx;
        FLc[5] = vx * Bz - vz * Bxx;
    }
    {
        real rho = qR[0], vx = fdiv(qR[1], rho), vy = fdiv(qR[2], rho), vz = fdiv(qR[3], rho), Bxx = qR[4], By = qR[5], Bz = qR[6];
        FRc[0] = rho * vx;
        FRc[1] = rho * (vx * vx + cs * cs) + real(0.5) * (By * By + Bz * Bz - Bxx * Bxx);
        FRc[2] = rho * vx * vy - Bxx * By;
        FRc[3] = rho * vx * vz - Bxx * Bz;
        FRc[4] = vx * By - vy * Bxx;
        FRc[5] = vx * Bz - vz * Bxx;
    }
    real FT[6] = {real(0), real(0), real(0), real(0), real(0), real(0)};
    const real dh[6] = {real(0.3), real(0), real(0.3), real(0.3), real(0), real(0.3)};  // vanleer_mod.f90:78
#pragma unroll
    for (int k = 0; k < 6; k++) {
        real a = fabs(ev[k]);
        real betak = (a >= dh[k]) ? a * ck[k] : fdiv(a + dh[k] * dh[k], real(2) * dh[k]) * ck[k];
#pragma unroll
        for (int m = 0; m < 6; m++) FT[m] = FT[m] - betak * R[m][k];
    }
    F[0] = real(0.5) * (FT[0] + FLc[0] + FRc[0]);
    F[1] = real(0.5) * (FT[1] + FLc[1] + FRc[1]);
    F[2] = real(0.5) * (FT[2] + FLc[2] + FRc[2]);
    F[3] = real(0.5) * (FT[3] + FLc[3] + FRc[3]);
    F[4] = real(0);
    F[5] = real(0.5) * (FT[4] + FLc[4] + FRc[4]);
    F[6] = real(0.5) * (FT[5] + FLc[5] + FRc[5]);
}

template <int RS, class real>
__device__ __forceinline__ void riemann(const real* uL, const real* uR, real cs, real* F, real& smax) {
    if constexpr (RS == RS_HLLD) hlld(uL, uR, cs * cs, F, smax);
    else if constexpr (RS == RS_HLLE) hlle(uL, uR, cs * cs, F, smax);
    else roe(uL, uR, cs, F, smax);
}

}  // namespace imhd
