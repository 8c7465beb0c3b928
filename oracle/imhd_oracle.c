/* CPU FP64 restatement of the reference's isothermal-MHD cell-update path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): used by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg.  The product library never links or calls it.
 *
 * It follows the reference (tsultan16/Isothermal_MHD, 3D_MPI/src) operation for
 * operation -- same index ranges, same evaluation order and parenthesisation, same
 * guards and quirks -- so that, built without FMA contraction, it reproduces the
 * compiled reference bit for bit (checked by tests/test_oracle_vs_reference.py).
 * Each function cites the reference file:line it restates.
 *
 * Parity status per function (how each one is pinned) is listed in the header of
 * tests/test_oracle_vs_reference.py and tests/test_oracle_properties.py.
 */
#define _GNU_SOURCE   /* sincos */
#include "imhd_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ZERO 1.0e-20      /* vanleer_mod.f90:75 */
#define MIN_DENS 1.0e-20  /* vanleer_mod.f90:76 */

struct orc_ctx {
    orc_cfg c;
    int NX, NY, NZ;
    long NC;
    double *q, *bf, *ec;   /* q_3d(7), bface_3d(3), emf_corner(3) */
    double *qw, *bfw;      /* qwork_3d, bface_work */
    double *Fs;            /* Fs(.,.,.,7,3) */
    double *emf;           /* emf_3d(.,.,.,3,3) */
    double *Fbuf;          /* TVD: Fs_buff(.,.,.,7,3) */
    /* flux protection (vanleer_mod.f90:93-110): Fs_prot, emf_3d_prot, emf_corner_prot,
     * q_3d_prot, protection_flags */
    double *Fsp, *emfp, *ecp, *qprot;
    unsigned char *pflag;
    double smax, delv_max, dx, dt;
    int floor_count;
    /* passive (passive_mod.f90:77-84) */
    int PX, PY, PZ;
    long PC;
    double *qp, *qpw;
    int VX, VY, VZ;
    long VC;
    double *vel, *vtavg, *Fp;
    /* turbulence forcing (turb_driver_mod.f90): delv_re on the padded grid, rho0 */
    double *delv, *rho0;
    double edot, dt_turb;
    double quad1, quad2, mavg[3];
    struct orc_ou_s *ou;
};

/* ----------------------------------------------------------------------------------- */
/* indexing                                                                            */

static inline long cidx(const orc_ctx *c, int i, int j, int k) {
    int nb = c->c.nb;
    return (long)(i + nb - 1) + (long)c->NX * ((long)(j + nb - 1) + (long)c->NY * (long)(k + nb - 1));
}
#define Q(f, i, j, k) c->q[(long)(f) * c->NC + cidx(c, i, j, k)]
#define QW(f, i, j, k) c->qw[(long)(f) * c->NC + cidx(c, i, j, k)]
#define BF(f, i, j, k) c->bf[(long)(f) * c->NC + cidx(c, i, j, k)]
#define BFW(f, i, j, k) c->bfw[(long)(f) * c->NC + cidx(c, i, j, k)]
#define EC(f, i, j, k) c->ec[(long)(f) * c->NC + cidx(c, i, j, k)]
#define FS(comp, dir, i, j, k) c->Fs[((long)(dir) * 7 + (comp)) * c->NC + cidx(c, i, j, k)]
#define FB(comp, dir, i, j, k) c->Fbuf[((long)(dir) * 7 + (comp)) * c->NC + cidx(c, i, j, k)]
#define EMF(comp, slot, i, j, k) c->emf[((long)(slot) * 3 + (comp)) * c->NC + cidx(c, i, j, k)]
#define FSP(comp, dir, i, j, k) c->Fsp[((long)(dir) * 7 + (comp)) * c->NC + cidx(c, i, j, k)]
#define EMFP(comp, slot, i, j, k) c->emfp[((long)(slot) * 3 + (comp)) * c->NC + cidx(c, i, j, k)]
#define ECP(f, i, j, k) c->ecp[(long)(f) * c->NC + cidx(c, i, j, k)]

static inline long pidx(const orc_ctx *c, int i, int j, int k) {
    int nb = c->c.nb;
    return (long)(i + nb - 3) + (long)c->PX * ((long)(j + nb - 3) + (long)c->PY * (long)(k + nb - 3));
}
static inline long vidx(const orc_ctx *c, int i, int j, int k) {
    return (long)(i + 2) + (long)c->VX * ((long)(j + 2) + (long)c->VY * (long)(k + 2));
}
#define QP(p, i, j, k) c->qp[(long)(p) * c->PC + pidx(c, i, j, k)]
#define VEL(d, i, j, k) c->vel[(long)(d) * c->VC + vidx(c, i, j, k)]
#define VTA(d, i, j, k) c->vtavg[(long)(d) * c->VC + vidx(c, i, j, k)]
#define FP(p, d, i, j, k) c->Fp[((long)(d) * c->c.npass + (p)) * c->VC + vidx(c, i, j, k)]

/* component permutation of a directional sweep: rotated slot r <- global field perm[d][r]
 * (vanleer_mod.f90:449-464, 610-625, 788-803; scatter :712-718, :889-895) */
static const int PERM[3][7] = {{0, 1, 2, 3, 4, 5, 6}, {0, 2, 3, 1, 5, 6, 4}, {0, 3, 1, 2, 6, 4, 5}};

static inline void dir_cell(int d, int n, int t1, int t2, int *i, int *j, int *k) {
    if (d == 0) { *i = n; *j = t1; *k = t2; }
    else if (d == 1) { *i = t1; *j = n; *k = t2; }
    else { *i = t1; *j = t2; *k = n; }
}

static inline double dmin3(double a, double b, double c) { double m = a < b ? a : b; return m < c ? m : c; }
static inline double dmax3(double a, double b, double c) { double m = a > b ? a : b; return m > c ? m : c; }
static inline double dmax2(double a, double b) { return a > b ? a : b; }
static inline double sgn(double x) { return copysign(1.0, x); } /* Fortran SIGN(1.d0, x) */

/* ----------------------------------------------------------------------------------- */
/* Riemann solvers (one interface).  u = (rho, mx, my, mz, Bx_face, By, Bz) rotated.   */

/* fast magnetosonic speed of a conserved state, vanleer_mod.f90:3048-3053 */
static inline double fast_speed(const double *u, double asqr) {
    double bsqr = (u[4] * u[4] + u[5] * u[5] + u[6] * u[6]) / u[0];
    double btsqr = (u[5] * u[5] + u[6] * u[6]) / u[0];
    double tmp1 = asqr + bsqr;
    double x = asqr - bsqr;
    double tmp2 = sqrt(x * x + 4.0 * asqr * btsqr);
    return sqrt(0.5 * (tmp1 + tmp2));
}

/* physical flux, vanleer_mod.f90:3029-3043 */
static inline void phys_flux(const double *u, double asqr, double vx, double vy, double vz, double *F) {
    F[0] = u[1];
    F[1] = u[0] * (vx * vx + asqr) + 0.5 * (u[5] * u[5] + u[6] * u[6] - u[4] * u[4]);
    F[2] = u[1] * vy - u[4] * u[5];
    F[3] = u[1] * vz - u[4] * u[6];
    F[4] = 0.0;
    F[5] = vx * u[5] - vy * u[4];
    F[6] = vx * u[6] - vz * u[4];
}

/* compute_HLLE_fluxes_1d, vanleer_mod.f90:2867-2981 (one interface) */
static void hlle_if(const double *uL, const double *uR, double asqr, double *F, double *smax) {
    double uavg[7], FL[7], FR[7];
    for (int m = 0; m < 7; m++) uavg[m] = 0.5 * (uL[m] + uR[m]);
    double vxL = uL[1] / uL[0], vyL = uL[2] / uL[0], vzL = uL[3] / uL[0];
    double vxR = uR[1] / uR[0], vyR = uR[2] / uR[0], vzR = uR[3] / uR[0];
    phys_flux(uL, asqr, vxL, vyL, vzL, FL);
    phys_flux(uR, asqr, vxR, vyR, vzR, FR);
    double cfL = fast_speed(uL, asqr), cfR = fast_speed(uR, asqr), cfa = fast_speed(uavg, asqr);
    double vxa = uavg[1] / uavg[0];
    double sL = dmin3(vxL - cfL, vxa - cfa, 0.0);
    double sR = dmax3(vxR + cfR, vxa + cfa, 0.0);
    *smax = dmax2(*smax, dmax3(fabs(vxL) + cfL, fabs(vxR) + cfR, fabs(vxa) + cfa));
    double isRL = 1.0 / (sR - sL);
    for (int m = 0; m < 7; m++) F[m] = (sR * FL[m] - sL * FR[m] + sR * sL * (uR[m] - uL[m])) * isRL;
}

/* compute_HLLD_fluxes_1d, vanleer_mod.f90:2986-3239 (one interface) */
static void hlld_if(const double *uL, const double *uR, double asqr, double *F, double *smax) {
    double uavg[7], FL[7], FR[7], uHLL[7], FHLL[7], uLs[7], uRs[7], uCs[7];
    for (int m = 0; m < 7; m++) uavg[m] = 0.5 * (uL[m] + uR[m]);
    double Bx = uL[4];
    double vxL = uL[1] / uL[0], vyL = uL[2] / uL[0], vzL = uL[3] / uL[0];
    double vxR = uR[1] / uR[0], vyR = uR[2] / uR[0], vzR = uR[3] / uR[0];
    phys_flux(uL, asqr, vxL, vyL, vzL, FL);
    phys_flux(uR, asqr, vxR, vyR, vzR, FR);
    double cfL = fast_speed(uL, asqr), cfR = fast_speed(uR, asqr), cfa = fast_speed(uavg, asqr);
    double vxa = uavg[1] / uavg[0];
    double sL = dmin3(vxL - cfL, vxa - cfa, 0.0);
    double sR = dmax3(vxR + cfR, vxa + cfa, 0.0);
    *smax = dmax2(*smax, dmax3(fabs(vxL) + cfL, fabs(vxR) + cfR, fabs(vxa) + cfa));
    if (sL >= 0.0) { for (int m = 0; m < 7; m++) F[m] = FL[m]; return; }
    if (sR <= 0.0) { for (int m = 0; m < 7; m++) F[m] = FR[m]; return; }
    double isRL = 1.0 / (sR - sL);
    for (int m = 0; m < 7; m++) {
        uHLL[m] = (sR * uR[m] - sL * uL[m] - FR[m] + FL[m]) * isRL;
        FHLL[m] = (sR * FL[m] - sL * FR[m] + sR * sL * (uR[m] - uL[m])) * isRL;
    }
    double rho_s = uHLL[0], momx_s = uHLL[1];
    double vx_s = FHLL[0] / rho_s;
    double irho_s = 1.0 / rho_s;
    double isrho_s = sqrt(irho_s);
    double sLs = vx_s - fabs(Bx) * isrho_s;
    double sRs = vx_s + fabs(Bx) * isrho_s;
    uLs[0] = rho_s; uLs[1] = momx_s;
    if (fabs(sL - sLs) < ZERO || fabs(sL - sRs) < ZERO) {
        for (int m = 2; m < 7; m++) uLs[m] = uL[m];
    } else {
        double is = 1.0 / ((sL - sLs) * (sL - sRs));
        double t1 = Bx * (vx_s - vxL) * is;
        double d = sL - vxL;
        double t2 = (uL[0] * (d * d) - Bx * Bx) * irho_s * is;
        uLs[2] = rho_s * vyL - t1 * uL[5];
        uLs[3] = rho_s * vzL - t1 * uL[6];
        uLs[4] = uL[4];
        uLs[5] = t2 * uL[5];
        uLs[6] = t2 * uL[6];
    }
    uRs[0] = rho_s; uRs[1] = momx_s;
    if (fabs(sR - sLs) < ZERO || fabs(sR - sRs) < ZERO) {
        for (int m = 2; m < 7; m++) uRs[m] = uR[m];
    } else {
        double is = 1.0 / ((sR - sLs) * (sR - sRs));
        double t1 = Bx * (vx_s - vxR) * is;
        double d = sR - vxR;
        double t2 = (uR[0] * (d * d) - Bx * Bx) * irho_s * is;
        uRs[2] = rho_s * vyR - t1 * uR[5];
        uRs[3] = rho_s * vzR - t1 * uR[6];
        uRs[4] = uR[4];
        uRs[5] = t2 * uR[5];
        uRs[6] = t2 * uR[6];
    }
    double X = sqrt(rho_s) * sgn(Bx);
    double iX = 1.0 / X;
    uCs[0] = rho_s; uCs[1] = momx_s;
    uCs[2] = 0.5 * (uLs[2] + uRs[2] + X * (uRs[5] - uLs[5]));
    uCs[3] = 0.5 * (uLs[3] + uRs[3] + X * (uRs[6] - uLs[6]));
    uCs[4] = uL[4];
    uCs[5] = 0.5 * (uLs[5] + uRs[5] + iX * (uRs[2] - uLs[2]));
    uCs[6] = 0.5 * (uLs[6] + uRs[6] + iX * (uRs[3] - uLs[3]));
    if (sLs >= 0.0) {
        for (int m = 0; m < 7; m++) F[m] = FL[m] + sL * (uLs[m] - uL[m]);
    } else if (sRs <= 0.0) {
        for (int m = 0; m < 7; m++) F[m] = FR[m] + sR * (uRs[m] - uR[m]);
    } else {
        F[0] = FHLL[0];
        F[1] = FHLL[1];
        F[2] = uCs[2] * vx_s - Bx * uCs[5];
        F[3] = uCs[3] * vx_s - Bx * uCs[6];
        F[4] = 0.0;
        F[5] = vx_s * uCs[5] - (uCs[2] * Bx) / rho_s;
        F[6] = vx_s * uCs[6] - (uCs[3] * Bx) / rho_s;
    }
}

/* Roe chain of the VL module (one interface): compute_roe_avg_state_1d :2264-2296,
 * compute_roe_eigenvalues_1d :2300-2353, compute_roe_eigenvectors_1d :2357-2745,
 * compute_ROE_fluxes_1d :2749-2863 (vanleer_mod.f90). */
static void roe_if(const double *qL, const double *qR, double cs, double *F, double *smax) {
    static const double delta_h[6] = {0.3, 0.0, 0.3, 0.3, 0.0, 0.3}; /* :78 */
    double asqr = cs * cs;
    double qa[7], R[6][6], ck[6], ev[6], Lm[6], Lp[6], delq[6];
    double rho2 = 0.5 * (qL[0] + qR[0]);
    double isrho = 1.0 / sqrt(rho2);
    qa[0] = rho2;
    qa[1] = 0.5 * (qL[1] / qL[0] + qR[1] / qR[0]);
    qa[2] = 0.5 * (qL[2] / qL[0] + qR[2] / qR[0]);
    qa[3] = 0.5 * (qL[3] / qL[0] + qR[3] / qR[0]);
    qa[4] = qL[4] * isrho;
    qa[5] = 0.5 * (qL[5] + qR[5]) * isrho;
    qa[6] = 0.5 * (qL[6] + qR[6]) * isrho;
    /* eigenvalues */
    double vx2 = qa[1], bx = qa[4];
    double bsqr = qa[4] * qa[4] + qa[5] * qa[5] + qa[6] * qa[6];
    double btsqr = qa[5] * qa[5] + qa[6] * qa[6];
    double tmp1 = asqr + bsqr, x = asqr - bsqr;
    double tmp2 = sqrt(x * x + 4.0 * asqr * btsqr);
    double cslow = sqrt(dmax2(0.0, 0.5 * (tmp1 - tmp2)));
    double ca = fabs(bx);
    double cf = sqrt(0.5 * (tmp1 + tmp2));
    ev[0] = vx2 - cf; ev[1] = vx2 - ca; ev[2] = vx2 - cslow;
    ev[3] = vx2 + cslow; ev[4] = vx2 + ca; ev[5] = vx2 + cf;
    *smax = dmax2(*smax, fabs(vx2) + cf);
    /* eigenvector renormalisation ("MODIFIED V1", :2392-2405) */
    double bt = sqrt(qa[5] * qa[5] + qa[6] * qa[6]);
    double af, as;
    if (fabs(cf * cf - cslow * cslow) < ZERO) { af = 1.0; as = 1.0; }
    else if (fabs(cf * cf - ca * ca) < ZERO) { af = 0.0; as = 1.0; }
    else if (fabs(cf * cf - asqr) < ZERO) { af = 1.0; as = 0.0; }
    else {
        double t = 1.0 / sqrt(dmax2(0.0, cf * cf - cslow * cslow));
        af = sqrt(dmax2(0.0, cf * cf - ca * ca)) * t;
        as = sqrt(dmax2(0.0, cf * cf - asqr)) * t;
    }
    double by_, bz_;
    if (bt < ZERO) { by_ = 1.0 / sqrt(2.0); bz_ = 1.0 / sqrt(2.0); }
    else { double t = 1.0 / bt; by_ = qa[5] * t; bz_ = qa[6] * t; }
    double signbt = (fabs(qa[5]) >= ZERO) ? sgn(qa[5]) : sgn(qa[6]);
    double amc = cs * cs - ca * ca;
    /* fast right eigenvectors :2449-2487 */
    {
        double t1 = af * qa[2], t2 = as * by_ * qa[4], t3 = af * qa[3], t4 = as * bz_ * qa[4];
        double t5 = as * cf * isrho;
        R[0][0] = af; R[1][0] = af * ev[0]; R[2][0] = t1 + t2; R[3][0] = t3 + t4;
        R[4][0] = t5 * by_; R[5][0] = t5 * bz_;
        R[0][5] = af; R[1][5] = af * ev[5]; R[2][5] = t1 - t2; R[3][5] = t3 - t4;
        R[4][5] = t5 * by_; R[5][5] = t5 * bz_;
        if (amc < ZERO) for (int m = 0; m < 6; m++) { R[m][0] = signbt * R[m][0]; R[m][5] = signbt * R[m][5]; }
    }
    /* Alfven right eigenvectors :2495-2512 */
    {
        double sbx = sgn(qa[4]);
        R[0][1] = 0.0; R[1][1] = 0.0; R[2][1] = sbx * bz_; R[3][1] = -sbx * by_;
        R[4][1] = isrho * bz_; R[5][1] = -isrho * by_;
        R[0][4] = 0.0; R[1][4] = 0.0; R[2][4] = -sbx * bz_; R[3][4] = sbx * by_;
        R[4][4] = isrho * bz_; R[5][4] = -isrho * by_;
    }
    /* slow right eigenvectors :2520-2557 */
    {
        double sbx = sgn(qa[4]);
        double t1 = as * qa[2], t2 = af * by_ * cs * sbx, t3 = as * qa[3], t4 = af * bz_ * cs * sbx;
        double t5 = af * (cs * cs) * isrho / cf;
        R[0][2] = as; R[1][2] = as * ev[2]; R[2][2] = t1 - t2; R[3][2] = t3 - t4;
        R[4][2] = -t5 * by_; R[5][2] = -t5 * bz_;
        R[0][3] = as; R[1][3] = as * ev[3]; R[2][3] = t1 + t2; R[3][3] = t3 + t4;
        R[4][3] = -t5 * by_; R[5][3] = -t5 * bz_;
        if (amc > ZERO) for (int m = 0; m < 6; m++) { R[m][2] = signbt * R[m][2]; R[m][3] = signbt * R[m][3]; }
    }
    for (int m = 0; m < 4; m++) delq[m] = qR[m] - qL[m];
    delq[4] = qR[5] - qL[5];
    delq[5] = qR[6] - qL[6];
    double sbx = sgn(qa[4]);
    double ith1 = 1.0 / (2.0 * (af * af * (cs * cs) + as * as * (cf * cf)));
    double ith2 = 1.0 / (2.0 * (af * af * cs * cf + as * as * cslow * ca));
    /* fast left eigenvectors / characteristic variables :2566-2628 */
    {
        double t1 = ith1 * af * (cs * cs);
        double t2 = ith2 * (-(af * cs * qa[1]) + as * cslow * sbx * (by_ * qa[2] + bz_ * qa[3]));
        double t3 = ith2 * as * cslow * sbx;
        double t4 = ith1 * as * cf * sqrt(qa[0]);
        Lm[0] = t1 - t2; Lp[0] = t1 + t2;
        Lm[1] = -(ith2 * af * cs); Lp[1] = -Lm[1];
        Lm[2] = t3 * by_; Lp[2] = -Lm[2];
        Lm[3] = t3 * bz_; Lp[3] = -Lm[3];
        Lm[4] = t4 * by_; Lp[4] = Lm[4];
        Lm[5] = t4 * bz_; Lp[5] = Lm[5];
        if (amc < ZERO) for (int m = 0; m < 6; m++) { Lm[m] = signbt * Lm[m]; Lp[m] = signbt * Lp[m]; }
        ck[0] = delq[0] * Lm[0] + delq[1] * Lm[1] + delq[2] * Lm[2] + delq[3] * Lm[3] + delq[4] * Lm[4] + delq[5] * Lm[5];
        ck[5] = delq[0] * Lp[0] + delq[1] * Lp[1] + delq[2] * Lp[2] + delq[3] * Lp[3] + delq[4] * Lp[4] + delq[5] * Lp[5];
    }
    /* Alfven :2633-2675 */
    {
        double t1 = 0.5 * (bz_ * qa[2] - by_ * qa[3]) * sbx;
        double t2 = 0.5 * sbx;
        double t3 = 0.5 * sqrt(qa[0]);
        Lm[0] = -t1; Lp[0] = t1;
        Lm[1] = 0.0; Lp[1] = 0.0;
        Lm[2] = t2 * bz_; Lp[2] = -Lm[2];
        Lm[3] = -(t2 * by_); Lp[3] = -Lm[3];
        Lm[4] = t3 * bz_; Lp[4] = Lm[4];
        Lm[5] = -(t3 * by_); Lp[5] = Lm[5];
        ck[1] = delq[0] * Lm[0] + delq[1] * Lm[1] + delq[2] * Lm[2] + delq[3] * Lm[3] + delq[4] * Lm[4] + delq[5] * Lm[5];
        ck[4] = delq[0] * Lp[0] + delq[1] * Lp[1] + delq[2] * Lp[2] + delq[3] * Lp[3] + delq[4] * Lp[4] + delq[5] * Lp[5];
    }
    /* slow :2678-2742 */
    {
        double t1 = ith1 * as * (cf * cf);
        double t2 = ith2 * (as * ca * qa[1] + af * cf * sbx * (by_ * qa[2] + bz_ * qa[3]));
        double t3 = ith2 * af * cf * sbx;
        double t4 = ith1 * af * cf * sqrt(qa[0]);
        Lm[0] = t1 + t2; Lp[0] = t1 - t2;
        Lm[1] = -(ith2 * as * ca); Lp[1] = -Lm[1];
        Lm[2] = -(t3 * by_); Lp[2] = -Lm[2];
        Lm[3] = -(t3 * bz_); Lp[3] = -Lm[3];
        Lm[4] = -(t4 * by_); Lp[4] = Lm[4];
        Lm[5] = -(t4 * bz_); Lp[5] = Lm[5];
        if (amc > ZERO) for (int m = 0; m < 6; m++) { Lm[m] = signbt * Lm[m]; Lp[m] = signbt * Lp[m]; }
        ck[2] = delq[0] * Lm[0] + delq[1] * Lm[1] + delq[2] * Lm[2] + delq[3] * Lm[3] + delq[4] * Lm[4] + delq[5] * Lm[5];
        ck[3] = delq[0] * Lp[0] + delq[1] * Lp[1] + delq[2] * Lp[2] + delq[3] * Lp[3] + delq[4] * Lp[4] + delq[5] * Lp[5];
    }
    /* flux :2764-2857 */
    double FLc[6], FRc[6], FT[6] = {0, 0, 0, 0, 0, 0};
    const double *us[2] = {qL, qR};
    double *Fc[2] = {FLc, FRc};
    for (int s = 0; s < 2; s++) {
        const double *u = us[s];
        double rho = u[0], vx = u[1] / rho, vy = u[2] / rho, vz = u[3] / rho;
        double Bxx = u[4], By = u[5], Bz = u[6];
        double *f = Fc[s];
        f[0] = rho * vx;
        f[1] = rho * (vx * vx + cs * cs) + 0.5 * (By * By + Bz * Bz - Bxx * Bxx);
        f[2] = rho * vx * vy - Bxx * By;
        f[3] = rho * vx * vz - Bxx * Bz;
        f[4] = vx * By - vy * Bxx;
        f[5] = vx * Bz - vz * Bxx;
    }
    for (int k = 0; k < 6; k++) {
        double betak, a = fabs(ev[k]);
        if (a >= delta_h[k]) betak = a * ck[k];
        else betak = ((a + delta_h[k] * delta_h[k]) / (2.0 * delta_h[k])) * ck[k];
        for (int m = 0; m < 6; m++) FT[m] = FT[m] - betak * R[m][k];
    }
    F[0] = 0.5 * (FT[0] + FLc[0] + FRc[0]);
    F[1] = 0.5 * (FT[1] + FLc[1] + FRc[1]);
    F[2] = 0.5 * (FT[2] + FLc[2] + FRc[2]);
    F[3] = 0.5 * (FT[3] + FLc[3] + FRc[3]);
    F[4] = 0.0;
    F[5] = 0.5 * (FT[4] + FLc[4] + FRc[4]);
    F[6] = 0.5 * (FT[5] + FLc[5] + FRc[5]);
}

static void vl_riemann(int kind, const double *uL, const double *uR, double cs, double *F, double *smax) {
    if (kind == 2) hlld_if(uL, uR, cs * cs, F, smax);
    else if (kind == 3) hlle_if(uL, uR, cs * cs, F, smax);
    else roe_if(uL, uR, cs, F, smax);
}

/* ----------------------------------------------------------------------------------- */
/* Reconstruction along one pencil.  w[n*6+k], wb[n] for n in [0, len); produces       */
/* qL/qR[n*7+m] for interfaces n (between cell n and n+1) in [*ilo, *ihi].            */

/* compute_interface_states_prim, PLM branch (vanleer_mod.f90:1160-1215) */
static void recon_plm_prim(int len, double beta, const double *w, const double *wb, double *delw,
                           double *qL, double *qR, int *ilo, int *ihi) {
    for (int n = 1; n <= len - 2; n++) {
        for (int k = 0; k < 6; k++) {
            double dL = w[n * 6 + k] - w[(n - 1) * 6 + k];
            double dR = w[(n + 1) * 6 + k] - w[n * 6 + k];
            double dC = 0.5 * (dL + dR);
            delw[n * 6 + k] = sgn(dC) * dmin3(beta * fabs(dL), beta * fabs(dR), fabs(dC));
        }
    }
    for (int n = 1; n <= len - 3; n++) {
        double *l = qL + n * 7, *r = qR + n * 7;
        const double *wc = w + n * 6, *wn = w + (n + 1) * 6, *dc = delw + n * 6, *dn = delw + (n + 1) * 6;
        l[0] = wc[0] + 0.5 * dc[0]; l[1] = wc[1] + 0.5 * dc[1]; l[2] = wc[2] + 0.5 * dc[2];
        l[3] = wc[3] + 0.5 * dc[3]; l[4] = wb[n]; l[5] = wc[4] + 0.5 * dc[4]; l[6] = wc[5] + 0.5 * dc[5];
        r[0] = wn[0] - 0.5 * dn[0]; r[1] = wn[1] - 0.5 * dn[1]; r[2] = wn[2] - 0.5 * dn[2];
        r[3] = wn[3] - 0.5 * dn[3]; r[4] = wb[n]; r[5] = wn[4] - 0.5 * dn[4]; r[6] = wn[5] - 0.5 * dn[5];
        l[1] = l[1] * l[0]; l[2] = l[2] * l[0]; l[3] = l[3] * l[0];
        r[1] = r[1] * r[0]; r[2] = r[2] * r[0]; r[3] = r[3] * r[0];
    }
    *ilo = 1;
    *ihi = len - 3;
}

/* CW84-style monotonicity of a parabola (vanleer_mod.f90:1313-1332, char: :1724-1743).
 * flat_to_linear selects the primitive branch (edges reset to w -/+ delw/2). */
static inline void ppm_monotone(double w, double dw, double *wL, double *wR, int flat_to_linear) {
    double tmp1 = (*wR - w) * (w - *wL);
    double tmp2 = 6.0 * (*wR - *wL) * (w - 0.5 * (*wL + *wR));
    double d = *wR - *wL;
    double tmp3 = d * d;
    if (tmp1 <= 0.0) {
        if (flat_to_linear) { *wL = w - 0.5 * dw; *wR = w + 0.5 * dw; }
        else { *wL = w; *wR = w; }
    } else if (tmp2 > tmp3) {
        *wL = 3.0 * w - 2.0 * *wR;
    } else if (tmp2 < -tmp3) {
        *wR = 3.0 * w - 2.0 * *wL;
    }
}

static void ppm_states(int len, const double *w, const double *wb, const double *wLa, const double *wRa,
                       double *qL, double *qR, int *ilo, int *ihi) {
    for (int n = 2; n <= len - 4; n++) {
        double *l = qL + n * 7, *r = qR + n * 7;
        const double *a = wRa + n * 6, *b = wLa + (n + 1) * 6;
        l[0] = a[0]; l[1] = a[1]; l[2] = a[2]; l[3] = a[3]; l[4] = wb[n]; l[5] = a[4]; l[6] = a[5];
        r[0] = b[0]; r[1] = b[1]; r[2] = b[2]; r[3] = b[3]; r[4] = wb[n]; r[5] = b[4]; r[6] = b[5];
        l[1] = l[1] * l[0]; l[2] = l[2] * l[0]; l[3] = l[3] * l[0];
        r[1] = r[1] * r[0]; r[2] = r[2] * r[0]; r[3] = r[3] * r[0];
    }
    *ilo = 2;
    *ihi = len - 4;
}

/* compute_interface_states_prim, PPM branch as compiled (vanleer_mod.f90:1217-1334,
 * the block after GO TO 1000 is dead code) */
static void recon_ppm_prim(int len, double beta, const double *w, const double *wb, double *delw,
                           double *wLa, double *wRa, double *qL, double *qR, int *ilo, int *ihi) {
    const double sixth = 1.0 / 6.0;
    for (int n = 1; n <= len - 2; n++) {
        for (int k = 0; k < 6; k++) {
            double dL = w[n * 6 + k] - w[(n - 1) * 6 + k];
            double dR = w[(n + 1) * 6 + k] - w[n * 6 + k];
            double dC = 0.5 * (dL + dR);
            double mabs = fmin(fabs(dL), fabs(dR));
            double lim = 0.0;
            if (dL * dR > 0.0) lim = sgn(dC) * fmin(beta * mabs, fabs(dC));
            delw[n * 6 + k] = lim;
        }
    }
    for (int n = 2; n <= len - 3; n++) {
        for (int k = 0; k < 6; k++) {
            wLa[n * 6 + k] = 0.5 * (w[(n - 1) * 6 + k] + w[n * 6 + k]) + sixth * (delw[(n - 1) * 6 + k] - delw[n * 6 + k]);
            wRa[n * 6 + k] = 0.5 * (w[n * 6 + k] + w[(n + 1) * 6 + k]) + sixth * (delw[n * 6 + k] - delw[(n + 1) * 6 + k]);
        }
    }
    for (int n = 2; n <= len - 3; n++)
        for (int k = 0; k < 6; k++)
            ppm_monotone(w[n * 6 + k], delw[n * 6 + k], &wLa[n * 6 + k], &wRa[n * 6 + k], 1);
    ppm_states(len, w, wb, wLa, wRa, qL, qR, ilo, ihi);
}

/* compute_primitive_eigenvectors (vanleer_mod.f90:941-1141) for one cell: R[m][k], L[k][m] */
static void prim_eigvecs(const double *w, double wb, double cs, double R[6][6], double L[6][6]) {
    double asqr = cs * cs;
    double rho = w[0], irho = 1.0 / rho, rhos = sqrt(rho), irhos = 1.0 / rhos;
    double bx = irhos * wb, by = irhos * w[4], bz = irhos * w[5];
    double bsqr = bx * bx + by * by + bz * bz;
    double btsqr = by * by + bz * bz;
    double tmp1 = asqr + bsqr, x = asqr - bsqr;
    double tmp2 = sqrt(x * x + 4.0 * asqr * btsqr);
    double cslow = sqrt(dmax2(0.0, 0.5 * (tmp1 - tmp2)));
    double cf = sqrt(0.5 * (tmp1 + tmp2));
    double af, as;
    if (fabs(cf * cf - cslow * cslow) < ZERO) { af = 1.0; as = 0.0; }
    else {
        af = sqrt(dmax2((asqr - cslow * cslow) / (cf * cf - cslow * cslow), 0.0));
        as = sqrt(dmax2((cf * cf - asqr) / (cf * cf - cslow * cslow), 0.0));
    }
    double by_, bz_;
    if (btsqr < ZERO) { by_ = 1.0 / sqrt(2.0); bz_ = 1.0 / sqrt(2.0); }
    else { by_ = by / sqrt(btsqr); bz_ = bz / sqrt(btsqr); }
    double S = sgn(bx);
    double Cff = cf * af, Css = cslow * as, Qf = cf * af * S, Qs = cslow * as * S;
    double Af = cs * af * rhos, As = cs * as * rhos;
    R[0][0] = rho * af; R[1][0] = -Cff; R[2][0] = Qs * by_; R[3][0] = Qs * bz_; R[4][0] = As * by_; R[5][0] = As * bz_;
    R[0][1] = 0.0; R[1][1] = 0.0; R[2][1] = -bz_; R[3][1] = by_; R[4][1] = -bz_ * S * rhos; R[5][1] = by_ * S * rhos;
    R[0][2] = rho * as; R[1][2] = -Css; R[2][2] = -Qf * by_; R[3][2] = -Qf * bz_; R[4][2] = -Af * by_; R[5][2] = -Af * bz_;
    R[0][3] = rho * as; R[1][3] = Css; R[2][3] = Qf * by_; R[3][3] = Qf * bz_; R[4][3] = -Af * by_; R[5][3] = -Af * bz_;
    R[0][4] = 0.0; R[1][4] = 0.0; R[2][4] = bz_; R[3][4] = -by_; R[4][4] = -bz_ * S * rhos; R[5][4] = by_ * S * rhos;
    R[0][5] = rho * af; R[1][5] = Cff; R[2][5] = -Qs * by_; R[3][5] = -Qs * bz_; R[4][5] = As * by_; R[5][5] = As * bz_;
    double Nf = 0.5 / asqr, Ns = Nf;
    L[0][0] = Nf * af * asqr * irho; L[0][1] = -(Nf * Cff); L[0][2] = Nf * Qs * by_; L[0][3] = Nf * Qs * bz_;
    L[0][4] = Nf * As * by_ * irho; L[0][5] = Nf * As * bz_ * irho;
    L[1][0] = 0.0; L[1][1] = 0.0; L[1][2] = -(0.5 * bz_); L[1][3] = 0.5 * by_;
    L[1][4] = -(0.5 * bz_ * S * irhos); L[1][5] = 0.5 * by_ * S * irhos;
    L[2][0] = Ns * as * asqr * irho; L[2][1] = -(Ns * Css); L[2][2] = -(Ns * Qf * by_); L[2][3] = -(Ns * Qf * bz_);
    L[2][4] = -(Ns * Af * by_ * irho); L[2][5] = -(Ns * Af * bz_ * irho);
    L[3][0] = Ns * as * asqr * irho; L[3][1] = Ns * Css; L[3][2] = Ns * Qf * by_; L[3][3] = Ns * Qf * bz_;
    L[3][4] = -(Ns * Af * by_ * irho); L[3][5] = -(Ns * Af * bz_ * irho);
    L[4][0] = 0.0; L[4][1] = 0.0; L[4][2] = 0.5 * bz_; L[4][3] = -(0.5 * by_);
    L[4][4] = -(0.5 * bz_ * S * irhos); L[4][5] = 0.5 * by_ * S * irhos;
    L[5][0] = Nf * af * asqr * irho; L[5][1] = Nf * Cff; L[5][2] = -(Nf * Qs * by_); L[5][3] = -(Nf * Qs * bz_);
    L[5][4] = Nf * As * by_ * irho; L[5][5] = Nf * As * bz_ * irho;
}

static inline double dot6(const double *a, const double *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3] + a[4] * b[4] + a[5] * b[5];
}
static inline double mc_lim(double dL, double dR, double dC, double beta) {
    if (dL * dR > 0.0) return sgn(dC) * dmin3(beta * fabs(dL), beta * fabs(dR), fabs(dC));
    return 0.0;
}

/* compute_interface_states_char, PPM branch (vanleer_mod.f90:1595-1818) */
static void recon_ppm_char(int len, double beta, double cs, const double *w, const double *wb, double *delw,
                           double *wLa, double *wRa, double *Rv, double *Lv, double *qL, double *qR, int *ilo,
                           int *ihi) {
    const double sixth = 1.0 / 6.0;
    for (int n = 0; n < len; n++)
        prim_eigvecs(w + n * 6, wb[n], cs, (double(*)[6])(Rv + n * 36), (double(*)[6])(Lv + n * 36));
    for (int n = 1; n <= len - 3; n++) {
        double(*Ln)[6] = (double(*)[6])(Lv + n * 36);
        double(*Rn)[6] = (double(*)[6])(Rv + n * 36);
        double aL1[6], aC[6], aR1[6], aR2[6], dm[6];
        for (int k = 0; k < 6; k++) {
            aL1[k] = dot6(Ln[k], w + (n - 1) * 6);
            aC[k] = dot6(Ln[k], w + n * 6);
            aR1[k] = dot6(Ln[k], w + (n + 1) * 6);
            aR2[k] = dot6(Ln[k], w + (n + 2) * 6);
        }
        for (int k = 0; k < 6; k++) {
            double dL1 = aC[k] - aL1[k], dR1 = aR1[k] - aC[k], dC1 = 0.5 * (dL1 + dR1);
            double dL2 = aR1[k] - aC[k], dR2 = aR2[k] - aR1[k], dC2 = 0.5 * (dL2 + dR2);
            double m1 = mc_lim(dL1, dR1, dC1, beta);
            double m2 = mc_lim(dL2, dR2, dC2, beta);
            dm[k] = m2 - m1;
        }
        for (int k = 0; k < 6; k++)
            delw[n * 6 + k] = dm[0] * Rn[k][0] + dm[1] * Rn[k][1] + dm[2] * Rn[k][2] + dm[3] * Rn[k][3] +
                              dm[4] * Rn[k][4] + dm[5] * Rn[k][5];
    }
    for (int n = 2; n <= len - 3; n++) {
        for (int k = 0; k < 6; k++) {
            wLa[n * 6 + k] = 0.5 * (w[(n - 1) * 6 + k] + w[n * 6 + k]) - sixth * delw[(n - 1) * 6 + k];
            wRa[n * 6 + k] = 0.5 * (w[n * 6 + k] + w[(n + 1) * 6 + k]) - sixth * delw[n * 6 + k];
        }
    }
    for (int n = 2; n <= len - 3; n++)
        for (int k = 0; k < 6; k++) ppm_monotone(w[n * 6 + k], 0.0, &wLa[n * 6 + k], &wRa[n * 6 + k], 0);
    ppm_states(len, w, wb, wLa, wRa, qL, qR, ilo, ihi);
}

/* ----------------------------------------------------------------------------------- */
/* VL2 + CT integrator                                                                  */

static int dim_n(const orc_ctx *c, int d) { return d == 0 ? c->c.nx : (d == 1 ? c->c.ny : c->c.nz); }

/* compute_{x,y,z}fluxes_3d (vanleer_mod.f90:429-938) */
static void vl_sweep(orc_ctx *c, int d, int offset, int corrector) {
    int nb = c->c.nb;
    int lo[3], hi[3];
    for (int a = 0; a < 3; a++) { lo[a] = 1 + offset; hi[a] = dim_n(c, a) - offset; }
    int t1d = (d == 0) ? 1 : 0, t2d = (d == 2) ? 1 : 2;
    int n0 = lo[d] - nb, n1 = hi[d] + nb; /* cell range of the pencil */
    int len = n1 - n0 + 1;
    double *w = malloc(sizeof(double) * len * 6), *wb = malloc(sizeof(double) * len);
    double *qL = calloc((size_t)len * 7, sizeof(double)), *qR = calloc((size_t)len * 7, sizeof(double));
    double *delw = calloc((size_t)len * 6, sizeof(double));
    double *wLa = calloc((size_t)len * 6, sizeof(double)), *wRa = calloc((size_t)len * 6, sizeof(double));
    double *Rv = NULL, *Lv = NULL;
    if (corrector && c->c.recon == 2) { Rv = malloc(sizeof(double) * len * 36); Lv = malloc(sizeof(double) * len * 36); }
    const int *perm = PERM[d];
    for (int t2 = lo[t2d] - nb; t2 <= hi[t2d] + nb; t2++) {
        for (int t1 = lo[t1d] - nb; t1 <= hi[t1d] + nb; t1++) {
            int ilo, ihi;
            if (!corrector) {
                for (int n = 0; n < len - 1; n++) {
                    int i, j, k, i2, j2, k2;
                    dir_cell(d, n0 + n, t1, t2, &i, &j, &k);
                    dir_cell(d, n0 + n + 1, t1, t2, &i2, &j2, &k2);
                    for (int r = 0; r < 7; r++) {
                        qL[n * 7 + r] = Q(perm[r], i, j, k);
                        qR[n * 7 + r] = Q(perm[r], i2, j2, k2);
                    }
                    qL[n * 7 + 4] = BF(d, i, j, k);
                    qR[n * 7 + 4] = BF(d, i, j, k);
                }
                ilo = 0; ihi = len - 2;
            } else {
                for (int n = 0; n < len; n++) {
                    int i, j, k;
                    dir_cell(d, n0 + n, t1, t2, &i, &j, &k);
                    double irho = 1.0 / Q(0, i, j, k);
                    w[n * 6 + 0] = Q(0, i, j, k);
                    w[n * 6 + 1] = Q(perm[1], i, j, k) * irho;
                    w[n * 6 + 2] = Q(perm[2], i, j, k) * irho;
                    w[n * 6 + 3] = Q(perm[3], i, j, k) * irho;
                    wb[n] = BF(d, i, j, k);
                    w[n * 6 + 4] = Q(perm[5], i, j, k);
                    w[n * 6 + 5] = Q(perm[6], i, j, k);
                }
                if (c->c.recon == 0) recon_plm_prim(len, c->c.mc_beta, w, wb, delw, qL, qR, &ilo, &ihi);
                else if (c->c.recon == 1) recon_ppm_prim(len, c->c.mc_beta, w, wb, delw, wLa, wRa, qL, qR, &ilo, &ihi);
                else recon_ppm_char(len, c->c.mc_beta, c->c.cs, w, wb, delw, wLa, wRa, Rv, Lv, qL, qR, &ilo, &ihi);
            }
            for (int n = ilo; n <= ihi; n++) {
                double F[7];
                vl_riemann(c->c.riemann, qL + n * 7, qR + n * 7, c->c.cs, F, &c->smax);
                int i, j, k;
                dir_cell(d, n0 + n, t1, t2, &i, &j, &k);
                for (int r = 0; r < 7; r++) FS(perm[r], d, i, j, k) = F[r];
                if (d == 0) { EMF(2, 0, i, j, k) = -FS(5, 0, i, j, k); EMF(1, 0, i, j, k) = FS(6, 0, i, j, k); }
                else if (d == 1) { EMF(0, 0, i, j, k) = -FS(6, 1, i, j, k); EMF(2, 1, i, j, k) = FS(4, 1, i, j, k); }
                else { EMF(1, 1, i, j, k) = -FS(4, 2, i, j, k); EMF(0, 1, i, j, k) = FS(5, 2, i, j, k); }
            }
            if (c->Fsp) {
                /* protection fluxes: compute_HLLE_fluxes_1d on the same interface states, after
                 * the Roe fluxes of the pencil (vanleer_mod.f90:540, 560-573, 703, 731-753, 880,
                 * 908-930); it raises smax too */
                for (int n = ilo; n <= ihi; n++) {
                    double F[7];
                    hlle_if(qL + n * 7, qR + n * 7, c->c.cs * c->c.cs, F, &c->smax);
                    int i, j, k;
                    dir_cell(d, n0 + n, t1, t2, &i, &j, &k);
                    for (int r = 0; r < 7; r++) FSP(perm[r], d, i, j, k) = F[r];
                    if (d == 0) { EMFP(2, 0, i, j, k) = -FSP(5, 0, i, j, k); EMFP(1, 0, i, j, k) = FSP(6, 0, i, j, k); }
                    else if (d == 1) { EMFP(0, 0, i, j, k) = -FSP(6, 1, i, j, k); EMFP(2, 1, i, j, k) = FSP(4, 1, i, j, k); }
                    else { EMFP(1, 1, i, j, k) = -FSP(4, 2, i, j, k); EMFP(0, 1, i, j, k) = FSP(5, 2, i, j, k); }
                }
            }
        }
    }
    free(w); free(wb); free(qL); free(qR); free(delw); free(wLa); free(wRa); free(Rv); free(Lv);
}

/* compute_cell_center_emf_3d (vanleer_mod.f90:1823-1862) */
static void vl_center_emf(orc_ctx *c, int off) {
    int nb = c->c.nb;
    for (int k = 1 + off - nb; k <= c->c.nz - off + nb; k++)
        for (int j = 1 + off - nb; j <= c->c.ny - off + nb; j++)
            for (int i = 1 + off - nb; i <= c->c.nx - off + nb; i++) {
                double irho = 1.0 / Q(0, i, j, k);
                EMF(0, 2, i, j, k) = (Q(5, i, j, k) * Q(3, i, j, k) - Q(6, i, j, k) * Q(2, i, j, k)) * irho;
                EMF(1, 2, i, j, k) = (Q(6, i, j, k) * Q(1, i, j, k) - Q(4, i, j, k) * Q(3, i, j, k)) * irho;
                EMF(2, 2, i, j, k) = (Q(4, i, j, k) * Q(2, i, j, k) - Q(5, i, j, k) * Q(1, i, j, k)) * irho;
                if (c->emfp)   /* vanleer_mod.f90:1848-1853 */
                    for (int m = 0; m < 3; m++) EMFP(m, 2, i, j, k) = EMF(m, 2, i, j, k);
            }
}

/* upwinded gradient selection of compute_cell_corner_emf_3d (vanleer_mod.f90:1916-1925) */
static inline void upw(double v, double L14, double R14, double L34, double R34, double *o14, double *o34) {
    if (v > ZERO) { *o14 = L14; *o34 = L34; }
    else if (v < -ZERO) { *o14 = R14; *o34 = R34; }
    else { *o14 = 0.5 * (L14 + R14); *o34 = 0.5 * (L34 + R34); }
}

/* compute_cell_corner_emf_3d (vanleer_mod.f90:1884-2076): Gardiner-Stone upwind corner EMF
 * from face EMFs E(comp, slot) and the state q (used for the upwinding velocities). */
static void corner_emf_arr(orc_ctx *c, int off, const double *qsrc, const double *emf_in, double *ec_out) {
    int nb = c->c.nb;
#define QS(f, i, j, k) qsrc[(long)(f) * c->NC + cidx(c, i, j, k)]
#define E(comp, slot, i, j, k) emf_in[((long)(slot) * 3 + (comp)) * c->NC + cidx(c, i, j, k)]
#undef EC
#define EC(f, i, j, k) ec_out[(long)(f) * c->NC + cidx(c, i, j, k)]
    for (int k = 1 + off - nb; k <= c->c.nz - off + nb - 1; k++)
        for (int j = 1 + off - nb; j <= c->c.ny - off + nb - 1; j++)
            for (int i = 1 + off - nb; i <= c->c.nx - off + nb - 1; i++) {
                double vx = 0.5 * (QS(1, i, j, k) / QS(0, i, j, k) + QS(1, i + 1, j, k) / QS(0, i + 1, j, k));
                double vy = 0.5 * (QS(2, i, j, k) / QS(0, i, j, k) + QS(2, i, j + 1, k) / QS(0, i, j + 1, k));
                double a14, a34, b14, b34;
                /* Ez at (i+1/2, j+1/2, k) */
                upw(vx, 2.0 * (E(2, 1, i, j, k) - E(2, 2, i, j, k)), 2.0 * (E(2, 1, i + 1, j, k) - E(2, 2, i + 1, j, k)),
                    2.0 * (E(2, 2, i, j + 1, k) - E(2, 1, i, j, k)), 2.0 * (E(2, 2, i + 1, j + 1, k) - E(2, 1, i + 1, j, k)),
                    &a14, &a34);
                upw(vy, 2.0 * (E(2, 0, i, j, k) - E(2, 2, i, j, k)), 2.0 * (E(2, 0, i, j + 1, k) - E(2, 2, i, j + 1, k)),
                    2.0 * (E(2, 2, i + 1, j, k) - E(2, 0, i, j, k)), 2.0 * (E(2, 2, i + 1, j + 1, k) - E(2, 0, i, j + 1, k)),
                    &b14, &b34);
                EC(2, i, j, k) = 0.25 * (E(2, 0, i, j, k) + E(2, 0, i, j + 1, k) + E(2, 1, i, j, k) + E(2, 1, i + 1, j, k)) +
                                 0.125 * ((a14 - a34) + (b14 - b34));
            }
    for (int k = 1 + off - nb; k <= c->c.nz - off + nb - 1; k++)
        for (int j = 1 + off - nb; j <= c->c.ny - off + nb - 1; j++)
            for (int i = 1 + off - nb; i <= c->c.nx - off + nb - 1; i++) {
                double vx = 0.5 * (QS(1, i, j, k) / QS(0, i, j, k) + QS(1, i + 1, j, k) / QS(0, i + 1, j, k));
                double vz = 0.5 * (QS(3, i, j, k) / QS(0, i, j, k) + QS(3, i, j, k + 1) / QS(0, i, j, k + 1));
                double a14, a34, b14, b34;
                /* Ey at (i+1/2, j, k+1/2) */
                upw(vx, 2.0 * (E(1, 1, i, j, k) - E(1, 2, i, j, k)), 2.0 * (E(1, 1, i + 1, j, k) - E(1, 2, i + 1, j, k)),
                    2.0 * (E(1, 2, i, j, k + 1) - E(1, 1, i, j, k)), 2.0 * (E(1, 2, i + 1, j, k + 1) - E(1, 1, i + 1, j, k)),
                    &a14, &a34);
                upw(vz, 2.0 * (E(1, 0, i, j, k) - E(1, 2, i, j, k)), 2.0 * (E(1, 0, i, j, k + 1) - E(1, 2, i, j, k + 1)),
                    2.0 * (E(1, 2, i + 1, j, k) - E(1, 0, i, j, k)), 2.0 * (E(1, 2, i + 1, j, k + 1) - E(1, 0, i, j, k + 1)),
                    &b14, &b34);
                EC(1, i, j, k) = 0.25 * (E(1, 0, i, j, k) + E(1, 0, i, j, k + 1) + E(1, 1, i, j, k) + E(1, 1, i + 1, j, k)) +
                                 0.125 * ((a14 - a34) + (b14 - b34));
            }
    for (int k = 1 + off - nb; k <= c->c.nz - off + nb - 1; k++)
        for (int j = 1 + off - nb; j <= c->c.ny - off + nb - 1; j++)
            for (int i = 1 + off - nb; i <= c->c.nx - off + nb - 1; i++) {
                double vy = 0.5 * (QS(2, i, j, k) / QS(0, i, j, k) + QS(2, i, j + 1, k) / QS(0, i, j + 1, k));
                double vz = 0.5 * (QS(3, i, j, k) / QS(0, i, j, k) + QS(3, i, j, k + 1) / QS(0, i, j, k + 1));
                double a14, a34, b14, b34;
                /* Ex at (i, j+1/2, k+1/2) */
                upw(vy, 2.0 * (E(0, 1, i, j, k) - E(0, 2, i, j, k)), 2.0 * (E(0, 1, i, j + 1, k) - E(0, 2, i, j + 1, k)),
                    2.0 * (E(0, 2, i, j, k + 1) - E(0, 1, i, j, k)), 2.0 * (E(0, 2, i, j + 1, k + 1) - E(0, 1, i, j + 1, k)),
                    &a14, &a34);
                upw(vz, 2.0 * (E(0, 0, i, j, k) - E(0, 2, i, j, k)), 2.0 * (E(0, 0, i, j, k + 1) - E(0, 2, i, j, k + 1)),
                    2.0 * (E(0, 2, i, j + 1, k) - E(0, 0, i, j, k)), 2.0 * (E(0, 2, i, j + 1, k + 1) - E(0, 0, i, j, k + 1)),
                    &b14, &b34);
                EC(0, i, j, k) = 0.25 * (E(0, 0, i, j, k) + E(0, 0, i, j, k + 1) + E(0, 1, i, j, k) + E(0, 1, i, j + 1, k)) +
                                 0.125 * ((a14 - a34) + (b14 - b34));
            }
#undef QS
#undef E
#undef EC
#define EC(f, i, j, k) c->ec[(long)(f) * c->NC + cidx(c, i, j, k)]
}

static void corner_emf(orc_ctx *c, int off, const double *qsrc) {
    corner_emf_arr(c, off, qsrc, c->emf, c->ec);
    if (c->ecp) corner_emf_arr(c, off, qsrc, c->emfp, c->ecp);   /* vanleer_mod.f90:336, 396 */
}

/* final_check_unphysical (vanleer_mod.f90:3243-3262) */
static void floor_check(orc_ctx *c, int ilo, int ihi, int jlo, int jhi, int klo, int khi) {
    for (int k = klo; k <= khi; k++)
        for (int j = jlo; j <= jhi; j++)
            for (int i = ilo; i <= ihi; i++)
                if (Q(0, i, j, k) < MIN_DENS) { Q(0, i, j, k) = MIN_DENS; c->floor_count++; }
}

/* update_hydro_state_3d (vanleer_mod.f90:2080-2207).  With flux protection: back up the
 * state, update, flag cells whose density fell below the threshold, give each flagged cell's
 * six faces the HLLE fluxes and the corner-EMF slots (i,j,k), (i-1,j,k), (i,j-1,k),
 * (i,j,k-1) of all three components the protection corner EMFs (the reference's index set,
 * :2155-2169), then redo the update of every flagged cell and face neighbour from the
 * backup with the modified fluxes (:2177-2198). */
#define UPD(f, i, j, k) (FS(f, 0, i, j, k) - FS(f, 0, i - 1, j, k)) + (FS(f, 1, i, j, k) - FS(f, 1, i, j - 1, k)) + \
                        (FS(f, 2, i, j, k) - FS(f, 2, i, j, k - 1))
static void vl_update_hydro(orc_ctx *c, double dtdx, int off) {
    int nb = c->c.nb;
    int ilo = 2 + off - nb, ihi = c->c.nx - off + nb - 1;
    int jlo = 2 + off - nb, jhi = c->c.ny - off + nb - 1;
    int klo = 2 + off - nb, khi = c->c.nz - off + nb - 1;
    if (c->pflag) {
        memset(c->pflag, 0, (size_t)c->NC);
        memcpy(c->qprot, c->q, sizeof(double) * 7 * c->NC);
    }
    for (int k = klo; k <= khi; k++)
        for (int j = jlo; j <= jhi; j++)
            for (int i = ilo; i <= ihi; i++)
                for (int f = 0; f < 4; f++)
                    Q(f, i, j, k) = Q(f, i, j, k) - dtdx * ((FS(f, 0, i, j, k) - FS(f, 0, i - 1, j, k)) +
                                                            (FS(f, 1, i, j, k) - FS(f, 1, i, j - 1, k)) +
                                                            (FS(f, 2, i, j, k) - FS(f, 2, i, j, k - 1)));
    if (c->pflag) {
        const double thr = c->c.prot_dens > 0.0 ? c->c.prot_dens : MIN_DENS;
        for (int k = klo; k <= khi; k++)
            for (int j = jlo; j <= jhi; j++)
                for (int i = ilo; i <= ihi; i++) {
                    if (!(Q(0, i, j, k) < thr)) continue;
                    c->pflag[cidx(c, i, j, k)] = 1;
                    for (int f = 0; f < 7; f++) {
                        FS(f, 0, i, j, k) = FSP(f, 0, i, j, k);
                        FS(f, 0, i - 1, j, k) = FSP(f, 0, i - 1, j, k);
                        FS(f, 1, i, j, k) = FSP(f, 1, i, j, k);
                        FS(f, 1, i, j - 1, k) = FSP(f, 1, i, j - 1, k);
                        FS(f, 2, i, j, k) = FSP(f, 2, i, j, k);
                        FS(f, 2, i, j, k - 1) = FSP(f, 2, i, j, k - 1);
                    }
                    for (int m = 0; m < 3; m++) {
                        EC(m, i, j, k) = ECP(m, i, j, k);
                        EC(m, i - 1, j, k) = ECP(m, i - 1, j, k);
                        EC(m, i, j - 1, k) = ECP(m, i, j - 1, k);
                        EC(m, i, j, k - 1) = ECP(m, i, j, k - 1);
                    }
                }
#define PF(i, j, k) c->pflag[cidx(c, i, j, k)]
        for (int k = klo; k <= khi; k++)
            for (int j = jlo; j <= jhi; j++)
                for (int i = ilo; i <= ihi; i++)
                    if (PF(i, j, k) || PF(i - 1, j, k) || PF(i + 1, j, k) || PF(i, j - 1, k) || PF(i, j + 1, k) ||
                        PF(i, j, k - 1) || PF(i, j, k + 1))
                        for (int f = 0; f < 4; f++)
                            Q(f, i, j, k) = c->qprot[(long)f * c->NC + cidx(c, i, j, k)] - dtdx * (UPD(f, i, j, k));
#undef PF
    }
    floor_check(c, ilo, ihi, jlo, jhi, klo, khi);
}
#undef UPD

/* update_CT_3d (vanleer_mod.f90:2211-2259) */
static void ct_update(orc_ctx *c, double dtdx, int off) {
    int nb = c->c.nb;
    for (int k = 2 + off - nb; k <= c->c.nz - off + nb - 1; k++)
        for (int j = 2 + off - nb; j <= c->c.ny - off + nb - 1; j++)
            for (int i = 2 + off - nb; i <= c->c.nx - off + nb - 1; i++) {
                BF(0, i, j, k) = BF(0, i, j, k) - dtdx * ((EC(2, i, j, k) - EC(2, i, j - 1, k)) - (EC(1, i, j, k) - EC(1, i, j, k - 1)));
                BF(1, i, j, k) = BF(1, i, j, k) - dtdx * ((EC(0, i, j, k) - EC(0, i, j, k - 1)) - (EC(2, i, j, k) - EC(2, i - 1, j, k)));
                BF(2, i, j, k) = BF(2, i, j, k) - dtdx * ((EC(1, i, j, k) - EC(1, i - 1, j, k)) - (EC(0, i, j, k) - EC(0, i, j - 1, k)));
            }
    for (int k = 3 + off - nb; k <= c->c.nz - off + nb - 1; k++)
        for (int j = 3 + off - nb; j <= c->c.ny - off + nb - 1; j++)
            for (int i = 3 + off - nb; i <= c->c.nx - off + nb - 1; i++) {
                Q(4, i, j, k) = 0.5 * (BF(0, i, j, k) + BF(0, i - 1, j, k));
                Q(5, i, j, k) = 0.5 * (BF(1, i, j, k) + BF(1, i, j - 1, k));
                Q(6, i, j, k) = 0.5 * (BF(2, i, j, k) + BF(2, i, j, k - 1));
            }
}

/* vanleer_solve_3d (vanleer_mod.f90:279-426) without the forcing hook */
static void vl_step(orc_ctx *c, double dt) {
    size_t n7 = sizeof(double) * 7 * c->NC, n3 = sizeof(double) * 3 * c->NC;
    memcpy(c->qw, c->q, n7);
    memcpy(c->bfw, c->bf, n3);
    memset(c->Fs, 0, sizeof(double) * 21 * c->NC);
    memset(c->emf, 0, sizeof(double) * 9 * c->NC);
    if (c->Fsp) memset(c->Fsp, 0, sizeof(double) * 21 * c->NC);   /* Fs_prot = 0 (vanleer_mod.f90:294) */
    memset(c->ec, 0, n3);
    c->smax = 0.0;
    c->delv_max = 0.0;
    int offset = 0;
    for (int d = 0; d < 3; d++) vl_sweep(c, d, offset, 0);
    vl_center_emf(c, offset);
    corner_emf(c, offset, c->q);
    double dtdx = 0.5 * dt / c->dx;
    vl_update_hydro(c, dtdx, offset);
    ct_update(c, dtdx, offset);
    offset = 2;
    for (int d = 0; d < 3; d++) vl_sweep(c, d, offset, 1);
    offset = (c->c.recon == 0) ? 3 : 4;
    vl_center_emf(c, offset);
    corner_emf(c, offset, c->q);
    memcpy(c->q, c->qw, n7);
    memcpy(c->bf, c->bfw, n3);
    dtdx = dt / c->dx;
    vl_update_hydro(c, dtdx, offset);
    ct_update(c, dtdx, offset);
}

/* One update_hydro_state_3d + update_CT_3d (vanleer_mod.f90:2080-2259) on the context's
 * current state with caller-given arrays: Fs and Fs_prot (21 padded fields each,
 * [dir][comp]), emf_corner and emf_corner_prot (3 each); the stage's active region is
 * shrunk by `off` as in vanleer_solve_3d.  Lets a test check the protection's flag, face,
 * corner and redo sets on arbitrary fluxes.  Needs a context created with protection. */
int orc_update_test(orc_ctx *c, double dtdx, int off, const double *Fs, const double *Fsp, const double *ec,
                    const double *ecp) {
    if (!c->Fsp) return -1;
    memcpy(c->Fs, Fs, sizeof(double) * 21 * c->NC);
    memcpy(c->Fsp, Fsp, sizeof(double) * 21 * c->NC);
    memcpy(c->ec, ec, sizeof(double) * 3 * c->NC);
    memcpy(c->ecp, ecp, sizeof(double) * 3 * c->NC);
    vl_update_hydro(c, dtdx, off);
    ct_update(c, dtdx, off);
    return 0;
}

#include "imhd_oracle_tvd.inc"

#define orc_ou struct orc_ou_s
#include "imhd_oracle_ou.inc"
#undef orc_ou

#include "imhd_oracle_force.inc"

/* ----------------------------------------------------------------------------------- */
/* forcing application of the reference's turbulence driver                            */

/* get_rho_turb (turb_driver_mod.f90:130-192): rho0 = rho^n on the padded grid; the
 * quadratic coefficients of the kinetic-energy gain and the momentum the field carries,
 * summed over interior cells in the reference's loop order. */
static void force_sums(orc_ctx *c) {
    for (long p = 0; p < c->NC; p++) c->rho0[p] = c->q[p];
    double q1 = 0.0, q2 = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
    for (int k = 1; k <= c->c.nz; k++)
        for (int j = 1; j <= c->c.ny; j++)
            for (int i = 1; i <= c->c.nx; i++) {
                long p = cidx(c, i, j, k);
                double d1 = c->delv[p], d2 = c->delv[c->NC + p], d3 = c->delv[2 * c->NC + p];
                double rho = Q(0, i, j, k);
                q1 = q1 + rho * (d1 * d1 + d2 * d2 + d3 * d3);
                q2 = q2 + d1 * Q(1, i, j, k) + d2 * Q(2, i, j, k) + d3 * Q(3, i, j, k);
                mx = mx + rho * d1;
                my = my + rho * d2;
                mz = mz + rho * d3;
            }
    c->quad1 = 0.5 * q1;
    c->quad2 = q2;
    c->mavg[0] = mx; c->mavg[1] = my; c->mavg[2] = mz;
}

/* add_random_force (turb_driver_mod.f90:395-457), one rank: constant-energy-input
 * normalisation (larger root of the KE quadratic) and zero net momentum injection */
static int force_apply(orc_ctx *c) {
    if (c->quad1 <= 1.0e-20) return -1;
    double dx3 = c->dx * c->dx * c->dx;
    double ma[3] = {c->mavg[0] * dx3, c->mavg[1] * dx3, c->mavg[2] * dx3};
    double tmp1 = c->edot * c->dt_turb / dx3;
    double tmp2 = sqrt(c->quad2 * c->quad2 + (4.0 * tmp1 * c->quad1));
    double constE = 0.5 * (-c->quad2 + tmp2) / c->quad1;
    for (int k = 1; k <= c->c.nz; k++)
        for (int j = 1; j <= c->c.ny; j++)
            for (int i = 1; i <= c->c.nx; i++) {
                long p = cidx(c, i, j, k);
                double r0 = c->rho0[p], dm[3];
                for (int d = 0; d < 3; d++) dm[d] = constE * (r0 * c->delv[d * c->NC + p] - ma[d]);
                Q(1, i, j, k) = Q(1, i, j, k) + dm[0];
                Q(2, i, j, k) = Q(2, i, j, k) + dm[1];
                Q(3, i, j, k) = Q(3, i, j, k) + dm[2];
                double m = dmax3(fabs(dm[0]) / r0, fabs(dm[1]) / r0, fabs(dm[2]) / r0);
                c->delv_max = dmax2(c->delv_max, m);
            }
    return 0;
}

void orc_set_force(orc_ctx *c, const double *delv, double edot, double dt_turb) {
    if (!c->delv) {
        c->delv = calloc((size_t)3 * c->NC, sizeof(double));
        c->rho0 = calloc((size_t)c->NC, sizeof(double));
    }
    memcpy(c->delv, delv, sizeof(double) * 3 * c->NC);
    c->edot = edot;
    c->dt_turb = dt_turb;
}

/* vanleer_solve_3d(dt, drive_next=.TRUE.) (vanleer_mod.f90:304-305, 419-421) */
int orc_step_drive(orc_ctx *c, double dt) {
    if (c->c.mhd != 2 || !c->delv) return -1;
    c->dt = dt;
    force_sums(c);
    vl_step(c, dt);
    return force_apply(c);
}

/* ----------------------------------------------------------------------------------- */
/* ghost fill, boundary_conditions (boundary_conditions_mod.f90:20-234), one rank       */

static void copy_plane(double *a, long NC, int nf, const orc_ctx *c, int d, int dst, int src, int is_pass) {
    (void)NC;
    int nb = c->c.nb;
    int lo0 = is_pass ? 3 - nb : 1 - nb;
    int hi[3] = {c->c.nx, c->c.ny, c->c.nz};
    for (int f = 0; f < nf; f++) {
        for (int t2 = lo0; t2 <= hi[d == 2 ? 1 : 2] + nb - (is_pass ? 2 : 0); t2++)
            for (int t1 = lo0; t1 <= hi[d == 0 ? 1 : 0] + nb - (is_pass ? 2 : 0); t1++) {
                int i, j, k, i2, j2, k2;
                dir_cell(d, dst, t1, t2, &i, &j, &k);
                dir_cell(d, src, t1, t2, &i2, &j2, &k2);
                if (is_pass) a[(long)f * c->PC + pidx(c, i, j, k)] = a[(long)f * c->PC + pidx(c, i2, j2, k2)];
                else a[(long)f * c->NC + cidx(c, i, j, k)] = a[(long)f * c->NC + cidx(c, i2, j2, k2)];
            }
    }
}

void orc_bc(orc_ctx *c) {
    int nb = c->c.nb, nbp = 3;
    int n[3] = {c->c.nx, c->c.ny, c->c.nz};
    for (int d = 0; d < 3; d++) {
        if (c->c.bc == 1) {
            for (int g = 0; g < nb; g++) {
                copy_plane(c->q, c->NC, 7, c, d, 1 - nb + g, 1, 0);
                copy_plane(c->bf, c->NC, 3, c, d, 1 - nb + g, 1, 0);
            }
            if (c->c.npass > 0)
                for (int g = 0; g < nbp; g++) copy_plane(c->qp, c->PC, c->c.npass, c, d, 1 - nbp + g, 1, 1);
            for (int g = 0; g < nb; g++) {
                copy_plane(c->q, c->NC, 7, c, d, n[d] + 1 + g, n[d], 0);
                copy_plane(c->bf, c->NC, 3, c, d, n[d] + 1 + g, n[d], 0);
            }
            if (c->c.npass > 0)
                for (int g = 0; g < nbp; g++) copy_plane(c->qp, c->PC, c->c.npass, c, d, n[d] + 1 + g, n[d], 1);
        } else if (c->c.bc == 2) {
            for (int g = 0; g < nb; g++) {
                copy_plane(c->q, c->NC, 7, c, d, -g, n[d] - g, 0);
                copy_plane(c->bf, c->NC, 3, c, d, -g, n[d] - g, 0);
            }
            if (c->c.npass > 0)
                for (int g = 0; g < nbp; g++) copy_plane(c->qp, c->PC, c->c.npass, c, d, -g, n[d] - g, 1);
            for (int g = 0; g < nb; g++) {
                copy_plane(c->q, c->NC, 7, c, d, n[d] + 1 + g, 1 + g, 0);
                copy_plane(c->bf, c->NC, 3, c, d, n[d] + 1 + g, 1 + g, 0);
            }
            if (c->c.npass > 0)
                for (int g = 0; g < nbp; g++) copy_plane(c->qp, c->PC, c->c.npass, c, d, n[d] + 1 + g, 1 + g, 1);
        }
    }
}

/* ----------------------------------------------------------------------------------- */
/* CFL: get_max_speed_vanleer (vanleer_mod.f90:3266-3334)                               */

static double vl_max_speed(orc_ctx *c) {
    double asqr = c->c.cs * c->c.cs, ms = 0.0;
    for (int k = 1; k <= c->c.nz; k++)
        for (int j = 1; j <= c->c.ny; j++)
            for (int i = 1; i <= c->c.nx; i++) {
                double rho = Q(0, i, j, k);
                double isrho = 1.0 / sqrt(rho);
                double bx = Q(4, i, j, k) * isrho, by = Q(5, i, j, k) * isrho, bz = Q(6, i, j, k) * isrho;
                double btsqr = by * by + bz * bz, bsqr = bx * bx + btsqr, x = asqr - bsqr;
                double tmp = sqrt(x * x + 4.0 * asqr * btsqr);
                double cfx = sqrt(0.5 * (asqr + bsqr + tmp));
                btsqr = bx * bx + bz * bz; bsqr = by * by + btsqr; x = asqr - bsqr;
                tmp = sqrt(x * x + 4.0 * asqr * btsqr);
                double cfy = sqrt(0.5 * (asqr + bsqr + tmp));
                btsqr = bx * bx + by * by; bsqr = bz * bz + btsqr; x = asqr - bsqr;
                tmp = sqrt(x * x + 4.0 * asqr * btsqr);
                double cfz = sqrt(0.5 * (asqr + bsqr + tmp));
                double vx = Q(1, i, j, k) / rho, vy = Q(2, i, j, k) / rho, vz = Q(3, i, j, k) / rho;
                double m = dmax3(fabs(vx) + cfx, fabs(vy) + cfy, fabs(vz) + cfz);
                ms = dmax2(ms, m);
            }
    return dmax2(ms, c->smax + c->delv_max);
}

double orc_max_speed(orc_ctx *c) {
    if (c->c.mhd == 1) return tvd_max_speed(c);
    return vl_max_speed(c);
}
double orc_smax(const orc_ctx *c) { return c->smax; }
int orc_floor_count(const orc_ctx *c) { return c->floor_count; }

void orc_step(orc_ctx *c, double dt) {
    c->dt = dt;
    if (c->c.mhd == 1) tvd_step(c, dt);
    else vl_step(c, dt);
}

/* ----------------------------------------------------------------------------------- */
/* passive scalars (passive_mod.f90)                                                    */

/* copy_velocity (passive_mod.f90:204-244) */
void orc_copy_velocity(orc_ctx *c, int stage) {
    for (int k = -2; k <= c->c.nz + 3; k++)
        for (int j = -2; j <= c->c.ny + 3; j++)
            for (int i = -2; i <= c->c.nx + 3; i++) {
                double irho = 1.0 / Q(0, i, j, k);
                if (!stage) {
                    VEL(0, i, j, k) = Q(1, i, j, k) * irho;
                    VEL(1, i, j, k) = Q(2, i, j, k) * irho;
                    VEL(2, i, j, k) = Q(3, i, j, k) * irho;
                } else {
                    VTA(0, i, j, k) = 0.5 * (VEL(0, i, j, k) + Q(1, i, j, k) * irho);
                    VTA(1, i, j, k) = 0.5 * (VEL(1, i, j, k) + Q(2, i, j, k) * irho);
                    VTA(2, i, j, k) = 0.5 * (VEL(2, i, j, k) + Q(3, i, j, k) * irho);
                }
            }
}

/* passive compute_{x,y,z}fluxes_3d (passive_mod.f90:248-506) + compute_interface_states
 * (:510-557) + compute_fluxes_1d (:594-627) */
static void pass_sweep(orc_ctx *c, int d, int offset, int stage2) {
    int nbp = 3;
    int lo[3], hi[3];
    for (int a = 0; a < 3; a++) { lo[a] = 1 + offset; hi[a] = dim_n(c, a) - offset; }
    int t1d = (d == 0) ? 1 : 0, t2d = (d == 2) ? 1 : 2;
    int n0 = lo[d] - nbp, n1 = hi[d] + nbp, len = n1 - n0 + 1;
    double *w = malloc(sizeof(double) * len), *dw = calloc((size_t)len, sizeof(double));
    double *sL = malloc(sizeof(double) * len), *sR = malloc(sizeof(double) * len), *v = malloc(sizeof(double) * len);
    for (int p = 0; p < c->c.npass; p++)
        for (int t2 = lo[t2d] - nbp; t2 <= hi[t2d] + nbp; t2++)
            for (int t1 = lo[t1d] - nbp; t1 <= hi[t1d] + nbp; t1++) {
                int ilo, ihi;
                for (int n = 0; n < len; n++) {
                    int i, j, k;
                    dir_cell(d, n0 + n, t1, t2, &i, &j, &k);
                    w[n] = QP(p, i, j, k);
                    v[n] = stage2 ? VTA(d, i, j, k) : VEL(d, i, j, k);
                }
                if (!stage2) {
                    for (int n = 0; n < len - 1; n++) { sL[n] = w[n]; sR[n] = w[n + 1]; }
                    ilo = 0; ihi = len - 2;
                } else {
                    for (int n = 1; n <= len - 2; n++) {
                        double dL = w[n] - w[n - 1], dR = w[n + 1] - w[n];
                        dw[n] = 0.5 * (sgn(dL) + sgn(dR)) * fmin(fabs(dL), fabs(dR));
                    }
                    for (int n = 1; n <= len - 3; n++) { sL[n] = w[n] + 0.5 * dw[n]; sR[n] = w[n + 1] - 0.5 * dw[n + 1]; }
                    ilo = 1; ihi = len - 3;
                }
                for (int n = ilo; n <= ihi; n++) {
                    double vavg = 0.5 * (v[n] + v[n + 1]);
                    double theta = (vavg >= 0.0) ? 1.0 : -1.0;
                    int i, j, k;
                    dir_cell(d, n0 + n, t1, t2, &i, &j, &k);
                    FP(p, d, i, j, k) = 0.5 * ((1.0 + theta) * sL[n] + (1.0 - theta) * sR[n]);
                }
            }
    free(w); free(dw); free(sL); free(sR); free(v);
}

/* update_passive_state_3d (passive_mod.f90:562-590) */
static void pass_update(orc_ctx *c, double dtdx, const double *vel, int off) {
    int nbp = 3;
#define VV(d, i, j, k) vel[(long)(d) * c->VC + vidx(c, i, j, k)]
    for (int p = 0; p < c->c.npass; p++)
        for (int k = 1 + off - nbp + 1; k <= c->c.nz - off + nbp - 1; k++)
            for (int j = 1 + off - nbp + 1; j <= c->c.ny - off + nbp - 1; j++)
                for (int i = 1 + off - nbp + 1; i <= c->c.nx - off + nbp - 1; i++)
                    QP(p, i, j, k) = QP(p, i, j, k) - dtdx * (VV(0, i, j, k) * (FP(p, 0, i, j, k) - FP(p, 0, i - 1, j, k)) +
                                                              VV(1, i, j, k) * (FP(p, 1, i, j, k) - FP(p, 1, i, j - 1, k)) +
                                                              VV(2, i, j, k) * (FP(p, 2, i, j, k) - FP(p, 2, i, j, k - 1)));
#undef VV
}

/* passive_solve_3d (passive_mod.f90:118-201) */
void orc_passive_step(orc_ctx *c, double dt) {
    if (c->c.npass <= 0) return;
    memcpy(c->qpw, c->qp, sizeof(double) * c->c.npass * c->PC);
    for (int d = 0; d < 3; d++) pass_sweep(c, d, 0, 0);
    pass_update(c, 0.5 * dt / c->dx, c->vel, 0);
    for (int d = 0; d < 3; d++) pass_sweep(c, d, 1, 1);
    memcpy(c->qp, c->qpw, sizeof(double) * c->c.npass * c->PC);
    pass_update(c, dt / c->dx, c->vtavg, 2);
}

/* ----------------------------------------------------------------------------------- */
/* diagnostics: compute_KE (driver.f90:518-553), compute_divB (driver.f90:492-515),     */
/* plus the max |div B| and total mass that the reference does not compute.             */

void orc_diagnostics(const orc_ctx *c, orc_diag *o) {
    double ke = 0, vr = 0, mx = 0, my = 0, mz = 0, dsum = 0, dmax = 0, mass = 0;
    for (int k = 1; k <= c->c.nz; k++)
        for (int j = 1; j <= c->c.ny; j++)
            for (int i = 1; i <= c->c.nx; i++) {
                double rho = fmin(fmax(Q(0, i, j, k), 1e-15), 1e10);
                double m1 = Q(1, i, j, k), m2 = Q(2, i, j, k), m3 = Q(3, i, j, k);
                double vsqr = fmin((m1 * m1 + m2 * m2 + m3 * m3) / (rho * rho), 1e5);
                vr = vr + vsqr;
                ke = ke + rho * vsqr;
                mx = mx + m1; my = my + m2; mz = mz + m3;
                mass = mass + Q(0, i, j, k);
                double db = (BF(0, i, j, k) - BF(0, i - 1, j, k)) + (BF(1, i, j, k) - BF(1, i, j - 1, k)) +
                            (BF(2, i, j, k) - BF(2, i, j, k - 1));
                dsum = dsum + db;
                if (fabs(db) > dmax) dmax = fabs(db);
            }
    double dx3 = c->dx * c->dx * c->dx;
    o->vrms_sq = vr * dx3;
    o->ke = 0.5 * ke * dx3;
    o->momx = mx * dx3; o->momy = my * dx3; o->momz = mz * dx3;
    o->divb_sum = dsum;
    o->divb_max = dmax;
    o->mass = mass * dx3;
}

/* ----------------------------------------------------------------------------------- */
/* context                                                                              */

orc_ctx *orc_create(const orc_cfg *cfg) {
    orc_ctx *c = calloc(1, sizeof(orc_ctx));
    c->c = *cfg;
    if (c->c.mc_beta == 0.0) c->c.mc_beta = 1.01;
    int nb = cfg->nb;
    c->NX = cfg->nx + 2 * nb; c->NY = cfg->ny + 2 * nb; c->NZ = cfg->nz + 2 * nb;
    c->NC = (long)c->NX * c->NY * c->NZ;
    c->q = calloc((size_t)7 * c->NC, sizeof(double));
    c->qw = calloc((size_t)7 * c->NC, sizeof(double));
    c->bf = calloc((size_t)3 * c->NC, sizeof(double));
    c->bfw = calloc((size_t)3 * c->NC, sizeof(double));
    c->ec = calloc((size_t)3 * c->NC, sizeof(double));
    c->Fs = calloc((size_t)21 * c->NC, sizeof(double));
    c->emf = calloc((size_t)9 * c->NC, sizeof(double));
    if (cfg->mhd == 1) c->Fbuf = calloc((size_t)21 * c->NC, sizeof(double));
    if (cfg->mhd == 2 && cfg->riemann == 1 && cfg->protection) {
        c->Fsp = calloc((size_t)21 * c->NC, sizeof(double));
        c->emfp = calloc((size_t)9 * c->NC, sizeof(double));
        c->ecp = calloc((size_t)3 * c->NC, sizeof(double));
        c->qprot = calloc((size_t)7 * c->NC, sizeof(double));
        c->pflag = calloc((size_t)c->NC, 1);
    }
    int nmax = cfg->nx > cfg->ny ? cfg->nx : cfg->ny;
    if (cfg->nz > nmax) nmax = cfg->nz;
    c->dx = 1.0 / (double)nmax;
    c->PX = cfg->nx + 2 * nb - 4; c->PY = cfg->ny + 2 * nb - 4; c->PZ = cfg->nz + 2 * nb - 4;
    c->PC = (long)c->PX * c->PY * c->PZ;
    c->VX = cfg->nx + 6; c->VY = cfg->ny + 6; c->VZ = cfg->nz + 6;
    c->VC = (long)c->VX * c->VY * c->VZ;
    int np = cfg->npass > 0 ? cfg->npass : 1;
    c->qp = calloc((size_t)np * c->PC, sizeof(double));
    c->qpw = calloc((size_t)np * c->PC, sizeof(double));
    c->vel = calloc((size_t)3 * c->VC, sizeof(double));
    c->vtavg = calloc((size_t)3 * c->VC, sizeof(double));
    c->Fp = calloc((size_t)3 * np * c->VC, sizeof(double));
    return c;
}

void orc_destroy(orc_ctx *c) {
    if (!c) return;
    free(c->q); free(c->qw); free(c->bf); free(c->bfw); free(c->ec); free(c->Fs); free(c->emf); free(c->Fbuf);
    free(c->qp); free(c->qpw); free(c->vel); free(c->vtavg); free(c->Fp);
    free(c->delv); free(c->rho0);
    free(c->Fsp); free(c->emfp); free(c->ecp); free(c->qprot); free(c->pflag);
    if (c->ou) { for (int d = 0; d < 3; d++) free(c->ou->tw[d]); free(c->ou); }
    free(c);
}

void orc_set_state(orc_ctx *c, const double *q, const double *bface, const double *qpass) {
    memcpy(c->q, q, sizeof(double) * 7 * c->NC);
    memcpy(c->bf, bface, sizeof(double) * 3 * c->NC);
    if (c->c.npass > 0 && qpass) memcpy(c->qp, qpass, sizeof(double) * c->c.npass * c->PC);
}

void orc_get_state(const orc_ctx *c, double *q, double *bface, double *qpass, double *emf_corner) {
    if (q) memcpy(q, c->q, sizeof(double) * 7 * c->NC);
    if (bface) memcpy(bface, c->bf, sizeof(double) * 3 * c->NC);
    if (qpass && c->c.npass > 0) memcpy(qpass, c->qp, sizeof(double) * c->c.npass * c->PC);
    if (emf_corner) memcpy(emf_corner, c->ec, sizeof(double) * 3 * c->NC);
}

/* ----------------------------------------------------------------------------------- */
/* per-interface pins                                                                   */

void orc_riemann_plane(int kind, int ni, int nj, double cs, const double *qL, const double *qR, double *F,
                       double *smax) {
    long np = (long)ni * nj;
    double sm = 0.0;
    for (long p = 0; p < np; p++) {
        double uL[7], uR[7], f[7];
        for (int m = 0; m < 7; m++) { uL[m] = qL[m * np + p]; uR[m] = qR[m * np + p]; }
        vl_riemann(kind, uL, uR, cs, f, &sm);
        for (int m = 0; m < 7; m++) F[m * np + p] = f[m];
    }
    *smax = sm;
}

/* kind 0 PLM prim, 1 PPM prim, 2 PPM char; pencils run along the first (i) axis. */
void orc_recon_plane(int kind, int ni, int nj, int nb, double cs, double mc_beta, const double *w,
                     const double *wb, double *qL, double *qR) {
    (void)nb;
    long np = (long)ni * nj;
    double *pw = malloc(sizeof(double) * ni * 6), *pwb = malloc(sizeof(double) * ni);
    double *pl = calloc((size_t)ni * 7, sizeof(double)), *pr = calloc((size_t)ni * 7, sizeof(double));
    double *dw = calloc((size_t)ni * 6, sizeof(double)), *a = calloc((size_t)ni * 6, sizeof(double));
    double *b = calloc((size_t)ni * 6, sizeof(double));
    double *Rv = malloc(sizeof(double) * ni * 36), *Lv = malloc(sizeof(double) * ni * 36);
    for (long p = 0; p < np * 7; p++) { qL[p] = 0.0; qR[p] = 0.0; }
    for (int j = 0; j < nj; j++) {
        for (int i = 0; i < ni; i++) {
            for (int k = 0; k < 6; k++) pw[i * 6 + k] = w[k * np + (long)j * ni + i];
            pwb[i] = wb[(long)j * ni + i];
        }
        int ilo, ihi;
        if (kind == 0) recon_plm_prim(ni, mc_beta, pw, pwb, dw, pl, pr, &ilo, &ihi);
        else if (kind == 1) recon_ppm_prim(ni, mc_beta, pw, pwb, dw, a, b, pl, pr, &ilo, &ihi);
        else recon_ppm_char(ni, mc_beta, cs, pw, pwb, dw, a, b, Rv, Lv, pl, pr, &ilo, &ihi);
        for (int i = ilo; i <= ihi; i++)
            for (int m = 0; m < 7; m++) {
                qL[m * np + (long)j * ni + i] = pl[i * 7 + m];
                qR[m * np + (long)j * ni + i] = pr[i * 7 + m];
            }
    }
    free(pw); free(pwb); free(pl); free(pr); free(dw); free(a); free(b); free(Rv); free(Lv);
}
