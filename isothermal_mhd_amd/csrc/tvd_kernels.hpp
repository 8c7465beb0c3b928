// Argument blocks and launch entry points of the K98 Roe-TVD / HLLE + CTU + CT integrator
// (isotvd_mod.f90 of the reference).
#pragma once
#include "aux_kernels.hpp"

namespace imhd {

// flux faces per wave of the fused x sweep (64 lanes: 62 faces + one neighbour face on
// each side)
constexpr int TVD_XSEG = 62;

// Directional sweep over cell states rotated so that the sweep direction is "x":
// w[r] = q[perm(d, r)], slot 4 unused (each cell's own upper face field comes from bn).
template <class real>
struct TvdSweepArgs {
    const real* w[7];
    const double* bn;
    real* gt[6];          // per face: limiter input g~_k = 0.5 (Q(chi) - chi^2) alpha_k
    real* ck[6];          // per face: characteristic amplitudes alpha_k
    real* F[4];           // rotated hydro fluxes
    real* Et1;            // face EMF = -flux(B_t1)
    real* Et2;            // face EMF = +flux(B_t2)
    const double* dt;     // per box
    double factor;        // 0.5 predictor, 1 corrector (dtdx = factor * dt / dx)
    double cs;
    int dir;
};

template <class real>
struct TvdCornerArgs {
    const real* q[4];     // t^n rho, m (upwinding velocity; predictor centre EMF)
    const real* qb[3];    // t^n cell-centred B (predictor centre EMF)
    const double* Ecc[3]; // corrector centre EMF (null in the predictor)
    const real* E[6];     // face EMFs, E[2*comp + slot]
    double* Ec[3];
};

template <class real>
struct TvdCenterArgs {
    const real* q[4];      // t^n rho, m
    const real* Fb[3][4];  // predictor fluxes
    const double* bh[3];   // half-step face fields
    double* Ecc[3];
    const double* dt;
};

template <class real>
struct TvdHalfArgs {
    const real* q[7];      // t^n state
    const double* b0[3];   // t^n face fields (source terms)
    const real* Fb[3][4];  // predictor fluxes
    const double* Ec[3];   // predictor corner EMFs
    real* dst[7];          // half state (the slot of the sweep's normal field is not written)
    const double* dt;
    unsigned int* floor_count;
    int dir;
};

template <class real>
struct TvdHalfDst {        // the three half states of k_tvd_half3, one set per sweep direction
    real* dst[3][7];
    // non-null: k_tvd_half3 also forms the corrector centre EMF (k_tvd_center's arithmetic)
    // on its range [4-nb, n+nb-3] from the half-step faces bh
    const double* bh[3];
    double* Ecc[3];
};

template <class real>
struct TvdCtArgs {         // predictor CT: whole box, half step inside [3-nb, n+nb-3]
    const double* b0[3];
    double* b1[3];
    const double* Ec[3];
    const double* dt;
};

template <class real>
struct TvdFinalArgs {      // whole box: interior update, t^n copy elsewhere
    StateP<real> src;
    StateP<real> dst;
    const real* F[3][4];
    const double* Ec[3];
    const double* dt;
    real* comp[4];
    unsigned int* floor_count;
};

template <class real>
void launch_tvd_eig(const Launch& L, const TvdSweepArgs<real>& a);
template <class real>
void launch_tvd_flux(const Launch& L, const TvdSweepArgs<real>& a);
// Fused eigen-system + flux sweep (plans: build_plan case 27 for x, 28 for y / z)
template <class real>
void launch_tvd_sweep(const Launch& L, const TvdSweepArgs<real>& a);
template <class real>
void launch_tvd_hlle(const Launch& L, const TvdSweepArgs<real>& a);
template <class real>
void launch_tvd_corner(const Launch& L, const TvdCornerArgs<real>& a);
template <class real>
void launch_tvd_center(const Launch& L, const TvdCenterArgs<real>& a);
template <class real>
void launch_tvd_half(const Launch& L, const TvdHalfArgs<real>& a);
template <class real>
void launch_tvd_half3(const Launch& L, const TvdHalfArgs<real>& a, const TvdHalfDst<real>& o);
template <class real>
void launch_tvd_ct_pred(const Launch& L, const TvdCtArgs<real>& a);
template <class real>
void launch_tvd_final(const Launch& L, const TvdFinalArgs<real>& a);
template <class real>
void launch_tvd_max_speed(const Launch& L, const QPtrs<real>& q, double cs, unsigned long long* cmax);

}  // namespace imhd
