// Argument blocks and launch entry points of the auxiliary kernels (ghost fill, CFL,
// diagnostics, passive scalars, conversion).
#pragma once
#include "kernels.hpp"

namespace imhd {

constexpr int DIAG_STRIDE = 8;  // per box: rho v^2, v^2, mx, my, mz, sum divB, mass, max|divB|
constexpr int DIAG_WG = 1024;   // partial-sum workgroups (and partial rows) per box
constexpr int DIAG_STATIC_MAXBOX = 64;   // boxes imhd_diagnostics_static serves
constexpr int DIAG_STATIC_MAXDEV = 64;   // devices whose scratch address is kept
// device address of this device's fixed scratch for imhd_diagnostics_static (partial rows
// at the start, results after DIAG_STATIC_MAXBOX x DIAG_WG rows)
hipError_t diag_static_scratch(double** out);

template <class real>
struct GhostArgs {
    real* r[7 + MAXPASS];
    double* d[3];
    int nreal, ndbl, depth;
};

template <class real>
struct QPtrs {
    const real* q[7];
};

template <class real>
struct DiagArgs {
    const real* q[4];
    const double* b[3];
    double* part;                  // DIAG_WG partial rows of DIAG_STRIDE doubles per box
    double* out;                   // DIAG_STRIDE doubles per box
};

template <class real>
struct VelArgs {
    const real* q[4];
    real* vel[3];
    real* vta[3];
    int stage;     // 0: v^n, 1: time-averaged 0.5 (v^n + v^{n+1})
};

template <class real>
struct PassFluxArgs {
    const real* s[MAXPASS];
    const real* v;   // velocity component along dir
    real* F[MAXPASS];
    int npass, dir, stage;
};

template <class real>
struct PassUpdArgs {
    const real* src[MAXPASS];
    real* dst[MAXPASS];
    const real* v[3];
    const real* F[3][MAXPASS];
    const double* dt;
    double factor;
    int npass;
};

// One stage of passive_solve_3d with the face values formed inside the update (each face is
// evaluated by both cells that share it, identically) and the velocities taken from the
// MHD states: stage 0 uses v^n (copy_velocity(.FALSE.)), stage 1 the time average
// 0.5 (v^n + v^{n+1}) (copy_velocity(.TRUE.)).
template <class real>
struct PassStageArgs {
    const real* src[MAXPASS];   // scalars at t^n (update base)
    const real* rec[MAXPASS];   // scalars the faces are built from (stage 1: half step)
    real* dst[MAXPASS];
    const real* qo[4];          // MHD rho, m at t^n
    const real* qn[4];          // MHD rho, m at t^{n+1} (stage 1)
    const real* vel[3];         // non-null: the stage's velocities, already formed per cell
    const double* dt;
    double factor;
    int npass, stage;
};

template <class real>
struct ShellArgs {
    const real* src[7];
    real* dst[7];
    const double* srcb[3];
    double* dstb[3];
    int lo, hi_off;   // cells with every index in [lo, n + hi_off] are left alone
};

// Interior planes k0 .. k0 + nk - 1 of one box, interleaved per cell (i fastest): the seven
// cell fields, then the first npo passive scalars -- the layout of the reference's xyz
// snapshot cube and parallel dump (io_mod.f90:249-277, 368-436).
template <class real>
struct CubeArgs {
    const real* q[7];
    const real* qp[MAXPASS];
    int npo, k0, nk, _pad;
};

// Flux protection of the VL Roe update (update_hydro_state_3d with flux_protection,
// vanleer_mod.f90:2080-2207), FP64 or FP32 fluxes (the corner EMFs are FP64 in both
// contexts).  k_prot_flag marks the cells of the update range whose
// first-pass density falls below thr; k_prot_fix gives every face of a marked cell the
// protection (HLLE) fluxes and the corner-EMF slots (i,j,k), (i-1,j,k), (i,j-1,k),
// (i,j,k-1) of a marked cell, all three components, the protection corner EMFs.  The
// ordinary update then recomputes the range from t^n with the modified arrays.
template <class real>
struct ProtArgs {
    const real* src[4];       // rho, m at the stage's base state (q_3d_prot)
    real* F[3][4];            // hydro fluxes per direction, global component order
    const real* Fp[3][4];     // protection fluxes, same layout
    double* Ec[3];            // corner EMFs
    const double* Ecp[3];     // protection corner EMFs
    unsigned char* flag;      // per cell of the field pool
    const double* dt;         // per box
    double factor;            // 0.5 predictor, 1 corrector
    double thr;               // density threshold (MIN_DENS in the reference)
    int lo, hi_off;           // update range [lo, n + hi_off] per axis
};
template <class real> void launch_prot_flag(const Launch& L, const ProtArgs<real>& a);
template <class real> void launch_prot_fix(const Launch& L, const ProtArgs<real>& a);

template <class real> void launch_restore_shell(const Launch& L, const ShellArgs<real>& a);
template <class real> void launch_fill_ghosts(const Launch& L, const GhostArgs<real>& a);
template <class real> void launch_fill_ghost_band(const Launch& L, const GhostArgs<real>& a);
template <class real> void launch_pass_stage(const Launch& L, const PassStageArgs<real>& a);
template <class real> void launch_max_speed(const Launch& L, const QPtrs<real>& q, double cs, unsigned long long* cmax);
template <class real> void launch_diag(hipStream_t st, const BoxDesc* boxes, int nbox, const DiagArgs<real>& a);
template <class real> void launch_copy_velocity(const Launch& L, const VelArgs<real>& a);
template <class real> void launch_pass_flux(const Launch& L, const PassFluxArgs<real>& a);
template <class real> void launch_pass_update(const Launch& L, const PassUpdArgs<real>& a);
void launch_new_dt(hipStream_t st, int nbox, const BoxDesc* boxes, const unsigned long long* cmax,
                   const unsigned long long* smax, const unsigned long long* dvmax, const double* cour, double t_end,
                   double* t, double* dt, double* dt_prev, int advance_time);
void launch_reset_u64(hipStream_t st, unsigned long long* p, int n);
template <class real> void launch_convert_in(hipStream_t st, const double* src, real* dst, long long n);
template <class real> void launch_convert_out(hipStream_t st, const real* src, double* dst, long long n);
template <class real, class out_t>
void launch_pack_cube(hipStream_t st, const BoxDesc& B, const CubeArgs<real>& a, out_t* out);
template <class real>
void launch_pass_layout(hipStream_t st, const BoxDesc& B, const double* hsrc, double* hdst, real* dev, int to_device,
                        long long g0, long long m);

}  // namespace imhd
