// Kernel argument blocks and host-side launch entry points of the cell-update kernels.
#pragma once
#include "imhd_common.hpp"

namespace imhd {

// A directional sweep sees the state rotated so that its direction is "x":
// w[r] = q[perm(d, r)] for r = 0..6 (w[4] unused; the normal field comes from bn).
template <class real>
struct SweepArgs {
    const real* w[7];
    const double* bn;
    const real* bnr;  // non-null: the normal face field in the working precision (the fused
                      // FP32 corrector's half-step faces), read instead of bn
    real* F[4];       // rotated hydro fluxes: rho, m_n, m_t1, m_t2
    real* Et1;        // face EMF = -flux(B_t1)
    real* Et2;        // face EMF = +flux(B_t2)
    int dir;
    int recon;        // corrector reconstruction (0 PLM primitive MC-beta)
    double cs;
    double mc_beta;
    unsigned long long* smax;  // per box, bit pattern of a non-negative double
};

// normal face field of a sweep at cell c, in the working precision
template <class real>
__device__ __forceinline__ real bn_at(const SweepArgs<real>& a, long long c) {
    return a.bnr != nullptr ? a.bnr[c] : (real)a.bn[c];
}

template <class real>
struct StateP {
    real* q[7];
    double* b[3];
};

// Face EMFs in the reference's emf_3d(comp, slot) layout (vanleer_mod.f90:1868-1883):
// E[2*comp + slot]; Ex: slot0 at y-face, slot1 at z-face; Ey: x-face, z-face;
// Ez: x-face, y-face.
template <class real>
struct CornerArgs {
    const real* q[7];      // state used for the centre EMF and the upwinding velocity
    const real* E[6];
    double* Ec[3];         // corner EMFs Ex(j+1/2,k+1/2), Ey(i+1/2,k+1/2), Ez(i+1/2,j+1/2)
};

template <class real>
struct UpdateArgs {
    StateP<real> src;      // state at t^n
    StateP<real> dst;      // updated state
    const real* F[3][4];   // hydro fluxes per direction, global component order
    const double* Ec[3];
    const double* dt;      // per-box time step
    double factor;         // 0.5 predictor, 1 corrector
    int ccb_lo_shift;      // cell-centred B is refreshed on [lo + shift, hi]
    real* comp[4];         // Kahan compensation of rho, m (FP32 only; null = off)
    unsigned int* floor_count;
    unsigned long long* cmax;  // non-null: also reduce the CFL speed of interior cells
    double cs;
};

struct Launch {
    const BoxDesc* boxes;
    Plan plan;
    hipStream_t stream;
    bool column = false;   // the plan is a column plan (z-marching kernel variant)
    unsigned blocks() const { return (unsigned)((plan.total + 255) / 256); }
};

// All three directional sweeps of a stage in one launch (see k_flux3): a cell computes
// its d-face when its index along d lies in [lo_off, n_d + hi_off].
template <class real>
struct Sweep3Args {
    SweepArgs<real> d[3];
    int lo_off, hi_off;
};
template <class real, int RS>
void launch_flux3(const Launch& L, const Sweep3Args<real>& a, int kind);

template <class real, int RS>
void launch_flux_predictor(const Launch& L, const SweepArgs<real>& a);
template <class real, int RS>
void launch_flux_corrector(const Launch& L, const SweepArgs<real>& a);
template <class real>
void launch_corner(const Launch& L, const CornerArgs<real>& a);
template <class real>
void launch_update(const Launch& L, const UpdateArgs<real>& a);

// One whole VL stage (fluxes, corner EMFs, update) in one launch: k_vl_stage.  A workgroup
// owns a tile of TX x TY (i, j) columns, of which the inner (TX-2) x (TY-2) are updated,
// and marches a z-segment; fluxes and EMFs stay in LDS and registers.  Tile shapes:
// 0 = 16 x 16 (256 threads), 1 = 32 x 16 (512 threads), 2 = 32 x 32 (1024 threads, FP32
// only: its 160 KB of LDS fill a CU, and a 30 x 30 update per tile solves 3.4 faces per
// updated cell where 30 x 14 solves 3.66).
constexpr int VS_SHAPES = 3;
constexpr int vs_tx(int shape) { return shape >= 1 ? 32 : 16; }
constexpr int vs_ty(int shape) { return shape == 2 ? 32 : 16; }
constexpr bool vs_shape_ok(int precision, int shape) { return shape < 2 || precision == 32; }
struct TileBox {
    int x0, y0;        // cell index of thread (0, 0) of tile (0, 0)
    int zlo, zhi;      // planes marched (updates masked to the update range)
    int nt[2];         // tiles along x, y (tile t starts at x0 + t (TX - 2))
    // z-segments: tiles 0 .. k1-1 (x fastest) have s+1 of them, the others s; segment g of
    // n covers planes zlo + g L / n .. zlo + (g+1) L / n - 1 (L = zhi - zlo + 1).
    // Workgroups are numbered segment-major: segment g of every tile, then g+1.
    int s, k1;
    long long start;   // first workgroup of this box
};
struct TilePlan {
    const TileBox* boxes;
    int nbox;
    long long total;   // workgroups
};
template <class real>
struct StageArgs {
    const real* qf[7];      // flux state, global order; direction d's rotated slot r is
                            // qf[perm(d, r)] (the kernel forms its sweep views from these)
    const double* bf[3];    // face field of the flux state
    const real* bfr[3];     // non-null: the flux state's face field in the working precision
                            // (FP32 half-step faces), read instead of bf
    double mc_beta;
    StateP<real> src, dst;  // update base and output
    const double* dt;
    double factor;          // 0.5 predictor, 1 corrector
    int ccb_lo_shift;
    real* comp[4];          // Kahan terms (final stage of the FP32 path) or null
    unsigned int* floor_count;
    unsigned long long* cmax;   // non-null: CFL speed of the new interior cells
    unsigned long long* smax;
    double cs;
    // index ranges per axis a, as [lo, n_a + hi]: faces normal / transverse, corners, updates
    int fn_lo, fn_hi, ft_lo, ft_hi, c_lo, c_hi, u_lo, u_hi;
    // non-null: the passive scalars' velocities of every updated cell, in the arithmetic of
    // copy_velocity -- the predictor writes m/rho at t^n, the corrector the average of t^n
    // and the new state
    real* pvel[3];
    // non-null (FP32 predictor): the half-step face field goes here in the working precision
    // instead of to dst.b -- the fused corrector reads it only as FP32 interface states
    real* bhw[3];
};
// kind: 0 predictor, 1 PLM corrector, 2 PPM-primitive corrector, 3 PPM-characteristic
// corrector (vs_kind), 4 FP32 PLM corrector reconstructing from LDS-held primitives with
// shared slopes, 5 FP32 predictor with LDS-held and lane-shifted neighbours
template <class real, int RS>
void launch_vl_stage(hipStream_t st, const BoxDesc* boxes, const TilePlan& plan, const StageArgs<real>& a,
                     int kind, int shape);
// workgroups of a fused stage resident per CU (occupancy of the kernel as compiled)
template <class real, int RS>
int vl_stage_blocks_per_cu(int kind, int shape);
constexpr int vs_kind(bool corrector, int recon) { return corrector ? 1 + recon : 0; }

}  // namespace imhd
