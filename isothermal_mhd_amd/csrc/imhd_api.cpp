// Host side of the imhd C ABI: device buffers, launch plans, step orchestration, native
// time loop with an optional hipGraph, per-kernel HIP-event timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include <hipfft/hipfft.h>

#include "../../include/imhd.h"
#include "aux_kernels.hpp"
#include "spectra.hpp"
#include "forcing.hpp"
#include "ic_kernels.hpp"
#include "kernels.hpp"
#include "tvd_kernels.hpp"

using namespace imhd;

namespace {

struct PlanDev {
    PlanBox* d = nullptr;
    long long total = 0;
    long long active = 0;  // work items (cells / faces) over all boxes
    bool column = false;   // (i, j, z-segment) enumeration for the z-marching kernels
    std::vector<PlanBox> h;
};

struct TilePlanDev {
    TileBox* d = nullptr;
    long long total = 0;
    long long active = 0;  // cells updated over all boxes
    std::vector<TileBox> h;
};

struct ProfRec {
    const char* name;
    long long items;
    hipEvent_t a, b;
};

}  // namespace

struct imhd_ctx {
    imhd_config cfg{};
    int nbox = 0;
    int prec = 64;
    size_t esz = 8;
    std::vector<BoxDesc> hb;
    BoxDesc* db = nullptr;
    long long pool = 0;  // padded cells over all boxes
    hipStream_t st = nullptr;
    // state buffers (two) and work arrays
    void* pvel[3] = {};    // passive velocities written by the fused VL stages
    void* q[4][7] = {};    // set 2: half-step state of the fused VL stages; TVD: sets 2, 3 hold
                           // the y and z half states of k_tvd_half3
    double* b[3][3] = {};
    void* bh[3] = {};      // FP32 fused VL: the half-step face field in FP32 (set 2's faces)
    void* qp[2][MAXPASS] = {};
    void* F[3][4] = {};
    void* E[6] = {};
    double* Ec[3] = {};
    void* vel[3] = {};
    void* vta[3] = {};
    void* Fp[3][MAXPASS] = {};
    void* comp[4] = {};
    // K98 TVD integrator: corrector fluxes, corrector centre EMF, per-face limiter inputs
    void* Fc[3][4] = {};
    double* Ecc[3] = {};
    void* gt[6] = {};
    void* ck[6] = {};
    int cur = 0, pcur = 0;
    bool have_state = false;
    // per-box device scalars
    unsigned long long* smax = nullptr;
    unsigned long long* cmax = nullptr;
    double* diag_part = nullptr;         // imhd_diagnostics: DIAG_WG partial rows per box (first call)
    double* diag_static = nullptr;       // imhd_diagnostics_static: this device's fixed scratch
    std::string diag_static_err;         // why the scratch address could not be had
    unsigned long long* dvmax = nullptr; // forcing: max |delta m| / rho0 of the last step (delv_max)
    // forcing application (driver-supplied field, turb_driver_mod.f90)
    double* delv[3] = {};
    double* force_part = nullptr;
    double* force_par = nullptr;
    unsigned int* force_err = nullptr;
    bool force_set = false;
    double edot = 0.0, dt_turb = 0.0;
    // the reference driver's impulsive forcing inside the native loop (driver.f90:57-70,
    // 102-111): taus88 state, next drive time, 0.5 COUR for the dt after a kick
    bool imp_on = false;
    int imp_kmax = 0;
    int imp_seeds[3] = {0, 0, 0};
    double imp_t_turb = 0.0;
    double* cour_half = nullptr;
    // in-situ spectra: the hipFFT R2C plan and its work buffers, kept between calls for
    // the cube edge fft_n (created on the first diagnostics call, freed by imhd_destroy)
    hipfftHandle fft_plan = 0;
    bool fft_have_plan = false;
    int fft_n = 0;
    unsigned long long fft_bytes = 0;   // plan work area and buffers (in dev_bytes too)
    void* fft_in = nullptr;         // n^3 floats (packed field)
    void* fft_out[3] = {};          // n^2 (n/2+1) complex each (mode decomposition uses 3)
    void* fft_acc = nullptr;        // (n/2)^3 floats (octant power)
    void* fft_bins = nullptr;       // 8 n doubles (bins) + 3 (mean field)
    // OU forcing of the native loop
    bool ou_on = false;
    imhd_ou ou_cfg{};
    std::vector<int> ou_k;        // 3 per mode (host copy)
    int* ou_n = nullptr;
    double* ou_are = nullptr;
    double* ou_aim = nullptr;
    unsigned int* ou_step = nullptr;
    double* ou_tw = nullptr;
    long long* ou_twoff = nullptr;
    double* ou_part = nullptr;
    double* ou_mean = nullptr;
    double* ou_col = nullptr;       // per-column mode sums of the separable field evaluation
    long long* ou_coloff = nullptr;
    int ou_maxcols = 0;
    int ou_K = 0;
    double* dt = nullptr;
    double* t = nullptr;
    double* dtprev = nullptr;
    double* cour = nullptr;
    double* diag = nullptr;
    unsigned int* floor_count = nullptr;
    unsigned long long dev_bytes = 0;   // field and work arrays allocated by alloc()
    double* stage = nullptr;  // FP64 staging for set/get
    size_t stage_n = 0;
    // plans
    PlanDev pl_pflux[3], pl_pcorner, pl_pupd, pl_cflux[3], pl_ccorner, pl_cupd;
    PlanDev pl_gband, pl_pgband;
    PlanDev pl_ghost, pl_pghost, pl_int, pl_vel, pl_ppflux[3], pl_ppupd, pl_pcflux[3], pl_pfupd;
    PlanDev pl_pstage[2];   // fused passive stages (column plans when zseg > 0)
    PlanDev pl_teig_p[3], pl_tflux_p[3], pl_teig_c[3], pl_tflux_c[3];
    PlanDev pl_tsweep_p[3], pl_tsweep_c[3];   // fused Roe-TVD sweeps
    // flux protection of the VL Roe update (vanleer_mod.f90:81, 93-110)
    bool prot = false;
    double prot_thr = MIN_DENS;
    void* Fpr[3][4] = {};
    void* Epr[6] = {};
    double* Ecpr[3] = {};
    unsigned char* pflag = nullptr;
    bool tvdfuse = true;                     // one launch per Roe-TVD sweep (IMHD_TVDFUSE=0: two)
    bool tvdhalf3 = true;                    // the three TVD half updates in one pass (IMHD_TVDHALF3=0: three)
    int zseg = 4;    // z-segment length of the column (z-marching) kernels
    int strip = 0;   // y-strip height of the work enumeration (0: whole planes)
    bool fuse3 = true;   // the three directional flux sweeps of a stage in one launch
    PlanDev pl_pflux3, pl_cflux3;
    bool vstage = true;  // VL stages as single fused launches (k_vl_stage)
    int vseg = 16;       // z-segment length of the fused stages when vrounds = 0
    int vrounds = 4;     // > 0: fused-stage workgroups = vrounds x resident workgroups of the GPU
    int ncu = 0;         // compute units of the device
    int vshape = 0;      // tile shape of the fused stages (0: 16 x 16, 1: 32 x 16; FP32 default 1)
    int vshape_c = 0;    // tile shape of the fused corrector (the autotuner may choose it apart)
    bool vsh = true;     // FP32 PLM corrector from LDS-held primitives with shared slopes (kind 4)
    TilePlanDev tp_pred, tp_corr;
    PlanDev pl_tcorner_p, pl_tcorner_c, pl_tcenter, pl_thalf, pl_tcfl;
    int ccb_shift = 1;
    // native loop
    hipGraphExec_t gexec[4] = {nullptr, nullptr, nullptr, nullptr};  // by state/passive buffer parity
    double graph_t_end = 0.0;
    bool prof = false;
    std::vector<ProfRec> prec_list;
    std::string prof_report;
    std::string err;
};

#define FAIL(ctx, code, ...)                                              \
    do {                                                                  \
        char _b[512];                                                     \
        snprintf(_b, sizeof(_b), __VA_ARGS__);                            \
        if (ctx) (ctx)->err = _b;                                         \
        return code;                                                      \
    } while (0)

#define HIPCK(ctx, call)                                                                          \
    do {                                                                                          \
        hipError_t _e = (call);                                                                   \
        if (_e != hipSuccess) FAIL(ctx, IMHD_EHIP, "%s failed: %s", #call, hipGetErrorString(_e)); \
    } while (0)

namespace {

// seg > 0: a column plan of the same cell range for the z-marching kernels, segments of
// seg cells along z
int build_plan(imhd_ctx* c, PlanDev& p, int which, int dir, int seg = 0) {
    // which: ranges as functions of (n, nb) per axis; dir = sweep direction for face plans
    p.h.clear();
    p.active = 0;
    p.column = seg > 0;
    if (p.d) { (void)hipFree(p.d); p.d = nullptr; }
    long long run = 0;
    const int e = c->cfg.recon == IMHD_PLM_PRIM ? 0 : 1;  // PPM needs one more cell (vanleer_mod.f90:379-383)
    for (int bx = 0; bx < c->nbox; bx++) {
        const BoxDesc& B = c->hb[bx];
        PlanBox pb{};
        for (int a = 0; a < 3; a++) {
            int n = B.n[a], nb = B.nb, lo = 0, hi = -1;
            bool normal = (a == dir);
            switch (which) {
                case 0:  // predictor fluxes (vanleer_mod.f90:437-446, offset 0)
                    lo = 1 - nb; hi = normal ? n + nb - 1 : n + nb; break;
                case 1:  // predictor corner EMF (vanleer_mod.f90:1900-1902)
                    lo = 1 - nb; hi = n + nb - 1; break;
                case 2:  // predictor update (vanleer_mod.f90:2102-2104)
                    lo = 2 - nb; hi = n + nb - 1; break;
                case 3:  // corrector fluxes, offset 2 (+1 PLM / +2 PPM for the reconstruction)
                    lo = normal ? 4 - nb + e : 3 - nb; hi = normal ? n + nb - 4 - e : n + nb - 2; break;
                case 4:  // corrector corner EMF, offset 3 (PLM) / 4 (PPM)
                    lo = 4 - nb + e; hi = n + nb - 4 - e; break;
                case 5:  // final update, offset 3 (PLM) / 4 (PPM)
                    lo = 5 - nb + e; hi = n + nb - 4 - e; break;
                case 6:  // whole padded box
                    lo = 1 - nb; hi = n + nb; break;
                case 7:  // passive ghost band
                    lo = -2; hi = n + 3; break;
                case 8:  // interior
                    lo = 1; hi = n; break;
                case 9:  // copy_velocity (passive_mod.f90:212-214)
                    lo = -2; hi = n + 3; break;
                case 10:  // passive predictor faces (passive_mod.f90:258-267, ilow = 1)
                    lo = -2; hi = normal ? n + 2 : n + 3; break;
                case 11:  // passive half update (passive_mod.f90:574-576, offset 0)
                    lo = -1; hi = n + 2; break;
                case 12:  // passive corrector faces (ilow = 2, reconstruction offset 1)
                    lo = normal ? 0 : -1; hi = normal ? n : n + 2; break;
                case 13:  // passive final update (offset 2)
                    lo = 1; hi = n; break;
                // K98 TVD integrator (isotvd_mod.f90); offset 0 predictor, 2 corrector
                case 25:  // union of the three predictor flux ranges (fused sweeps)
                    lo = 1 - nb; hi = n + nb; break;
                case 26:  // union of the three corrector flux ranges (their transverse range)
                    lo = 3 - nb; hi = n + nb - 2; break;
                case 16:  // predictor Roe average / HLLE faces: pencil cells [1-nb, n+nb]
                    lo = 1 - nb; hi = normal ? n + nb - 1 : n + nb; break;
                case 17:  // predictor TVD fluxes and face EMFs (isotvd_mod.f90:2413, 446)
                    lo = normal ? 2 - nb : 1 - nb; hi = normal ? n + nb - 2 : n + nb; break;
                case 18:  // corrector Roe average / HLLE faces: pencil cells [3-nb, n+nb-2]
                    lo = 3 - nb; hi = normal ? n + nb - 3 : n + nb - 2; break;
                case 19:  // corrector TVD fluxes and face EMFs
                    lo = normal ? 4 - nb : 3 - nb; hi = normal ? n + nb - 4 : n + nb - 2; break;
                case 27:  // fused predictor sweep: the faces of case 17
                    lo = normal ? 2 - nb : 1 - nb; hi = normal ? n + nb - 2 : n + nb; break;
                case 28:  // fused corrector sweep: the faces of case 19
                    lo = normal ? 4 - nb : 3 - nb; hi = normal ? n + nb - 4 : n + nb - 2; break;
                case 20:  // predictor corner EMF (isotvd_mod.f90:1006-1008)
                    lo = 2 - nb; hi = n + nb - 2; break;
                case 21:  // corrector corner EMF
                    lo = 4 - nb; hi = n + nb - 4; break;
                case 22:  // corrector centre EMF (isotvd_mod.f90:880-882)
                    lo = 4 - nb; hi = n + nb - 3; break;
                case 23:  // CTU half updates (isotvd_mod.f90:1203-1205)
                    lo = 3 - nb; hi = n + nb - 2; break;
                case 24:  // get_max_speed_tvd (isotvd_mod.f90:2620-2622)
                    lo = 0; hi = n; break;
            }
            pb.lo[a] = lo;
            pb.r[a] = std::max(0, hi - lo + 1);
        }
        if (which == 14 || which == 15) {  // ghost band of depth w, enumerated linearly
            long long w = which == 14 ? B.nb : 3;
            long long E0 = B.n[0] + 2 * w, E1 = B.n[1] + 2 * w;
            long long cnt = E0 * E1 * 2 * w + E0 * 2 * w * B.n[2] + 2 * w * (long long)B.n[1] * B.n[2];
            pb.lo[0] = pb.lo[1] = pb.lo[2] = 0;
            pb.r[0] = (int)cnt;
            pb.r[1] = pb.r[2] = 1;
        }
        if (seg > 0) {
            pb.zlo = pb.lo[2];
            pb.zn = pb.r[2];
            pb.seg = seg;
            pb.lo[2] = 0;
            pb.r[2] = (pb.zn + seg - 1) / seg;
        }
        pb.start = run;
        long long cnt = (long long)pb.r[0] * pb.r[1] * pb.r[2];
        if (which == 27 || which == 28) {
            // fused Roe-TVD sweep: x -- one wave per (row, TVD_XSEG faces);
            // y, z -- one thread per (x, transverse, segment of seg faces along the sweep)
            if (dir == 0) {
                pb.seg = TVD_XSEG;
                pb.zn = (pb.r[0] + TVD_XSEG - 1) / TVD_XSEG;
                cnt = (long long)pb.r[1] * pb.r[2] * pb.zn * 64;
            } else {
                const int ta = dir == 1 ? 2 : 1;
                pb.seg = pb.r[dir] >= 128 ? 32 : 8;
                pb.zn = (pb.r[dir] + pb.seg - 1) / pb.seg;
                cnt = (long long)pb.r[0] * pb.r[ta] * pb.zn;
            }
        }
        p.active += seg > 0 ? (long long)pb.r[0] * pb.r[1] * pb.zn
                            : (which == 27 || which == 28) ? (long long)pb.r[0] * pb.r[1] * pb.r[2] : cnt;   // cells
        run += (cnt + 255) / 256 * 256;
        p.h.push_back(pb);
    }
    p.total = run;
    HIPCK(c, hipMalloc(&p.d, sizeof(PlanBox) * p.h.size()));
    HIPCK(c, hipMemcpy(p.d, p.h.data(), sizeof(PlanBox) * p.h.size(), hipMemcpyHostToDevice));
    return 0;
}

// Index ranges of one VL stage per axis, as [lo, n + hi]: the flux sweeps' normal and
// transverse ranges, the corner EMFs and the updated cells (the ranges of build_plan's
// cases 0-2 for the predictor and 3-5 for the corrector).
struct StageRanges {
    int fn_lo, fn_hi, ft_lo, ft_hi, c_lo, c_hi, u_lo, u_hi;
};
StageRanges stage_ranges(const imhd_ctx* c, bool corr) {
    const int nb = c->hb[0].nb, e = c->cfg.recon == IMHD_PLM_PRIM ? 0 : 1;
    if (!corr) return {1 - nb, nb - 1, 1 - nb, nb, 1 - nb, nb - 1, 2 - nb, nb - 1};
    return {4 - nb + e, nb - 4 - e, 3 - nb, nb - 2, 4 - nb + e, nb - 4 - e, 5 - nb + e, nb - 4 - e};
}

// Tiles and z-segments of a fused stage.  Thread t of tile T sits at lo + T w + t (w = tile
// width - 2); threads 1 .. width-2 update, threads 0 .. width-1 solve faces, threads
// 0 .. width-2 form corners.  The tile count covers every face, corner and cell of the
// stage's ranges; the march covers planes zlo-1 .. zhi+1.
// kind of a fused stage's kernel: vs_kind, with the FP32 PLM corrector (4) and the FP32
// predictor (5) on their LDS-neighbour forms unless IMHD_VSH=0
int stage_kind(const imhd_ctx* c, bool corr) {
    const int k = vs_kind(corr, c->cfg.recon);
    if (c->prec == 32 && c->vsh) return k == 1 ? 4 : (k == 0 ? 5 : k);
    return k;
}
int stage_shape(const imhd_ctx* c, bool corr) { return corr ? c->vshape_c : c->vshape; }

template <class real>
int stage_blocks_per_cu(const imhd_ctx* c, bool corr) {
    switch (c->cfg.riemann) {
        case IMHD_ROE: return vl_stage_blocks_per_cu<real, RS_ROE>(stage_kind(c, corr), stage_shape(c, corr));
        case IMHD_HLLE: return vl_stage_blocks_per_cu<real, RS_HLLE>(stage_kind(c, corr), stage_shape(c, corr));
        default: return vl_stage_blocks_per_cu<real, RS_HLLD>(stage_kind(c, corr), stage_shape(c, corr));
    }
}

// Tiles and z-segments of a fused stage.  Thread t of tile T sits at lo + T w + t (w = tile
// width - 2); threads 1 .. width-2 update, threads 0 .. width-1 solve faces, threads
// 0 .. width-2 form corners.  The tile count covers every face, corner and cell of the
// stage's ranges; a segment's march covers planes ks-1 .. ke.  Segments per tile: with
// vrounds > 0 the launch has vrounds x (resident workgroups of the GPU) workgroups, split
// over boxes by work and over a box's tiles as s or s+1 segments, so every round of
// resident workgroups is full (more rounds if segments would exceed VS_MAXSEG planes);
// with vrounds = 0, segments of at most vseg planes.
constexpr long long VS_MAXSEG = 32;

int build_tiles(imhd_ctx* c, TilePlanDev& p, bool corr) {
    p.h.clear();
    p.active = 0;
    if (p.d) { (void)hipFree(p.d); p.d = nullptr; }
    const StageRanges R = stage_ranges(c, corr);
    const int lo = std::min(std::min(R.ft_lo, R.fn_lo), R.u_lo - 1);
    const int fhi = std::max(R.ft_hi, R.fn_hi);
    const int wx[2] = {vs_tx(stage_shape(c, corr)) - 2, vs_ty(stage_shape(c, corr)) - 2};
    std::vector<long long> tiles, planes;
    long long work = 0;
    for (int bx = 0; bx < c->nbox; bx++) {
        const BoxDesc& B = c->hb[bx];
        TileBox tb{};
        tb.x0 = lo;
        tb.y0 = lo;
        for (int a = 0; a < 2; a++) {
            const int need = std::max(B.n[a] + fhi - 1, B.n[a] + R.u_hi) - lo;
            tb.nt[a] = (need + wx[a] - 1) / wx[a];
        }
        tb.zlo = lo + 1;
        tb.zhi = B.n[2] + std::max(std::max(R.ft_hi - 1, R.fn_hi), R.u_hi);
        tiles.push_back((long long)tb.nt[0] * tb.nt[1]);
        planes.push_back(tb.zhi - tb.zlo + 1);
        work += tiles.back() * planes.back();
        long long cells = 1;
        for (int a = 0; a < 3; a++) cells *= std::max(0, B.n[a] + R.u_hi - R.u_lo + 1);
        p.active += cells;
        p.h.push_back(tb);
    }
    long long target = 0;
    if (c->vrounds > 0) {
        const int bpc = c->prec == 64 ? stage_blocks_per_cu<double>(c, corr) : stage_blocks_per_cu<float>(c, corr);
        const long long slots = (long long)bpc * std::max(1, c->ncu);
        target = (long long)c->vrounds * slots;
        // segments of at most VS_MAXSEG planes (a whole number of rounds): many small boxes
        // would otherwise get one long segment per tile, of unequal lengths across boxes
        const long long need = (work + VS_MAXSEG - 1) / VS_MAXSEG;
        if (target < need) target = (need + slots - 1) / slots * slots;
    }
    long long run = 0;
    for (int bx = 0; bx < c->nbox; bx++) {
        TileBox& tb = p.h[bx];
        const long long T = tiles[bx], L = planes[bx];
        long long n = target > 0 ? std::llround((double)target * (double)(T * L) / (double)work)
                                 : T * ((L + std::max(1, c->vseg) - 1) / std::max(1, c->vseg));
        n = std::min(std::max(n, T), T * L);
        tb.s = (int)(n / T);
        tb.k1 = (int)(n % T);
        tb.start = run;
        run += n;
    }
    p.total = run;
    HIPCK(c, hipMalloc(&p.d, sizeof(TileBox) * p.h.size()));
    HIPCK(c, hipMemcpy(p.d, p.h.data(), sizeof(TileBox) * p.h.size(), hipMemcpyHostToDevice));
    return 0;
}

Launch mk(imhd_ctx* c, const PlanDev& p) {
    Launch L;
    L.boxes = c->db;
    L.plan.boxes = p.d;
    L.plan.nbox = c->nbox;
    L.plan.total = p.total;
    L.stream = c->st;
    L.column = p.column;
    L.plan.sy = c->strip;
    return L;
}

void prof_begin(imhd_ctx* c, const char* name, long long items) {
    if (!c->prof) return;
    ProfRec r{name, items, nullptr, nullptr};
    (void)hipEventCreate(&r.a);
    (void)hipEventCreate(&r.b);
    (void)hipEventRecord(r.a, c->st);
    c->prec_list.push_back(r);
}
void prof_end(imhd_ctx* c) {
    if (!c->prof || c->prec_list.empty()) return;
    (void)hipEventRecord(c->prec_list.back().b, c->st);
}
struct ProfScope {
    imhd_ctx* c;
    ProfScope(imhd_ctx* c_, const char* n, long long items) : c(c_) { prof_begin(c, n, items); }
    ~ProfScope() { prof_end(c); }
};

int alloc(imhd_ctx* c, void** p, size_t bytes) {
    HIPCK(c, hipMalloc(p, bytes));
    HIPCK(c, hipMemset(*p, 0, bytes));
    c->dev_bytes += bytes;
    return 0;
}

void set_scalar_dt(imhd_ctx* c, double dt);

template <class real>
real* R(void* p) { return static_cast<real*>(p); }

constexpr int ET1[3] = {4, 0, 3};  // -flux(B_t1) -> Ez@x, Ex@y, Ey@z
constexpr int ET2[3] = {2, 5, 1};  // +flux(B_t2) -> Ey@x, Ez@y, Ex@z

template <class real>
SweepArgs<real> sweep_args(imhd_ctx* c, int s, int d) {
    SweepArgs<real> a{};
    for (int r = 0; r < 7; r++) a.w[r] = (r == 4) ? nullptr : R<real>(c->q[s][perm(d, r)]);
    a.bn = c->b[s][d];
    for (int r = 0; r < 4; r++) a.F[r] = R<real>(c->F[d][perm(d, r)]);
    a.Et1 = R<real>(c->E[ET1[d]]);
    a.Et2 = R<real>(c->E[ET2[d]]);
    a.dir = d;
    a.recon = c->cfg.recon;
    a.cs = c->cfg.cs;
    a.mc_beta = c->cfg.mc_beta;
    a.smax = c->smax;
    return a;
}

template <class real, int RS>
void sweeps(imhd_ctx* c, int s, bool corrector) {
    if (c->fuse3) {   // the three directions in one launch over their union range
        const PlanDev& p = corrector ? c->pl_cflux3 : c->pl_pflux3;
        ProfScope ps(c, corrector ? "flux_corr" : "flux_pred", p.active);
        Sweep3Args<real> a{};
        for (int d = 0; d < 3; d++) a.d[d] = sweep_args<real>(c, s, d);
        const int nb = c->hb[0].nb, e = c->cfg.recon == IMHD_PLM_PRIM ? 0 : 1;
        a.lo_off = corrector ? 4 - nb + e : 1 - nb;
        a.hi_off = corrector ? nb - 4 - e : nb - 1;
        const int kind = !corrector ? 0 : (c->cfg.recon == IMHD_PLM_PRIM ? 1 : (c->cfg.recon == IMHD_PPM_PRIM ? 2 : 3));
        launch_flux3<real, RS>(mk(c, p), a, kind);
        return;
    }
    static const char* nm[2][3] = {{"flux_pred_x", "flux_pred_y", "flux_pred_z"},
                                   {"flux_corr_x", "flux_corr_y", "flux_corr_z"}};
    for (int d = 0; d < 3; d++) {
        ProfScope ps(c, nm[corrector][d], (corrector ? c->pl_cflux[d] : c->pl_pflux[d]).active);
        SweepArgs<real> a = sweep_args<real>(c, s, d);
        if (!corrector) launch_flux_predictor<real, RS>(mk(c, c->pl_pflux[d]), a);
        else launch_flux_corrector<real, RS>(mk(c, c->pl_cflux[d]), a);
    }
}

template <class real>
void corner(imhd_ctx* c, int s, const PlanDev& p, const char* name) {
    ProfScope ps(c, name, p.active);
    CornerArgs<real> a{};
    for (int f = 0; f < 7; f++) a.q[f] = R<real>(c->q[s][f]);
    for (int e = 0; e < 6; e++) a.E[e] = R<real>(c->E[e]);
    for (int d = 0; d < 3; d++) a.Ec[d] = c->Ec[d];
    launch_corner<real>(mk(c, p), a);
}

template <class real>
void update(imhd_ctx* c, int src, int dst, double factor, const PlanDev& p, bool final, const char* name,
            bool fuse_cmax = false) {
    ProfScope ps(c, name, p.active);
    UpdateArgs<real> a{};
    for (int f = 0; f < 7; f++) { a.src.q[f] = R<real>(c->q[src][f]); a.dst.q[f] = R<real>(c->q[dst][f]); }
    for (int d = 0; d < 3; d++) { a.src.b[d] = c->b[src][d]; a.dst.b[d] = c->b[dst][d]; a.Ec[d] = c->Ec[d]; }
    for (int d = 0; d < 3; d++)
        for (int f = 0; f < 4; f++) a.F[d][f] = R<real>(c->F[d][f]);
    a.dt = c->dt;
    a.factor = factor;
    a.ccb_lo_shift = c->ccb_shift;
    for (int f = 0; f < 4; f++) a.comp[f] = (final && c->comp[0]) ? R<real>(c->comp[f]) : nullptr;
    a.floor_count = c->floor_count;
    a.cmax = fuse_cmax ? c->cmax : nullptr;
    a.cs = c->cfg.cs;
    launch_update<real>(mk(c, p), a);
}

// One whole stage in one launch (k_vl_stage): fluxes from set sflux, update src -> dst.
// The predictor's update base is its flux state (sflux == src): the kernel reads that
// state once per cell for both.
template <class real>
void vl_stage(imhd_ctx* c, int sflux, int src, int dst, bool corr, bool fuse_cmax, const char* name,
              bool pvel = false) {
    if (!corr) src = sflux;
    const TilePlanDev& p = corr ? c->tp_corr : c->tp_pred;
    ProfScope ps(c, name, p.active);
    StageArgs<real> a{};
    for (int f = 0; f < 7; f++) {
        a.qf[f] = R<real>(c->q[sflux][f]);
        a.src.q[f] = R<real>(c->q[src][f]);
        a.dst.q[f] = R<real>(c->q[dst][f]);
    }
    for (int d = 0; d < 3; d++) {
        a.bf[d] = c->b[sflux][d];
        a.src.b[d] = c->b[src][d];
        a.dst.b[d] = c->b[dst][d];
    }
    if (c->bh[0] != nullptr) {   // FP32: the half-step faces live in bh
        for (int d = 0; d < 3; d++) {
            if (dst == 2) a.bhw[d] = R<real>(c->bh[d]);
            if (sflux == 2) a.bfr[d] = R<real>(c->bh[d]);
        }
    }
    a.mc_beta = c->cfg.mc_beta;
    a.dt = c->dt;
    a.factor = corr ? 1.0 : 0.5;
    a.ccb_lo_shift = c->ccb_shift;
    for (int f = 0; f < 4; f++) a.comp[f] = (corr && c->comp[0]) ? R<real>(c->comp[f]) : nullptr;
    a.floor_count = c->floor_count;
    a.cmax = fuse_cmax ? c->cmax : nullptr;
    a.smax = c->smax;
    a.cs = c->cfg.cs;
    const StageRanges r = stage_ranges(c, corr);
    a.fn_lo = r.fn_lo; a.fn_hi = r.fn_hi; a.ft_lo = r.ft_lo; a.ft_hi = r.ft_hi;
    a.c_lo = r.c_lo; a.c_hi = r.c_hi; a.u_lo = r.u_lo; a.u_hi = r.u_hi;
    for (int d = 0; d < 3; d++) a.pvel[d] = pvel ? R<real>(c->pvel[d]) : nullptr;
    const TilePlan tp{p.d, c->nbox, p.total};
    switch (c->cfg.riemann) {
        case IMHD_ROE: launch_vl_stage<real, RS_ROE>(c->st, c->db, tp, a, stage_kind(c, corr), stage_shape(c, corr)); break;
        case IMHD_HLLE: launch_vl_stage<real, RS_HLLE>(c->st, c->db, tp, a, stage_kind(c, corr), stage_shape(c, corr)); break;
        default: launch_vl_stage<real, RS_HLLD>(c->st, c->db, tp, a, stage_kind(c, corr), stage_shape(c, corr)); break;
    }
}

template <class real>
void passive_stage(imhd_ctx* c, int qold, int qnew, int stage, bool vel);

// with_passive (both stages fused, passives, no forcing kick between the MHD step and the
// passive step): the passive stages run inline -- the predictor right after the MHD
// predictor (it needs only t^n data), the corrector after the MHD corrector -- and read
// the velocities the fused MHD stages write, instead of re-deriving them from rho and m
// ---- flux protection (vanleer_mod.f90:540, 703, 880, 336, 396, 2084-2200) ----
// The protection fluxes are the VL module's HLLE on the same interface states (the flux
// kernels run again with the HLLE solver and write to the protection arrays; smax takes
// both solvers' speeds, as compute_HLLE_fluxes_1d raises it too).
template <class real>
void protection_fluxes(imhd_ctx* c, int s, bool corrector) {
    ProfScope ps(c, corrector ? "prot_flux_corr" : "prot_flux_pred", 0);
    for (int d = 0; d < 3; d++)
        for (int f = 0; f < 4; f++) std::swap(c->F[d][f], c->Fpr[d][f]);
    for (int e = 0; e < 6; e++) std::swap(c->E[e], c->Epr[e]);
    sweeps<real, RS_HLLE>(c, s, corrector);
    for (int d = 0; d < 3; d++)
        for (int f = 0; f < 4; f++) std::swap(c->F[d][f], c->Fpr[d][f]);
    for (int e = 0; e < 6; e++) std::swap(c->E[e], c->Epr[e]);
}

template <class real>
void protection_corner(imhd_ctx* c, int s, const PlanDev& p) {
    for (int e = 0; e < 6; e++) std::swap(c->E[e], c->Epr[e]);
    for (int d = 0; d < 3; d++) std::swap(c->Ec[d], c->Ecpr[d]);
    corner<real>(c, s, p, "prot_corner");
    for (int e = 0; e < 6; e++) std::swap(c->E[e], c->Epr[e]);
    for (int d = 0; d < 3; d++) std::swap(c->Ec[d], c->Ecpr[d]);
}

// Mark the cells whose first-pass density falls below the threshold and swap in the
// protection fluxes and corner EMFs around them; the ordinary update then recomputes the
// range from the stage's base state src.
template <class real>
void protection_fix(imhd_ctx* c, int src, bool corrector) {
    ProfScope ps(c, corrector ? "prot_fix_corr" : "prot_fix_pred", c->pl_ghost.active);
    ProtArgs<real> a{};
    for (int f = 0; f < 4; f++) a.src[f] = (const real*)c->q[src][f];
    for (int d = 0; d < 3; d++)
        for (int f = 0; f < 4; f++) {
            a.F[d][f] = (real*)c->F[d][f];
            a.Fp[d][f] = (const real*)c->Fpr[d][f];
        }
    for (int d = 0; d < 3; d++) { a.Ec[d] = c->Ec[d]; a.Ecp[d] = c->Ecpr[d]; }
    a.flag = c->pflag;
    a.dt = c->dt;
    a.factor = corrector ? 1.0 : 0.5;
    a.thr = c->prot_thr;
    const int nb = c->hb[0].nb, e = c->cfg.recon == IMHD_PLM_PRIM ? 0 : 1;
    a.lo = corrector ? 5 - nb + e : 2 - nb;          // build_plan cases 5 and 2
    a.hi_off = corrector ? nb - 4 - e : nb - 1;
    launch_prot_flag(mk(c, c->pl_ghost), a);
    launch_prot_fix(mk(c, c->pl_ghost), a);
}

template <class real>
void vl_step(imhd_ctx* c, bool fuse_cmax = false, bool with_passive = false) {
    launch_reset_u64(c->st, c->smax, (int)stripe_slots(c->nbox));
    launch_reset_u64(c->st, c->dvmax, (int)stripe_slots(c->nbox));   // delv_max = 0 (vanleer_mod.f90:301)
    const int s0 = c->cur, s1 = 1 - c->cur;
    if (with_passive) {
        vl_stage<real>(c, s0, s0, 2, false, false, "stage_pred", true);
        passive_stage<real>(c, s0, s0, 0, true);
        vl_stage<real>(c, 2, s0, s1, true, fuse_cmax, fuse_cmax ? "stage_corr_cfl" : "stage_corr", true);
        passive_stage<real>(c, s0, s1, 1, true);
        c->cur = s1;
        return;
    }
    if (c->vstage) {
        // Fused stages.  The corrector's fluxes read the half step while its update writes
        // the new state, so the half step goes to its own set (2) when the corrector is
        // fused too; with IMHD_VPPM=0 a PPM corrector keeps the three-kernel path on set s1.
        const bool fcorr = c->q[2][0] != nullptr;
        const int sh = fcorr ? 2 : s1;
        vl_stage<real>(c, s0, s0, sh, false, false, "stage_pred");
        if (fcorr) {
            vl_stage<real>(c, 2, s0, s1, true, fuse_cmax, fuse_cmax ? "stage_corr_cfl" : "stage_corr");
        } else {
            switch (c->cfg.riemann) {
                case IMHD_ROE: sweeps<real, RS_ROE>(c, s1, true); break;
                case IMHD_HLLE: sweeps<real, RS_HLLE>(c, s1, true); break;
                default: sweeps<real, RS_HLLD>(c, s1, true); break;
            }
            corner<real>(c, s1, c->pl_ccorner, "corner_corr");
            update<real>(c, s0, s1, 1.0, c->pl_cupd, true, fuse_cmax ? "update_corr_cfl" : "update_corr", fuse_cmax);
        }
        c->cur = s1;
        return;
    }
    switch (c->cfg.riemann) {
        case IMHD_ROE: sweeps<real, RS_ROE>(c, s0, false); break;
        case IMHD_HLLE: sweeps<real, RS_HLLE>(c, s0, false); break;
        default: sweeps<real, RS_HLLD>(c, s0, false); break;
    }
    if (c->prot) protection_fluxes<real>(c, s0, false);
    corner<real>(c, s0, c->pl_pcorner, "corner_pred");
    if (c->prot) protection_corner<real>(c, s0, c->pl_pcorner);
    if (c->prot) protection_fix<real>(c, s0, false);
    update<real>(c, s0, s1, 0.5, c->pl_pupd, false, "update_pred");
    switch (c->cfg.riemann) {
        case IMHD_ROE: sweeps<real, RS_ROE>(c, s1, true); break;
        case IMHD_HLLE: sweeps<real, RS_HLLE>(c, s1, true); break;
        default: sweeps<real, RS_HLLD>(c, s1, true); break;
    }
    if (c->prot) protection_fluxes<real>(c, s1, true);
    corner<real>(c, s1, c->pl_ccorner, "corner_corr");
    if (c->prot) protection_corner<real>(c, s1, c->pl_ccorner);
    if (c->prot) protection_fix<real>(c, s0, true);
    update<real>(c, s0, s1, 1.0, c->pl_cupd, true, fuse_cmax ? "update_corr_cfl" : "update_corr", fuse_cmax);
    c->cur = s1;
}

// ---- K98 TVD integrator (tvd_solve_3d, isotvd_mod.f90:226-364) ----

template <class real>
TvdSweepArgs<real> tvd_args(imhd_ctx* c, int sq, int sb, bool corrector, int d) {
    TvdSweepArgs<real> a{};
    for (int r = 0; r < 7; r++) a.w[r] = (r == 4) ? nullptr : R<real>(c->q[sq][perm(d, r)]);
    a.bn = c->b[sb][d];
    for (int w = 0; w < 6; w++) { a.gt[w] = R<real>(c->gt[w]); a.ck[w] = R<real>(c->ck[w]); }
    for (int r = 0; r < 4; r++) a.F[r] = R<real>(corrector ? c->Fc[d][perm(d, r)] : c->F[d][perm(d, r)]);
    a.Et1 = R<real>(c->E[ET1[d]]);
    a.Et2 = R<real>(c->E[ET2[d]]);
    a.dt = c->dt;
    a.factor = corrector ? 1.0 : 0.5;
    a.cs = c->cfg.cs;
    a.dir = d;
    return a;
}

// compute_{x,y,z}fluxes_3d (isotvd_mod.f90:368-863): Roe-TVD (eig + flux) or HLLE
template <class real>
void tvd_sweep(imhd_ctx* c, int sq, int sb, bool corr, int d) {
    static const char* nm[2][3][3] = {
        {{"tvd_eig_pred_x", "tvd_eig_pred_y", "tvd_eig_pred_z"},
         {"tvd_flux_pred_x", "tvd_flux_pred_y", "tvd_flux_pred_z"},
         {"tvd_hlle_pred_x", "tvd_hlle_pred_y", "tvd_hlle_pred_z"}},
        {{"tvd_eig_corr_x", "tvd_eig_corr_y", "tvd_eig_corr_z"},
         {"tvd_flux_corr_x", "tvd_flux_corr_y", "tvd_flux_corr_z"},
         {"tvd_hlle_corr_x", "tvd_hlle_corr_y", "tvd_hlle_corr_z"}}};
    TvdSweepArgs<real> a = tvd_args<real>(c, sq, sb, corr, d);
    const PlanDev& pe = corr ? c->pl_teig_c[d] : c->pl_teig_p[d];
    if (c->cfg.riemann == IMHD_HLLE) {
        ProfScope ps(c, nm[corr][2][d], pe.active);
        launch_tvd_hlle<real>(mk(c, pe), a);
        return;
    }
    if (c->tvdfuse) {
        static const char* nf[2][3] = {{"tvd_sweep_pred_x", "tvd_sweep_pred_y", "tvd_sweep_pred_z"},
                                       {"tvd_sweep_corr_x", "tvd_sweep_corr_y", "tvd_sweep_corr_z"}};
        const PlanDev& ps_ = corr ? c->pl_tsweep_c[d] : c->pl_tsweep_p[d];
        ProfScope ps(c, nf[corr][d], ps_.active);
        launch_tvd_sweep<real>(mk(c, ps_), a);
        return;
    }
    {
        ProfScope ps(c, nm[corr][0][d], pe.active);
        launch_tvd_eig<real>(mk(c, pe), a);
    }
    const PlanDev& pf = corr ? c->pl_tflux_c[d] : c->pl_tflux_p[d];
    ProfScope ps(c, nm[corr][1][d], pf.active);
    launch_tvd_flux<real>(mk(c, pf), a);
}

template <class real>
void tvd_corner(imhd_ctx* c, int s0, bool corr) {
    const PlanDev& p = corr ? c->pl_tcorner_c : c->pl_tcorner_p;
    ProfScope ps(c, corr ? "tvd_corner_corr" : "tvd_corner_pred", p.active);
    TvdCornerArgs<real> a{};
    for (int f = 0; f < 4; f++) a.q[f] = R<real>(c->q[s0][f]);
    for (int f = 0; f < 3; f++) a.qb[f] = R<real>(c->q[s0][4 + f]);
    for (int d = 0; d < 3; d++) a.Ecc[d] = corr ? c->Ecc[d] : nullptr;
    for (int e = 0; e < 6; e++) a.E[e] = R<real>(c->E[e]);
    for (int d = 0; d < 3; d++) a.Ec[d] = c->Ec[d];
    launch_tvd_corner<real>(mk(c, p), a);
}

// One tvd_solve_3d.  Buffers: t^n in state/face set s0; the CT half-step faces, the CTU
// half states and finally the new state in set s1.
template <class real>
void tvd_step(imhd_ctx* c) {
    launch_reset_u64(c->st, c->smax, (int)stripe_slots(c->nbox));  // no smax in the TVD module
    const int s0 = c->cur, s1 = 1 - c->cur;
    for (int d = 0; d < 3; d++) tvd_sweep<real>(c, s0, s0, false, d);
    tvd_corner<real>(c, s0, false);
    {
        ProfScope ps(c, "tvd_ct_pred", c->pl_ghost.active);
        TvdCtArgs<real> a{};
        for (int d = 0; d < 3; d++) { a.b0[d] = c->b[s0][d]; a.b1[d] = c->b[s1][d]; a.Ec[d] = c->Ec[d]; }
        a.dt = c->dt;
        launch_tvd_ct_pred<real>(mk(c, c->pl_ghost), a);
    }
    if (c->q[3][0] == nullptr) {   // else k_tvd_half3 forms the centre EMF from its own loads
        ProfScope ps(c, "tvd_center_corr", c->pl_tcenter.active);
        TvdCenterArgs<real> a{};
        for (int f = 0; f < 4; f++) a.q[f] = R<real>(c->q[s0][f]);
        for (int d = 0; d < 3; d++) {
            for (int f = 0; f < 4; f++) a.Fb[d][f] = R<real>(c->F[d][f]);
            a.bh[d] = c->b[s1][d];
            a.Ecc[d] = c->Ecc[d];
        }
        a.dt = c->dt;
        launch_tvd_center<real>(mk(c, c->pl_tcenter), a);
    }
    if (c->q[3][0] != nullptr) {   // the three half states (and centre EMF) in one pass, then the sweeps
        {
            ProfScope ps(c, "tvd_half3", c->pl_thalf.active);
            TvdHalfArgs<real> a{};
            TvdHalfDst<real> o{};
            const int hs[3] = {s1, 2, 3};
            for (int f = 0; f < 7; f++) a.q[f] = R<real>(c->q[s0][f]);
            for (int e = 0; e < 3; e++) {
                a.b0[e] = c->b[s0][e];
                a.Ec[e] = c->Ec[e];
                for (int f = 0; f < 4; f++) a.Fb[e][f] = R<real>(c->F[e][f]);
                for (int f = 0; f < 7; f++) o.dst[e][f] = R<real>(c->q[hs[e]][f]);
                o.bh[e] = c->b[s1][e];
                o.Ecc[e] = c->Ecc[e];
            }
            a.dt = c->dt;
            a.floor_count = c->floor_count;
            launch_tvd_half3<real>(mk(c, c->pl_thalf), a, o);
        }
        tvd_sweep<real>(c, s1, s1, true, 0);
        tvd_sweep<real>(c, 2, s1, true, 1);
        tvd_sweep<real>(c, 3, s1, true, 2);
    } else {
    static const char* hn[3] = {"tvd_half_x", "tvd_half_y", "tvd_half_z"};
    for (int d = 0; d < 3; d++) {
        {
            ProfScope ps(c, hn[d], c->pl_thalf.active);
            TvdHalfArgs<real> a{};
            for (int f = 0; f < 7; f++) { a.q[f] = R<real>(c->q[s0][f]); a.dst[f] = R<real>(c->q[s1][f]); }
            for (int e = 0; e < 3; e++) {
                a.b0[e] = c->b[s0][e];
                a.Ec[e] = c->Ec[e];
                for (int f = 0; f < 4; f++) a.Fb[e][f] = R<real>(c->F[e][f]);
            }
            a.dt = c->dt;
            a.floor_count = c->floor_count;
            a.dir = d;
            launch_tvd_half<real>(mk(c, c->pl_thalf), a);
        }
        tvd_sweep<real>(c, s1, s1, true, d);
    }
    }
    tvd_corner<real>(c, s0, true);
    {
        ProfScope ps(c, "tvd_final", c->pl_ghost.active);
        TvdFinalArgs<real> a{};
        for (int f = 0; f < 7; f++) { a.src.q[f] = R<real>(c->q[s0][f]); a.dst.q[f] = R<real>(c->q[s1][f]); }
        for (int d = 0; d < 3; d++) {
            a.src.b[d] = c->b[s0][d];
            a.dst.b[d] = c->b[s1][d];
            a.Ec[d] = c->Ec[d];
            for (int f = 0; f < 4; f++) a.F[d][f] = R<real>(c->Fc[d][f]);
        }
        a.dt = c->dt;
        for (int f = 0; f < 4; f++) a.comp[f] = c->comp[0] ? R<real>(c->comp[f]) : nullptr;
        a.floor_count = c->floor_count;
        launch_tvd_final<real>(mk(c, c->pl_ghost), a);
    }
    c->cur = s1;
}

template <class real>
void tvd_max_speed_launch(imhd_ctx* c) {
    ProfScope ps(c, "max_speed_tvd", c->pl_tcfl.active);
    launch_reset_u64(c->st, c->cmax, (int)stripe_slots(c->nbox));
    QPtrs<real> q{};
    for (int f = 0; f < 7; f++) q.q[f] = R<real>(c->q[c->cur][f]);
    launch_tvd_max_speed<real>(mk(c, c->pl_tcfl), q, c->cfg.cs, c->cmax);
}

bool is_tvd(const imhd_ctx* c) { return c->cfg.integrator == IMHD_TVD; }

// get_rho_turb (before the step, on the t^n state s0) and add_random_force (after it, on
// the new state) with the field set by imhd_set_force (vanleer_mod.f90:304-305, 419-421)
template <class real>
void force_sums(imhd_ctx* c, int s0) {
    ProfScope ps(c, "force_sums", c->pl_int.active);
    ForceSumArgs<real> a{};
    for (int f = 0; f < 4; f++) a.q[f] = R<real>(c->q[s0][f]);
    for (int d = 0; d < 3; d++) a.delv[d] = c->delv[d];
    a.part = c->force_part;
    launch_force_sums<real>(mk(c, c->pl_int), a);
    ForceFinArgs f{c->force_part, c->force_par, c->edot, c->dt_turb, c->force_err};
    launch_force_finalize(c->st, c->db, c->pl_int.d, c->nbox, f);
}

template <class real>
void force_apply(imhd_ctx* c, int s0, int s1) {
    ProfScope ps(c, "force_apply", c->pl_int.active);
    ForceApplyArgs<real> a{};
    a.rho0 = R<real>(c->q[s0][0]);
    for (int d = 0; d < 3; d++) { a.m[d] = R<real>(c->q[s1][1 + d]); a.delv[d] = c->delv[d]; }
    a.par = c->force_par;
    a.dvmax = c->dvmax;
    launch_force_apply<real>(mk(c, c->pl_int), a);
}

OuDev ou_dev(imhd_ctx* c) {
    OuDev o{};
    o.n = c->ou_n;
    o.are = c->ou_are;
    o.aim = c->ou_aim;
    o.step = c->ou_step;
    o.tw = c->ou_tw;
    o.twoff = c->ou_twoff;
    o.nmode = (int)(c->ou_k.size() / 3);
    o.K = c->ou_K;
    o.t_corr = c->ou_cfg.t_corr;
    o.sigma = c->ou_cfg.sigma;
    o.seed_lo = (unsigned)(c->ou_cfg.seed & 0xffffffffull);
    o.seed_hi = (unsigned)(c->ou_cfg.seed >> 32);
    o.solenoidal = c->ou_cfg.solenoidal;
    o.dt = c->dt;
    o.col = c->ou_col;
    o.coloff = c->ou_coloff;
    return o;
}

// OU step on the current (new) state: mode update, <rho a>, kick (+ the CFL speed of the
// kicked cells when fuse_cmax)
template <class real>
void ou_step(imhd_ctx* c, bool fuse_cmax) {
    ProfScope ps(c, "ou_forcing", c->pl_int.active);
    OuDev o = ou_dev(c);
    launch_ou_update(c->st, c->nbox, c->db, o, 0);
    launch_ou_columns(c->st, c->nbox, c->ou_maxcols, c->db, o);
    OuApplyArgs<real> a{};
    a.ou = o;
    for (int f = 0; f < 7; f++) a.q[f] = R<real>(c->q[c->cur][f]);
    a.part = c->ou_part;
    a.mean = c->ou_mean;
    a.cs = c->cfg.cs;
    int maxnx = 0;
    for (const BoxDesc& B : c->hb) maxnx = std::max(maxnx, B.n[0]);
    launch_ou_sums<real>(c->st, c->nbox, maxnx, c->db, a);
    launch_ou_finalize(c->st, c->db, c->nbox, maxnx, a.ou.K, c->ou_part, c->ou_mean);
    a.cmax = fuse_cmax ? c->cmax : nullptr;
    launch_ou_kick<real>(c->st, c->nbox, maxnx, c->db, a);
}

// Drop-in step only: cells outside the VL final-update range take their t^n values, as
// the reference's restore leaves them (the native loop's ghost fill makes this moot).
template <class real>
void restore_shell(imhd_ctx* c) {
    ProfScope ps(c, "restore_shell", c->pl_ghost.active);
    const int e = c->cfg.recon == IMHD_PLM_PRIM ? 0 : 1;
    const int s0 = 1 - c->cur, s1 = c->cur;   // after the step: cur is the new state
    ShellArgs<real> a{};
    for (int f = 0; f < 7; f++) { a.src[f] = R<real>(c->q[s0][f]); a.dst[f] = R<real>(c->q[s1][f]); }
    for (int d = 0; d < 3; d++) { a.srcb[d] = c->b[s0][d]; a.dstb[d] = c->b[s1][d]; }
    a.lo = 5 - c->hb[0].nb + e;       // final update range [5-nb+e, n+nb-4-e] (build_plan case 5)
    a.hi_off = c->hb[0].nb - 4 - e;
    launch_restore_shell<real>(mk(c, c->pl_ghost), a);
}

template <class real>
void fill_ghosts(imhd_ctx* c) {
    {
        ProfScope ps(c, "fill_ghosts", c->pl_gband.active);
        GhostArgs<real> g{};
        for (int f = 0; f < 7; f++) g.r[f] = R<real>(c->q[c->cur][f]);
        for (int d = 0; d < 3; d++) g.d[d] = c->b[c->cur][d];
        g.nreal = 7;
        g.ndbl = 3;
        g.depth = c->hb[0].nb;
        launch_fill_ghost_band<real>(mk(c, c->pl_gband), g);
    }
    if (c->cfg.npass > 0) {
        ProfScope ps(c, "fill_ghosts_pass", c->pl_pgband.active);
        GhostArgs<real> g{};
        for (int p = 0; p < c->cfg.npass; p++) g.r[p] = R<real>(c->qp[c->pcur][p]);
        g.nreal = c->cfg.npass;
        g.ndbl = 0;
        g.depth = 3;
        launch_fill_ghost_band<real>(mk(c, c->pl_pgband), g);
    }
}

template <class real>
void copy_velocity(imhd_ctx* c, int stage) {
    ProfScope ps(c, stage ? "copy_velocity_tavg" : "copy_velocity", c->pl_vel.active);
    VelArgs<real> a{};
    for (int f = 0; f < 4; f++) a.q[f] = R<real>(c->q[c->cur][f]);
    for (int d = 0; d < 3; d++) { a.vel[d] = R<real>(c->vel[d]); a.vta[d] = R<real>(c->vta[d]); }
    a.stage = stage;
    launch_copy_velocity<real>(mk(c, c->pl_vel), a);
}

template <class real>
void passive_step(imhd_ctx* c) {
    const int np = c->cfg.npass;
    const int p0 = c->pcur, p1 = 1 - c->pcur;
    for (int stage = 0; stage < 2; stage++) {
        const int src = stage ? p1 : p0;
        void** v = stage ? c->vta : c->vel;
        for (int d = 0; d < 3; d++) {
            ProfScope ps(c, stage ? "pass_flux_corr" : "pass_flux_pred",
                         (stage ? c->pl_pcflux[d] : c->pl_ppflux[d]).active);
            PassFluxArgs<real> a{};
            for (int p = 0; p < np; p++) { a.s[p] = R<real>(c->qp[src][p]); a.F[p] = R<real>(c->Fp[d][p]); }
            a.v = R<real>(v[d]);
            a.npass = np;
            a.dir = d;
            a.stage = stage;
            launch_pass_flux<real>(mk(c, stage ? c->pl_pcflux[d] : c->pl_ppflux[d]), a);
        }
        ProfScope ps(c, stage ? "pass_update_corr" : "pass_update_pred", (stage ? c->pl_pfupd : c->pl_ppupd).active);
        PassUpdArgs<real> u{};
        for (int p = 0; p < np; p++) {
            u.src[p] = R<real>(c->qp[p0][p]);
            u.dst[p] = R<real>(c->qp[p1][p]);
            for (int d = 0; d < 3; d++) u.F[d][p] = R<real>(c->Fp[d][p]);
        }
        for (int d = 0; d < 3; d++) u.v[d] = R<real>(v[d]);
        u.dt = c->dt;
        u.factor = stage ? 1.0 : 0.5;
        u.npass = np;
        launch_pass_update<real>(mk(c, stage ? c->pl_pfupd : c->pl_ppupd), u);
    }
    c->pcur = p1;
}

template <class real>
void max_speed_launch(imhd_ctx* c) {
    if (is_tvd(c)) { tvd_max_speed_launch<real>(c); return; }
    ProfScope ps(c, "max_speed", c->pl_int.active);
    launch_reset_u64(c->st, c->cmax, (int)stripe_slots(c->nbox));
    QPtrs<real> q{};
    for (int f = 0; f < 7; f++) q.q[f] = R<real>(c->q[c->cur][f]);
    launch_max_speed<real>(mk(c, c->pl_int), q, c->cfg.cs, c->cmax);
}

// passive_solve_3d with the velocities taken from the MHD states before (qold) and after
// (cur) the step: same arithmetic as copy_velocity + passive_step, no velocity arrays.
// Stage 0 writes the half step into the spare buffer; stage 1 reconstructs from it and
// updates the t^n buffer in place (a cell reads only its own t^n value), so the current
// passive buffer does not change.
// One passive stage: velocities from the MHD states qold (t^n) and qnew (t^{n+1}), or, with
// vel, read from the per-cell velocity set the fused MHD stage has just written.
template <class real>
void passive_stage(imhd_ctx* c, int qold, int qnew, int stage, bool vel) {
    const int np = c->cfg.npass;
    const int p0 = c->pcur, p1 = 1 - c->pcur;
    ProfScope ps(c, stage ? "pass_stage_corr" : "pass_stage_pred", c->pl_pstage[stage].active);
    PassStageArgs<real> a{};
    for (int p = 0; p < np; p++) {
        a.src[p] = R<real>(c->qp[p0][p]);
        a.rec[p] = R<real>(c->qp[stage ? p1 : p0][p]);
        a.dst[p] = R<real>(c->qp[stage ? p0 : p1][p]);
    }
    for (int f = 0; f < 4; f++) {
        a.qo[f] = R<real>(c->q[qold][f]);
        a.qn[f] = R<real>(c->q[qnew][f]);
    }
    for (int d = 0; d < 3; d++) a.vel[d] = vel ? R<real>(c->pvel[d]) : nullptr;
    a.dt = c->dt;
    a.factor = stage ? 1.0 : 0.5;
    a.npass = np;
    a.stage = stage;
    launch_pass_stage<real>(mk(c, c->pl_pstage[stage]), a);
}

template <class real>
void passive_fused(imhd_ctx* c, int qold) {
    for (int stage = 0; stage < 2; stage++) passive_stage<real>(c, qold, c->cur, stage, false);
}

template <class real>
void native_step(imhd_ctx* c, double t_end) {
    const int qold = c->cur;
    if (is_tvd(c)) {  // driver.f90:35-160 order: solve, passive, new dt, then ghost fill
        tvd_step<real>(c);
        if (c->ou_on) ou_step<real>(c, false);
        if (c->cfg.npass > 0) passive_fused<real>(c, qold);
        tvd_max_speed_launch<real>(c);
        fill_ghosts<real>(c);
        launch_new_dt(c->st, c->nbox, c->db, c->cmax, c->smax, c->dvmax, c->cour, t_end, c->t, c->dt, c->dtprev, 1);
        return;
    }
    launch_reset_u64(c->st, c->cmax, (int)stripe_slots(c->nbox));
    const bool inl = c->pvel[0] != nullptr && !c->ou_on;
    vl_step<real>(c, !c->ou_on, inl);
    if (c->ou_on) ou_step<real>(c, true);   // the CFL speed is taken after the kick
    if (c->cfg.npass > 0 && !inl) passive_fused<real>(c, qold);
    fill_ghosts<real>(c);
    launch_new_dt(c->st, c->nbox, c->db, c->cmax, c->smax, c->dvmax, c->cour, t_end, c->t, c->dt, c->dtprev, 1);
}

// A native VL step with the reference's impulsive kick (vanleer_solve_3d with drive_next:
// get_rho_turb before the step, add_random_force after it), then the passive step, the
// ghost fill and the next dt with COUR halved (driver.f90:102-111).  The caller has halved
// dt (driver.f90:68).  The CFL speed is taken from the kicked state, with delv_max.
template <class real>
void native_step_drive(imhd_ctx* c, double t_end) {
    const int s0 = c->cur;
    force_sums<real>(c, s0);
    vl_step<real>(c, false, false);
    force_apply<real>(c, s0, c->cur);
    max_speed_launch<real>(c);
    if (c->cfg.npass > 0) passive_fused<real>(c, s0);
    fill_ghosts<real>(c);
    launch_new_dt(c->st, c->nbox, c->db, c->cmax, c->smax, c->dvmax, c->cour_half, t_end, c->t, c->dt, c->dtprev, 1);
}

__global__ void k_set_dt(double* dt, int n, double v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dt[i] = v;
}

void set_scalar_dt(imhd_ctx* c, double dt) {
    hipLaunchKernelGGL(k_set_dt, dim3((c->nbox + 63) / 64), dim3(64), 0, c->st, c->dt, c->nbox, dt);
}

int prof_collect(imhd_ctx* c) {
    if (c->prec_list.empty()) return 0;
    HIPCK(c, hipStreamSynchronize(c->st));
    std::map<std::string, std::pair<double, int>> acc;
    std::map<std::string, long long> items;
    std::vector<std::string> order;
    for (auto& r : c->prec_list) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, r.a, r.b);
        if (!acc.count(r.name)) order.push_back(r.name);
        acc[r.name].first += ms;
        acc[r.name].second += 1;
        items[r.name] += r.items;
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    c->prec_list.clear();
    std::string s;
    for (auto& n : order) {
        char b[160];
        snprintf(b, sizeof(b), "%s=%.6f/%d/%lld;", n.c_str(), acc[n].first, acc[n].second, items[n]);
        s += b;
    }
    c->prof_report = s;
    return 0;
}

}  // namespace

namespace {

static int ensure_stage(imhd_ctx* c, size_t n) {
    if (c->stage_n >= n) return 0;
    if (c->stage) HIPCK(c, hipFree(c->stage));
    c->stage = nullptr;
    HIPCK(c, hipMalloc(&c->stage, n * 8));
    c->stage_n = n;
    return 0;
}

// Host FP64 <-> device arrays through a bounded staging buffer (STAGE_CHUNK doubles), so a
// 1024^3 box needs no whole-field FP64 copy on the device.  FP64 contexts copy directly.
constexpr long long STAGE_CHUNK = 1LL << 25;   // 256 MB

template <class real>
static int upload_field(imhd_ctx* c, const double* h, real* d0, real* d1, long long n) {
    if (sizeof(real) == 8) {
        HIPCK(c, hipMemcpy(d0, h, (size_t)n * 8, hipMemcpyHostToDevice));
        if (d1) HIPCK(c, hipMemcpy(d1, d0, (size_t)n * 8, hipMemcpyDeviceToDevice));
        return 0;
    }
    int rc;
    if ((rc = ensure_stage(c, (size_t)std::min(n, STAGE_CHUNK)))) return rc;
    for (long long o = 0; o < n; o += STAGE_CHUNK) {
        const long long m = std::min(STAGE_CHUNK, n - o);
        HIPCK(c, hipMemcpy(c->stage, h + o, (size_t)m * 8, hipMemcpyHostToDevice));
        launch_convert_in<real>(c->st, c->stage, d0 + o, m);
        if (d1) launch_convert_in<real>(c->st, c->stage, d1 + o, m);
        HIPCK(c, hipStreamSynchronize(c->st));
    }
    return 0;
}

template <class real>
static int download_field(imhd_ctx* c, const real* d, double* h, long long n) {
    if (sizeof(real) == 8) {
        HIPCK(c, hipMemcpy(h, d, (size_t)n * 8, hipMemcpyDeviceToHost));
        return 0;
    }
    int rc;
    if ((rc = ensure_stage(c, (size_t)std::min(n, STAGE_CHUNK)))) return rc;
    for (long long o = 0; o < n; o += STAGE_CHUNK) {
        const long long m = std::min(STAGE_CHUNK, n - o);
        launch_convert_out<real>(c->st, d + o, c->stage, m);
        HIPCK(c, hipStreamSynchronize(c->st));
        HIPCK(c, hipMemcpy(h + o, c->stage, (size_t)m * 8, hipMemcpyDeviceToHost));
    }
    return 0;
}

template <class real>
static int set_state_t(imhd_ctx* c, int bx, const double* q, const double* bf, const double* qp) {
    const BoxDesc& B = c->hb[bx];
    const long long nc = (long long)B.P[0] * B.P[1] * B.P[2];
    int rc;
    HIPCK(c, hipStreamSynchronize(c->st));
    for (int f = 0; f < 7; f++)
        if ((rc = upload_field<real>(c, q + f * nc, R<real>(c->q[0][f]) + B.off, R<real>(c->q[1][f]) + B.off, nc)))
            return rc;
    for (int s = 0; s < 2; s++)
        for (int d = 0; d < 3; d++)
            HIPCK(c, hipMemcpy(c->b[s][d] + B.off, bf + d * nc, (size_t)nc * 8, hipMemcpyHostToDevice));
    if (c->cfg.npass > 0 && qp) {
        const long long pn = (long long)(B.n[0] + 2 * B.nb - 4) * (B.n[1] + 2 * B.nb - 4) * (B.n[2] + 2 * B.nb - 4);
        if ((rc = ensure_stage(c, (size_t)std::min(pn, STAGE_CHUNK)))) return rc;
        for (int p = 0; p < c->cfg.npass; p++)
            for (long long o = 0; o < pn; o += STAGE_CHUNK) {
                const long long m = std::min(STAGE_CHUNK, pn - o);
                HIPCK(c, hipMemcpy(c->stage, qp + p * pn + o, (size_t)m * 8, hipMemcpyHostToDevice));
                for (int s = 0; s < 2; s++)
                    launch_pass_layout<real>(c->st, B, c->stage, nullptr, R<real>(c->qp[s][p]), 1, o, m);
                HIPCK(c, hipStreamSynchronize(c->st));
            }
    }
    HIPCK(c, hipStreamSynchronize(c->st));
    HIPCK(c, hipGetLastError());
    return 0;
}

template <class real>
static int get_state_t(imhd_ctx* c, int bx, double* q, double* bf, double* qp) {
    const BoxDesc& B = c->hb[bx];
    const long long nc = (long long)B.P[0] * B.P[1] * B.P[2];
    int rc;
    HIPCK(c, hipStreamSynchronize(c->st));
    if (q)
        for (int f = 0; f < 7; f++)
            if ((rc = download_field<real>(c, R<real>(c->q[c->cur][f]) + B.off, q + f * nc, nc))) return rc;
    if (bf)
        for (int d = 0; d < 3; d++)
            HIPCK(c, hipMemcpy(bf + d * nc, c->b[c->cur][d] + B.off, (size_t)nc * 8, hipMemcpyDeviceToHost));
    if (c->cfg.npass > 0 && qp) {
        const long long pn = (long long)(B.n[0] + 2 * B.nb - 4) * (B.n[1] + 2 * B.nb - 4) * (B.n[2] + 2 * B.nb - 4);
        if ((rc = ensure_stage(c, (size_t)std::min(pn, STAGE_CHUNK)))) return rc;
        for (int p = 0; p < c->cfg.npass; p++)
            for (long long o = 0; o < pn; o += STAGE_CHUNK) {
                const long long m = std::min(STAGE_CHUNK, pn - o);
                launch_pass_layout<real>(c->st, B, nullptr, c->stage, R<real>(c->qp[c->pcur][p]), 0, o, m);
                HIPCK(c, hipStreamSynchronize(c->st));
                HIPCK(c, hipMemcpy(qp + p * pn + o, c->stage, (size_t)m * 8, hipMemcpyDeviceToHost));
            }
    }
    HIPCK(c, hipGetLastError());
    return 0;
}

// Work arrays of the drop-in passive path (copy_velocity + passive_solve_3d), allocated on
// its first use: the native loop's fused passive stages never need them.
static int ensure_passive_work(imhd_ctx* c) {
    if (c->cfg.npass <= 0 || c->vel[0]) return 0;
    const size_t nr = (size_t)c->pool * c->esz;
    int rc;
    for (int d = 0; d < 3; d++) {
        if ((rc = alloc(c, &c->vel[d], nr))) return rc;
        if ((rc = alloc(c, &c->vta[d], nr))) return rc;
        for (int p = 0; p < c->cfg.npass; p++) if ((rc = alloc(c, &c->Fp[d][p], nr))) return rc;
    }
    return 0;
}

// ---------------------------------------------------------------------------------
// Restart and snapshot files (io_mod.f90).  Fortran sequential-unformatted records as
// gfortran frames them: a 4-byte length before and after the data; a record longer than
// MAX_SUBRECORD bytes is split into subrecords whose leading marker is negated when another
// subrecord follows and whose trailing marker is negated when one precedes.

constexpr long long MAX_SUBRECORD = 2147483639LL;  // gfortran's default and maximum (2^31 - 9)
constexpr long long IO_CHUNK = 1LL << 22;            // doubles per staged transfer (32 MB)

// IMHD_MAX_SUBRECORD (bytes, 1 .. MAX_SUBRECORD) lowers the split point so that tests can
// exercise subrecords on small boxes.
static long long max_subrecord() {
    const char* e = getenv("IMHD_MAX_SUBRECORD");
    const long long v = e ? atoll(e) : 0;
    return (v >= 1 && v <= MAX_SUBRECORD) ? v : MAX_SUBRECORD;
}

struct SeqWriter {
    FILE* f;
    long long rec_left = 0, sub_left = 0, sub_n = 0, max_sub = max_subrecord();
    bool first = true, ok = true;
    explicit SeqWriter(FILE* fp) : f(fp) {}
    void marker(long long m) {
        const int32_t v = (int32_t)m;
        ok = ok && fwrite(&v, 4, 1, f) == 1;
    }
    void open_sub() {
        sub_n = std::min(max_sub, rec_left);
        sub_left = sub_n;
        marker(sub_n < rec_left ? -sub_n : sub_n);
    }
    void close_sub() {
        marker(first ? sub_n : -sub_n);
        first = false;
    }
    void begin(long long len) {
        rec_left = len;
        first = true;
        open_sub();
    }
    void put(const void* p, long long n) {
        const char* c = (const char*)p;
        while (n > 0 && ok) {
            if (sub_left == 0) {
                close_sub();
                open_sub();
            }
            const long long m = std::min(n, sub_left);
            ok = ok && fwrite(c, 1, (size_t)m, f) == (size_t)m;
            c += m;
            n -= m;
            sub_left -= m;
            rec_left -= m;
        }
    }
    void end() {
        ok = ok && rec_left == 0 && sub_left == 0;
        close_sub();
    }
};

struct SeqReader {
    FILE* f;
    long long rec_left = 0, sub_left = 0, sub_n = 0;
    bool chain = false, more = false, first = true;
    int32_t lead = 0;
    std::string err;
    explicit SeqReader(FILE* fp) : f(fp) {}
    bool fail(const char* m) {
        if (err.empty()) err = m;
        return false;
    }
    bool rd(void* p, long long n) { return fread(p, 1, (size_t)n, f) == (size_t)n; }
    // Opens a record that must hold len bytes: gfortran subrecords, a plain record, or one
    // record whose markers hold len modulo 2^32.
    bool begin(long long len) {
        const long long pos = (long long)ftello(f);
        if (!rd(&lead, 4)) return fail("truncated file (record marker)");
        rec_left = len;
        first = true;
        const bool wrapped = len > INT32_MAX && (uint32_t)lead == (uint32_t)(len & 0xffffffffLL);
        if (lead >= 0) {
            if ((long long)lead != len && !wrapped) return fail("record length does not match the context's sizes");
            chain = more = false;
            sub_n = sub_left = len;
            return true;
        }
        const long long n = -(long long)lead;
        int32_t t = 0;
        if (fseeko(f, (off_t)(pos + 4 + n), SEEK_SET) == 0 && rd(&t, 4) && t == n) {
            if (fseeko(f, (off_t)(pos + 4), SEEK_SET) != 0) return fail("seek failed");
            chain = more = true;
            sub_n = sub_left = n;
            return true;
        }
        if (fseeko(f, (off_t)(pos + 4), SEEK_SET) != 0) return fail("seek failed");
        if (!wrapped) return fail("record markers do not match the context's sizes");
        chain = more = false;
        sub_n = sub_left = len;
        return true;
    }
    bool next_sub() {
        if (!chain || !more) return fail("record shorter than the context's sizes");
        int32_t t = 0, l = 0;
        if (!rd(&t, 4) || t != (int32_t)(first ? sub_n : -sub_n)) return fail("bad subrecord marker");
        first = false;
        if (!rd(&l, 4)) return fail("truncated file (subrecord marker)");
        more = l < 0;
        sub_n = sub_left = l < 0 ? -(long long)l : (long long)l;
        return true;
    }
    // dst == nullptr: skip (the validation pass)
    bool get(void* dst, long long n) {
        char* c = (char*)dst;
        while (n > 0) {
            if (sub_left == 0 && !next_sub()) return false;
            const long long m = std::min(n, sub_left);
            if (c) {
                if (!rd(c, m)) return fail("truncated file (record data)");
                c += m;
            } else if (fseeko(f, (off_t)m, SEEK_CUR) != 0) {
                return fail("seek failed");
            }
            n -= m;
            sub_left -= m;
            rec_left -= m;
        }
        return true;
    }
    bool end() {
        if (rec_left != 0 || sub_left != 0 || (chain && more)) return fail("record longer than the context's sizes");
        int32_t t = 0;
        if (!rd(&t, 4)) return fail("truncated file (trailing marker)");
        if (t != (chain ? (int32_t)(first ? sub_n : -sub_n) : lead)) return fail("bad trailing record marker");
        return true;
    }
    bool record(void* dst, long long len) { return begin(len) && get(dst, len) && end(); }
};

struct FileCloser {
    FILE* f;
    ~FileCloser() {
        if (f) fclose(f);
    }
};

static long long pass_cells(const BoxDesc& B) {
    return (long long)(B.n[0] + 2 * B.nb - 4) * (B.n[1] + 2 * B.nb - 4) * (B.n[2] + 2 * B.nb - 4);
}

template <class real>
static int write_restart_t(imhd_ctx* c, int bx, FILE* f, double t_sim, int dump, const int* seeds) {
    const BoxDesc& B = c->hb[bx];
    const long long nc = (long long)B.P[0] * B.P[1] * B.P[2];
    const int np = c->cfg.npass;
    std::vector<double> buf((size_t)IO_CHUNK);
    SeqWriter w(f);
    int rc;
    HIPCK(c, hipStreamSynchronize(c->st));
    w.begin(8);
    w.put(&t_sim, 8);
    w.end();
    w.begin(7 * nc * 8);
    for (int fl = 0; fl < 7 && w.ok; fl++)
        for (long long o = 0; o < nc && w.ok; o += IO_CHUNK) {
            const long long m = std::min(IO_CHUNK, nc - o);
            if ((rc = download_field<real>(c, R<real>(c->q[c->cur][fl]) + B.off + o, buf.data(), m))) return rc;
            w.put(buf.data(), m * 8);
        }
    w.end();
    w.begin(3 * nc * 8);
    for (int d = 0; d < 3 && w.ok; d++)
        for (long long o = 0; o < nc && w.ok; o += IO_CHUNK) {
            const long long m = std::min(IO_CHUNK, nc - o);
            HIPCK(c, hipMemcpy(buf.data(), c->b[c->cur][d] + B.off + o, (size_t)m * 8, hipMemcpyDeviceToHost));
            w.put(buf.data(), m * 8);
        }
    w.end();
    if (np > 0) {
        const long long pn = pass_cells(B);
        if ((rc = ensure_stage(c, (size_t)std::min(pn, IO_CHUNK)))) return rc;
        w.begin(np * pn * 8);
        for (int p = 0; p < np && w.ok; p++)
            for (long long o = 0; o < pn && w.ok; o += IO_CHUNK) {
                const long long m = std::min(IO_CHUNK, pn - o);
                launch_pass_layout<real>(c->st, B, nullptr, c->stage, R<real>(c->qp[c->pcur][p]), 0, o, m);
                HIPCK(c, hipStreamSynchronize(c->st));
                HIPCK(c, hipMemcpy(buf.data(), c->stage, (size_t)m * 8, hipMemcpyDeviceToHost));
                w.put(buf.data(), m * 8);
            }
        w.end();
    }
    const int32_t ints[4] = {dump, seeds[0], seeds[1], seeds[2]};
    for (int i = 0; i < 4; i++) {
        w.begin(4);
        w.put(&ints[i], 4);
        w.end();
    }
    if (!w.ok) FAIL(c, IMHD_EIO, "write_restart: write failed");
    HIPCK(c, hipGetLastError());
    return 0;
}

template <class real>
static int read_restart_t(imhd_ctx* c, int bx, FILE* f, double* t_sim, int* dump, int* seeds) {
    const BoxDesc& B = c->hb[bx];
    const long long nc = (long long)B.P[0] * B.P[1] * B.P[2];
    const int np = c->cfg.npass;
    const long long pn = pass_cells(B);
    std::vector<long long> len = {8, 7 * nc * 8, 3 * nc * 8};
    if (np > 0) len.push_back(np * pn * 8);
    for (int i = 0; i < 4; i++) len.push_back(4);
    // pass 1: every marker against the context's sizes, before any state changes
    {
        SeqReader r(f);
        for (long long L : len)
            if (!r.record(nullptr, L)) FAIL(c, IMHD_EIO, "read_restart: %s", r.err.c_str());
        char extra;
        if (fread(&extra, 1, 1, f) == 1) FAIL(c, IMHD_EIO, "read_restart: data after the last record");
        if (fseeko(f, 0, SEEK_SET) != 0) FAIL(c, IMHD_EIO, "read_restart: seek failed");
    }
    SeqReader r(f);
    std::vector<double> buf((size_t)IO_CHUNK);
    int rc;
    HIPCK(c, hipStreamSynchronize(c->st));
#define RD(call) \
    if (!(call)) FAIL(c, IMHD_EIO, "read_restart: %s", r.err.c_str())
    RD(r.record(t_sim, 8));
    RD(r.begin(7 * nc * 8));
    for (int fl = 0; fl < 7; fl++)
        for (long long o = 0; o < nc; o += IO_CHUNK) {
            const long long m = std::min(IO_CHUNK, nc - o);
            RD(r.get(buf.data(), m * 8));
            if ((rc = upload_field<real>(c, buf.data(), R<real>(c->q[0][fl]) + B.off + o,
                                         R<real>(c->q[1][fl]) + B.off + o, m)))
                return rc;
        }
    RD(r.end());
    RD(r.begin(3 * nc * 8));
    for (int d = 0; d < 3; d++)
        for (long long o = 0; o < nc; o += IO_CHUNK) {
            const long long m = std::min(IO_CHUNK, nc - o);
            RD(r.get(buf.data(), m * 8));
            for (int s = 0; s < 2; s++)
                HIPCK(c, hipMemcpy(c->b[s][d] + B.off + o, buf.data(), (size_t)m * 8, hipMemcpyHostToDevice));
        }
    RD(r.end());
    if (np > 0) {
        if ((rc = ensure_stage(c, (size_t)std::min(pn, IO_CHUNK)))) return rc;
        RD(r.begin(np * pn * 8));
        for (int p = 0; p < np; p++)
            for (long long o = 0; o < pn; o += IO_CHUNK) {
                const long long m = std::min(IO_CHUNK, pn - o);
                RD(r.get(buf.data(), m * 8));
                HIPCK(c, hipMemcpy(c->stage, buf.data(), (size_t)m * 8, hipMemcpyHostToDevice));
                for (int s = 0; s < 2; s++)
                    launch_pass_layout<real>(c->st, B, c->stage, nullptr, R<real>(c->qp[s][p]), 1, o, m);
                HIPCK(c, hipStreamSynchronize(c->st));
            }
        RD(r.end());
    }
    int32_t ints[4];
    for (int i = 0; i < 4; i++) RD(r.record(&ints[i], 4));
#undef RD
    *dump = ints[0];
    for (int i = 0; i < 3; i++) seeds[i] = ints[1 + i];
    HIPCK(c, hipStreamSynchronize(c->st));
    HIPCK(c, hipGetLastError());
    return 0;
}

template <class real, class out_t>
static int write_cube_t(imhd_ctx* c, int bx, FILE* f, int npo) {
    const BoxDesc& B = c->hb[bx];
    const int V = 7 + npo;
    const long long plane = (long long)B.n[0] * B.n[1];
    // planes per chunk: the packed chunk fits the staging buffer (IO_CHUNK doubles)
    const long long per = std::max(1LL, (IO_CHUNK * 8) / (plane * V * (long long)sizeof(out_t)));
    const long long nk = std::min(per, (long long)B.n[2]);
    const long long chunk_bytes = nk * plane * V * (long long)sizeof(out_t);
    int rc;
    if ((rc = ensure_stage(c, (size_t)((chunk_bytes + 7) / 8)))) return rc;
    std::vector<char> buf((size_t)chunk_bytes);
    CubeArgs<real> a{};
    for (int v = 0; v < 7; v++) a.q[v] = R<real>(c->q[c->cur][v]);
    for (int p = 0; p < npo; p++) a.qp[p] = R<real>(c->qp[c->pcur][p]);
    a.npo = npo;
    HIPCK(c, hipStreamSynchronize(c->st));
    for (int k0 = 0; k0 < B.n[2]; k0 += (int)nk) {
        a.k0 = k0;
        a.nk = (int)std::min(nk, (long long)B.n[2] - k0);
        const size_t bytes = (size_t)(a.nk * plane * V) * sizeof(out_t);
        launch_pack_cube<real, out_t>(c->st, B, a, (out_t*)c->stage);
        HIPCK(c, hipGetLastError());
        HIPCK(c, hipStreamSynchronize(c->st));
        HIPCK(c, hipMemcpy(buf.data(), c->stage, bytes, hipMemcpyDeviceToHost));
        if (fwrite(buf.data(), 1, bytes, f) != bytes) FAIL(c, IMHD_EIO, "write_cube: write failed");
    }
    return 0;
}

// Force-field buffers of the turbulence driver, allocated on first use.
static int ensure_force(imhd_ctx* c) {
    if (c->delv[0]) return 0;
    for (int d = 0; d < 3; d++) HIPCK(c, hipMalloc(&c->delv[d], (size_t)c->pool * 8));
    for (int d = 0; d < 3; d++) HIPCK(c, hipMemset(c->delv[d], 0, (size_t)c->pool * 8));
    HIPCK(c, hipMalloc(&c->force_part, (size_t)(c->pl_int.total / 256 + 1) * FORCE_NSUM * 8));
    HIPCK(c, hipMalloc(&c->force_par, (size_t)c->nbox * FORCE_NPAR * 8));
    HIPCK(c, hipMalloc(&c->force_err, 4));
    HIPCK(c, hipMemset(c->force_err, 0, 4));
    return 0;
}

// rand_taus (turb_driver_mod.f90:104-126): taus88 on the driver's signed 32-bit state.
static double taus88_next(int32_t* st) {
    uint32_t s1 = (uint32_t)st[0], s2 = (uint32_t)st[1], s3 = (uint32_t)st[2], b;
    b = ((s1 << 13) ^ s1) >> 19;
    s1 = ((s1 & 0xFFFFFFFEu) << 12) ^ b;
    b = ((s2 << 2) ^ s2) >> 25;
    s2 = ((s2 & 0xFFFFFFF8u) << 4) ^ b;
    b = ((s3 << 3) ^ s3) >> 11;
    s3 = ((s3 & 0xFFFFFFF0u) << 17) ^ b;
    st[0] = (int32_t)s1;
    st[1] = (int32_t)s2;
    st[2] = (int32_t)s3;
    return 0.5 + (double)(int32_t)(s1 ^ s2 ^ s3) * 2.3283064365e-10;
}

struct DevBuf {
    double* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

// ---------------------------------------------------------------------------------
// Seeded random-field start (imhd_init_turbulence).  Host: the modes of the half space
// inside the shell, their Philox draws, the exact rms scaling (Parseval) and the sum over
// n_x for every (n_y, n_z) pair; device: the sums over n_y (k_ic_ypass) and n_z, rho and
// the stores (k_ic_cells), then the ghost fill and a copy of the new state into the other
// state set, as imhd_set_state leaves both sets.

// (cos, sin) of 2 pi r / n for r = 0 .. n-1
static std::vector<double> twiddle_table(int n) {
    std::vector<double> t(2 * (size_t)n);
    for (int r = 0; r < n; r++) {
        const double a = 6.283185307179586 * (double)r / (double)n;
        t[2 * r] = cos(a);
        t[2 * r + 1] = sin(a);
    }
    return t;
}

template <class real>
static int init_turbulence_t(imhd_ctx* c, int bx, const imhd_turb_ic* ic) {
    const BoxDesc& B = c->hb[bx];
    const int nx = B.n[0], ny = B.n[1], nz = B.n[2];
    const int K = ic->kmax >= 1.0 ? (int)floor(ic->kmax) : 0;
    if (K > 0 && (2 * K >= nx || 2 * K >= ny || 2 * K >= nz))
        FAIL(c, IMHD_EINVAL, "init_turbulence: floor(kmax) = %d needs fewer than n/2 modes per axis (box %d x %d x %d)",
             K, nx, ny, nz);
    const unsigned slo = (unsigned)ic->seed, shi = (unsigned)(ic->seed >> 32);
    // modes of the half space inside the shell, in the documented order, and their draws
    struct Mode { int n[3]; double re[3], im[3]; };
    std::vector<Mode> modes;
    double S = 0.0;
    for (int kz = 0; kz <= K; kz++)
        for (int ky = -K; ky <= K; ky++)
            for (int kx = -K; kx <= K; kx++) {
                if (!(kz > 0 || (kz == 0 && ky > 0) || (kz == 0 && ky == 0 && kx > 0))) continue;
                const double n2 = (double)kx * kx + (double)ky * ky + (double)kz * kz, kn = sqrt(n2);
                if (kn < ic->kmin || kn > ic->kmax) continue;
                Mode md{{kx, ky, kz}, {}, {}};
                const double amp = pow(kn, -0.5 * (ic->slope + 2.0));
                const unsigned m = (unsigned)modes.size();
                for (int d = 0; d < 3; d++) {
                    unsigned o[4];
                    philox4x32(m, (unsigned)d, 1u, 0u, slo, shi, o);
                    md.re[d] = amp * box_muller(o[0], o[1]);
                    md.im[d] = amp * box_muller(o[2], o[3]);
                }
                if (ic->solenoidal) {   // normal to the physical wave vector k_d = n_d / n_cells,d
                    const double k[3] = {(double)kx / nx, (double)ky / ny, (double)kz / nz};
                    const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
                    const double dr = k[0] * md.re[0] + k[1] * md.re[1] + k[2] * md.re[2];
                    const double di = k[0] * md.im[0] + k[1] * md.im[1] + k[2] * md.im[2];
                    for (int d = 0; d < 3; d++) {
                        md.re[d] = md.re[d] - k[d] * dr / k2;
                        md.im[d] = md.im[d] - k[d] * di / k2;
                    }
                }
                for (int d = 0; d < 3; d++) S = S + 2.0 * (md.re[d] * md.re[d] + md.im[d] * md.im[d]);
                modes.push_back(md);
            }
    const bool vel = S > 0.0 && ic->mach != 0.0;
    const double scale = vel ? ic->mach * c->cfg.cs / sqrt(S) : 0.0;
    const int W = 2 * K + 1;
    const size_t nA = vel ? (size_t)3 * (K + 1) * W * nx * 2 : 0;
    const size_t nB = vel ? (size_t)3 * (K + 1) * ny * nx * 2 : 0;
    const size_t nty = vel ? 2 * (size_t)ny : 0, ntz = vel ? (size_t)2 * nz * (K + 1) : 0;
    std::vector<double> host(nA + nty + ntz, 0.0);
    if (vel) {
        const std::vector<double> tx = twiddle_table(nx), ty = twiddle_table(ny), tz = twiddle_table(nz);
        double* A = host.data();
        for (const Mode& md : modes) {
            const int iy = md.n[1] + K, kz = md.n[2];
            for (int d = 0; d < 3; d++) {
                const double vr = scale * md.re[d], vi = scale * md.im[d];
                double* a = A + (((size_t)d * (K + 1) + kz) * W + iy) * nx * 2;
                for (int i = 0; i < nx; i++) {
                    long long r = ((long long)md.n[0] * i) % nx;
                    if (r < 0) r += nx;
                    const double cr = tx[2 * r], sr = tx[2 * r + 1];
                    a[2 * i] = a[2 * i] + (vr * cr - vi * sr);
                    a[2 * i + 1] = a[2 * i + 1] + (vr * sr + vi * cr);
                }
            }
        }
        std::copy(ty.begin(), ty.end(), host.begin() + nA);
        double* twz = host.data() + nA + nty;
        for (int k = 0; k < nz; k++)
            for (int kz = 0; kz <= K; kz++) {
                const long long r = ((long long)kz * k) % nz;
                twz[((size_t)k * (K + 1) + kz) * 2] = tz[2 * r];
                twz[((size_t)k * (K + 1) + kz) * 2 + 1] = tz[2 * r + 1];
            }
    }
    double* dbuf = nullptr;
    if (vel) {
        HIPCK(c, hipMalloc(&dbuf, (host.size() + nB) * sizeof(double)));
        hipError_t e = hipMemcpy(dbuf, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            hipFree(dbuf);
            FAIL(c, IMHD_EHIP, "init_turbulence: upload failed: %s", hipGetErrorString(e));
        }
        launch_ic_ypass(c->st, nx, ny, K, dbuf, dbuf + nA, dbuf + host.size());
    }
    IcArgs<real> a{};
    for (int f = 0; f < 7; f++) a.q[f] = R<real>(c->q[c->cur][f]);
    for (int d = 0; d < 3; d++) a.b[d] = c->b[c->cur][d];
    for (int p = 0; p < c->cfg.npass; p++) a.qp[p] = R<real>(c->qp[c->pcur][p]);
    a.Bv = vel ? dbuf + host.size() : nullptr;
    a.twz = vel ? dbuf + nA + nty : nullptr;
    a.rho0 = ic->rho0;
    a.sigma = ic->sigma_ln;
    for (int d = 0; d < 3; d++) a.b0[d] = ic->b0[d];
    a.seed_lo = slo;
    a.seed_hi = shi;
    a.K = K;
    a.npass = c->cfg.npass;
    a.passive = ic->passive;
    launch_ic_cells<real>(c->st, B, a);
    hipError_t e = hipStreamSynchronize(c->st);
    if (dbuf) hipFree(dbuf);
    if (e != hipSuccess) FAIL(c, IMHD_EHIP, "init_turbulence: %s", hipGetErrorString(e));
    HIPCK(c, hipGetLastError());
    if (c->comp[0])
        for (int f = 0; f < 4; f++) HIPCK(c, hipMemsetAsync(c->comp[f], 0, (size_t)c->pool * c->esz, c->st));
    fill_ghosts<real>(c);
    const long long nc = (long long)B.P[0] * B.P[1] * B.P[2];
    for (int f = 0; f < 7; f++)
        HIPCK(c, hipMemcpyAsync(R<real>(c->q[1 - c->cur][f]) + B.off, R<real>(c->q[c->cur][f]) + B.off,
                                (size_t)nc * sizeof(real), hipMemcpyDeviceToDevice, c->st));
    for (int d = 0; d < 3; d++)
        HIPCK(c, hipMemcpyAsync(c->b[1 - c->cur][d] + B.off, c->b[c->cur][d] + B.off, (size_t)nc * 8,
                                hipMemcpyDeviceToDevice, c->st));
    for (int p = 0; p < c->cfg.npass; p++)
        HIPCK(c, hipMemcpyAsync(R<real>(c->qp[1 - c->pcur][p]) + B.off, R<real>(c->qp[c->pcur][p]) + B.off,
                                (size_t)nc * sizeof(real), hipMemcpyDeviceToDevice, c->st));
    HIPCK(c, hipStreamSynchronize(c->st));
    HIPCK(c, hipGetLastError());
    return 0;
}

template <class real>
static int get_field_t(imhd_ctx* c, int bx, int field, double* out) {
    const BoxDesc& B = c->hb[bx];
    const long long nc = (long long)B.P[0] * B.P[1] * B.P[2];
    HIPCK(c, hipStreamSynchronize(c->st));
    if (field < 7) return download_field<real>(c, R<real>(c->q[c->cur][field]) + B.off, out, nc);
    if (field < 10) {
        HIPCK(c, hipMemcpy(out, c->b[c->cur][field - 7] + B.off, (size_t)nc * 8, hipMemcpyDeviceToHost));
        return 0;
    }
    const int p = field - 10;
    const long long pn = (long long)(B.n[0] + 2 * B.nb - 4) * (B.n[1] + 2 * B.nb - 4) * (B.n[2] + 2 * B.nb - 4);
    int rc;
    if ((rc = ensure_stage(c, (size_t)std::min(pn, STAGE_CHUNK)))) return rc;
    for (long long o = 0; o < pn; o += STAGE_CHUNK) {
        const long long m = std::min(STAGE_CHUNK, pn - o);
        launch_pass_layout<real>(c->st, B, nullptr, c->stage, R<real>(c->qp[c->pcur][p]), 0, o, m);
        HIPCK(c, hipStreamSynchronize(c->st));
        HIPCK(c, hipMemcpy(out + o, c->stage, (size_t)m * 8, hipMemcpyDeviceToHost));
    }
    HIPCK(c, hipGetLastError());
    return 0;
}

// generate_rand_force (turb_driver_mod.f90:197-391) for one box, with the modes the
// reference never assigns (mixed-sign wave vectors) zero.  Host: the k_max^3 octant modes
// (six taus88 draws each: Box-Muller amplitude and phase per component, P(k) = k^6
// exp(-8k/k_pk), projection normal to khat, which is zero on the axis planes), their
// Hermitian mirrors, the per-axis twiddles and the small x pass.  Device: the y and z
// passes over the padded box (k_rand_force_y/z), into the force field.  Every COS/SIN
// pair of one argument is one sincos call, as the reference's compilers emit it (glibc's
// sincos can differ from separate sin and cos in the last bit).
static int generate_force(imhd_ctx* c, int bx, int K, int* seeds) {
    const BoxDesc& B = c->hb[bx];
    const double twopi = 8.0 * atan(1.0);   // TWOPI (constants_mod.f90)
    const int W = 2 * K + 1, W3 = W * W * W, S = 2 * K, J = 2 * K - 1;
    const int* n = B.n;
    const int* P = B.P;
    const double L[3] = {(double)n[0] * B.dx, (double)n[1] * B.dx, (double)n[2] * B.dx};
    const double k0[3] = {twopi / L[0], twopi / L[1], twopi / L[2]};
    const double k_pk = twopi / std::max(L[0], std::max(L[1], L[2]));
    const double const_sol = sqrt(1.0), const_comp = sqrt(2.0 * (1.0 - 1.0));   // pure solenoidal driving
    std::vector<double> re((size_t)3 * W3, 0.0), im((size_t)3 * W3, 0.0);
    auto mi = [&](int x, int y, int z) { return ((z + K) * W + (y + K)) * W + (x + K); };
    int32_t st[3] = {seeds[0], seeds[1], seeds[2]};
    for (int kz = 0; kz < K; kz++)
        for (int ky = 0; ky < K; ky++)
            for (int kx = 0; kx < K; kx++) {
                const int kv[3] = {kx, ky, kz};
                const double a0 = (double)kx * k0[0], a1 = (double)ky * k0[1], a2 = (double)kz * k0[2];
                const double k = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
                const double k2 = k * k;
                const double Pk = (k2 * (k2 * k2)) * exp(-8.0 * k / k_pk);
                double rn[6], amp[3], ph[3], dre[3], dim[3], kh[3] = {0.0, 0.0, 0.0};
                for (int r = 0; r < 6; r++) rn[r] = taus88_next(st);
                for (int d = 0; d < 3; d++) {
                    amp[d] = sqrt(-2.0 * Pk * log(rn[d]));
                    ph[d] = twopi * rn[3 + d];
                }
                for (int d = 0; d < 3; d++) {
                    double cs, sn;
                    sincos(ph[d], &sn, &cs);
                    dre[d] = amp[d] * cs;
                    dim[d] = amp[d] * sn;
                }
                if (kx >= 1 && ky >= 1 && kz >= 1)
                    for (int d = 0; d < 3; d++) kh[d] = k0[d] * (double)kv[d] / k;
                const double dot_re = kh[0] * dre[0] + kh[1] * dre[1] + kh[2] * dre[2];
                const double dot_im = kh[0] * dim[0] + kh[1] * dim[1] + kh[2] * dim[2];
                for (int d = 0; d < 3; d++) {
                    const double cre = kh[d] * dot_re, cim = kh[d] * dot_im;
                    const double sre = dre[d] - cre, sim = dim[d] - cim;
                    const int p = d * W3 + mi(kx, ky, kz), q = d * W3 + mi(-kx, -ky, -kz);
                    re[p] = const_sol * sre + const_comp * cre;
                    im[p] = const_sol * sim + const_comp * cim;
                    re[q] = re[p];
                    im[q] = -im[p];
                }
            }
    // twiddles per axis: theta = theta0 * i * j, i the global cell index 1-nb .. n+nb
    std::vector<double> tw[3];
    for (int a = 0; a < 3; a++) {
        tw[a].resize((size_t)2 * P[a] * J);
        const double theta0 = twopi / (double)n[a];
        for (int ii = 0; ii < P[a]; ii++)
            for (int j = -(K - 1); j <= K - 1; j++) {
                const double th = theta0 * (double)(ii + 1 - B.nb) * (double)j;
                sincos(th, &tw[a][2 * (ii * J + j + K - 1) + 1], &tw[a][2 * (ii * J + j + K - 1)]);
            }
    }
    // x pass on the host: X[d][kz slot][ky slot][ix]
    const size_t nX = (size_t)3 * S * S * P[0];
    std::vector<double> xr(nX), xi(nX);
    const double sg = -1.0;
    for (int d = 0; d < 3; d++)
        for (int kz = -K; kz < K; kz++)
            for (int ky = -K; ky < K; ky++)
                for (int ix = 0; ix < P[0]; ix++) {
                    double sr = 0.0, si = 0.0;
                    for (int j = -(K - 1); j <= K - 1; j++) {
                        const double ar = re[d * W3 + mi(j, ky, kz)], ai = im[d * W3 + mi(j, ky, kz)];
                        const double cs = tw[0][2 * (ix * J + j + K - 1)], sn = tw[0][2 * (ix * J + j + K - 1) + 1];
                        sr = sr + (ar * cs + sg * ai * sn);
                        si = si + (ai * cs - sg * ar * sn);
                    }
                    const size_t o = ((size_t)(d * S + kz + K) * S + ky + K) * P[0] + ix;
                    xr[o] = sr;
                    xi[o] = si;
                }
    int rc;
    if ((rc = ensure_force(c))) return rc;
    const size_t nY = (size_t)3 * S * P[1] * P[0];
    DevBuf dxr, dxi, dyr, dyi, dty, dtz;
    HIPCK(c, hipMalloc(&dxr.p, nX * 8));
    HIPCK(c, hipMalloc(&dxi.p, nX * 8));
    HIPCK(c, hipMalloc(&dyr.p, nY * 8));
    HIPCK(c, hipMalloc(&dyi.p, nY * 8));
    HIPCK(c, hipMalloc(&dty.p, tw[1].size() * 8));
    HIPCK(c, hipMalloc(&dtz.p, tw[2].size() * 8));
    HIPCK(c, hipStreamSynchronize(c->st));
    HIPCK(c, hipMemcpy(dxr.p, xr.data(), nX * 8, hipMemcpyHostToDevice));
    HIPCK(c, hipMemcpy(dxi.p, xi.data(), nX * 8, hipMemcpyHostToDevice));
    HIPCK(c, hipMemcpy(dty.p, tw[1].data(), tw[1].size() * 8, hipMemcpyHostToDevice));
    HIPCK(c, hipMemcpy(dtz.p, tw[2].data(), tw[2].size() * 8, hipMemcpyHostToDevice));
    RandForceArgs a{};
    a.xr = dxr.p;
    a.xi = dxi.p;
    a.twy = dty.p;
    a.twz = dtz.p;
    a.yr = dyr.p;
    a.yi = dyi.p;
    for (int d = 0; d < 3; d++) a.delv[d] = c->delv[d] + B.off;
    a.K = K;
    for (int d = 0; d < 3; d++) a.P[d] = P[d];
    a.idft_pre = 1.0 / (double)((long long)n[0] * n[1] * n[2]);
    launch_rand_force_y(c->st, a);
    launch_rand_force_z(c->st, a);
    HIPCK(c, hipGetLastError());
    HIPCK(c, hipStreamSynchronize(c->st));
    for (int i = 0; i < 3; i++) seeds[i] = st[i];
    return 0;
}

// ---------------------------------------------------------------------------------
// Power spectra and density PDF of the reference's analysis program (3D_MPI/Analysis/src/
// analysis_main.f90:543-698), in situ: REAL*4 field components packed from the device
// state, unnormalised 3-D R2C transforms (hipFFT), octant |F|^2 summed over components in
// REAL*4, binned on the device.

static void fft_release(imhd_ctx* c) {
    if (c->fft_have_plan) (void)hipfftDestroy(c->fft_plan);
    c->fft_have_plan = false;
    for (void* p : {c->fft_in, c->fft_out[0], c->fft_out[1], c->fft_out[2], c->fft_acc, c->fft_bins})
        if (p) (void)hipFree(p);
    c->fft_in = c->fft_acc = c->fft_bins = nullptr;
    c->fft_out[0] = c->fft_out[1] = c->fft_out[2] = nullptr;
    c->fft_n = 0;
    c->dev_bytes -= c->fft_bytes;
    c->fft_bytes = 0;
}

static int fft_malloc(imhd_ctx* c, void** p, size_t bytes) {
    HIPCK(c, hipMalloc(p, bytes));
    c->fft_bytes += bytes;
    c->dev_bytes += bytes;
    return IMHD_OK;
}

// The plan and buffers for cube edge n, with nout transform outputs.  All or nothing: on
// any failure everything made so far is released (fft_n stays 0), so the next call plans
// again instead of launching on missing buffers.  The bytes (plan work area included)
// count in imhd_device_bytes.
static int fft_setup(imhd_ctx* c, int n, int nout) {
    int rc = IMHD_OK;
    if (c->fft_n != n) {
        HIPCK(c, hipStreamSynchronize(c->st));
        fft_release(c);
        const int h = n / 2;
        const long long nc = (long long)n * n * n, no = (long long)h * h * h;
        size_t work = 0;
        if (hipfftPlan3d(&c->fft_plan, n, n, n, HIPFFT_R2C) != HIPFFT_SUCCESS) FAIL(c, IMHD_EHIP, "hipfftPlan3d failed");
        c->fft_have_plan = true;
        if (hipfftSetStream(c->fft_plan, c->st) != HIPFFT_SUCCESS || hipfftGetSize(c->fft_plan, &work) != HIPFFT_SUCCESS) {
            fft_release(c);
            FAIL(c, IMHD_EHIP, "hipfftSetStream / hipfftGetSize failed");
        }
        c->fft_bytes += work;
        c->dev_bytes += work;
        if ((rc = fft_malloc(c, &c->fft_in, (size_t)nc * 4)) || (rc = fft_malloc(c, &c->fft_acc, (size_t)no * 4)) ||
            (rc = fft_malloc(c, &c->fft_bins, (size_t)(8 * n + 3) * 8))) {
            fft_release(c);
            return rc;
        }
        if (getenv("IMHD_TEST_FFT_SETUP_FAIL")) {   // tests: a set-up that fails after allocating
            fft_release(c);
            FAIL(c, IMHD_EHIP, "spectrum set-up failed (IMHD_TEST_FFT_SETUP_FAIL)");
        }
        c->fft_n = n;
    }
    if (!c->fft_in || !c->fft_acc || !c->fft_bins) FAIL(c, IMHD_ESTATE, "spectrum buffers missing");
    const long long nk = (long long)n * n * (n / 2 + 1);
    for (int d = 0; d < nout; d++)
        if (!c->fft_out[d] && (rc = fft_malloc(c, &c->fft_out[d], (size_t)nk * 8))) {
            c->fft_out[d] = nullptr;   // the next call allocates it again
            return rc;
        }
    return IMHD_OK;
}

template <class real>
static int spectrum_t(imhd_ctx* c, int bx, int which, double* k_out, double* pk_out) {
    const BoxDesc& B = c->hb[bx];
    const int n = B.n[0], h = n / 2;
    int rc = fft_setup(c, n, 1);
    if (rc) return rc;
    struct Raw { double* p; };   // the cached buffers, in the form the launch calls take
    const Raw fin{(double*)c->fft_in}, fout{(double*)c->fft_out[0]}, acc{(double*)c->fft_acc},
        pk{(double*)c->fft_bins};
    const hipfftHandle plan = c->fft_plan;
    SpecPackArgs<real> a{};
    for (int f = 0; f < 7; f++) a.q[f] = R<real>(c->q[c->cur][f]);
    a.which = which;
    a.out = (float*)fin.p;
    const int ncomp = which == SPEC_DENSITY ? 1 : 3;
    for (int comp = 0; comp < ncomp; comp++) {
        a.comp = comp;
        launch_spec_pack<real>(c->st, B, a);
        if (hipfftExecR2C(plan, (hipfftReal*)fin.p, (hipfftComplex*)fout.p) != HIPFFT_SUCCESS)
            FAIL(c, IMHD_EHIP, "hipfftExecR2C failed");
        launch_spec_accum(c->st, (const float*)fout.p, n, (float*)acc.p, comp == 0);
    }
    // compute_pk's bin geometry in REAL*4 (analysis_main.f90:573-579, 585-593)
    const float k0 = (float)(8.0 * atan(1.0));
    const float kmax = (float)(sqrt(3.0) * (double)k0 * (double)h);
    const float dk = kmax / (float)n;
    launch_reset_u64(c->st, (unsigned long long*)pk.p, n);
    launch_spec_bin(c->st, (const float*)acc.p, n, k0, dk, pk.p);
    HIPCK(c, hipGetLastError());
    HIPCK(c, hipStreamSynchronize(c->st));
    HIPCK(c, hipMemcpy(pk_out, pk.p, (size_t)n * 8, hipMemcpyDeviceToHost));
    const double dx = 1.0 / (double)n;
    for (int b = 0; b < n; b++) {
        pk_out[b] *= dx * dx * dx;
        k_out[b] = (double)b * (double)dk;
    }
    return 0;
}

// mode_decomposition (analysis_main.f90:305-540): Alfven / fast / slow spectra of the v_perp
// transforms (the velocity transforms the program holds when it calls it, :238-249)
template <class real>
static int mode_spectra_t(imhd_ctx* c, int bx, double* k_out, double* P_out) {
    const BoxDesc& B = c->hb[bx];
    const int n = B.n[0], h = n / 2;
    int rc = fft_setup(c, n, 3);
    if (rc) return rc;
    struct Raw { double* p; };
    const Raw fin{(double*)c->fft_in},
        f[3] = {{(double*)c->fft_out[0]}, {(double*)c->fft_out[1]}, {(double*)c->fft_out[2]}},
        P{(double*)c->fft_bins}, b0{(double*)c->fft_bins + 8 * n};
    const hipfftHandle plan = c->fft_plan;
    SpecPackArgs<real> a{};
    for (int q = 0; q < 7; q++) a.q[q] = R<real>(c->q[c->cur][q]);
    a.which = SPEC_VPERP;
    a.out = (float*)fin.p;
    for (int d = 0; d < 3; d++) {
        a.comp = d;
        launch_spec_pack<real>(c->st, B, a);
        if (hipfftExecR2C(plan, (hipfftReal*)fin.p, (hipfftComplex*)f[d].p) != HIPFFT_SUCCESS)
            FAIL(c, IMHD_EHIP, "hipfftExecR2C failed");
    }
    launch_reset_u64(c->st, (unsigned long long*)b0.p, 3);
    launch_mean_b<real>(c->st, B, R<real>(c->q[c->cur][4]), R<real>(c->q[c->cur][5]), R<real>(c->q[c->cur][6]), b0.p);
    const float k0 = (float)(8.0 * atan(1.0));
    const float kmax = (float)(sqrt(3.0) * (double)k0 * (double)h);
    const float dk = kmax / (float)n;
    launch_reset_u64(c->st, (unsigned long long*)P.p, 8 * n);
    launch_mode_bin(c->st, (const float*)f[0].p, (const float*)f[1].p, (const float*)f[2].p, n, b0.p, k0, dk, P.p);
    HIPCK(c, hipGetLastError());
    HIPCK(c, hipStreamSynchronize(c->st));
    HIPCK(c, hipMemcpy(P_out, P.p, (size_t)8 * n * 8, hipMemcpyDeviceToHost));
    const double dx3 = 1.0 / ((double)n * n * n);
    for (int m = 0; m < 8 * n; m++) P_out[m] *= dx3;
    for (int b = 0; b < n; b++) k_out[b] = (double)b * (double)dk;
    return 0;
}

template <class real>
static int density_pdf_t(imhd_ctx* c, int bx, int nbins, double rho_max, double* rho_out, double* pdf_out) {
    const BoxDesc& B = c->hb[bx];
    DevBuf cnt;
    HIPCK(c, hipMalloc(&cnt.p, (size_t)nbins * 8));
    launch_reset_u64(c->st, (unsigned long long*)cnt.p, nbins);
    const float drho = (float)rho_max / (float)nbins;
    launch_density_hist<real>(c->st, B, R<real>(c->q[c->cur][0]), drho, nbins, (unsigned long long*)cnt.p);
    HIPCK(c, hipGetLastError());
    HIPCK(c, hipStreamSynchronize(c->st));
    std::vector<unsigned long long> h((size_t)nbins);
    HIPCK(c, hipMemcpy(h.data(), cnt.p, (size_t)nbins * 8, hipMemcpyDeviceToHost));
    unsigned long long tot = 0;
    for (auto v : h) tot += v;
    for (int b = 0; b < nbins; b++) {
        pdf_out[b] = tot ? (double)h[b] / ((double)drho * (double)tot) : 0.0;
        rho_out[b] = ((double)b + 0.5) * (double)drho;
    }
    return 0;
}

}  // namespace

extern "C" {

const char* imhd_version(void) { return "imhd 0.1 (gfx950)"; }

const char* imhd_last_error(const imhd_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int imhd_create(const imhd_config* cfg, imhd_ctx** out) { return imhd_create_ensemble(cfg, 1, out); }

int imhd_create_ensemble(const imhd_config* cfg, int nbox, imhd_ctx** out) {
    if (!cfg || !out || nbox < 1 || nbox > MAXBOX) return IMHD_EINVAL;
    *out = nullptr;
    imhd_ctx* c = new imhd_ctx();
    c->cfg = cfg[0];
    c->nbox = nbox;
    const imhd_config& g = cfg[0];
    auto bad = [&](const char* m) { c->err = m; return IMHD_EINVAL; };
    if (g.integrator != IMHD_VL && g.integrator != IMHD_TVD) { int r = bad("integrator must be 1 (TVD) or 2 (VL)"); *out = c; return r; }
    if (g.riemann < 1 || g.riemann > 3) { int r = bad("riemann must be 1 (ROE), 2 (HLLD) or 3 (HLLE)"); *out = c; return r; }
    if (g.integrator == IMHD_TVD && g.riemann == IMHD_HLLD) { int r = bad("the TVD integrator takes riemann 1 (K98 Roe-TVD) or 3 (HLLE)"); *out = c; return r; }
    if (g.recon < 0 || g.recon > 2) { int r = bad("recon must be 0 (PLM primitive), 1 (PPM primitive) or 2 (PPM characteristic)"); *out = c; return r; }
    if (g.precision != 32 && g.precision != 64) { int r = bad("precision must be 32 or 64"); *out = c; return r; }
    if (g.npass < 0 || g.npass > MAXPASS) { int r = bad("npass must be in 0..4"); *out = c; return r; }
    if (g.bc != 1 && g.bc != 2) { int r = bad("bc must be 1 (open) or 2 (periodic)"); *out = c; return r; }
    if (c->cfg.mc_beta == 0.0) c->cfg.mc_beta = 1.01;
    if (c->cfg.cour == 0.0) c->cfg.cour = 0.35;
    if (c->cfg.cs <= 0.0) { int r = bad("cs must be positive"); *out = c; return r; }
    c->prec = g.precision;
    c->esz = g.precision == 64 ? 8 : 4;
    if (const char* e = getenv("IMHD_ZSEG")) c->zseg = std::max(0, atoi(e));   // 0: per-cell kernels
    if (const char* e = getenv("IMHD_FUSE3")) c->fuse3 = atoi(e) != 0;      // 0: one launch per direction
    if (const char* e = getenv("IMHD_STRIP")) c->strip = std::max(0, atoi(e));
    if (const char* e = getenv("IMHD_VSTAGE")) c->vstage = atoi(e) != 0;    // 0: three kernels per stage
    if (const char* e = getenv("IMHD_TVDFUSE")) c->tvdfuse = atoi(e) != 0;  // 0: eig + flux kernels
    if (const char* e = getenv("IMHD_TVDHALF3")) c->tvdhalf3 = atoi(e) != 0;
    if (const char* e = getenv("IMHD_VROUNDS")) c->vrounds = std::max(0, atoi(e));
    if (const char* e = getenv("IMHD_VSEG")) { c->vseg = std::max(1, atoi(e)); c->vrounds = 0; }
    // FP32: 32 x 16 tiles (2 workgroups of 80 KB LDS per CU); FP64: 16 x 16 (a 32 x 16
    // FP64 tile would leave one workgroup per CU)
    c->vshape = g.precision == 32 ? 1 : 0;
    if (const char* e = getenv("IMHD_VSHAPE")) {
        const int v = std::min(VS_SHAPES - 1, std::max(0, atoi(e)));
        if (vs_shape_ok(g.precision, v)) c->vshape = v;
    }
    if (const char* e = getenv("IMHD_VSH")) c->vsh = atoi(e) != 0;
    c->vshape_c = c->vshape;
    if (const char* e = getenv("IMHD_VSHAPE_C")) {   // the corrector's alone
        const int v = std::min(VS_SHAPES - 1, std::max(0, atoi(e)));
        if (vs_shape_ok(g.precision, v)) c->vshape_c = v;
    }
    if (g.flux_protection) {
        // flux_protection of vanleer_mod.f90 (:81): Roe with HLLE protection fluxes (FP64 or
        // FP32; FP32 contexts run the three-kernel FP32 stages with FP64 corner EMFs)
        if (g.integrator != IMHD_VL || g.riemann != IMHD_ROE) {
            int r = bad("flux_protection needs the VL integrator and the Roe solver");
            *out = c;
            return r;
        }
        c->prot = true;
        c->vstage = false;   // the protected update works on the stored flux arrays
        // test knob: a higher density threshold marks more cells (the reference uses MIN_DENS)
        if (const char* e = getenv("IMHD_PROT_DENS")) c->prot_thr = atof(e);
    }
    long long run = 0;
    for (int bx = 0; bx < nbox; bx++) {
        const imhd_config& k = cfg[bx];
        if (k.nx < 4 || k.ny < 4 || k.nz < 4) { int r = bad("each box needs at least 4 cells per axis"); *out = c; return r; }
        if (k.nb != g.nb || k.nb < 6) { int r = bad("nb must be >= 6 and equal for all boxes"); *out = c; return r; }
        BoxDesc B{};
        B.n[0] = k.nx; B.n[1] = k.ny; B.n[2] = k.nz;
        B.nb = k.nb;
        for (int a = 0; a < 3; a++) B.P[a] = B.n[a] + 2 * k.nb;
        B.bc = g.bc;
        B.off = run;
        int nmax = std::max(k.nx, std::max(k.ny, k.nz));
        B.dx = 1.0 / (double)nmax;
        run += (long long)B.P[0] * B.P[1] * B.P[2];
        c->hb.push_back(B);
    }
    c->pool = run;
    *out = c;
    HIPCK(c, hipStreamCreate(&c->st));
    {
        int dev = 0;
        HIPCK(c, hipGetDevice(&dev));
        HIPCK(c, hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, dev));
        // set-up of the allocation-free diagnostics: this device's scratch address (a
        // failure only disables imhd_diagnostics_static)
        const hipError_t e = diag_static_scratch(&c->diag_static);
        if (e != hipSuccess) {
            c->diag_static = nullptr;
            c->diag_static_err = hipGetErrorString(e);
        }
    }
    HIPCK(c, hipMalloc(&c->db, sizeof(BoxDesc) * nbox));
    HIPCK(c, hipMemcpy(c->db, c->hb.data(), sizeof(BoxDesc) * nbox, hipMemcpyHostToDevice));
    const size_t nr = (size_t)run * c->esz, nd = (size_t)run * 8;
    int rc;
    for (int s = 0; s < 2; s++) {
        for (int f = 0; f < 7; f++) if ((rc = alloc(c, &c->q[s][f], nr))) return rc;
        for (int d = 0; d < 3; d++) if ((rc = alloc(c, (void**)&c->b[s][d], nd))) return rc;
        for (int p = 0; p < g.npass; p++) if ((rc = alloc(c, &c->qp[s][p], nr))) return rc;
    }
    // Both VL stages fused: a half-step set, and no flux / face-EMF / corner-EMF work
    // arrays (18 real + 3 FP64 per cell, the largest share of a 1024^3 box).  Otherwise
    // (TVD, IMHD_VSTAGE=0, a PPM corrector with IMHD_VPPM=0, flux protection) the
    // three-kernel stages need the work arrays.
    bool vppm = true;
    if (const char* e = getenv("IMHD_VPPM")) vppm = atoi(e) != 0;   // 0: PPM corrector on three kernels
    const bool fused = g.integrator == IMHD_VL && c->vstage && (g.recon == IMHD_PLM_PRIM || vppm);
    if (fused) {
        for (int f = 0; f < 7; f++) if ((rc = alloc(c, &c->q[2][f], nr))) return rc;
        // the half-step face field: FP64 in FP64 contexts; FP32 in FP32 contexts, whose
        // fused corrector converts it to FP32 interface states and never reads it otherwise
        for (int d = 0; d < 3; d++)
            if ((rc = g.precision == 64 ? alloc(c, (void**)&c->b[2][d], nd) : alloc(c, &c->bh[d], nr))) return rc;
        if (g.npass > 0)
            for (int d = 0; d < 3; d++) if ((rc = alloc(c, &c->pvel[d], nr))) return rc;
    } else {
        if (g.integrator == IMHD_TVD && c->tvdhalf3)   // per-direction half states (k_tvd_half3)
            for (int s = 2; s < 4; s++)
                for (int f = 0; f < 7; f++) if ((rc = alloc(c, &c->q[s][f], nr))) return rc;
        for (int d = 0; d < 3; d++)
            for (int f = 0; f < 4; f++) if ((rc = alloc(c, &c->F[d][f], nr))) return rc;
        for (int e = 0; e < 6; e++) if ((rc = alloc(c, &c->E[e], nr))) return rc;
        for (int d = 0; d < 3; d++) if ((rc = alloc(c, (void**)&c->Ec[d], nd))) return rc;
    }
    if (g.precision == 32 && g.compensated)
        for (int f = 0; f < 4; f++) if ((rc = alloc(c, &c->comp[f], nr))) return rc;
    if (c->prot) {
        for (int d = 0; d < 3; d++)
            for (int f = 0; f < 4; f++) if ((rc = alloc(c, &c->Fpr[d][f], nd))) return rc;
        for (int e = 0; e < 6; e++) if ((rc = alloc(c, &c->Epr[e], nd))) return rc;
        for (int d = 0; d < 3; d++) if ((rc = alloc(c, (void**)&c->Ecpr[d], nd))) return rc;
        if ((rc = alloc(c, (void**)&c->pflag, (size_t)run))) return rc;
    }
    if (g.integrator == IMHD_TVD) {
        for (int d = 0; d < 3; d++) {
            for (int f = 0; f < 4; f++) if ((rc = alloc(c, &c->Fc[d][f], nr))) return rc;
            if ((rc = alloc(c, (void**)&c->Ecc[d], nd))) return rc;
        }
        if (g.riemann == IMHD_ROE && !c->tvdfuse)
            for (int w = 0; w < 6; w++) {
                if ((rc = alloc(c, &c->gt[w], nr))) return rc;
                if ((rc = alloc(c, &c->ck[w], nr))) return rc;
            }
    }
    if ((rc = alloc(c, (void**)&c->smax, 8 * stripe_slots(nbox)))) return rc;
    if ((rc = alloc(c, (void**)&c->cmax, 8 * stripe_slots(nbox)))) return rc;
    if ((rc = alloc(c, (void**)&c->dvmax, 8 * stripe_slots(nbox)))) return rc;
    if ((rc = alloc(c, (void**)&c->dt, 8 * nbox))) return rc;
    if ((rc = alloc(c, (void**)&c->t, 8 * nbox))) return rc;
    if ((rc = alloc(c, (void**)&c->dtprev, 8 * nbox))) return rc;
    if ((rc = alloc(c, (void**)&c->cour, 8 * nbox))) return rc;
    if ((rc = alloc(c, (void**)&c->diag, 8 * DIAG_STRIDE * nbox))) return rc;
    if ((rc = alloc(c, (void**)&c->floor_count, 4))) return rc;
    std::vector<double> v(nbox);
    for (int bx = 0; bx < nbox; bx++) v[bx] = cfg[bx].cour != 0.0 ? cfg[bx].cour : c->cfg.cour;
    HIPCK(c, hipMemcpy(c->cour, v.data(), 8 * nbox, hipMemcpyHostToDevice));
    std::fill(v.begin(), v.end(), 1.0e-15);  // driver.f90:249
    HIPCK(c, hipMemcpy(c->dt, v.data(), 8 * nbox, hipMemcpyHostToDevice));
    for (int d = 0; d < 3; d++) {
        if ((rc = build_plan(c, c->pl_pflux[d], 0, d))) return rc;
        if ((rc = build_plan(c, c->pl_cflux[d], 3, d))) return rc;
        if ((rc = build_plan(c, c->pl_ppflux[d], 10, d))) return rc;
        if ((rc = build_plan(c, c->pl_pcflux[d], 12, d))) return rc;
    }
    if ((rc = build_plan(c, c->pl_pflux3, 25, -1))) return rc;
    if ((rc = build_plan(c, c->pl_cflux3, 26, -1))) return rc;
    if ((rc = build_plan(c, c->pl_pcorner, 1, -1, c->zseg))) return rc;
    if ((rc = build_plan(c, c->pl_pupd, 2, -1, c->zseg))) return rc;
    if ((rc = build_plan(c, c->pl_ccorner, 4, -1, c->zseg))) return rc;
    if ((rc = build_plan(c, c->pl_cupd, 5, -1, c->zseg))) return rc;
    if ((rc = build_plan(c, c->pl_ghost, 6, -1))) return rc;
    if ((rc = build_plan(c, c->pl_gband, 14, -1))) return rc;
    if ((rc = build_plan(c, c->pl_pgband, 15, -1))) return rc;
    if ((rc = build_plan(c, c->pl_pghost, 7, -1))) return rc;
    if ((rc = build_plan(c, c->pl_int, 8, -1))) return rc;
    if ((rc = build_plan(c, c->pl_vel, 9, -1))) return rc;
    if ((rc = build_plan(c, c->pl_ppupd, 11, -1))) return rc;
    if ((rc = build_plan(c, c->pl_pfupd, 13, -1))) return rc;
    if ((rc = build_plan(c, c->pl_pstage[0], 11, -1, c->zseg))) return rc;
    if ((rc = build_plan(c, c->pl_pstage[1], 13, -1, c->zseg))) return rc;
    if (g.integrator == IMHD_VL) {
        if ((rc = build_tiles(c, c->tp_pred, false))) return rc;
        if ((rc = build_tiles(c, c->tp_corr, true))) return rc;
    }
    if (g.integrator == IMHD_TVD) {
        for (int d = 0; d < 3; d++) {
            if ((rc = build_plan(c, c->pl_teig_p[d], 16, d))) return rc;
            if ((rc = build_plan(c, c->pl_tflux_p[d], 17, d))) return rc;
            if ((rc = build_plan(c, c->pl_teig_c[d], 18, d))) return rc;
            if ((rc = build_plan(c, c->pl_tflux_c[d], 19, d))) return rc;
            if ((rc = build_plan(c, c->pl_tsweep_p[d], 27, d))) return rc;
            if ((rc = build_plan(c, c->pl_tsweep_c[d], 28, d))) return rc;
        }
        if ((rc = build_plan(c, c->pl_tcorner_p, 20, -1))) return rc;
        if ((rc = build_plan(c, c->pl_tcorner_c, 21, -1))) return rc;
        if ((rc = build_plan(c, c->pl_tcenter, 22, -1))) return rc;
        if ((rc = build_plan(c, c->pl_thalf, 23, -1))) return rc;
        if ((rc = build_plan(c, c->pl_tcfl, 24, -1))) return rc;
    }
    return IMHD_OK;
}

int imhd_destroy(imhd_ctx* c) {
    if (!c) return IMHD_EINVAL;
    if (c->st) (void)hipStreamSynchronize(c->st);
    for (auto& g : c->gexec) if (g) (void)hipGraphExecDestroy(g);
    auto fr = [](void* p) { if (p) (void)hipFree(p); };
    for (int s = 0; s < 2; s++) {
        for (auto p : c->q[s]) fr(p);
        for (auto p : c->b[s]) fr(p);
        for (auto p : c->qp[s]) fr(p);
    }
    for (auto p : c->q[2]) fr(p);
    for (auto p : c->q[3]) fr(p);
    for (auto p : c->pvel) fr(p);
    for (auto p : c->b[2]) fr(p);
    for (auto p : c->bh) fr(p);
    fr(c->tp_pred.d);
    fr(c->tp_corr.d);
    for (auto& a : c->F) for (auto p : a) fr(p);
    for (auto p : c->E) fr(p);
    for (auto p : c->Ec) fr(p);
    for (auto p : c->vel) fr(p);
    for (auto p : c->vta) fr(p);
    for (auto& a : c->Fp) for (auto p : a) fr(p);
    for (auto p : c->comp) fr(p);
    for (auto& a : c->Fc) for (auto p : a) fr(p);
    for (auto p : c->Ecc) fr(p);
    for (auto p : c->gt) fr(p);
    for (auto p : c->ck) fr(p);
    for (auto& a : c->Fpr) for (auto p : a) fr(p);
    for (auto p : c->Epr) fr(p);
    for (auto p : c->Ecpr) fr(p);
    fr(c->pflag);
    for (auto p : c->delv) fr(p);
    for (void* p : {(void*)c->ou_n, (void*)c->ou_are, (void*)c->ou_aim, (void*)c->ou_step, (void*)c->ou_tw,
                    (void*)c->ou_twoff, (void*)c->ou_part, (void*)c->ou_mean, (void*)c->ou_col,
                    (void*)c->ou_coloff})
        fr(p);
    for (void* p : {(void*)c->force_part, (void*)c->force_par, (void*)c->force_err, (void*)c->dvmax,
                    (void*)c->cour_half})
        fr(p);
    fft_release(c);
    for (void* p : {(void*)c->smax, (void*)c->cmax, (void*)c->diag_part, (void*)c->dt, (void*)c->t, (void*)c->dtprev, (void*)c->cour,
                    (void*)c->diag, (void*)c->floor_count, (void*)c->stage, (void*)c->db})
        fr(p);
    PlanDev* pls[] = {&c->pl_pflux[0], &c->pl_pflux[1], &c->pl_pflux[2], &c->pl_pcorner, &c->pl_pupd,
                      &c->pl_cflux[0], &c->pl_cflux[1], &c->pl_cflux[2], &c->pl_ccorner, &c->pl_cupd,
                      &c->pl_ghost, &c->pl_pghost, &c->pl_int, &c->pl_vel, &c->pl_ppflux[0], &c->pl_ppflux[1],
                      &c->pl_ppflux[2], &c->pl_ppupd, &c->pl_pcflux[0], &c->pl_pcflux[1], &c->pl_pcflux[2],
                      &c->pl_pfupd, &c->pl_gband, &c->pl_pgband, &c->pl_teig_p[0], &c->pl_teig_p[1],
                      &c->pl_teig_p[2], &c->pl_tflux_p[0], &c->pl_tflux_p[1], &c->pl_tflux_p[2],
                      &c->pl_teig_c[0], &c->pl_teig_c[1], &c->pl_teig_c[2], &c->pl_tflux_c[0],
                      &c->pl_tflux_c[1], &c->pl_tflux_c[2], &c->pl_tcorner_p, &c->pl_tcorner_c,
                      &c->pl_tcenter, &c->pl_thalf, &c->pl_tcfl, &c->pl_pflux3, &c->pl_cflux3,
                      &c->pl_pstage[0], &c->pl_pstage[1], &c->pl_tsweep_p[0], &c->pl_tsweep_p[1],
                      &c->pl_tsweep_p[2], &c->pl_tsweep_c[0], &c->pl_tsweep_c[1], &c->pl_tsweep_c[2]};
    for (auto p : pls) fr(p->d);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
    return IMHD_OK;
}

int imhd_set_state(imhd_ctx* c, int box, const double* q, const double* bf, const double* qp) {
    if (!c || !q || !bf || box < 0 || box >= c->nbox) return IMHD_EINVAL;
    if (c->cfg.npass > 0 && !qp) FAIL(c, IMHD_EINVAL, "qpass_3d is required when npass > 0");
    int rc = c->prec == 64 ? set_state_t<double>(c, box, q, bf, qp) : set_state_t<float>(c, box, q, bf, qp);
    if (rc) return rc;
    if (c->comp[0]) for (int f = 0; f < 4; f++) HIPCK(c, hipMemset(c->comp[f], 0, (size_t)c->pool * c->esz));
    c->have_state = true;
    return IMHD_OK;
}

int imhd_get_state(imhd_ctx* c, int box, double* q, double* bf, double* qp) {
    if (!c || box < 0 || box >= c->nbox) return IMHD_EINVAL;
    if (!c->have_state) FAIL(c, IMHD_ESTATE, "no state set");
    return c->prec == 64 ? get_state_t<double>(c, box, q, bf, qp) : get_state_t<float>(c, box, q, bf, qp);
}

int imhd_get_field(imhd_ctx* c, int box, int field, double* out) {
    if (!c || !out || box < 0 || box >= c->nbox || field < 0 || field >= 10 + c->cfg.npass) return IMHD_EINVAL;
    if (!c->have_state) FAIL(c, IMHD_ESTATE, "no state set");
    return c->prec == 64 ? get_field_t<double>(c, box, field, out) : get_field_t<float>(c, box, field, out);
}

int imhd_init_turbulence(imhd_ctx* c, int box, const imhd_turb_ic* ic) {
    if (!c || !ic || box < 0 || box >= c->nbox) return IMHD_EINVAL;
    if (!(ic->rho0 > 0.0) || !(ic->sigma_ln >= 0.0) || !(ic->kmax <= 1024.0) || !(ic->mach >= 0.0))
        FAIL(c, IMHD_EINVAL, "init_turbulence: needs rho0 > 0, sigma_ln >= 0, mach >= 0, kmax <= 1024");
    int rc = c->prec == 64 ? init_turbulence_t<double>(c, box, ic) : init_turbulence_t<float>(c, box, ic);
    if (rc) return rc;
    c->have_state = true;
    return IMHD_OK;
}

int imhd_step(imhd_ctx* c, double dt, int drive_next) {
    if (!c) return IMHD_EINVAL;
    if (!c->have_state) FAIL(c, IMHD_ESTATE, "no state set");
    if (!(dt > 0.0)) FAIL(c, IMHD_EINVAL, "dt must be positive");
    if (drive_next && is_tvd(c)) FAIL(c, IMHD_EINVAL, "tvd_solve_3d has no forcing hook (isotvd_mod.f90:226)");
    if (drive_next && !c->force_set) FAIL(c, IMHD_ESTATE, "drive_next needs a force field (imhd_set_force)");
    set_scalar_dt(c, dt);
    const int s0 = c->cur, s1 = 1 - c->cur;
    if (is_tvd(c)) {
        if (c->prec == 64) tvd_step<double>(c); else tvd_step<float>(c);
    } else if (c->prec == 64) {
        if (drive_next) force_sums<double>(c, s0);
        vl_step<double>(c);
        restore_shell<double>(c);
        if (drive_next) force_apply<double>(c, s0, s1);
    } else {
        if (drive_next) force_sums<float>(c, s0);
        vl_step<float>(c);
        restore_shell<float>(c);
        if (drive_next) force_apply<float>(c, s0, s1);
    }
    HIPCK(c, hipGetLastError());
    if (drive_next) {
        unsigned int err = 0;
        HIPCK(c, hipMemcpyAsync(&err, c->force_err, 4, hipMemcpyDeviceToHost, c->st));
        HIPCK(c, hipStreamSynchronize(c->st));
        if (err) {
            HIPCK(c, hipMemsetAsync(c->force_err, 0, 4, c->st));
            FAIL(c, IMHD_EINVAL, "bad value in turbulence driver constant energy normalization "
                                 "(turb_driver_mod.f90:420-425)");
        }
    }
    return prof_collect(c);
}

// Launch-shape autotuner: times `steps` native steps for each candidate (z-segment length
// of the column kernels x y-strip height of the work enumeration) from a snapshot of the
// current state, keeps the fastest, and restores the state and time.
int imhd_autotune(imhd_ctx* c, int steps, char* report, int report_len) {
    if (!c || steps < 1) return IMHD_EINVAL;
    if (!c->have_state) FAIL(c, IMHD_ESTATE, "no state set");
    if (c->prof) FAIL(c, IMHD_EINVAL, "switch profiling off before autotuning");
    HIPCK(c, hipStreamSynchronize(c->st));
    const size_t nr = (size_t)c->pool * c->esz, nd = (size_t)c->pool * 8;
    const int cur0 = c->cur, pcur0 = c->pcur;
    std::vector<std::pair<void*, void*>> saved;   // (live buffer, snapshot)
    std::vector<size_t> sizes;
    auto keep = [&](void* live, size_t bytes) -> int {
        if (!live) return 0;
        void* snap = nullptr;
        HIPCK(c, hipMalloc(&snap, bytes));
        HIPCK(c, hipMemcpy(snap, live, bytes, hipMemcpyDeviceToDevice));
        saved.push_back({live, snap});
        sizes.push_back(bytes);
        return 0;
    };
    int rc = 0;
    for (int f = 0; f < 7 && !rc; f++) rc = keep(c->q[cur0][f], nr);
    for (int d = 0; d < 3 && !rc; d++) rc = keep(c->b[cur0][d], nd);
    for (int p = 0; p < c->cfg.npass && !rc; p++) rc = keep(c->qp[pcur0][p], nr);
    for (int f = 0; f < 4 && !rc; f++) rc = keep(c->comp[f], nr);
    if (!rc) rc = keep(c->t, 8 * c->nbox);
    if (!rc) rc = keep(c->dt, 8 * c->nbox);
    if (!rc) rc = keep(c->dtprev, 8 * c->nbox);
    if (!rc) rc = keep(c->floor_count, 4);
    if (!rc && c->ou_on) {
        const size_t na = 8 * (size_t)c->nbox * (c->ou_k.size() / 3) * 3;
        rc = keep(c->ou_are, na);
        if (!rc) rc = keep(c->ou_aim, na);
        if (!rc) rc = keep(c->ou_step, 4 * (size_t)c->nbox);
    }
    auto restore = [&]() -> int {
        for (size_t n = 0; n < saved.size(); n++)
            HIPCK(c, hipMemcpy(saved[n].first, saved[n].second, sizes[n], hipMemcpyDeviceToDevice));
        c->cur = cur0;
        c->pcur = pcur0;
        return 0;
    };
    // launch shape: z-segment of the column kernels, y-strip of the work enumeration, and
    // tile shape and z-segment of the fused VL stages
    struct Shape {
        int zseg, strip, vshape, vseg, vrounds, vshape_c;
    };
    const bool vl = c->tp_pred.d != nullptr;
    auto set_shape = [&](const Shape& sh) -> int {
        c->zseg = sh.zseg;
        c->strip = sh.strip;
        c->vshape = sh.vshape;
        c->vshape_c = sh.vshape_c;
        c->vseg = sh.vseg;
        c->vrounds = sh.vrounds;
        int r = build_plan(c, c->pl_pupd, 2, -1, sh.zseg);
        if (!r) r = build_plan(c, c->pl_cupd, 5, -1, sh.zseg);
        if (!r) r = build_plan(c, c->pl_pcorner, 1, -1, sh.zseg);
        if (!r) r = build_plan(c, c->pl_ccorner, 4, -1, sh.zseg);
        if (!r) r = build_plan(c, c->pl_pstage[0], 11, -1, sh.zseg);
        if (!r) r = build_plan(c, c->pl_pstage[1], 13, -1, sh.zseg);
        if (!r && vl) r = build_tiles(c, c->tp_pred, false);
        if (!r && vl) r = build_tiles(c, c->tp_corr, true);
        for (auto& g : c->gexec) if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
        return r;
    };
    std::string rep;
    Shape bsh{c->zseg, c->strip, c->vshape, c->vseg, c->vrounds, c->vshape_c};
    const Shape orig = bsh;
    float best = 1e30f;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!rc) { HIPCK(c, hipEventCreate(&e0)); HIPCK(c, hipEventCreate(&e1)); }
    auto trial = [&](const Shape& sh) -> int {
        int r = set_shape(sh);
        if (!r) r = restore();
        if (r) return r;
        for (int s = 0; s < 1; s++) { if (c->prec == 64) native_step<double>(c, 1e300); else native_step<float>(c, 1e300); }
        HIPCK(c, hipEventRecord(e0, c->st));
        for (int s = 0; s < steps; s++) { if (c->prec == 64) native_step<double>(c, 1e300); else native_step<float>(c, 1e300); }
        HIPCK(c, hipEventRecord(e1, c->st));
        HIPCK(c, hipEventSynchronize(e1));
        HIPCK(c, hipGetLastError());
        float ms = 0.f;
        HIPCK(c, hipEventElapsedTime(&ms, e0, e1));
        ms /= (float)steps;
        char b[160];
        snprintf(b, sizeof(b), "zseg=%d strip=%d tile=%d ctile=%d vrounds=%d vseg=%d: %.4f ms/step; ", sh.zseg,
                 sh.strip, sh.vshape, sh.vshape_c, sh.vrounds, sh.vseg, ms);
        rep += b;
        if (ms < best) { best = ms; bsh = sh; }
        return 0;
    };
    // coordinate search, one parameter at a time from the current best
    static const int strips[] = {0, 8, 16, 32, 64};
    static const int zsegs[] = {0, 2, 4, 8, 16};
    static const int vsegs[] = {8, 12, 16};
    static const int vrounds[] = {1, 2, 3, 4, 6, 8};
    const Shape s0 = bsh;
    for (int s : strips) if (!rc) rc = trial({s0.zseg, s, s0.vshape, s0.vseg, s0.vrounds, s0.vshape_c});
    { const Shape b0 = bsh; for (int z : zsegs) if (!rc && z != b0.zseg) rc = trial({z, b0.strip, b0.vshape, b0.vseg, b0.vrounds, b0.vshape_c}); }
    if (vl) {
        // both stages' tile shape, then the corrector's alone
        { const Shape b0 = bsh; for (int v = 0; v < VS_SHAPES; v++) if (!rc && (v != b0.vshape || v != b0.vshape_c) && vs_shape_ok(c->prec, v)) rc = trial({b0.zseg, b0.strip, v, b0.vseg, b0.vrounds, v}); }
        { const Shape b0 = bsh; for (int v = 0; v < VS_SHAPES; v++) if (!rc && v != b0.vshape_c && vs_shape_ok(c->prec, v)) rc = trial({b0.zseg, b0.strip, b0.vshape, b0.vseg, b0.vrounds, v}); }
        { const Shape b0 = bsh; for (int r : vrounds) if (!rc && r != b0.vrounds) rc = trial({b0.zseg, b0.strip, b0.vshape, b0.vseg, r, b0.vshape_c}); }
        { const Shape b0 = bsh; for (int v : vsegs) if (!rc && (b0.vrounds != 0 || v != b0.vseg)) rc = trial({b0.zseg, b0.strip, b0.vshape, v, 0, b0.vshape_c}); }
    }
    if (!rc) rc = set_shape(bsh);
    if (!rc) rc = restore();
    if (rc) {
        // a failed trial (or snapshot) leaves the launch shape and state as they were before
        // the call: the original shape and every snapshot taken are put back; the first
        // error is the one returned
        const std::string first = c->err;
        (void)set_shape(orig);
        (void)hipStreamSynchronize(c->st);
        (void)restore();
        c->err = first;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (auto& sv : saved) (void)hipFree(sv.second);
    if (rc) return rc;
    char b[160];
    snprintf(b, sizeof(b), "chosen zseg=%d strip=%d tile=%d ctile=%d vrounds=%d vseg=%d (%.4f ms/step)", bsh.zseg,
             bsh.strip, bsh.vshape, bsh.vshape_c, bsh.vrounds, bsh.vseg, best);
    rep += b;
    c->prof_report = rep;
    if (report && report_len > 0) {
        strncpy(report, rep.c_str(), (size_t)report_len - 1);
        report[report_len - 1] = 0;
    }
    HIPCK(c, hipStreamSynchronize(c->st));
    return IMHD_OK;
}

int imhd_set_ou_forcing(imhd_ctx* c, const imhd_ou* ou) {
    if (!c) return IMHD_EINVAL;
    HIPCK(c, hipStreamSynchronize(c->st));
    for (auto& g : c->gexec) if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }   // the step changes shape
    if (!ou || !(ou->sigma > 0.0)) { c->ou_on = false; return IMHD_OK; }
    if (c->imp_on) FAIL(c, IMHD_EINVAL, "impulsive forcing is on: the native loop takes one kind of forcing");
    if (!(ou->t_corr > 0.0) || !(ou->kmax >= ou->kmin) || !(ou->kmax >= 1.0) || ou->kmax > 16.0)
        FAIL(c, IMHD_EINVAL, "OU forcing needs t_corr > 0 and 1 <= kmax <= 16, kmin <= kmax");
    const int K = (int)floor(ou->kmax + 1e-12);
    std::vector<int> kv;
    for (int nz = -K; nz <= K; nz++)
        for (int ny = -K; ny <= K; ny++)
            for (int nx = -K; nx <= K; nx++) {
                bool half = nx > 0 || (nx == 0 && ny > 0) || (nx == 0 && ny == 0 && nz > 0);
                double mag = sqrt((double)(nx * nx + ny * ny + nz * nz));
                if (half && mag >= ou->kmin - 1e-12 && mag <= ou->kmax + 1e-12) {
                    kv.push_back(nx); kv.push_back(ny); kv.push_back(nz);
                }
            }
    const int nm = (int)kv.size() / 3;
    if (nm == 0 || nm > OU_MAXMODE) FAIL(c, IMHD_EINVAL, "OU forcing: %d modes in the shell (1..%d allowed)", nm, OU_MAXMODE);
    for (int b = 0; b < c->nbox; b++)
        for (int d = 0; d < 3; d++)
            if (2 * K >= c->hb[b].n[d]) FAIL(c, IMHD_EINVAL, "OU forcing: kmax must be below n/2 on every axis");
    auto fr = [](void* p) { if (p) (void)hipFree(p); };
    for (void* p : {(void*)c->ou_n, (void*)c->ou_are, (void*)c->ou_aim, (void*)c->ou_step, (void*)c->ou_tw,
                    (void*)c->ou_twoff, (void*)c->ou_part, (void*)c->ou_mean, (void*)c->ou_col,
                    (void*)c->ou_coloff})
        fr(p);
    c->ou_col = nullptr;
    c->ou_coloff = nullptr;
    // per-axis twiddles e^{i 2 pi n (i - 1/2) / N}, n in [-K, K], interior i (host libm)
    std::vector<double> tw;
    std::vector<long long> off;
    const double two_pi = 6.283185307179586;
    for (int b = 0; b < c->nbox; b++) {
        off.push_back((long long)tw.size());
        for (int d = 0; d < 3; d++) {
            const int N = c->hb[b].n[d];
            for (int n = -K; n <= K; n++)
                for (int i = 1; i <= N; i++) {
                    double th = two_pi * (double)n * ((double)i - 0.5) / (double)N;
                    tw.push_back(cos(th));
                    tw.push_back(sin(th));
                }
        }
    }
    HIPCK(c, hipMalloc(&c->ou_n, sizeof(int) * kv.size()));
    HIPCK(c, hipMemcpy(c->ou_n, kv.data(), sizeof(int) * kv.size(), hipMemcpyHostToDevice));
    HIPCK(c, hipMalloc(&c->ou_are, 8 * (size_t)c->nbox * nm * 3));
    HIPCK(c, hipMalloc(&c->ou_aim, 8 * (size_t)c->nbox * nm * 3));
    HIPCK(c, hipMalloc(&c->ou_step, 4 * (size_t)c->nbox));
    HIPCK(c, hipMemset(c->ou_step, 0, 4 * (size_t)c->nbox));
    HIPCK(c, hipMalloc(&c->ou_tw, 8 * tw.size()));
    HIPCK(c, hipMemcpy(c->ou_tw, tw.data(), 8 * tw.size(), hipMemcpyHostToDevice));
    HIPCK(c, hipMalloc(&c->ou_twoff, 8 * off.size()));
    HIPCK(c, hipMemcpy(c->ou_twoff, off.data(), 8 * off.size(), hipMemcpyHostToDevice));
    {
        int maxnx = 0;
        for (const BoxDesc& B : c->hb) maxnx = std::max(maxnx, B.n[0]);
        HIPCK(c, hipMalloc(&c->ou_part, 8 * 3 * ou_part_rows(c->nbox, maxnx)));
    }
    HIPCK(c, hipMalloc(&c->ou_mean, 8 * 3 * (size_t)c->nbox));
    std::vector<long long> coff;
    long long ncol = 0;
    c->ou_maxcols = 0;
    for (int b = 0; b < c->nbox; b++) {
        coff.push_back(ncol);
        const int cols = c->hb[b].n[1] * c->hb[b].n[2];
        ncol += (long long)cols * (2 * K + 1) * 6;
        c->ou_maxcols = std::max(c->ou_maxcols, cols);
    }
    HIPCK(c, hipMalloc(&c->ou_col, 8 * (size_t)ncol));
    HIPCK(c, hipMalloc(&c->ou_coloff, 8 * coff.size()));
    HIPCK(c, hipMemcpy(c->ou_coloff, coff.data(), 8 * coff.size(), hipMemcpyHostToDevice));
    c->ou_cfg = *ou;
    c->ou_k = kv;
    c->ou_K = K;
    c->ou_on = true;
    launch_ou_update(c->st, c->nbox, c->db, ou_dev(c), 1);   // stationary start
    HIPCK(c, hipGetLastError());
    HIPCK(c, hipStreamSynchronize(c->st));
    return IMHD_OK;
}

int imhd_get_ou_modes(imhd_ctx* c, int box, int* nmode, int* kvec, double* amp) {
    if (!c || !nmode || box < 0 || box >= c->nbox) return IMHD_EINVAL;
    const int nm = c->ou_on ? (int)c->ou_k.size() / 3 : 0;
    *nmode = nm;
    if (nm == 0) return IMHD_OK;
    if (kvec) memcpy(kvec, c->ou_k.data(), sizeof(int) * 3 * nm);
    if (amp) {
        std::vector<double> re(3 * nm), im(3 * nm);
        HIPCK(c, hipStreamSynchronize(c->st));
        HIPCK(c, hipMemcpy(re.data(), c->ou_are + (size_t)box * nm * 3, 8 * 3 * nm, hipMemcpyDeviceToHost));
        HIPCK(c, hipMemcpy(im.data(), c->ou_aim + (size_t)box * nm * 3, 8 * 3 * nm, hipMemcpyDeviceToHost));
        for (int m = 0; m < nm; m++)
            for (int d = 0; d < 3; d++) {
                amp[6 * m + d] = re[3 * m + d];
                amp[6 * m + 3 + d] = im[3 * m + d];
            }
    }
    return IMHD_OK;
}

int imhd_set_force(imhd_ctx* c, int box, const double* delv, double edot, double dt_turb) {
    if (!c || !delv || box < 0 || box >= c->nbox) return IMHD_EINVAL;
    int rc;
    if ((rc = ensure_force(c))) return rc;
    const BoxDesc& B = c->hb[box];
    const long long nc = (long long)B.P[0] * B.P[1] * B.P[2];
    HIPCK(c, hipStreamSynchronize(c->st));
    for (int d = 0; d < 3; d++)
        HIPCK(c, hipMemcpy(c->delv[d] + B.off, delv + d * nc, (size_t)nc * 8, hipMemcpyHostToDevice));
    c->edot = edot;
    c->dt_turb = dt_turb;
    c->force_set = true;
    return IMHD_OK;
}

int imhd_generate_force(imhd_ctx* c, int box, int k_max, double edot, double dt_turb, int seeds[3]) {
    if (!c || !seeds || box < 0 || box >= c->nbox) return IMHD_EINVAL;
    const BoxDesc& B = c->hb[box];
    // init_turb_driver's check (turb_driver_mod.f90:31-34)
    if (k_max < 1 || k_max >= std::min(B.n[0], std::min(B.n[1], B.n[2])) / 2)
        FAIL(c, IMHD_EINVAL, "invalid k_max %d (needs 1 <= k_max < min(nx,ny,nz)/2)", k_max);
    int rc;
    if ((rc = generate_force(c, box, k_max, seeds))) return rc;
    c->edot = edot;
    c->dt_turb = dt_turb;
    c->force_set = true;
    return IMHD_OK;
}

int imhd_get_force(imhd_ctx* c, int box, double* delv) {
    if (!c || !delv || box < 0 || box >= c->nbox) return IMHD_EINVAL;
    if (!c->force_set) FAIL(c, IMHD_ESTATE, "no force field (imhd_set_force / imhd_generate_force)");
    const BoxDesc& B = c->hb[box];
    const long long nc = (long long)B.P[0] * B.P[1] * B.P[2];
    HIPCK(c, hipStreamSynchronize(c->st));
    for (int d = 0; d < 3; d++)
        HIPCK(c, hipMemcpy(delv + d * nc, c->delv[d] + B.off, (size_t)nc * 8, hipMemcpyDeviceToHost));
    return IMHD_OK;
}

int imhd_spectrum(imhd_ctx* c, int box, int field, double* k, double* pk) {
    if (!c || !k || !pk || box < 0 || box >= c->nbox) return IMHD_EINVAL;
    if (!c->have_state) FAIL(c, IMHD_ESTATE, "no state set");
    const BoxDesc& B = c->hb[box];
    if (B.n[0] != B.n[1] || B.n[0] != B.n[2] || (B.n[0] & 1))
        FAIL(c, IMHD_EINVAL, "spectra need a cubic box with an even number of cells per axis");
    if (field < 0 || field > 4) FAIL(c, IMHD_EINVAL, "field must be 0 (density) .. 4 (v_perp)");
    if (B.n[0] > 4096) FAIL(c, IMHD_EINVAL, "spectra: at most 4096 cells per axis");
    HIPCK(c, hipStreamSynchronize(c->st));
    return c->prec == 64 ? spectrum_t<double>(c, box, field, k, pk) : spectrum_t<float>(c, box, field, k, pk);
}

int imhd_mode_spectra(imhd_ctx* c, int box, double* k, double* P) {
    if (!c || !k || !P || box < 0 || box >= c->nbox) return IMHD_EINVAL;
    if (!c->have_state) FAIL(c, IMHD_ESTATE, "no state set");
    const BoxDesc& B = c->hb[box];
    if (B.n[0] != B.n[1] || B.n[0] != B.n[2] || (B.n[0] & 1))
        FAIL(c, IMHD_EINVAL, "spectra need a cubic box with an even number of cells per axis");
    HIPCK(c, hipStreamSynchronize(c->st));
    return c->prec == 64 ? mode_spectra_t<double>(c, box, k, P) : mode_spectra_t<float>(c, box, k, P);
}

int imhd_density_pdf(imhd_ctx* c, int box, int nbins, double rho_max, double* rho, double* pdf) {
    if (!c || !rho || !pdf || box < 0 || box >= c->nbox || nbins < 1 || !(rho_max > 0.0)) return IMHD_EINVAL;
    if (!c->have_state) FAIL(c, IMHD_ESTATE, "no state set");
    HIPCK(c, hipStreamSynchronize(c->st));
    return c->prec == 64 ? density_pdf_t<double>(c, box, nbins, rho_max, rho, pdf)
                         : density_pdf_t<float>(c, box, nbins, rho_max, rho, pdf);
}

int imhd_fill_ghosts(imhd_ctx* c) {
    if (!c) return IMHD_EINVAL;
    if (!c->have_state) FAIL(c, IMHD_ESTATE, "no state set");
    if (c->prec == 64) fill_ghosts<double>(c); else fill_ghosts<float>(c);
    HIPCK(c, hipGetLastError());
    return prof_collect(c);
}

int imhd_copy_velocity(imhd_ctx* c, int stage) {
    if (!c) return IMHD_EINVAL;
    if (!c->have_state) FAIL(c, IMHD_ESTATE, "no state set");
    if (c->cfg.npass <= 0) return IMHD_OK;
    int rc;
    if ((rc = ensure_passive_work(c))) return rc;
    if (c->prec == 64) copy_velocity<double>(c, stage != 0); else copy_velocity<float>(c, stage != 0);
    HIPCK(c, hipGetLastError());
    return prof_collect(c);
}

int imhd_passive_step(imhd_ctx* c, double dt) {
    if (!c) return IMHD_EINVAL;
    if (!c->have_state) FAIL(c, IMHD_ESTATE, "no state set");
    if (c->cfg.npass <= 0) return IMHD_OK;
    if (!(dt > 0.0)) FAIL(c, IMHD_EINVAL, "dt must be positive");
    int rc;
    if ((rc = ensure_passive_work(c))) return rc;
    set_scalar_dt(c, dt);
    if (c->prec == 64) passive_step<double>(c); else passive_step<float>(c);
    HIPCK(c, hipGetLastError());
    return prof_collect(c);
}

int imhd_max_speed(imhd_ctx* c, double* cmax) {
    if (!c || !cmax) return IMHD_EINVAL;
    if (!c->have_state) FAIL(c, IMHD_ESTATE, "no state set");
    if (c->prec == 64) max_speed_launch<double>(c); else max_speed_launch<float>(c);
    HIPCK(c, hipGetLastError());
    const size_t ns = stripe_slots(c->nbox);
    std::vector<unsigned long long> cm(ns), sm(ns), dv(ns);
    HIPCK(c, hipStreamSynchronize(c->st));
    HIPCK(c, hipMemcpy(cm.data(), c->cmax, 8 * ns, hipMemcpyDeviceToHost));
    HIPCK(c, hipMemcpy(sm.data(), c->smax, 8 * ns, hipMemcpyDeviceToHost));
    HIPCK(c, hipMemcpy(dv.data(), c->dvmax, 8 * ns, hipMemcpyDeviceToHost));
    for (int b = 0; b < c->nbox; b++) {
        unsigned long long am = 0, smx = 0, dvx = 0;
        for (int st = 0; st < NSTRIPE; st++) {
            am = std::max(am, cm[((size_t)b * NSTRIPE + st) * STRIPE_STRIDE]);
            smx = std::max(smx, sm[((size_t)b * NSTRIPE + st) * STRIPE_STRIDE]);
            dvx = std::max(dvx, dv[((size_t)b * NSTRIPE + st) * STRIPE_STRIDE]);
        }
        double a, s, d;
        memcpy(&a, &am, 8);
        memcpy(&s, &smx, 8);
        memcpy(&d, &dvx, 8);
        cmax[b] = std::max(a, s + d);  // MAX(max_speed, smax + delv_max) (vanleer_mod.f90:3323)
        if (cmax[b] < ZERO) FAIL(c, IMHD_ENOSPEED, "zero max speed encountered in box %d", b);
    }
    prof_collect(c);
    return IMHD_OK;
}

static void diag_unpack(const imhd_ctx* c, const double* h, imhd_diag* out) {
    for (int b = 0; b < c->nbox; b++) {
        double dx = c->hb[b].dx, dx3 = dx * dx * dx;
        const double* v = &h[b * DIAG_STRIDE];
        out[b].ke = 0.5 * v[0] * dx3;
        out[b].vrms_sq = v[1] * dx3;
        out[b].momx = v[2] * dx3;
        out[b].momy = v[3] * dx3;
        out[b].momz = v[4] * dx3;
        out[b].divb_sum = v[5];
        out[b].mass = v[6] * dx3;
        out[b].divb_max = v[7];
    }
}

// partial rows into `part` (DIAG_WG * DIAG_STRIDE doubles per box), results into `res`
// (DIAG_STRIDE per box), both on the device, then the results to the host
static int diagnostics_into(imhd_ctx* c, double* part, double* res, imhd_diag* out) {
    if (c->prec == 64) {
        DiagArgs<double> a{};
        for (int f = 0; f < 4; f++) a.q[f] = R<double>(c->q[c->cur][f]);
        for (int d = 0; d < 3; d++) a.b[d] = c->b[c->cur][d];
        a.part = part;
        a.out = res;
        launch_diag<double>(c->st, c->db, c->nbox, a);
    } else {
        DiagArgs<float> a{};
        for (int f = 0; f < 4; f++) a.q[f] = R<float>(c->q[c->cur][f]);
        for (int d = 0; d < 3; d++) a.b[d] = c->b[c->cur][d];
        a.part = part;
        a.out = res;
        launch_diag<float>(c->st, c->db, c->nbox, a);
    }
    HIPCK(c, hipGetLastError());
    double h[DIAG_STRIDE * MAXBOX];
    HIPCK(c, hipMemcpyAsync(h, res, 8 * (size_t)DIAG_STRIDE * c->nbox, hipMemcpyDeviceToHost, c->st));
    HIPCK(c, hipStreamSynchronize(c->st));
    diag_unpack(c, h, out);
    return IMHD_OK;
}

int imhd_diagnostics(imhd_ctx* c, imhd_diag* out) {
    if (!c || !out) return IMHD_EINVAL;
    if (!c->have_state) FAIL(c, IMHD_ESTATE, "no state set");
    if (!c->diag_part) {
        int rc = alloc(c, (void**)&c->diag_part, sizeof(double) * DIAG_WG * DIAG_STRIDE * (size_t)c->nbox);
        if (rc) { c->diag_part = nullptr; return rc; }
    }
    return diagnostics_into(c, c->diag_part, c->diag, out);
}

int imhd_diagnostics_static(imhd_ctx* c, imhd_diag* out) {
    if (!c || !out) return IMHD_EINVAL;
    if (!c->have_state) FAIL(c, IMHD_ESTATE, "no state set");
    if (c->nbox > DIAG_STATIC_MAXBOX)
        FAIL(c, IMHD_EINVAL, "imhd_diagnostics_static serves at most %d boxes (this context has %d): use imhd_diagnostics",
             DIAG_STATIC_MAXBOX, c->nbox);
    if (!c->diag_static)
        FAIL(c, IMHD_EHIP, "imhd_diagnostics_static: no device scratch on this device (%s)", c->diag_static_err.c_str());
    double* res = c->diag_static + (size_t)DIAG_STATIC_MAXBOX * DIAG_WG * DIAG_STRIDE;
    return diagnostics_into(c, c->diag_static, res, out);
}

int imhd_floor_count(imhd_ctx* c, unsigned int* count) {
    if (!c || !count) return IMHD_EINVAL;
    HIPCK(c, hipStreamSynchronize(c->st));
    HIPCK(c, hipMemcpy(count, c->floor_count, 4, hipMemcpyDeviceToHost));
    return IMHD_OK;
}

int imhd_device_bytes(imhd_ctx* c, unsigned long long* bytes) {
    if (!c || !bytes) return IMHD_EINVAL;
    *bytes = c->dev_bytes;
    return IMHD_OK;
}

int imhd_write_restart(imhd_ctx* c, int box, const char* path, double t_sim, int dump_count, const int seeds[3]) {
    if (!c || !path || !seeds || box < 0 || box >= c->nbox) return IMHD_EINVAL;
    if (!c->have_state) FAIL(c, IMHD_ESTATE, "no state set");
    FileCloser fc{fopen(path, "wb")};
    if (!fc.f) FAIL(c, IMHD_EIO, "write_restart: cannot open %s", path);
    int rc = c->prec == 64 ? write_restart_t<double>(c, box, fc.f, t_sim, dump_count, seeds)
                           : write_restart_t<float>(c, box, fc.f, t_sim, dump_count, seeds);
    if (rc) return rc;
    FILE* f = fc.f;
    fc.f = nullptr;
    if (fclose(f) != 0) FAIL(c, IMHD_EIO, "write_restart: close failed");
    return IMHD_OK;
}

int imhd_read_restart(imhd_ctx* c, int box, const char* path, double* t_sim, int* dump_count, int seeds[3]) {
    if (!c || !path || !t_sim || !dump_count || !seeds || box < 0 || box >= c->nbox) return IMHD_EINVAL;
    FileCloser fc{fopen(path, "rb")};
    if (!fc.f) FAIL(c, IMHD_EIO, "read_restart: cannot open %s", path);
    int rc = c->prec == 64 ? read_restart_t<double>(c, box, fc.f, t_sim, dump_count, seeds)
                           : read_restart_t<float>(c, box, fc.f, t_sim, dump_count, seeds);
    if (rc) return rc;
    if (c->comp[0]) for (int f = 0; f < 4; f++) HIPCK(c, hipMemset(c->comp[f], 0, (size_t)c->pool * c->esz));
    c->have_state = true;
    return IMHD_OK;
}

int imhd_write_cube(imhd_ctx* c, int box, const char* path, int double_prec, int npass_out) {
    if (!c || !path || box < 0 || box >= c->nbox) return IMHD_EINVAL;
    if (npass_out < 0 || npass_out > c->cfg.npass) FAIL(c, IMHD_EINVAL, "npass_out must be 0..npass");
    if (!c->have_state) FAIL(c, IMHD_ESTATE, "no state set");
    FileCloser fc{fopen(path, "wb")};
    if (!fc.f) FAIL(c, IMHD_EIO, "write_cube: cannot open %s", path);
    int rc;
    if (c->prec == 64)
        rc = double_prec ? write_cube_t<double, double>(c, box, fc.f, npass_out)
                         : write_cube_t<double, float>(c, box, fc.f, npass_out);
    else
        rc = double_prec ? write_cube_t<float, double>(c, box, fc.f, npass_out)
                         : write_cube_t<float, float>(c, box, fc.f, npass_out);
    if (rc) return rc;
    FILE* f = fc.f;
    fc.f = nullptr;
    if (fclose(f) != 0) FAIL(c, IMHD_EIO, "write_cube: close failed");
    return IMHD_OK;
}

int imhd_set_time(imhd_ctx* c, const double* t, const double* dt) {
    if (!c) return IMHD_EINVAL;
    HIPCK(c, hipStreamSynchronize(c->st));
    if (t) HIPCK(c, hipMemcpy(c->t, t, 8 * c->nbox, hipMemcpyHostToDevice));
    if (dt) HIPCK(c, hipMemcpy(c->dt, dt, 8 * c->nbox, hipMemcpyHostToDevice));
    return IMHD_OK;
}

int imhd_get_time(imhd_ctx* c, double* t, double* dt) {
    if (!c) return IMHD_EINVAL;
    HIPCK(c, hipStreamSynchronize(c->st));
    if (t) HIPCK(c, hipMemcpy(t, c->t, 8 * c->nbox, hipMemcpyDeviceToHost));
    if (dt) HIPCK(c, hipMemcpy(dt, c->dt, 8 * c->nbox, hipMemcpyDeviceToHost));
    return IMHD_OK;
}

int imhd_set_profiling(imhd_ctx* c, int on) {
    if (!c) return IMHD_EINVAL;
    c->prof = on != 0;
    return IMHD_OK;
}

const char* imhd_profile_report(imhd_ctx* c) { return c ? c->prof_report.c_str() : ""; }

int imhd_sync(imhd_ctx* c) {
    if (!c) return IMHD_EINVAL;
    HIPCK(c, hipStreamSynchronize(c->st));
    return IMHD_OK;
}

// nsteps native steps, eager or replayed from the captured step graph (one per buffer
// parity), with no host synchronisation.
static int run_native(imhd_ctx* c, int nsteps, double t_end, int use_graph) {
    auto one = [&](void) {
        if (c->prec == 64) native_step<double>(c, t_end); else native_step<float>(c, t_end);
    };
    if (!use_graph) {
        for (int s = 0; s < nsteps; s++) one();
        HIPCK(c, hipGetLastError());
        return prof_collect(c);
    }
    // A step swaps the state and passive buffers; capture one step per buffer parity
    // once and replay it.
    if (c->graph_t_end != t_end) {
        for (auto& g : c->gexec) if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
        c->graph_t_end = t_end;
    }
    for (int s = 0; s < nsteps; s++) {
        const int par = c->cur * 2 + c->pcur;
        if (!c->gexec[par]) {
            hipGraph_t g;
            HIPCK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
            int cur0 = c->cur, pcur0 = c->pcur;
            one();
            hipError_t e = hipStreamEndCapture(c->st, &g);
            if (e != hipSuccess) FAIL(c, IMHD_EHIP, "graph capture failed: %s", hipGetErrorString(e));
            HIPCK(c, hipGraphInstantiate(&c->gexec[par], g, nullptr, nullptr, 0));
            HIPCK(c, hipGraphDestroy(g));
            c->cur = cur0;
            c->pcur = pcur0;
        }
        HIPCK(c, hipGraphLaunch(c->gexec[par], c->st));
        c->cur = 1 - c->cur;  // the native step keeps the passive buffer in place
    }
    HIPCK(c, hipGetLastError());
    return IMHD_OK;
}

// The native loop with the reference driver's impulsive forcing: per step the host reads
// t and dt and decides the kick as driver.f90:57-70 does; a kick step runs eagerly
// (native_step_drive) and is followed by the next field (generate_rand_force,
// driver.f90:102-111); the other steps run as in run_native.
static int run_impulsive(imhd_ctx* c, int nsteps, double t_end, int use_graph) {
    for (int s = 0; s < nsteps; s++) {
        double td[2];
        HIPCK(c, hipMemcpyAsync(&td[0], c->t, 8, hipMemcpyDeviceToHost, c->st));
        HIPCK(c, hipMemcpyAsync(&td[1], c->dt, 8, hipMemcpyDeviceToHost, c->st));
        HIPCK(c, hipStreamSynchronize(c->st));
        const double t_old = td[0], dt = td[1];
        if (!(c->imp_t_turb > t_old && c->imp_t_turb <= t_old + dt)) {
            int rc = run_native(c, 1, t_end, use_graph);
            if (rc) return rc;
            continue;
        }
        c->imp_t_turb = std::max(c->imp_t_turb + c->dt_turb, t_old + 1e-2 * dt);
        set_scalar_dt(c, 0.5 * dt);
        if (c->prec == 64) native_step_drive<double>(c, t_end); else native_step_drive<float>(c, t_end);
        HIPCK(c, hipGetLastError());
        unsigned int err = 0;
        HIPCK(c, hipMemcpyAsync(&err, c->force_err, 4, hipMemcpyDeviceToHost, c->st));
        HIPCK(c, hipStreamSynchronize(c->st));
        if (err) {
            HIPCK(c, hipMemsetAsync(c->force_err, 0, 4, c->st));
            FAIL(c, IMHD_EINVAL, "bad value in turbulence driver constant energy normalization "
                                 "(turb_driver_mod.f90:420-425)");
        }
        int rc = generate_force(c, 0, c->imp_kmax, c->imp_seeds);
        if (rc) return rc;
        if ((rc = prof_collect(c))) return rc;
    }
    return IMHD_OK;
}

int imhd_run(imhd_ctx* c, int nsteps, double t_end, int use_graph) {
    if (!c || nsteps < 0) return IMHD_EINVAL;
    if (!c->have_state) FAIL(c, IMHD_ESTATE, "no state set");
    if (nsteps == 0) return IMHD_OK;
    if (use_graph && c->prof) FAIL(c, IMHD_EINVAL, "profiling needs eager launches (use_graph = 0)");
    if (c->imp_on) return run_impulsive(c, nsteps, t_end, use_graph);
    return run_native(c, nsteps, t_end, use_graph);
}

int imhd_set_impulsive_forcing(imhd_ctx* c, int k_max, double edot, double dt_turb, const int* seeds) {
    if (!c) return IMHD_EINVAL;
    HIPCK(c, hipStreamSynchronize(c->st));
    if (k_max <= 0) { c->imp_on = false; return IMHD_OK; }
    if (!seeds) return IMHD_EINVAL;
    if (is_tvd(c)) FAIL(c, IMHD_EINVAL, "tvd_solve_3d has no forcing hook (isotvd_mod.f90:226)");
    if (c->nbox != 1) FAIL(c, IMHD_EINVAL, "impulsive forcing runs on single-box contexts");
    if (c->ou_on) FAIL(c, IMHD_EINVAL, "OU forcing is on: the native loop takes one kind of forcing");
    if (!c->have_state) FAIL(c, IMHD_ESTATE, "no state set");
    if (!(dt_turb > 0.0)) FAIL(c, IMHD_EINVAL, "dt_turb must be positive");
    const BoxDesc& B = c->hb[0];
    if (k_max >= std::min(B.n[0], std::min(B.n[1], B.n[2])) / 2)   // turb_driver_mod.f90:31-34
        FAIL(c, IMHD_EINVAL, "invalid k_max %d (needs 1 <= k_max < min(nx,ny,nz)/2)", k_max);
    if (!c->cour_half) {
        HIPCK(c, hipMalloc(&c->cour_half, 8));
    }
    double cour = 0.0;
    HIPCK(c, hipMemcpy(&cour, c->cour, 8, hipMemcpyDeviceToHost));
    const double half = 0.5 * cour;
    HIPCK(c, hipMemcpy(c->cour_half, &half, 8, hipMemcpyHostToDevice));
    int sd[3] = {seeds[0], seeds[1], seeds[2]};
    int rc = generate_force(c, 0, k_max, sd);   // driver.f90:336-339
    if (rc) return rc;
    double t = 0.0;
    HIPCK(c, hipMemcpy(&t, c->t, 8, hipMemcpyDeviceToHost));
    c->edot = edot;
    c->dt_turb = dt_turb;
    c->force_set = true;
    c->imp_kmax = k_max;
    for (int i = 0; i < 3; i++) c->imp_seeds[i] = sd[i];
    c->imp_t_turb = t + dt_turb;                  // driver.f90:344
    c->imp_on = true;
    return IMHD_OK;
}

int imhd_get_impulsive_forcing(imhd_ctx* c, int* seeds, double* t_turb) {
    if (!c) return IMHD_EINVAL;
    if (!c->imp_on) FAIL(c, IMHD_ESTATE, "impulsive forcing is off");
    if (seeds) for (int i = 0; i < 3; i++) seeds[i] = c->imp_seeds[i];
    if (t_turb) *t_turb = c->imp_t_turb;
    return IMHD_OK;
}

}  // extern "C"
