// Shared device-side definitions of the MI355X isothermal-MHD cell-update library.
//
// Memory layout in HBM (per box): every field is its own ghost-padded 3-D array with the
// reference's index meaning (grid_data_mod.f90:38-42): cell (i,j,k), i in [1-nb, nx+nb],
// x fastest.  Fields of all boxes of an ensemble live in one pool per field; a box is a
// contiguous slice at BoxDesc::off.  Work items of a launch are enumerated box by box in
// x-fastest order (a "plan"), so one launch serves one box or an uneven ensemble.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace imhd {

constexpr int MAXBOX = 256;
constexpr int MAXPASS = 4;
constexpr double ZERO = 1.0e-20;      // vanleer_mod.f90:75
constexpr double MIN_DENS = 1.0e-20;  // vanleer_mod.f90:76

// Riemann solver ids, reference numbering of the VL module (constants_mod.f90:57)
enum { RS_ROE = 1, RS_HLLD = 2, RS_HLLE = 3 };

struct BoxDesc {
    int n[3];        // interior cells
    int nb;          // ghost layers
    int P[3];        // padded extents n + 2 nb
    int bc;          // 1 open, 2 periodic
    long long off;   // first padded cell of this box in each field pool
    double dx;       // 1 / max(n)  (driver.f90:254)
};

// One box's share of a launch: cells lo .. lo + r - 1 (box index space), work items
// start .. start + r0*r1*r2 - 1 of the launch.  Column plans (z-marching kernels) enumerate
// (i, j, segment) instead: r[2] segments of seg cells each cover z cells zlo .. zlo + zn - 1.
struct PlanBox {
    int lo[3];
    int r[3];
    long long start;
    int zlo, zn, seg, _pad;
};

struct Plan {
    const PlanBox* boxes;  // device
    int nbox;
    long long total;
    int sy;                // > 0: rows enumerated in y-strips of sy rows, all k per strip
};

// The (i, j) part is formed in 32 bits (a padded plane holds far fewer than 2^31 cells), the
// plane part in 64: for a uniform k that is scalar work, and the per-lane part of a
// z-marching loop is loop-invariant.
__device__ __forceinline__ long long cell_index(const BoxDesc& b, int i, int j, int k) {
    const int in_plane = (i + b.nb - 1) + b.P[0] * (j + b.nb - 1);
    return b.off + (long long)b.P[0] * (long long)b.P[1] * (long long)(k + b.nb - 1) + (long long)in_plane;
}

// Workgroup ids are dealt round-robin over the 8 XCDs; renumber them so that each XCD
// walks one contiguous chunk of the work list (neighbouring stencil rows then share
// that XCD's L2).
__device__ __forceinline__ unsigned xcd_block(unsigned w, unsigned G) {
    unsigned q = G >> 3, r = G & 7u, x = w & 7u, idx = w >> 3;
    unsigned base = (x < r) ? x * (q + 1) : r * (q + 1) + (x - r) * q;
    return base + idx;
}

// Maps this thread to (box, i, j, k).  Returns false for padding threads.
__device__ __forceinline__ bool plan_item(const Plan& p, int& box, int& i, int& j, int& k) {
    unsigned blk = xcd_block(blockIdx.x, gridDim.x);
    long long g = (long long)blk * blockDim.x + threadIdx.x;
    if (g >= p.total) return false;
    int b = 0;
    if (p.nbox > 1) {
        int lo = 0, hi = p.nbox - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (p.boxes[mid].start <= g) lo = mid; else hi = mid - 1;
        }
        b = lo;
    }
    const PlanBox& pb = p.boxes[b];
    long long loc = g - pb.start;
    const long long r0 = pb.r[0], r1 = pb.r[1], r2 = pb.r[2];
    if (loc >= r0 * r1 * r2) return false;
    long long t = loc / r0;
    i = pb.lo[0] + (int)(loc - t * r0);
    const long long sy = p.sy;
    if (sy > 0 && r1 > sy) {
        // y-strip order: strip s holds rows [s sy, (s+1) sy) of every k, k outer within it,
        // so a strip's k-1 / k+1 planes stay in the XCD's L2 while the strip is swept
        const long long full = r1 / sy, per = sy * r2;
        const long long s = t / per;
        long long jj, kk;
        if (s < full) {
            const long long rem = t - s * per;
            kk = rem / sy;
            jj = s * sy + (rem - kk * sy);
        } else {
            const long long h = r1 - full * sy, rem = t - full * per;
            kk = rem / h;
            jj = full * sy + (rem - kk * h);
        }
        j = pb.lo[1] + (int)jj;
        k = pb.lo[2] + (int)kk;
    } else {
        long long u = t / r1;
        j = pb.lo[1] + (int)(t - u * r1);
        k = pb.lo[2] + (int)u;
    }
    box = b;
    return true;
}

// Division and square root.  FP64: IEEE correctly rounded (the reference's arithmetic).
// FP32 path: hardware reciprocal / square root (about 1 ulp), as the FP32 path is held to
// an L1 tolerance against FP64 rather than to bitwise parity.
__device__ __forceinline__ double fdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double fsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Plan lookup that also reports the box for inactive (padding) lanes so that every lane
// of a wavefront joins the smax reduction; box starts are 64-aligned.
__device__ __forceinline__ bool plan_item_box(const Plan& p, int& box, int& i, int& j, int& k) {
    box = 0;
    bool act = plan_item(p, box, i, j, k);
    if (!act && p.nbox > 1) {
        unsigned blk = xcd_block(blockIdx.x, gridDim.x);
        long long g = (long long)blk * blockDim.x + threadIdx.x;
        int lo = 0, hi = p.nbox - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (p.boxes[mid].start <= g) lo = mid; else hi = mid - 1;
        }
        box = lo;
    }
    return act;
}

// Largest |v_d| + c_f,d of one cell (vanleer_mod.f90:3282-3311), FP64.
__device__ __forceinline__ double cell_fast_speed(const double* q, double cs) {
    double asqr = cs * cs;
    double rho = q[0];
    double isrho = 1.0 / sqrt(rho);
    double bx = q[4] * isrho, by = q[5] * isrho, bz = q[6] * isrho;
    double btsqr = by * by + bz * bz, bsqr = bx * bx + btsqr, x = asqr - bsqr;
    double tmp = sqrt(x * x + 4.0 * asqr * btsqr);
    double cfx = sqrt(0.5 * (asqr + bsqr + tmp));
    btsqr = bx * bx + bz * bz; bsqr = by * by + btsqr; x = asqr - bsqr;
    tmp = sqrt(x * x + 4.0 * asqr * btsqr);
    double cfy = sqrt(0.5 * (asqr + bsqr + tmp));
    btsqr = bx * bx + by * by; bsqr = bz * bz + btsqr; x = asqr - bsqr;
    tmp = sqrt(x * x + 4.0 * asqr * btsqr);
    double cfz = sqrt(0.5 * (asqr + bsqr + tmp));
    double vx = q[1] / rho, vy = q[2] / rho, vz = q[3] / rho;
    double m = fabs(vx) + cfx;
    m = m > fabs(vy) + cfy ? m : fabs(vy) + cfy;
    return m > fabs(vz) + cfz ? m : fabs(vz) + cfz;
}

// The same speed in FP32 with the hardware reciprocal, reciprocal square root and square
// root: the FP32 path's per-cell CFL speed (its maximum over cells is still taken in FP64).
__device__ __forceinline__ float cell_fast_speed_f(const float* q, float cs) {
    float asqr = cs * cs;
    float ir = __builtin_amdgcn_rcpf(q[0]);
    float isrho = __builtin_amdgcn_rsqf(q[0]);
    float bx = q[4] * isrho, by = q[5] * isrho, bz = q[6] * isrho;
    float btsqr = by * by + bz * bz, bsqr = bx * bx + btsqr, x = asqr - bsqr;
    float tmp = __builtin_amdgcn_sqrtf(x * x + 4.0f * asqr * btsqr);
    float cfx = __builtin_amdgcn_sqrtf(0.5f * (asqr + bsqr + tmp));
    btsqr = bx * bx + bz * bz; bsqr = by * by + btsqr; x = asqr - bsqr;
    tmp = __builtin_amdgcn_sqrtf(x * x + 4.0f * asqr * btsqr);
    float cfy = __builtin_amdgcn_sqrtf(0.5f * (asqr + bsqr + tmp));
    btsqr = bx * bx + by * by; bsqr = bz * bz + btsqr; x = asqr - bsqr;
    tmp = __builtin_amdgcn_sqrtf(x * x + 4.0f * asqr * btsqr);
    float cfz = __builtin_amdgcn_sqrtf(0.5f * (asqr + bsqr + tmp));
    float m = fabsf(q[1] * ir) + cfx;
    m = fmaxf(m, fabsf(q[2] * ir) + cfy);
    return fmaxf(m, fabsf(q[3] * ir) + cfz);
}

template <class real>
__device__ __forceinline__ real sgn1(real x) { return copysign(real(1), x); }  // Fortran SIGN(1.d0, x)

// Minima and maxima.  FP64 keeps the reference's comparisons, so that ties between +0 and
// -0 resolve as there; FP32 uses the hardware min/max (v_min3_f32 and v_max3_f32 take three
// operands in one instruction), which differ from the comparisons only in the sign of a
// zero result and for NaN operands.
template <class real>
__device__ __forceinline__ real min3(real a, real b, real c) {
    if constexpr (sizeof(real) == 4) return __builtin_fminf(__builtin_fminf(a, b), c);
    real m = a < b ? a : b;
    return m < c ? m : c;
}
template <class real>
__device__ __forceinline__ real max3(real a, real b, real c) {
    if constexpr (sizeof(real) == 4) return __builtin_fmaxf(__builtin_fmaxf(a, b), c);
    real m = a > b ? a : b;
    return m > c ? m : c;
}
template <class real>
__device__ __forceinline__ real max2(real a, real b) {
    if constexpr (sizeof(real) == 4) return __builtin_fmaxf(a, b);
    return a > b ? a : b;
}
template <class real>
__device__ __forceinline__ real min2(real a, real b) {
    if constexpr (sizeof(real) == 4) return __builtin_fminf(a, b);
    return a < b ? a : b;
}

// Non-negative doubles order like their bit patterns, so the CFL/smax maxima use an
// integer atomicMax.  A wave reduces with shuffles; its lane 0 then updates one of
// NSTRIPE per-box slots (each on its own 128-byte line) and skips the atomic when the slot
// already holds a larger value.  One address per box would serialise ~1e5 same-address
// atomics per launch in L2.
constexpr int NSTRIPE = 64;
constexpr int STRIPE_STRIDE = 16;  // u64 per slot (128 B)
__host__ __device__ constexpr size_t stripe_slots(int nbox) { return (size_t)nbox * NSTRIPE * STRIPE_STRIDE; }

__device__ __forceinline__ double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ void atomic_max_nonneg(unsigned long long* base, int box, double v) {
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0 && v > 0.0) {
        unsigned stripe = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) % NSTRIPE;
        unsigned long long* a = base + ((size_t)box * NSTRIPE + stripe) * STRIPE_STRIDE;
        unsigned long long bits = (unsigned long long)__double_as_longlong(v);
        if (bits > __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(a, bits);
    }
}
__device__ __forceinline__ double striped_max(const unsigned long long* base, int box) {
    unsigned long long m = 0;
    for (int s = 0; s < NSTRIPE; s++) {
        unsigned long long v = base[((size_t)box * NSTRIPE + s) * STRIPE_STRIDE];
        m = v > m ? v : m;
    }
    return __longlong_as_double((long long)m);
}

// Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11): 10 rounds of two 32x32->64
// multiplies with the Weyl key schedule.  Host and device run the same code, so the OU
// noise and the seeded starts are identical wherever they are drawn.
__host__ __device__ __forceinline__ void philox4x32(unsigned c0, unsigned c1, unsigned c2, unsigned c3,
                                                    unsigned k0, unsigned k1, unsigned* o) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        unsigned long long p0 = (unsigned long long)0xD2511F53u * c0;
        unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c2;
        unsigned y0 = (unsigned)(p1 >> 32) ^ c1 ^ k0, y1 = (unsigned)p1;
        unsigned y2 = (unsigned)(p0 >> 32) ^ c3 ^ k1, y3 = (unsigned)p0;
        c0 = y0; c1 = y1; c2 = y2; c3 = y3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    o[0] = c0; o[1] = c1; o[2] = c2; o[3] = c3;
}

// Box-Muller (cosine branch) from two 32-bit words mapped to (0, 1)
__host__ __device__ __forceinline__ double box_muller(unsigned a, unsigned b) {
    const double s = 2.3283064365386963e-10;   // 2^-32
    double u1 = ((double)a + 0.5) * s, u2 = ((double)b + 0.5) * s;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// rotated-component permutation of a directional sweep (vanleer_mod.f90:449-464,
// 610-625, 788-803): rotated slot r <- global field perm(d, r)
__host__ __device__ constexpr int perm(int d, int r) {
    return d == 0 ? r
                  : (d == 1 ? (r == 1 ? 2 : r == 2 ? 3 : r == 3 ? 1 : r == 4 ? 5 : r == 5 ? 6 : r == 6 ? 4 : 0)
                            : (r == 1 ? 3 : r == 2 ? 1 : r == 3 ? 2 : r == 4 ? 6 : r == 5 ? 4 : r == 6 ? 5 : 0));
}

}  // namespace imhd
