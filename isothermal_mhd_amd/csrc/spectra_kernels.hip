// Power spectra and density PDF of the reference's analysis program
// (3D_MPI/Analysis/src/analysis_main.f90:105-226, 543-698) on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include "spectra.hpp"

namespace imhd {

// One REAL*4 component of the analysed field on the interior, as the analysis program
// forms it from the snapshot: v = (rho v)/rho; v_par = (v.B/|B|) B/|B|; v_perp = v - v_par.
template <class real>
__global__ __launch_bounds__(256) void k_spec_pack(BoxDesc B, SpecPackArgs<real> a) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)B.n[0] * B.n[1] * B.n[2];
    if (g >= n) return;
    const int i = (int)(g % B.n[0]);
    const long long t = g / B.n[0];
    const int j = (int)(t % B.n[1]), k = (int)(t / B.n[1]);
    const long long c = cell_index(B, i + 1, j + 1, k + 1);
    float f[7];
    for (int m = 0; m < 7; m++) f[m] = (float)a.q[m][c];
    const float rho = f[0];
    const float v[3] = {f[1] / rho, f[2] / rho, f[3] / rho};
    float out;
    if (a.which == SPEC_DENSITY) {
        out = rho;
    } else if (a.which == SPEC_VELOCITY) {
        out = v[a.comp];
    } else if (a.which == SPEC_BFIELD) {
        out = f[4 + a.comp];
    } else {
        const float bmag = sqrtf(f[4] * f[4] + f[5] * f[5] + f[6] * f[6]);
        const float bhat = f[4 + a.comp] / bmag;
        const float vdotb = (v[0] * f[4] + v[1] * f[5] + v[2] * f[6]) / bmag;
        out = a.which == SPEC_VPAR ? vdotb * bhat : v[a.comp] - vdotb * bhat;
    }
    a.out[g] = out;
}

__global__ __launch_bounds__(256) void k_spec_accum(const float* __restrict__ F, int n, float* __restrict__ s,
                                                    int first) {
    const int h = n / 2;
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long long)h * h * h) return;
    const int kx = (int)(g % h);
    const long long t = g / h;
    const int ky = (int)(t % h), kz = (int)(t / h);
    const long long o = ((long long)kz * n + ky) * (h + 1) + kx;   // R2C output [z][y][x/2+1]
    const float re = F[2 * o], im = F[2 * o + 1];
    float acc = first ? 0.0f : s[g];
    acc = acc + re * re;
    acc = acc + im * im;
    s[g] = acc;
}

// Block histogram in LDS (n <= SPEC_MAXBINS), then one global add per bin.
constexpr int SPEC_MAXBINS = 4096;

__global__ __launch_bounds__(256) void k_spec_bin(const float* __restrict__ s, int n, float k0, float dk,
                                                  double* __restrict__ pk) {
    __shared__ double hist[SPEC_MAXBINS];
    for (int b = threadIdx.x; b < n; b += blockDim.x) hist[b] = 0.0;
    __syncthreads();
    const int h = n / 2;
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < (long long)h * h * h) {
        const int kx = (int)(g % h);
        const long long t = g / h;
        const int ky = (int)(t % h), kz = (int)(t / h);
        float fk = 2.0f * s[g];
        if (kx == 0 && ky == 0 && kz == 0) fk = 0.5f * fk;
        const float k = k0 * sqrtf((float)(kx * kx + ky * ky + kz * kz));
        const int ik = (int)(k / dk);
        if (ik >= 0 && ik < n) atomicAdd(&hist[ik], (double)fk);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < n; b += blockDim.x)
        if (hist[b] != 0.0) atomicAdd(&pk[b], hist[b]);
}

// Density histogram.  A narrow distribution (a nearly uniform density puts every cell in
// one bin) must not serialise on one counter: each wave adds once per distinct bin among
// its lanes (leader's bin by ballot, count of lanes sharing it), into a per-workgroup LDS
// histogram (nbins <= HIST_LDS_BINS) that is flushed once per block, or straight into the
// global counts for more bins.  Counts are exact integers, so the order does not matter.
constexpr int HIST_LDS_BINS = 4096;
template <class real, bool LDS>
__global__ __launch_bounds__(256) void k_density_hist(BoxDesc B, const real* __restrict__ rho, float drho, int nbins,
                                                      unsigned long long* __restrict__ counts) {
    __shared__ unsigned int sh[LDS ? HIST_LDS_BINS : 1];
    if (LDS) {
        for (int b = threadIdx.x; b < nbins; b += blockDim.x) sh[b] = 0u;
        __syncthreads();
    }
    const long long n = (long long)B.n[0] * B.n[1] * B.n[2];
    const int lane = threadIdx.x & 63;
    // the trip count depends on the block only, so every lane of a wave takes part in
    // every ballot
    for (long long g0 = (long long)blockIdx.x * blockDim.x; g0 < n; g0 += (long long)gridDim.x * blockDim.x) {
        const long long g = g0 + threadIdx.x;
        int ib = -1;
        if (g < n) {
            const int i = (int)(g % B.n[0]);
            const long long t = g / B.n[0];
            const int j = (int)(t % B.n[1]), k = (int)(t / B.n[1]);
            const float r = (float)rho[cell_index(B, i + 1, j + 1, k + 1)];
            const float x = r / drho;
            if (x >= 0.0f) {
                const int v = (int)x;
                if (v < nbins) ib = v;
            }
        }
        bool pending = ib >= 0;
        while (true) {
            const unsigned long long m = __ballot(pending);
            if (m == 0ull) break;
            const int leader = __ffsll((long long)m) - 1;
            const int bin = __shfl(ib, leader);
            const bool mine = pending && ib == bin;
            const unsigned long long same = __ballot(mine);
            if (lane == leader) {
                if (LDS) atomicAdd(&sh[bin], (unsigned int)__popcll(same));
                else atomicAdd(&counts[bin], (unsigned long long)__popcll(same));
            }
            if (mine) pending = false;
        }
    }
    if (LDS) {
        __syncthreads();
        for (int b = threadIdx.x; b < nbins; b += blockDim.x)
            if (sh[b]) atomicAdd(&counts[b], (unsigned long long)sh[b]);
    }
}

template <class real>
__global__ __launch_bounds__(256) void k_mean_b(BoxDesc B, const real* __restrict__ bx, const real* __restrict__ by,
                                                const real* __restrict__ bz, double* __restrict__ sum) {
    __shared__ double sh[3][256];
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)B.n[0] * B.n[1] * B.n[2];
    double v[3] = {0.0, 0.0, 0.0};
    if (g < n) {
        const int i = (int)(g % B.n[0]);
        const long long t = g / B.n[0];
        const int j = (int)(t % B.n[1]), k = (int)(t / B.n[1]);
        const long long c = cell_index(B, i + 1, j + 1, k + 1);
        v[0] = (double)(float)bx[c];
        v[1] = (double)(float)by[c];
        v[2] = (double)(float)bz[c];
    }
    for (int d = 0; d < 3; d++) sh[d][threadIdx.x] = v[d];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int d = 0; d < 3; d++) sh[d][threadIdx.x] += sh[d][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int d = 0; d < 3; d++) atomicAdd(&sum[d], sh[d][0]);
}

// One octant mode of mode_decomposition (analysis_main.f90:362-527).  The geometry is
// evaluated in FP64.
constexpr int MODE_LDS_BINS = 512;

template <bool LDS>
__global__ __launch_bounds__(256) void k_mode_bin(const float* __restrict__ F0, const float* __restrict__ F1,
                                                  const float* __restrict__ F2, int n, const double* __restrict__ b0sum,
                                                  float k0f, float dkf, double* __restrict__ P) {
    __shared__ double hist[LDS ? 8 * MODE_LDS_BINS : 1];
    if (LDS) {
        for (int b = threadIdx.x; b < 8 * n; b += blockDim.x) hist[b] = 0.0;
        __syncthreads();
    }
    const int h = n / 2;
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double amp[8];
    int ik = -1;
    if (g < (long long)h * h * h && g > 0) {
        const int kx = (int)(g % h);
        const long long t = g / h;
        const int ky = (int)(t % h), kz = (int)(t / h);
        const double nc = (double)n * n * n;
        const double B0[3] = {b0sum[0] / nc, b0sum[1] / nc, b0sum[2] / nc};
        const double bsq = B0[0] * B0[0] + B0[1] * B0[1] + B0[2] * B0[2];
        const double bmag = sqrt(bsq), alpha = 1.0 / bsq;
        const double k0 = (double)k0f;
        const double kv[3] = {k0 * kx, k0 * ky, k0 * kz};
        const double k = k0 * sqrt((double)(kx * kx + ky * ky + kz * kz));
        const double khat[3] = {kv[0] / k, kv[1] / k, kv[2] / k};
        const double ph[3] = {B0[0] / bmag, B0[1] / bmag, B0[2] / bmag};
        const double kd = kv[0] * ph[0] + kv[1] * ph[1] + kv[2] * ph[2];
        double kpar[3], kperp[3], kph[3] = {0.0, 0.0, 0.0};
        for (int d = 0; d < 3; d++) { kpar[d] = kd * ph[d]; kperp[d] = kv[d] - kpar[d]; }
        const double kpn = sqrt(kperp[0] * kperp[0] + kperp[1] * kperp[1] + kperp[2] * kperp[2]);
        if (kpn > 0.0) for (int d = 0; d < 3; d++) kph[d] = kperp[d] / kpn;
        const double kpar2 = kpar[0] * kpar[0] + kpar[1] * kpar[1] + kpar[2] * kpar[2];
        const double D = fmax(0.0, (1.0 + alpha) * (1.0 + alpha) - 4.0 * alpha * kpar2 / (k * k));
        const double sD = sqrt(D);
        const double za[3] = {kph[1] * ph[2] - kph[2] * ph[1], kph[2] * ph[0] - kph[0] * ph[2],
                              kph[0] * ph[1] - kph[1] * ph[0]};
        double zf[3], zs[3];
        for (int d = 0; d < 3; d++) {
            zf[d] = (-1.0 + alpha + sD) * kpar[d] + (1.0 + alpha + sD) * kperp[d];
            zs[d] = (-1.0 + alpha - sD) * kpar[d] + (1.0 + alpha - sD) * kperp[d];
        }
        const double nf = sqrt(zf[0] * zf[0] + zf[1] * zf[1] + zf[2] * zf[2]);
        const double ns = sqrt(zs[0] * zs[0] + zs[1] * zs[1] + zs[2] * zs[2]);
        if (nf > 0.0) for (int d = 0; d < 3; d++) zf[d] /= nf;
        if (ns > 0.0) for (int d = 0; d < 3; d++) zs[d] /= ns;
        const long long o = ((long long)kz * n + ky) * (h + 1) + kx;
        const double fr[3] = {F0[2 * o], F1[2 * o], F2[2 * o]};
        const double fi[3] = {F0[2 * o + 1], F1[2 * o + 1], F2[2 * o + 1]};
        double va[2] = {0, 0}, vf[2] = {0, 0}, vs[2] = {0, 0};
        for (int d = 0; d < 3; d++) {
            va[0] += fr[d] * za[d]; va[1] += fi[d] * za[d];
            vf[0] += fr[d] * zf[d]; vf[1] += fi[d] * zf[d];
            vs[0] += fr[d] * zs[d]; vs[1] += fi[d] * zs[d];
        }
        const double cf = sqrt(0.5 * bsq * (1.0 + alpha + sD));
        const double csl = sqrt(fmax(0.0, 0.5 * bsq * (1.0 + alpha - sD)));
        const double t1 = khat[0] * zf[0] + khat[1] * zf[1] + khat[2] * zf[2];
        const double t2 = khat[0] * zs[0] + khat[1] * zs[1] + khat[2] * zs[2];
        const double dvf = 2.0 * vf[1], dvs = 2.0 * vs[1];
        auto cross_norm = [&](const double* z) {
            const double c0 = B0[1] * z[2] - B0[2] * z[1], c1 = B0[2] * z[0] - B0[0] * z[2],
                         c2 = B0[0] * z[1] - B0[1] * z[0];
            return sqrt(c0 * c0 + c1 * c1 + c2 * c2);
        };
        // a zero slow speed (k along B0) leaves the reference's slow density and field
        // amplitudes undefined (0/0); they are not counted here
        const double rf = dvf * t1 / cf, rs = csl > 0.0 ? dvs * t2 / csl : 0.0;
        const double bf = dvf * cross_norm(zf) / cf, bs = csl > 0.0 ? dvs * cross_norm(zs) / csl : 0.0;
        amp[0] = 2.0 * (va[0] * va[0] + va[1] * va[1]);
        amp[1] = 2.0 * (vf[0] * vf[0] + vf[1] * vf[1]);
        amp[2] = 2.0 * (vs[0] * vs[0] + vs[1] * vs[1]);
        amp[3] = amp[0];
        amp[4] = 2.0 * bf * bf;
        amp[5] = 2.0 * bs * bs;
        amp[6] = 2.0 * rf * rf;
        amp[7] = 2.0 * rs * rs;
        const float kf = k0f * sqrtf((float)(kx * kx + ky * ky + kz * kz));
        ik = (int)(kf / dkf);
        if (ik >= n) ik = -1;
    }
    if (LDS) {
        if (ik >= 0)
            for (int m = 0; m < 8; m++) atomicAdd(&hist[m * n + ik], amp[m]);
        __syncthreads();
        for (int b = threadIdx.x; b < 8 * n; b += blockDim.x)
            if (hist[b] != 0.0) atomicAdd(&P[b], hist[b]);
    } else if (ik >= 0) {
        for (int m = 0; m < 8; m++) atomicAdd(&P[m * n + ik], amp[m]);
    }
}

template <class real>
void launch_mean_b(hipStream_t st, const BoxDesc& B, const real* bx, const real* by, const real* bz, double* sum) {
    const long long n = (long long)B.n[0] * B.n[1] * B.n[2];
    hipLaunchKernelGGL((k_mean_b<real>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, B, bx, by, bz, sum);
}
void launch_mode_bin(hipStream_t st, const float* F0, const float* F1, const float* F2, int n, const double* b0sum,
                     float k0, float dk, double* P) {
    const long long m = (long long)(n / 2) * (n / 2) * (n / 2);
    const dim3 grid((unsigned)((m + 255) / 256));
    if (n <= MODE_LDS_BINS)
        hipLaunchKernelGGL(k_mode_bin<true>, grid, dim3(256), 0, st, F0, F1, F2, n, b0sum, k0, dk, P);
    else
        hipLaunchKernelGGL(k_mode_bin<false>, grid, dim3(256), 0, st, F0, F1, F2, n, b0sum, k0, dk, P);
}

template <class real>
void launch_spec_pack(hipStream_t st, const BoxDesc& B, const SpecPackArgs<real>& a) {
    const long long n = (long long)B.n[0] * B.n[1] * B.n[2];
    hipLaunchKernelGGL((k_spec_pack<real>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, B, a);
}
void launch_spec_accum(hipStream_t st, const float* F, int n, float* s, int first) {
    const long long m = (long long)(n / 2) * (n / 2) * (n / 2);
    hipLaunchKernelGGL(k_spec_accum, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, F, n, s, first);
}
void launch_spec_bin(hipStream_t st, const float* s, int n, float k0, float dk, double* pk) {
    const long long m = (long long)(n / 2) * (n / 2) * (n / 2);
    hipLaunchKernelGGL(k_spec_bin, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, s, n, k0, dk, pk);
}
template <class real>
void launch_density_hist(hipStream_t st, const BoxDesc& B, const real* rho, float drho, int nbins,
                         unsigned long long* counts) {
    const long long n = (long long)B.n[0] * B.n[1] * B.n[2];
    const unsigned blocks = (unsigned)std::min<long long>((n + 255) / 256, 4096);
    if (nbins <= HIST_LDS_BINS)
        hipLaunchKernelGGL((k_density_hist<real, true>), dim3(blocks), dim3(256), 0, st, B, rho, drho, nbins, counts);
    else
        hipLaunchKernelGGL((k_density_hist<real, false>), dim3(blocks), dim3(256), 0, st, B, rho, drho, nbins, counts);
}

template void launch_spec_pack<double>(hipStream_t, const BoxDesc&, const SpecPackArgs<double>&);
template void launch_spec_pack<float>(hipStream_t, const BoxDesc&, const SpecPackArgs<float>&);
template void launch_mean_b<double>(hipStream_t, const BoxDesc&, const double*, const double*, const double*, double*);
template void launch_mean_b<float>(hipStream_t, const BoxDesc&, const float*, const float*, const float*, double*);
template void launch_density_hist<double>(hipStream_t, const BoxDesc&, const double*, float, int,
                                          unsigned long long*);
template void launch_density_hist<float>(hipStream_t, const BoxDesc&, const float*, float, int, unsigned long long*);

}  // namespace imhd
