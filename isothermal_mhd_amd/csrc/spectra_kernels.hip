// Power spectra and density PDF of the reference's analysis program
// (3D_MPI/Analysis/src/analysis_main.f90:105-226, 543-698) on the device.
#include <hip/hip_runtime.h>

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

template <class real>
__global__ __launch_bounds__(256) void k_density_hist(BoxDesc B, const real* __restrict__ rho, float drho, int nbins,
                                                      unsigned long long* __restrict__ counts) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)B.n[0] * B.n[1] * B.n[2];
    if (g >= n) return;
    const int i = (int)(g % B.n[0]);
    const long long t = g / B.n[0];
    const int j = (int)(t % B.n[1]), k = (int)(t / B.n[1]);
    const float r = (float)rho[cell_index(B, i + 1, j + 1, k + 1)];
    const float x = r / drho;
    if (!(x >= 0.0f)) return;
    const int ib = (int)x;
    if (ib < nbins) atomicAdd(&counts[ib], 1ull);
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
    hipLaunchKernelGGL((k_density_hist<real>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, B, rho, drho,
                       nbins, counts);
}

template void launch_spec_pack<double>(hipStream_t, const BoxDesc&, const SpecPackArgs<double>&);
template void launch_spec_pack<float>(hipStream_t, const BoxDesc&, const SpecPackArgs<float>&);
template void launch_density_hist<double>(hipStream_t, const BoxDesc&, const double*, float, int,
                                          unsigned long long*);
template void launch_density_hist<float>(hipStream_t, const BoxDesc&, const float*, float, int, unsigned long long*);

}  // namespace imhd
