// Seeded random-field start of a box, generated on the device (imhd_init_turbulence).
//
// The velocity is a sum of Fourier modes with |n_d| <= K = floor(kmax) on each axis,
// evaluated separably: the host sums over n_x for every (n_y, n_z) pair (a few hundred KB),
// k_ic_ypass sums over n_y for every (i, j, n_z), and k_ic_cells sums over n_z per cell,
// forms rho from its per-cell Gaussian and writes rho, rho v, B and the passive scalars of
// the interior.  Everything is FP64 (built without FMA contraction); FP32 contexts store
// the FP64 values rounded once.
#include "ic_kernels.hpp"

namespace imhd {

// Bout[d][kz][j][i] = sum_{iy} A[d][kz][iy][i] e^{2 pi i (iy - K) j / ny}, complex
// (re, im interleaved).  One thread per (i, j).
__global__ __launch_bounds__(256) void k_ic_ypass(int nx, int ny, int K, const double* __restrict__ A,
                                                  const double* __restrict__ twy, double* __restrict__ Bout) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x), j = (int)blockIdx.y;
    if (i >= nx) return;
    const int W = 2 * K + 1;
    for (int d = 0; d < 3; d++)
        for (int kz = 0; kz <= K; kz++) {
            double sr = 0.0, si = 0.0;
            for (int iy = 0; iy < W; iy++) {
                const long long r = ((long long)(iy - K) * j) % ny;
                const int t = (int)(r < 0 ? r + ny : r);
                const double c = twy[2 * t], s = twy[2 * t + 1];
                const size_t a = ((((size_t)d * (K + 1) + kz) * W + iy) * nx + i) * 2;
                const double ar = A[a], ai = A[a + 1];
                sr = sr + (ar * c - ai * s);
                si = si + (ar * s + ai * c);
            }
            const size_t o = ((((size_t)d * (K + 1) + kz) * ny + j) * nx + i) * 2;
            Bout[o] = sr;
            Bout[o + 1] = si;
        }
}

// One thread per (i, j, component d, z-segment of ICSEG planes): v_d(k) = 2 sum_kz
// Re(Bout[d][kz][j][i] e^{2 pi i kz k / nz}) for the segment's planes (accumulated in
// registers), rho from the cell's Philox draw, then rho v_d; component 0 also writes rho,
// the cell-centred and face field and the passive scalars.
template <class real>
__global__ __launch_bounds__(256) void k_ic_cells(BoxDesc B, IcArgs<real> a) {
    const int nx = B.n[0], ny = B.n[1], nz = B.n[2];
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x), j = (int)blockIdx.y;
    const int d = (int)(blockIdx.z % 3), k0 = (int)(blockIdx.z / 3) * ICSEG;
    if (i >= nx) return;
    const int K = a.K;
    double v[ICSEG];
#pragma unroll
    for (int s = 0; s < ICSEG; s++) v[s] = 0.0;
    if (a.Bv) {
        for (int kz = 0; kz <= K; kz++) {
            const size_t o = ((((size_t)d * (K + 1) + kz) * ny + j) * nx + i) * 2;
            const double br = a.Bv[o], bi = a.Bv[o + 1];
#pragma unroll
            for (int s = 0; s < ICSEG; s++) {
                const int k = k0 + s < nz ? k0 + s : nz - 1;
                const double* t = a.twz + ((size_t)k * (K + 1) + kz) * 2;
                v[s] = v[s] + (br * t[0] - bi * t[1]);
            }
        }
    }
#pragma unroll
    for (int s = 0; s < ICSEG; s++) {
        const int k = k0 + s;
        if (k >= nz) break;
        const long long cell = (long long)i + (long long)nx * ((long long)j + (long long)ny * k);
        const unsigned clo = (unsigned)cell, chi = (unsigned)(cell >> 32);
        double rho = a.rho0;
        if (a.sigma != 0.0) {
            unsigned o[4];
            philox4x32(clo, chi, 2u, 0u, a.seed_lo, a.seed_hi, o);
            const double g = box_muller(o[0], o[1]);
            rho = a.rho0 * exp(a.sigma * g - 0.5 * a.sigma * a.sigma);
        }
        const long long c = cell_index(B, i + 1, j + 1, k + 1);
        a.q[1 + d][c] = (real)(rho * (2.0 * v[s]));
        if (d == 0) {
            a.q[0][c] = (real)rho;
            for (int m = 0; m < 3; m++) {
                a.q[4 + m][c] = (real)a.b0[m];
                a.b[m][c] = a.b0[m];
            }
            for (int p = 0; p < a.npass; p++) {
                double u = 0.0;
                if (a.passive) {
                    unsigned o[4];
                    philox4x32(clo, chi, 3u + (unsigned)p, 0u, a.seed_lo, a.seed_hi, o);
                    // 53 random bits: [0, 1)
                    u = (double)((((unsigned long long)o[0] << 32) | o[1]) >> 11) * 1.1102230246251565e-16;
                }
                a.qp[p][c] = (real)u;
            }
        }
    }
}

void launch_ic_ypass(hipStream_t st, int nx, int ny, int K, const double* A, const double* twy, double* Bout) {
    hipLaunchKernelGGL(k_ic_ypass, dim3((unsigned)((nx + 255) / 256), (unsigned)ny), dim3(256), 0, st, nx, ny, K, A,
                       twy, Bout);
}

template <class real>
void launch_ic_cells(hipStream_t st, const BoxDesc& B, const IcArgs<real>& a) {
    const unsigned nseg = (unsigned)((B.n[2] + ICSEG - 1) / ICSEG);
    hipLaunchKernelGGL((k_ic_cells<real>), dim3((unsigned)((B.n[0] + 255) / 256), (unsigned)B.n[1], 3 * nseg),
                       dim3(256), 0, st, B, a);
}

template void launch_ic_cells<double>(hipStream_t, const BoxDesc&, const IcArgs<double>&);
template void launch_ic_cells<float>(hipStream_t, const BoxDesc&, const IcArgs<float>&);

}  // namespace imhd
