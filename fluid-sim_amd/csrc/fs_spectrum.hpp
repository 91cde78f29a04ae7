// fs_spectrum.hpp -- isotropic kinetic-energy spectrum on the device
// (energySpectrum, src/core/fft.js:196-234).
//
// The centre velocities (cu, cv) are packed into one complex field z = cu + i cv, so
// one 2D FFT serves both components: with Z = FFT(z), |U_k|^2 + |V_k|^2 =
// (|Z_k|^2 + |Z_-k|^2) / 2, and since k and -k lie in the same shell the shell sum of
// 0.5 (|U_k|^2 + |V_k|^2) / N^2 equals the shell sum of |Z_k|^2 / (2 N^2).
// FP64 throughout. Rows, then columns, each an in-LDS radix-2 FFT of one line per
// workgroup with a twiddle table in LDS; the column pass bins its line into
// integer shells round(|k|) <= n/2 with exact integer bounds, in a fixed order, and
// a final pass adds the per-column partials in a fixed order.
#pragma once
#include "fs_kernels.hpp"

namespace fsk {

struct cplx {
  double re, im;
};

// In-LDS radix-2 FFT (forward, e^{-2 pi i k j / n}) of x[0..n) given in bit-reversed
// order; tw[m] = e^{-2 pi i m / n} for m < n/2.
__device__ __forceinline__ void lds_fft(cplx* x, const cplx* tw, int n) {
  for (int half = 1; half < n; half <<= 1) {
    int stride = n / (2 * half);
    for (int b = threadIdx.x; b < n / 2; b += blockDim.x) {
      int pos = b & (half - 1), grp = b / half;
      int i0 = grp * 2 * half + pos, i1 = i0 + half;
      cplx w = tw[pos * stride], a = x[i0], c = x[i1];
      cplx t = {w.re * c.re - w.im * c.im, w.re * c.im + w.im * c.re};
      x[i0] = {a.re + t.re, a.im + t.im};
      x[i1] = {a.re - t.re, a.im - t.im};
    }
    __syncthreads();
  }
}

__device__ __forceinline__ void lds_twiddles(cplx* tw, int n) {
  for (int m = threadIdx.x; m < n / 2; m += blockDim.x) {
    double s, c;
    sincospi(-2.0 * (double)m / (double)n, &s, &c);
    tw[m] = {c, s};
  }
}

__device__ __forceinline__ int bitrev(int i, int lg) { return (int)(__brev((unsigned)i) >> (32 - lg)); }

// Row pass: z(j, i) = cu + i cv at cell centres, FFT along i, into Z (row-major).
__global__ void __launch_bounds__(256)
k_spec_rows(const float* __restrict__ u, const float* __restrict__ v, cplx* __restrict__ Z, int n, int lg) {
  extern __shared__ __align__(16) unsigned char smem[];
  cplx* x = (cplx*)smem;
  cplx* tw = x + n;
  const int j = blockIdx.x, jp = j + 1 == n ? 0 : j + 1;
  lds_twiddles(tw, n);
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    int ip = i + 1 == n ? 0 : i + 1;
    size_t k = (size_t)j * n + i;
    double cu = 0.5 * ((double)u[k] + (double)u[(size_t)j * n + ip]);
    double cv = 0.5 * ((double)v[k] + (double)v[(size_t)jp * n + i]);
    x[bitrev(i, lg)] = {cu, cv};
  }
  __syncthreads();
  lds_fft(x, tw, n);
  for (int i = threadIdx.x; i < n; i += blockDim.x) Z[(size_t)j * n + i] = x[i];
}

// Column pass: FFT of column i, then per-shell sums of |Z|^2 / (2 N^2) into part[i][s].
__global__ void __launch_bounds__(256)
k_spec_cols(const cplx* __restrict__ Z, double* __restrict__ part, int n, int lg) {
  extern __shared__ __align__(16) unsigned char smem[];
  cplx* x = (cplx*)smem;
  cplx* tw = x + n;
  double* pw = (double*)(tw + n / 2);
  const int i = blockIdx.x, kmax = n / 2;
  lds_twiddles(tw, n);
  for (int j = threadIdx.x; j < n; j += blockDim.x) x[bitrev(j, lg)] = Z[(size_t)j * n + i];
  __syncthreads();
  lds_fft(x, tw, n);
  const double N = (double)n * (double)n, norm = 1.0 / (2.0 * N * N);
  for (int j = threadIdx.x; j < n; j += blockDim.x) pw[j] = (x[j].re * x[j].re + x[j].im * x[j].im) * norm;
  __syncthreads();
  const long kx = i <= n / 2 ? i : i - n;
  const long kx2 = kx * kx;
  for (int s = threadIdx.x; s <= kmax; s += blockDim.x) {
    // ky with (2s - 1)^2 <= 4 (kx^2 + ky^2) < (2s + 1)^2, i.e. round(|k|) = s
    const long lo4 = (2L * s - 1) * (2L * s - 1), hi4 = (2L * s + 1) * (2L * s + 1);
    double acc = 0.0;
    if (4 * kx2 < hi4) {
      long a = lo4 > 4 * kx2 ? (long)floor(sqrt((double)(lo4 - 4 * kx2) / 4.0)) - 1 : 0;
      if (a < 0) a = 0;
      for (long ky = a; ky <= n / 2; ky++) {
        long q = 4 * (kx2 + ky * ky);
        if (q >= hi4) break;
        if (q < lo4) continue;
        acc += pw[ky];                                 // +ky (row ky)
        if (ky != 0 && ky != n / 2) acc += pw[n - ky];  // -ky
      }
    }
    part[(size_t)i * (kmax + 1) + s] = acc;
  }
}

// E[s] = sum over columns of part[i][s], in column order.
__global__ void __launch_bounds__(256)
k_spec_final(const double* __restrict__ part, double* __restrict__ E, int n) {
  const int kmax = n / 2;
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > kmax) return;
  double acc = 0.0;
  for (int i = 0; i < n; i++) acc += part[(size_t)i * (kmax + 1) + s];
  E[s] = acc;
}

}  // namespace fsk
