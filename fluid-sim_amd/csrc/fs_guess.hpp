// fs_guess.hpp -- learned initial pressure guess: a 4-level U-Net evaluated with
// BF16 matrix cores (v_mfma_f32_16x16x16_bf16, FP32 accumulation).
//
// The network maps (r / s_r, p_prev / s_p), where r = div u* - L p_prev is the
// divergence the previous pressure leaves and s_r, s_p are their maxima, to an
// increment: p_guess = p_prev + s_r h^2 o. Layers (channels 16/32/64/64, 3x3
// periodic convolutions, ReLU, 2x2 average pooling down, nearest-neighbour
// upsampling and skip concatenation [upsampled, skip] up, 1x1 output):
//
//   enc0a 2->16 (VALU)  enc0b 16->16            at n
//   enc1a 16->32 (pool) enc1b 32->32            at n/2
//   enc2a 32->64 (pool) enc2b 64->64            at n/4
//   bot_a 64->64 (pool) bot_b 64->64            at n/8
//   dec2a 128->64 (up + enc2b) dec2b 64->64     at n/4
//   dec1a 96->32 (up + enc1b)  dec1b 32->32     at n/2
//   dec0a 48->16 (up + enc0b)  dec0b 16->16     at n
//   out 1x1 16->1, p_guess = p_prev + s_r h^2 out
//
// Activations are NHWC BF16 (channels of a pixel contiguous). A 3x3 layer is an
// implicit GEMM per workgroup: 8 x 16 output pixels, all output channels; the
// input tile with its one-pixel periodic halo is staged in LDS (pooling and
// upsampling + concatenation happen in that load), weights are read as MFMA B
// fragments from L2. Each wavefront owns two 16-pixel rows.
#pragma once
#include "fs_kernels.hpp"

namespace fsk {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int GT_H = 8, GT_W = 16, GT_NT = 256;  // output tile and threads
constexpr int GT_LH = GT_H + 2, GT_LW = GT_W + 2;

enum ConvIn { CONV_DIRECT = 0, CONV_POOL = 1, CONV_UPCAT = 2 };

__device__ __forceinline__ uint16_t bf_bits(float f) { return st<bf16>(f).b; }
__device__ __forceinline__ float bf_val(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ int wrapm(int i, int m) { return i < 0 ? i + m : (i >= m ? i - m : i); }

// 3x3 periodic convolution + bias + ReLU, CIN -> COUT channels at an m x m level.
//   CONV_DIRECT: x1 holds CIN channels at m.
//   CONV_POOL:   x1 holds CIN channels at 2m; the input is its 2x2 average.
//   CONV_UPCAT:  channels [0, C2) are x2 (C2 channels at m/2, nearest-upsampled),
//                channels [C2, CIN) are x1 (CIN - C2 channels at m).
// w: BF16 [COUT][9][CIN] (tap = 3 (dy + 1) + dx + 1); bias FP32 [COUT].
template <int CIN, int COUT, int MODE, int C2>
__global__ void __launch_bounds__(GT_NT)
k_conv3x3(const uint16_t* __restrict__ x1, const uint16_t* __restrict__ x2, const uint16_t* __restrict__ w,
          const float* __restrict__ bias, uint16_t* __restrict__ y, int m) {
  static_assert(CIN % 16 == 0 && COUT % 16 == 0, "channel counts are multiples of 16");
  constexpr int PS = CIN + 4;  // LDS pixel stride (bf16): breaks bank conflicts of the A reads
  constexpr int CH4 = CIN / 4;
  constexpr int NB = COUT / 16;
  __shared__ __align__(16) uint16_t tile[GT_LH * GT_LW * PS];
  const int tiles_x = m / GT_W;
  const int tx0 = (blockIdx.x % tiles_x) * GT_W, ty0 = (blockIdx.x / tiles_x) * GT_H;
  // ---- stage the input tile (with halo) in LDS, 4 channels (8 bytes) per item
  for (int q = threadIdx.x; q < GT_LH * GT_LW * CH4; q += GT_NT) {
    int pix = q / CH4, cg = q - pix * CH4;
    int ly = pix / GT_LW, lx = pix - ly * GT_LW;
    int gy = wrapm(ty0 - 1 + ly, m), gx = wrapm(tx0 - 1 + lx, m);
    int c = cg * 4;
    uint2 v;
    if (MODE == CONV_DIRECT) {
      v = *(const uint2*)(x1 + ((size_t)gy * m + gx) * CIN + c);
    } else if (MODE == CONV_POOL) {
      const int m2 = 2 * m;
      float a[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int aa = 0; aa < 2; aa++) {
          uint2 t = *(const uint2*)(x1 + ((size_t)(2 * gy + b) * m2 + (2 * gx + aa)) * CIN + c);
          a[0] += bf_val(t.x & 0xffff);
          a[1] += bf_val(t.x >> 16);
          a[2] += bf_val(t.y & 0xffff);
          a[3] += bf_val(t.y >> 16);
        }
      v.x = (uint32_t)bf_bits(0.25f * a[0]) | ((uint32_t)bf_bits(0.25f * a[1]) << 16);
      v.y = (uint32_t)bf_bits(0.25f * a[2]) | ((uint32_t)bf_bits(0.25f * a[3]) << 16);
    } else {
      if (c < C2) {
        const int mh = m >> 1;
        v = *(const uint2*)(x2 + ((size_t)(gy >> 1) * mh + (gx >> 1)) * C2 + c);
      } else {
        constexpr int C1 = CIN - C2;
        v = *(const uint2*)(x1 + ((size_t)gy * m + gx) * C1 + (c - C2));
      }
    }
    *(uint2*)(tile + pix * PS + c) = v;
  }
  __syncthreads();
  // ---- implicit GEMM: M = 16 pixels of a row, N = 16 output channels, K = 16 input channels of one tap
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = (lane >> 4) * 4;
  f32x4 acc[2][NB];
#pragma unroll
  for (int r = 0; r < 2; r++)
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[r][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int tap = 0; tap < 9; tap++) {
    const int dy = tap / 3, dx = tap % 3;  // LDS offsets (halo included)
#pragma unroll
    for (int kc = 0; kc < CIN / 16; kc++) {
      bf16x4 bfrag[NB];
#pragma unroll
      for (int nb = 0; nb < NB; nb++) {
        const uint16_t* wp = w + ((size_t)(nb * 16 + li) * 9 + tap) * CIN + kc * 16 + lk;
        bfrag[nb] = *(const bf16x4*)wp;
      }
#pragma unroll
      for (int r = 0; r < 2; r++) {
        int row = 2 * wave + r;
        const uint16_t* ap = tile + ((row + dy) * GT_LW + (li + dx)) * PS + kc * 16 + lk;
        bf16x4 afrag = *(const bf16x4*)ap;
#pragma unroll
        for (int nb = 0; nb < NB; nb++)
          acc[r][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(afrag, bfrag[nb], acc[r][nb], 0, 0, 0);
      }
    }
  }
  // ---- epilogue: lane holds D[pixel 4*(lane/16) + q][channel lane%16] of each block
#pragma unroll
  for (int r = 0; r < 2; r++) {
    int py = ty0 + 2 * wave + r;
#pragma unroll
    for (int nb = 0; nb < NB; nb++) {
      int co = nb * 16 + li;
      float b = bias[co];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        int px = tx0 + lk + q;
        float val = fmaxf(acc[r][nb][q] + b, 0.f);
        y[((size_t)py * m + px) * COUT + co] = bf_bits(val);
      }
    }
  }
}

// First layer, 2 -> 16 channels at n, FP32 on the vector ALUs: the inputs are
// formed on the fly from the FP32 residual r and the FP64 previous pressure.
// w: FP32 [16][9][2]; scales[0] = max|r|, scales[1] = max|p_prev|.
__global__ void __launch_bounds__(256)
k_guess_first(const float* __restrict__ r, const double* __restrict__ pprev, const double* __restrict__ scales,
              const float* __restrict__ w, const float* __restrict__ bias, uint16_t* __restrict__ y, int n) {
  int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= n || j >= n) return;
  double sr = scales[0], sp = scales[1];
  float ir = sr > 0 ? (float)(1.0 / sr) : 0.f, ip = sp > 0 ? (float)(1.0 / sp) : 0.f;
  float in0[9], in1[9];
#pragma unroll
  for (int t = 0; t < 9; t++) {
    int yy = wrapm(j + t / 3 - 1, n), xx = wrapm(i + t % 3 - 1, n);
    size_t k = (size_t)yy * n + xx;
    in0[t] = r[k] * ir;
    in1[t] = (float)pprev[k] * ip;
  }
  uint32_t packed[8];
#pragma unroll
  for (int co = 0; co < 16; co += 2) {
    float a0 = bias[co], a1 = bias[co + 1];
#pragma unroll
    for (int t = 0; t < 9; t++) {
      a0 = fmaf(w[(co * 9 + t) * 2 + 0], in0[t], a0);
      a0 = fmaf(w[(co * 9 + t) * 2 + 1], in1[t], a0);
      a1 = fmaf(w[((co + 1) * 9 + t) * 2 + 0], in0[t], a1);
      a1 = fmaf(w[((co + 1) * 9 + t) * 2 + 1], in1[t], a1);
    }
    packed[co / 2] = (uint32_t)bf_bits(fmaxf(a0, 0.f)) | ((uint32_t)bf_bits(fmaxf(a1, 0.f)) << 16);
  }
  uint4* dst = (uint4*)(y + ((size_t)j * n + i) * 16);
  dst[0] = uint4{packed[0], packed[1], packed[2], packed[3]};
  dst[1] = uint4{packed[4], packed[5], packed[6], packed[7]};
}

// Output layer (1x1, 16 -> 1) and the guess: p = p_prev + s_r h^2 (b + w . x).
__global__ void __launch_bounds__(256)
k_guess_out(const uint16_t* __restrict__ x, const float* __restrict__ w, const float* __restrict__ bias,
            const double* __restrict__ scales, double* __restrict__ p, int n, double h2) {
  size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= (size_t)n * n) return;
  const uint4* src = (const uint4*)(x + k * 16);
  uint4 a = src[0], b = src[1];
  uint32_t v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  float o = bias[0];
#pragma unroll
  for (int c = 0; c < 8; c++) {
    o = fmaf(w[2 * c], bf_val(v[c] & 0xffff), o);
    o = fmaf(w[2 * c + 1], bf_val(v[c] >> 16), o);
  }
  p[k] += scales[0] * h2 * (double)o;
}

// Per-block max |p| (FP64) for the input scale.
__global__ void __launch_bounds__(256)
k_absmax64(const double* __restrict__ a, size_t count, double* __restrict__ part) {
  double m = 0.0;
  for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < count; k += (size_t)gridDim.x * 256)
    m = fmax(m, nan_to_inf(fabs(a[k])));
  __shared__ double sm[4];
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
}

// scales[0] = max of the residual partial row, scales[1] = max of the |p| partials.
__global__ void __launch_bounds__(256)
k_guess_scales(const double* __restrict__ rrow, int nr, const double* __restrict__ ppart, int np,
               double* __restrict__ scales) {
  double a = 0.0, b = 0.0;
  for (int k = threadIdx.x; k < nr; k += 256) a = fmax(a, rrow[k]);
  for (int k = threadIdx.x; k < np; k += 256) b = fmax(b, ppart[k]);
  __shared__ double sa[4], sb[4];
  a = wave_max(a);
  b = wave_max(b);
  if ((threadIdx.x & 63) == 0) {
    sa[threadIdx.x >> 6] = a;
    sb[threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    scales[0] = fmax(fmax(sa[0], sa[1]), fmax(sa[2], sa[3]));
    scales[1] = fmax(fmax(sb[0], sb[1]), fmax(sb[2], sb[3]));
  }
}

}  // namespace fsk
