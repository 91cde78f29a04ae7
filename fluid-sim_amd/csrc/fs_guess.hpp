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
// implicit GEMM per workgroup over a tile of 16-pixel rows and a block of output
// channels; the input tile with its one-pixel periodic halo (pooling, upsampling
// and concatenation happen in that load) and the weight slice are staged in LDS.
#pragma once
#include "fs_kernels.hpp"

namespace fsk {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int GT_W = 16, GT_NT = 256;  // tile width (one MFMA M-block) and threads

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
// A workgroup computes TH rows x 16 columns of pixels and NBW blocks of 16 output
// channels (blockIdx.y selects which); each wavefront owns TH/4 rows (one MFMA
// M-block each). The workgroup's weight slice and its input tile (with the
// one-pixel halo) are staged in LDS, so every B fragment feeds TH/4 MFMAs from LDS.
template <int CIN, int COUT, int MODE, int C2, int TH, int NBW>
__global__ void __launch_bounds__(GT_NT)
k_conv3x3(const uint16_t* __restrict__ x1, const uint16_t* __restrict__ x2, const uint16_t* __restrict__ w,
          const float* __restrict__ bias, uint16_t* __restrict__ y, int m) {
  static_assert(CIN % 16 == 0 && COUT % (16 * NBW) == 0 && TH % 4 == 0, "tile shape");
  constexpr int PS = CIN + 8;          // LDS pixel stride (bf16): 16-byte aligned, breaks A-read bank conflicts
  constexpr int WS = 9 * CIN + 4;      // LDS weight row stride (bf16): breaks B-read bank conflicts
  constexpr int CH8 = CIN / 8;
  constexpr int LH = TH + 2, LW = GT_W + 2;
  constexpr int MB = TH / 4;           // M-blocks (rows) per wavefront
  constexpr int OC = NBW * 16, OCP = OC + 8;  // output channels of the workgroup, staged stride
  __shared__ __align__(16) uint16_t tile[LH * LW * PS];
  __shared__ __align__(16) uint16_t wl[NBW * 16 * WS];
  __shared__ __align__(16) uint16_t otile[TH * GT_W * OCP];
  const int tiles_x = m / GT_W, ntiles = tiles_x * (m / TH);
  const int co0 = blockIdx.y * NBW * 16;
  // ---- weights of this workgroup's output channels, once for all its tiles
  {
    constexpr int RW8 = 9 * CIN / 8, WI = NBW * 16 * RW8;  // 16-byte items per row, in total
    auto wsrc = [&](int q) { return (const uint4*)(w + (size_t)(co0 + q / RW8) * 9 * CIN + (q % RW8) * 8); };
    auto wdst = [&](int q) { return (uint2*)(wl + (q / RW8) * WS + (q % RW8) * 8); };
    for (int q0 = threadIdx.x; q0 < WI; q0 += 4 * GT_NT) {
      const int q1 = q0 + GT_NT, q2 = q1 + GT_NT, q3 = q2 + GT_NT;
      uint4 v0 = *wsrc(q0), v1, v2, v3;
      if (q1 < WI) v1 = *wsrc(q1);
      if (q2 < WI) v2 = *wsrc(q2);
      if (q3 < WI) v3 = *wsrc(q3);
      // the weight rows are 8-byte aligned in LDS (stride 9 CIN + 4), so two 8-byte stores
      wdst(q0)[0] = uint2{v0.x, v0.y};
      wdst(q0)[1] = uint2{v0.z, v0.w};
      if (q1 < WI) { wdst(q1)[0] = uint2{v1.x, v1.y}; wdst(q1)[1] = uint2{v1.z, v1.w}; }
      if (q2 < WI) { wdst(q2)[0] = uint2{v2.x, v2.y}; wdst(q2)[1] = uint2{v2.z, v2.w}; }
      if (q3 < WI) { wdst(q3)[0] = uint2{v3.x, v3.y}; wdst(q3)[1] = uint2{v3.z, v3.w}; }
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = (lane >> 4) * 4;
  float bv[NBW];
#pragma unroll
  for (int nb = 0; nb < NBW; nb++) bv[nb] = bias[co0 + nb * 16 + li];
  // ---- input tiles (with halo), 8 channels (16 bytes) per item; the next tile's
  // items are loaded into registers while the current one is multiplied
  constexpr int TI = LH * LW * CH8, TK = (TI + GT_NT - 1) / GT_NT;
  auto load_item = [&](int q, int tx0, int ty0) -> uint4 {
    int pix = q / CH8, cg = q - pix * CH8;
    int ly = pix / LW, lx = pix - ly * LW;
    int gy = wrapm(ty0 - 1 + ly, m), gx = wrapm(tx0 - 1 + lx, m);
    int c = cg * 8;
    if (MODE == CONV_DIRECT) return *(const uint4*)(x1 + ((size_t)gy * m + gx) * CIN + c);
    if (MODE == CONV_POOL) {
      const int m2 = 2 * m;
      float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f, a4 = 0.f, a5 = 0.f, a6 = 0.f, a7 = 0.f;
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int aa = 0; aa < 2; aa++) {
          uint4 tt = *(const uint4*)(x1 + ((size_t)(2 * gy + b) * m2 + (2 * gx + aa)) * CIN + c);
          a0 += bf_val(tt.x & 0xffff);
          a1 += bf_val(tt.x >> 16);
          a2 += bf_val(tt.y & 0xffff);
          a3 += bf_val(tt.y >> 16);
          a4 += bf_val(tt.z & 0xffff);
          a5 += bf_val(tt.z >> 16);
          a6 += bf_val(tt.w & 0xffff);
          a7 += bf_val(tt.w >> 16);
        }
      auto pk = [](float lo, float hi) {
        return (uint32_t)bf_bits(0.25f * lo) | ((uint32_t)bf_bits(0.25f * hi) << 16);
      };
      return uint4{pk(a0, a1), pk(a2, a3), pk(a4, a5), pk(a6, a7)};
    }
    if (c < C2) return *(const uint4*)(x2 + ((size_t)(gy >> 1) * (m >> 1) + (gx >> 1)) * C2 + c);
    return *(const uint4*)(x1 + ((size_t)gy * m + gx) * (CIN - C2) + (c - C2));
  };
  uint4 pre[TK];
  auto prefetch = [&](int t) {
    const int tx0 = (t % tiles_x) * GT_W, ty0 = (t / tiles_x) * TH;
#pragma unroll
    for (int k = 0; k < TK; k++) {
      int q = threadIdx.x + k * GT_NT;
      if (q < TI) pre[k] = load_item(q, tx0, ty0);
    }
  };
  if (blockIdx.x < ntiles) prefetch(blockIdx.x);
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int tx0 = (t % tiles_x) * GT_W, ty0 = (t / tiles_x) * TH;
#pragma unroll
    for (int k = 0; k < TK; k++) {
      int q = threadIdx.x + k * GT_NT;
      if (q < TI) {
        int pix = q / CH8, cg = q - pix * CH8;
        *(uint4*)(tile + pix * PS + cg * 8) = pre[k];
      }
    }
    __syncthreads();
    if (t + (int)gridDim.x < ntiles) prefetch(t + gridDim.x);
    // ---- implicit GEMM: M = 16 pixels of a row, N = 16 output channels, K = 16 input
    // channels of one tap. A fragment of LDS row lr (shifted by dx) serves output rows
    // lr, lr-1, lr-2 through taps dy = 0, 1, 2, so each is read once per (dx, kc).
    f32x4 acc[MB][NBW];
#pragma unroll
    for (int r = 0; r < MB; r++)
#pragma unroll
      for (int nb = 0; nb < NBW; nb++) acc[r][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kc = 0; kc < CIN / 16; kc++) {
#pragma unroll
      for (int dx = 0; dx < 3; dx++) {
        bf16x4 bfrag[3][NBW];
#pragma unroll
        for (int dy = 0; dy < 3; dy++)
#pragma unroll
          for (int nb = 0; nb < NBW; nb++)
            bfrag[dy][nb] = *(const bf16x4*)(wl + (nb * 16 + li) * WS + (dy * 3 + dx) * CIN + kc * 16 + lk);
#pragma unroll
        for (int lr = 0; lr < MB + 2; lr++) {
          bf16x4 afrag = *(const bf16x4*)(tile + ((wave * MB + lr) * LW + (li + dx)) * PS + kc * 16 + lk);
#pragma unroll
          for (int dy = 0; dy < 3; dy++) {
            const int r = lr - dy;
            if (r >= 0 && r < MB) {
#pragma unroll
              for (int nb = 0; nb < NBW; nb++)
                acc[r][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(afrag, bfrag[dy][nb], acc[r][nb], 0, 0, 0);
            }
          }
        }
      }
    }
    // ---- epilogue: lane holds D[pixel 4*(lane/16) + q][channel lane%16] of each block;
    // staged through LDS so that each pixel's channels leave as 16-byte stores
#pragma unroll
    for (int r = 0; r < MB; r++) {
      int row = wave * MB + r;
#pragma unroll
      for (int nb = 0; nb < NBW; nb++)
#pragma unroll
        for (int q = 0; q < 4; q++)
          otile[(row * GT_W + lk + q) * OCP + nb * 16 + li] = bf_bits(fmaxf(acc[r][nb][q] + bv[nb], 0.f));
    }
    __syncthreads();
    constexpr int OI = TH * GT_W * (OC / 8);
    for (int q = threadIdx.x; q < OI; q += GT_NT) {
      int pix = q / (OC / 8), cg = q - pix * (OC / 8);
      int py = ty0 + pix / GT_W, px = tx0 + (pix % GT_W);
      *(uint4*)(y + ((size_t)py * m + px) * COUT + co0 + cg * 8) = *(const uint4*)(otile + pix * OCP + cg * 8);
    }
    __syncthreads();  // the next tile overwrites the staged input
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
