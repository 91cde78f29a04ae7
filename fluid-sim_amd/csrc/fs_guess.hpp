// fs_guess.hpp -- learned initial pressure guess: a 4-level U-Net evaluated with
// BF16 matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulation).
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
// Inference runs enc0a and enc0b as one kernel (k_guess_enc0) and the 1x1 output
// layer in dec0b's epilogue (EPI_OUT), so neither enc0a's nor dec0b's activations
// are stored; the trainer, which needs them, runs k_guess_first and EPI_RELU.
#pragma once
#include "fs_kernels.hpp"

namespace fsk {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8v __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
typedef float f32x2v __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

constexpr int GT_W = 16, GT_NT = 256;  // tile width (the 16 pixels of one MFMA block) and threads

enum ConvIn { CONV_DIRECT = 0, CONV_POOL = 1, CONV_UPCAT = 2 };

__device__ __forceinline__ uint16_t bf_bits(float f) { return st<bf16>(f).b; }
__device__ __forceinline__ float bf_val(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ int wrapm(int i, int m) { return i < 0 ? i + m : (i >= m ? i - m : i); }

// One 16-byte item (8 channels of one pixel) of a 3x3 layer's input at pixel (gy, gx)
// of an m x m level, channels [c, c + 8):
//   CONV_DIRECT: x1 holds CIN channels at m.
//   CONV_POOL:   x1 holds CIN channels at 2m; the input is its 2x2 average, rounded to BF16.
//   CONV_UPCAT:  channels [0, C2) are x2 (C2 channels at m/2, nearest-upsampled),
//                channels [C2, CIN) are x1 (CIN - C2 channels at m).
// x1 and x2 point at the image's own tensors.
template <int CIN, int MODE, int C2>
__device__ __forceinline__ uint4 conv_in_item(const uint16_t* __restrict__ x1, const uint16_t* __restrict__ x2, int m,
                                              int gy, int gx, int c) {
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
}

// Element counts of one image's input tensors of a 3x3 layer at level m.
template <int CIN, int MODE, int C2>
__device__ __forceinline__ size_t conv_x1_image(int m) {
  return MODE == CONV_POOL ? (size_t)4 * m * m * CIN : (size_t)m * m * (CIN - C2);
}
template <int C2>
__device__ __forceinline__ size_t conv_x2_image(int m) {
  return (size_t)(m >> 1) * (m >> 1) * C2;
}

// What a 3x3 layer does with its accumulators (the last three are the backward pass,
// the kernel then run with the transposed, tap-flipped weights on the gradient):
//   EPI_RELU:   forward, y = BF16(relu(acc + bias)).
//   EPI_MASK:   y = BF16(acc) where the activation e.act > 0, else 0 (all at m).
//   EPI_UNPOOL: the accumulators are the gradient of a 2x2-averaged input; each of the
//               four fine pixels f gets g = acc / 4 + e.gin[f] (FP32, may be null) and
//               y[f] = BF16(g) where e.act[f] > 0, else 0 (y, e.act, e.gin at 2m).
//   EPI_UPCAT:  the accumulators are the gradient of [up(coarse), skip]: channels
//               [0, OS) are summed over each 2x2 block into the coarse pixel and
//               y = BF16(sum) where e.act > 0 (y, e.act: OS channels at m/2); channels
//               [OS, COUT) go unrounded to e.gout (FP32, COUT - OS channels at m).
//   EPI_OUT:    the last 3x3 layer of inference (COUT = 16) with the 1x1 output layer
//               after it: the BF16 activations relu(acc + bias) are not stored; each
//               pixel's o = e.bout + sum_c e.wout[c] a_c (channels in order) updates the
//               pressure, e.p += e.scales[0] e.h2 o (FP64, one image).
enum ConvEpi { EPI_RELU = 0, EPI_MASK = 1, EPI_UNPOOL = 2, EPI_UPCAT = 3, EPI_OUT = 4 };
struct ConvEpiArgs {
  const uint16_t* act;
  const float* gin;
  float* gout;
  const float* wout;  // EPI_OUT: 1x1 weights [16] and bias [1]
  const float* bout;
  const double* scales;
  double* p;
  double h2;
};

__device__ __forceinline__ uint32_t bf_pair_masked(float lo, float hi, uint32_t act2) {
  uint32_t a = (act2 & 0x8000u) || !(act2 & 0x7fffu) ? 0u : (uint32_t)bf_bits(lo);
  uint32_t b = (act2 & 0x80000000u) || !(act2 & 0x7fff0000u) ? 0u : (uint32_t)bf_bits(hi);
  return a | (b << 16);
}

// Two FP32 values to a packed BF16 pair, round to nearest even (v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint32_t bf_pack(float lo, float hi) {
  bf16x2v b = __builtin_convertvector(f32x2v{lo, hi}, bf16x2v);
  return __builtin_bit_cast(uint32_t, b);
}

// LDS bytes of k_conv3x3: the input tile with its halo and the weight slice.
template <int CIN, int TH, int NBW>
constexpr int conv_lds_bytes() {
  return 2 * ((TH + 2) * (GT_W + 2) * (CIN + 8) + NBW * 16 * (9 * CIN + 8));
}

// 3x3 periodic convolution, CIN -> COUT channels at an m x m level, over nimg images
// (tensors of consecutive images follow each other). Input modes as in conv_in_item.
// w: BF16 [COUT][9][CIN] (tap = 3 (dy + 1) + dx + 1); bias FP32 [COUT] (EPI_RELU).
//
// A workgroup of 4 WC wavefronts computes TH rows x 16 columns of pixels and NBW
// blocks of 16 output channels (blockIdx.y selects which); wavefront (wr, wc) owns
// rows [wr TH/4, (wr + 1) TH/4) and channel blocks [wc NBW/WC, (wc + 1) NBW/WC). The
// weight slice stays in LDS for all of the workgroup's tiles; each tile is staged in
// LDS with its one-pixel periodic halo (the next tile's loads are issued while the
// current one is multiplied).
//
// Implicit GEMM per tap: D[channel][pixel] += W[channel][k] X[k][pixel], with the
// weights as the MFMA A operand and 16 pixels of a row as B, on
// v_mfma_f32_16x16x32_bf16 (32 input channels per step; a 16-channel remainder on
// v_mfma_f32_16x16x16_bf16). Lane l holds channels 4 (l / 16) + q of pixel l % 16, so
// the epilogue writes 4 channels of a pixel per lane straight to memory. An LDS row
// of pixels (shifted by dx) serves output rows lr, lr - 1, lr - 2 through taps
// dy = 0, 1, 2, so it is read once per (dx, channel step).
template <int CIN, int COUT, int MODE, int C2, int TH, int NBW, int EPI = EPI_RELU, int OS = 0, int WC = 1>
__global__ void __launch_bounds__(GT_NT * WC)
k_conv3x3(const uint16_t* __restrict__ x1, const uint16_t* __restrict__ x2, const uint16_t* __restrict__ w,
          const float* __restrict__ bias, uint16_t* __restrict__ y, int m, int nimg, ConvEpiArgs e) {
  constexpr int NT = GT_NT * WC, R = TH / 4, N = NBW / WC;
  static_assert(CIN % 16 == 0 && COUT % (16 * NBW) == 0 && TH % 4 == 0 && NBW % WC == 0, "tile shape");
  static_assert(EPI != EPI_UPCAT || (OS % (16 * NBW) == 0 && OS < COUT && R % 2 == 0), "split on a channel block");
  static_assert(EPI != EPI_OUT || (COUT == 16 && NBW == 1), "the output layer reads 16 channels");
  constexpr int PS = CIN + 8;      // LDS pixel stride (bf16): 16-byte rows, 16 pixels hit distinct banks
  constexpr int WS = 9 * CIN + 8;  // LDS weight row stride (bf16): the same for 16 channel rows
  constexpr int LH = TH + 2, LW = GT_W + 2;
  constexpr int OC = NBW * 16;
  constexpr int K32 = CIN / 32, K16 = (CIN % 32) / 16;
  __shared__ __align__(16) uint16_t tile[LH * LW * PS];
  __shared__ __align__(16) uint16_t wl[OC * WS];
  const int tiles_x = m / GT_W, tpi = tiles_x * (m / TH), ntiles = tpi * nimg;
  const int co0 = blockIdx.y * OC;
  // ---- weights of this workgroup's output channels, once for all its tiles
  {
    constexpr int RW8 = 9 * CIN / 8, WI = OC * RW8;  // 16-byte items per row, in total
#pragma unroll 4
    for (int q = threadIdx.x; q < WI; q += NT) {
      const int o = q / RW8, k8 = q - o * RW8;
      *(uint4*)(wl + o * WS + k8 * 8) = *(const uint4*)(w + (size_t)(co0 + o) * 9 * CIN + k8 * 8);
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = wave & 3, wc = wave >> 2;
  const int li = lane & 15, lg = lane >> 4;
  const int cb0 = wc * N;  // first channel block of this wavefront
  float bv[N][4];
#pragma unroll
  for (int nb = 0; nb < N; nb++)
#pragma unroll
    for (int q = 0; q < 4; q++)
      bv[nb][q] = EPI == EPI_RELU || EPI == EPI_OUT ? bias[co0 + (cb0 + nb) * 16 + 4 * lg + q] : 0.f;
  // ---- input tile items (8 channels, 16 bytes). An upsample-concatenation input is
  // staged in two passes, the C2 upsampled channels (part A) and the rest (part B), so
  // each item loop reads one tensor; the other modes have part A only.
  constexpr int CA = MODE == CONV_UPCAT ? C2 : CIN, CB = CIN - CA;
  constexpr int IA = LH * LW * (CA / 8), IB = LH * LW * (CB / 8);
  constexpr int KA = (IA + NT - 1) / NT, KB = CB > 0 ? (IB + NT - 1) / NT : 0;
  constexpr int CB8 = CB > 0 ? CB / 8 : 1;  // items per pixel of part B
  // staged in native vectors (the HIP uint4 struct made the compiler keep them in scratch)
  u32x4 preA[KA], preB[KB > 0 ? KB : 1];
  auto prefetch = [&](int t) {
    const int z = t / tpi, tt = t - z * tpi;
    const int tx0 = (tt % tiles_x) * GT_W, ty0 = (tt / tiles_x) * TH;
    const uint16_t* x1z = x1 + z * conv_x1_image<CIN, MODE, C2>(m);
#pragma unroll
    for (int k = 0; k < KA; k++) {
      const int q = threadIdx.x + k * NT;
      if (q < IA) {
        const int pix = q / (CA / 8), cg = q - pix * (CA / 8);
        const int ly = pix / LW, lx = pix - ly * LW;
        const int gy = wrapm(ty0 - 1 + ly, m), gx = wrapm(tx0 - 1 + lx, m);
        if (MODE == CONV_UPCAT)
          preA[k] = *(const u32x4*)(x2 + z * conv_x2_image<C2>(m) + ((size_t)(gy >> 1) * (m >> 1) + (gx >> 1)) * C2 +
                                    cg * 8);
        else {
          const uint4 v = conv_in_item<CIN, MODE, C2>(x1z, nullptr, m, gy, gx, cg * 8);
          preA[k] = u32x4{v.x, v.y, v.z, v.w};
        }
      }
    }
#pragma unroll
    for (int k = 0; k < KB; k++) {
      const int q = threadIdx.x + k * NT;
      if (q < IB) {
        const int pix = q / CB8, cg = q - pix * CB8;
        const int ly = pix / LW, lx = pix - ly * LW;
        const int gy = wrapm(ty0 - 1 + ly, m), gx = wrapm(tx0 - 1 + lx, m);
        preB[k] = *(const u32x4*)(x1z + ((size_t)gy * m + gx) * CB + cg * 8);
      }
    }
  };
  if ((int)blockIdx.x < ntiles) prefetch(blockIdx.x);
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int z = t / tpi, tt = t - z * tpi;
    const int tx0 = (tt % tiles_x) * GT_W, ty0 = (tt / tiles_x) * TH;
#pragma unroll
    for (int k = 0; k < KA; k++) {
      const int q = threadIdx.x + k * NT;
      if (q < IA) {
        const int pix = q / (CA / 8), cg = q - pix * (CA / 8);
        *(u32x4*)(tile + pix * PS + cg * 8) = preA[k];
      }
    }
#pragma unroll
    for (int k = 0; k < KB; k++) {
      const int q = threadIdx.x + k * NT;
      if (q < IB) {
        const int pix = q / CB8, cg = q - pix * CB8;
        *(u32x4*)(tile + pix * PS + CA + cg * 8) = preB[k];
      }
    }
    __syncthreads();
    if (t + (int)gridDim.x < ntiles) prefetch(t + gridDim.x);
    // EPI_OUT: the pressure this tile updates, loaded before the multiply
    double pold[EPI == EPI_OUT ? R : 1];
    if (EPI == EPI_OUT) {
#pragma unroll
      for (int r = 0; r < R; r++) pold[r] = e.p[(size_t)(ty0 + wr * R + r) * m + tx0 + li];
    }
    f32x4 acc[R][N];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
      for (int nb = 0; nb < N; nb++) acc[r][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
    const uint16_t* wrow = wl + (cb0 * 16 + li) * WS + lg * 8;
    const uint16_t* trow = tile + (wr * R * LW + li) * PS + lg * 8;
#pragma unroll
    for (int kc = 0; kc < K32; kc++) {
#pragma unroll
      for (int dx = 0; dx < 3; dx++) {
        bf16x8v af[3][N];
#pragma unroll
        for (int dy = 0; dy < 3; dy++)
#pragma unroll
          for (int nb = 0; nb < N; nb++)
            af[dy][nb] = *(const bf16x8v*)(wrow + nb * 16 * WS + (dy * 3 + dx) * CIN + kc * 32);
#pragma unroll
        for (int lr = 0; lr < R + 2; lr++) {
          const bf16x8v bx = *(const bf16x8v*)(trow + (lr * LW + dx) * PS + kc * 32);
#pragma unroll
          for (int dy = 0; dy < 3; dy++) {
            const int r = lr - dy;
            if (r >= 0 && r < R) {
#pragma unroll
              for (int nb = 0; nb < N; nb++)
                acc[r][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[dy][nb], bx, acc[r][nb], 0, 0, 0);
            }
          }
        }
      }
    }
    if constexpr (K16 > 0) {
      // the last 16 channels: lane l holds k = 4 (l / 16) + j
      const uint16_t* wrow16 = wl + (cb0 * 16 + li) * WS + K32 * 32 + lg * 4;
      const uint16_t* trow16 = tile + (wr * R * LW + li) * PS + K32 * 32 + lg * 4;
#pragma unroll
      for (int dx = 0; dx < 3; dx++) {
        bf16x4 af[3][N];
#pragma unroll
        for (int dy = 0; dy < 3; dy++)
#pragma unroll
          for (int nb = 0; nb < N; nb++) af[dy][nb] = *(const bf16x4*)(wrow16 + nb * 16 * WS + (dy * 3 + dx) * CIN);
#pragma unroll
        for (int lr = 0; lr < R + 2; lr++) {
          const bf16x4 bx = *(const bf16x4*)(trow16 + (lr * LW + dx) * PS);
#pragma unroll
          for (int dy = 0; dy < 3; dy++) {
            const int r = lr - dy;
            if (r >= 0 && r < R) {
#pragma unroll
              for (int nb = 0; nb < N; nb++)
                acc[r][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[dy][nb], bx, acc[r][nb], 0, 0, 0);
            }
          }
        }
      }
    }
    __syncthreads();  // the next tile overwrites the staged input
    // ---- epilogue: lane holds channels c .. c + 3 of pixel (py, px) of each (row, block)
    const int px = tx0 + li;
    if (EPI == EPI_RELU || EPI == EPI_MASK) {
      const size_t base = (size_t)z * m * m * COUT;
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int py = ty0 + wr * R + r;
#pragma unroll
        for (int nb = 0; nb < N; nb++) {
          const size_t k = base + ((size_t)py * m + px) * COUT + co0 + (cb0 + nb) * 16 + 4 * lg;
          const f32x4 a = acc[r][nb];
          uint2 o;
          if (EPI == EPI_RELU) {
            o = uint2{bf_pack(fmaxf(a[0] + bv[nb][0], 0.f), fmaxf(a[1] + bv[nb][1], 0.f)),
                      bf_pack(fmaxf(a[2] + bv[nb][2], 0.f), fmaxf(a[3] + bv[nb][3], 0.f))};
          } else {
            const uint2 act = *(const uint2*)(e.act + k);
            o = uint2{bf_pair_masked(a[0], a[1], act.x), bf_pair_masked(a[2], a[3], act.y)};
          }
          *(uint2*)(y + k) = o;
        }
      }
    } else if (EPI == EPI_OUT) {
      // lane l % 16 gathers its pixel's 16 channels from lanes l % 16 + 16 j, in order
      const double sc = e.scales[0] * e.h2;
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int py = ty0 + wr * R + r;
        float v[4];
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = bf_val(bf_bits(fmaxf(acc[r][0][q] + bv[0][q], 0.f)));
        float a[16];
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
          for (int q = 0; q < 4; q++) a[4 * j + q] = __shfl(v[q], li + 16 * j);
        if (lg == 0) {
          float o = e.bout[0];
#pragma unroll
          for (int c = 0; c < 16; c++) o = fmaf(e.wout[c], a[c], o);
          e.p[(size_t)py * m + px] = pold[r] + sc * (double)o;
        }
      }
    } else if (EPI == EPI_UNPOOL) {
      // the gradient of a 2x2-averaged input: a quarter to each fine pixel, plus the skip
      const int m2 = 2 * m;
      const size_t base = (size_t)z * m2 * m2 * COUT;
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int py = ty0 + wr * R + r;
#pragma unroll
        for (int nb = 0; nb < N; nb++) {
          const f32x4 a = acc[r][nb];
#pragma unroll
          for (int f = 0; f < 4; f++) {
            const size_t k =
                base + ((size_t)(2 * py + (f >> 1)) * m2 + 2 * px + (f & 1)) * COUT + co0 + (cb0 + nb) * 16 + 4 * lg;
            float g[4];
            if (e.gin) {
              const float4 g0 = *(const float4*)(e.gin + k);
              g[0] = 0.25f * a[0] + g0.x;
              g[1] = 0.25f * a[1] + g0.y;
              g[2] = 0.25f * a[2] + g0.z;
              g[3] = 0.25f * a[3] + g0.w;
            } else {
#pragma unroll
              for (int q = 0; q < 4; q++) g[q] = 0.25f * a[q];
            }
            const uint2 act = *(const uint2*)(e.act + k);
            *(uint2*)(y + k) = uint2{bf_pair_masked(g[0], g[1], act.x), bf_pair_masked(g[2], g[3], act.y)};
          }
        }
      }
    } else if (EPI == EPI_UPCAT) {
      if (co0 < OS) {
        // the upsampled channels: sum over each 2x2 block, (a00 + a01) + (a10 + a11); the
        // column pair sits in lanes l, l ^ 1 and the row pair in rows r, r + 1
        const int mc = m >> 1;
        const size_t base = (size_t)z * mc * mc * OS;
#pragma unroll
        for (int r = 0; r < R; r += 2) {
          const int cy = (ty0 + wr * R + r) >> 1, cx = px >> 1;
#pragma unroll
          for (int nb = 0; nb < N; nb++) {
            float g[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
              const float a0 = acc[r][nb][q], a1 = acc[r + 1][nb][q];
              g[q] = (a0 + __shfl_xor(a0, 1)) + (a1 + __shfl_xor(a1, 1));
            }
            if ((li & 1) == 0) {
              const size_t k = base + ((size_t)cy * mc + cx) * OS + co0 + (cb0 + nb) * 16 + 4 * lg;
              const uint2 act = *(const uint2*)(e.act + k);
              *(uint2*)(y + k) = uint2{bf_pair_masked(g[0], g[1], act.x), bf_pair_masked(g[2], g[3], act.y)};
            }
          }
        }
      } else {
        // the skip channels: FP32, unrounded
        constexpr int CS = COUT - OS;
        const size_t base = (size_t)z * m * m * CS;
#pragma unroll
        for (int r = 0; r < R; r++) {
          const int py = ty0 + wr * R + r;
#pragma unroll
          for (int nb = 0; nb < N; nb++) {
            const size_t k = base + ((size_t)py * m + px) * CS + (co0 - OS) + (cb0 + nb) * 16 + 4 * lg;
            const f32x4 a = acc[r][nb];
            *(float4*)(e.gout + k) = float4{a[0], a[1], a[2], a[3]};
          }
        }
      }
    }
  }
}

// First layer, 2 -> 16 channels at n, FP32 on the vector ALUs: the inputs are
// formed on the fly from the residual r (stored as S: FP32, or BF16 under BF16
// multigrid storage) and the FP64 previous pressure.
// w: FP32 [16][9][2]; scales[0] = max|r|, scales[1] = max|p_prev|. Image blockIdx.z of
// a batch reads r, pprev at offset z n^2, its scales at 2z and writes y at z n^2 16.
template <typename S>
__global__ void __launch_bounds__(256)
k_guess_first(const S* __restrict__ r, const double* __restrict__ pprev, const double* __restrict__ scales,
              const float* __restrict__ w, const float* __restrict__ bias, uint16_t* __restrict__ y, int n) {
  int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= n || j >= n) return;
  {
    const size_t zo = (size_t)blockIdx.z * n * n;
    r += zo;
    pprev += zo;
    y += zo * 16;
    scales += 2 * blockIdx.z;
  }
  double sr = scales[0], sp = scales[1];
  float ir = sr > 0 ? (float)(1.0 / sr) : 0.f, ip = sp > 0 ? (float)(1.0 / sp) : 0.f;
  float in0[9], in1[9];
#pragma unroll
  for (int t = 0; t < 9; t++) {
    int yy = wrapm(j + t / 3 - 1, n), xx = wrapm(i + t % 3 - 1, n);
    size_t k = (size_t)yy * n + xx;
    in0[t] = ld(r[k]) * ir;
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

// The first two layers of inference in one kernel: enc0a (2 -> 16, FP32 on the vector
// ALUs, exactly as k_guess_first) evaluated into LDS on the tile and its one-pixel
// halo, then enc0b (16 -> 16, BF16 MFMA, as k_conv3x3) from there; enc0a's output never
// reaches memory. Per tile of TH x 16 pixels: the scaled inputs (r / s_r, p_prev / s_p)
// on a (TH + 4) x 20 window, enc0a on the (TH + 2) x 18 pixels around the tile, the
// 3x3 convolution, y = BF16(relu(acc + b1)) (16 channels at n, one image).
// w0: FP32 [16][9][2], b0 [16]; w1: BF16 [16][9][16], b1 [16].
template <typename S, int TH>
__global__ void __launch_bounds__(GT_NT)
k_guess_enc0(const S* __restrict__ r, const double* __restrict__ pprev, const double* __restrict__ scales,
             const float* __restrict__ w0, const float* __restrict__ b0, const uint16_t* __restrict__ w1,
             const float* __restrict__ b1, uint16_t* __restrict__ y, int n) {
  constexpr int R = TH / 4, PS = 24, WS = 9 * 16 + 8;
  constexpr int LH = TH + 2, LW = GT_W + 2, IH = TH + 4, IW = GT_W + 4;
  static_assert(TH % 4 == 0 && LH * LW <= GT_NT, "one enc0a pixel per thread");
  __shared__ __align__(16) float2 in2[IH * IW];
  __shared__ __align__(16) uint16_t tile[LH * LW * PS];
  __shared__ __align__(16) uint16_t wl[16 * WS];
  for (int q = threadIdx.x; q < 16 * 18; q += GT_NT) {
    const int o = q / 18, k8 = q - o * 18;
    *(uint4*)(wl + o * WS + k8 * 8) = *(const uint4*)(w1 + (size_t)o * 144 + k8 * 8);
  }
  const double sr = scales[0], sp = scales[1];
  const float ir = sr > 0 ? (float)(1.0 / sr) : 0.f, ip = sp > 0 ? (float)(1.0 / sp) : 0.f;
  const int lane = threadIdx.x & 63, wr = threadIdx.x >> 6;
  const int li = lane & 15, lg = lane >> 4;
  float bv[4];
#pragma unroll
  for (int q = 0; q < 4; q++) bv[q] = b1[4 * lg + q];
  const int tiles_x = n / GT_W, ntiles = tiles_x * (n / TH);
  // ---- scaled inputs on the window (origin 2 cells up and left of the tile); the next
  // tile's window is loaded into registers while the current tile is multiplied
  constexpr int CNT = IH * IW, K = (CNT + GT_NT - 1) / GT_NT;
  float wa[K], wb[K];
  auto window = [&](int t) {
    const int tx0 = (t % tiles_x) * GT_W, ty0 = (t / tiles_x) * TH;
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int q = threadIdx.x + k * GT_NT;
      if (q < CNT) {
        const int ly = q / IW, lx = q - ly * IW;
        const size_t g = (size_t)wrapm(ty0 - 2 + ly, n) * n + wrapm(tx0 - 2 + lx, n);
        wa[k] = ld(r[g]) * ir;
        wb[k] = (float)pprev[g] * ip;
      }
    }
  };
  if ((int)blockIdx.x < ntiles) window(blockIdx.x);
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int tx0 = (t % tiles_x) * GT_W, ty0 = (t / tiles_x) * TH;
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int q = threadIdx.x + k * GT_NT;
      if (q < CNT) in2[q] = make_float2(wa[k], wb[k]);
    }
    __syncthreads();
    // ---- enc0a on the tile and its halo: one pixel per thread, 16 channels
    if (threadIdx.x < LH * LW) {
      const int ly = threadIdx.x / LW, lx = threadIdx.x - ly * LW;
      float in0[9], in1[9];
#pragma unroll
      for (int tp = 0; tp < 9; tp++) {
        const float2 v = in2[(ly + tp / 3) * IW + lx + tp % 3];
        in0[tp] = v.x;
        in1[tp] = v.y;
      }
      uint32_t packed[8];
#pragma unroll
      for (int co = 0; co < 16; co += 2) {
        float a0 = b0[co], a1 = b0[co + 1];
#pragma unroll
        for (int tp = 0; tp < 9; tp++) {
          a0 = fmaf(w0[(co * 9 + tp) * 2 + 0], in0[tp], a0);
          a0 = fmaf(w0[(co * 9 + tp) * 2 + 1], in1[tp], a0);
          a1 = fmaf(w0[((co + 1) * 9 + tp) * 2 + 0], in0[tp], a1);
          a1 = fmaf(w0[((co + 1) * 9 + tp) * 2 + 1], in1[tp], a1);
        }
        packed[co / 2] = (uint32_t)bf_bits(fmaxf(a0, 0.f)) | ((uint32_t)bf_bits(fmaxf(a1, 0.f)) << 16);
      }
      uint4* dst = (uint4*)(tile + threadIdx.x * PS);
      dst[0] = uint4{packed[0], packed[1], packed[2], packed[3]};
      dst[1] = uint4{packed[4], packed[5], packed[6], packed[7]};
    }
    __syncthreads();
    if (t + (int)gridDim.x < ntiles) window(t + gridDim.x);
    // ---- enc0b: 16 -> 16 on v_mfma_f32_16x16x16_bf16, weights as A, pixels as B
    f32x4 acc[R];
#pragma unroll
    for (int rr = 0; rr < R; rr++) acc[rr] = f32x4{0.f, 0.f, 0.f, 0.f};
    const uint16_t* wrow = wl + li * WS + lg * 4;
    const uint16_t* trow = tile + (wr * R * LW + li) * PS + lg * 4;
#pragma unroll
    for (int dx = 0; dx < 3; dx++) {
      bf16x4 af[3];
#pragma unroll
      for (int dy = 0; dy < 3; dy++) af[dy] = *(const bf16x4*)(wrow + (dy * 3 + dx) * 16);
#pragma unroll
      for (int lr = 0; lr < R + 2; lr++) {
        const bf16x4 bx = *(const bf16x4*)(trow + (lr * LW + dx) * PS);
#pragma unroll
        for (int dy = 0; dy < 3; dy++) {
          const int rr = lr - dy;
          if (rr >= 0 && rr < R) acc[rr] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[dy], bx, acc[rr], 0, 0, 0);
        }
      }
    }
    // (the next tile's window goes to in2, which every wavefront has finished reading;
    // its enc0a waits at the barrier after that store until every wavefront is done here)
#pragma unroll
    for (int rr = 0; rr < R; rr++) {
      const int py = ty0 + wr * R + rr, px = tx0 + li;
      const f32x4 a = acc[rr];
      *(uint2*)(y + ((size_t)py * n + px) * 16 + 4 * lg) =
          uint2{bf_pack(fmaxf(a[0] + bv[0], 0.f), fmaxf(a[1] + bv[1], 0.f)),
                bf_pack(fmaxf(a[2] + bv[2], 0.f), fmaxf(a[3] + bv[3], 0.f))};
    }
  }
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

// ---- the network at a coarser resolution m = n / 2^k (k = 1, 2)
// Inputs averaged over 2^k x 2^k cells: r (at its storage precision) to FP32 r_c,
// p_prev to FP64 p_c; the increment d_c cleared (the output epilogue adds to it);
// per-block maxima of |r_c| and |p_c| to part[block] and part[nblocks + block] for the
// input scales. One thread per coarse cell; m * m is a multiple of 256.
template <typename S>
__global__ void __launch_bounds__(256)
k_guess_restrict(const S* __restrict__ r, const double* __restrict__ p, int n, int k, float* __restrict__ rc,
                 double* __restrict__ pc, double* __restrict__ dc, double* __restrict__ part) {
  const int m = n >> k, f = 1 << k;
  const int q = blockIdx.x * 256 + threadIdx.x;
  const int J = q / m, I = q - J * m;
  const size_t base = (size_t)(f * J) * n + f * I;
  float a;
  double b;
  if (k == 1) {
    a = 0.25f * ((ld(r[base]) + ld(r[base + 1])) + (ld(r[base + n]) + ld(r[base + n + 1])));
    b = 0.25 * ((p[base] + p[base + 1]) + (p[base + n] + p[base + n + 1]));
  } else {
    float sa = 0.f;
    double sb = 0.0;
    for (int y = 0; y < f; y++)
      for (int x = 0; x < f; x++) {
        sa += ld(r[base + (size_t)y * n + x]);
        sb += p[base + (size_t)y * n + x];
      }
    a = sa / (float)(f * f);
    b = sb / (double)(f * f);
  }
  rc[q] = a;
  pc[q] = b;
  dc[q] = 0.0;
  double ma = wave_max(nan_to_inf(fabs((double)a))), mb = wave_max(nan_to_inf(fabs(b)));
  __shared__ double sa4[4], sb4[4];
  if ((threadIdx.x & 63) == 0) {
    sa4[threadIdx.x >> 6] = ma;
    sb4[threadIdx.x >> 6] = mb;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[blockIdx.x] = fmax(fmax(sa4[0], sa4[1]), fmax(sa4[2], sa4[3]));
    part[gridDim.x + blockIdx.x] = fmax(fmax(sb4[0], sb4[1]), fmax(sb4[2], sb4[3]));
  }
}

// p += the bilinear interpolation of the coarse increment d_c (cell-centred,
// periodic; for k = 1 the geometry of the multigrid's prolongation): fine centre
// (i + 1/2) sits at coarse coordinate (i + 1/2) / 2^k - 1/2. One thread per pair of
// fine cells of a row (a 16-byte read-modify-write of p); n * n / 2 is a multiple of 256.
__global__ void __launch_bounds__(256)
k_guess_prolong_add(const double* __restrict__ dc, double* __restrict__ p, int n, int k) {
  const int m = n >> k;
  const int q = blockIdx.x * 256 + threadIdx.x;
  const int hn = n >> 1;
  const int j = q / hn, i2 = 2 * (q - j * hn);
  const double sc = 1.0 / (double)(1 << k);
  const double gy = (j + 0.5) * sc - 0.5, fj = floor(gy), fy = gy - fj;  // exact
  const int y0 = wrapi((int)fj, m), y1 = nxt(y0, m);
  double add[2];
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int i = i2 + h;
    const double gx = (i + 0.5) * sc - 0.5, fi = floor(gx), fx = gx - fi;
    const int x0 = wrapi((int)fi, m), x1 = nxt(x0, m);
    const double a = dc[(size_t)y0 * m + x0], b = dc[(size_t)y0 * m + x1];
    const double c = dc[(size_t)y1 * m + x0], d = dc[(size_t)y1 * m + x1];
    const double top = a * (1.0 - fx) + b * fx, bot = c * (1.0 - fx) + d * fx;
    add[h] = top * (1.0 - fy) + bot * fy;
  }
  double2* pp = (double2*)(p + (size_t)j * n + i2);
  double2 v = *pp;
  v.x += add[0];
  v.y += add[1];
  *pp = v;
}

}  // namespace fsk
