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
enum ConvEpi { EPI_RELU = 0, EPI_MASK = 1, EPI_UNPOOL = 2, EPI_UPCAT = 3 };
struct ConvEpiArgs {
  const uint16_t* act;
  const float* gin;
  float* gout;
};

__device__ __forceinline__ uint32_t bf_pair_masked(float lo, float hi, uint32_t act2) {
  uint32_t a = (act2 & 0x8000u) || !(act2 & 0x7fffu) ? 0u : (uint32_t)bf_bits(lo);
  uint32_t b = (act2 & 0x80000000u) || !(act2 & 0x7fff0000u) ? 0u : (uint32_t)bf_bits(hi);
  return a | (b << 16);
}
__device__ __forceinline__ uint4 bf8_masked(const float* v, uint4 act) {
  return uint4{bf_pair_masked(v[0], v[1], act.x), bf_pair_masked(v[2], v[3], act.y),
               bf_pair_masked(v[4], v[5], act.z), bf_pair_masked(v[6], v[7], act.w)};
}

// Backward epilogues of k_conv3x3 (EPI_MASK / EPI_UNPOOL / EPI_UPCAT) over one staged
// tile: ofl[pixel][channel] holds the FP32 accumulators of TH x 16 pixels and the OC
// channels [co0, co0 + OC) of image z; 8 channels per item.
template <int COUT, int TH, int OC, int OFP, int EPI, int OS>
__device__ __forceinline__ void conv_bwd_epilogue(const float* ofl, uint16_t* __restrict__ y, const ConvEpiArgs& e,
                                                  int m, int z, int tx0, int ty0, int co0) {
  constexpr int C8 = OC / 8;
  if (EPI == EPI_MASK) {
    const size_t base = (size_t)z * m * m * COUT;
    for (int q = threadIdx.x; q < TH * GT_W * C8; q += GT_NT) {
      int pix = q / C8, cg = q - pix * C8;
      size_t k = base + ((size_t)(ty0 + pix / GT_W) * m + tx0 + pix % GT_W) * COUT + co0 + cg * 8;
      uint4 act = *(const uint4*)(e.act + k);
      *(uint4*)(y + k) = bf8_masked(ofl + pix * OFP + cg * 8, act);
    }
  } else if (EPI == EPI_UNPOOL) {
    const int m2 = 2 * m;
    const size_t base = (size_t)z * m2 * m2 * COUT;
    for (int q = threadIdx.x; q < 4 * TH * GT_W * C8; q += GT_NT) {
      int f = q / C8, cg = q - f * C8;                  // fine pixel of the 2TH x 32 block
      int fy = f / (2 * GT_W), fx = f - fy * (2 * GT_W);
      int pix = (fy >> 1) * GT_W + (fx >> 1);
      size_t k = base + ((size_t)(2 * ty0 + fy) * m2 + 2 * tx0 + fx) * COUT + co0 + cg * 8;
      float g[8];
      const float* a = ofl + pix * OFP + cg * 8;
      if (e.gin) {
        float4 g0 = *(const float4*)(e.gin + k), g1 = *(const float4*)(e.gin + k + 4);
        g[0] = 0.25f * a[0] + g0.x; g[1] = 0.25f * a[1] + g0.y; g[2] = 0.25f * a[2] + g0.z; g[3] = 0.25f * a[3] + g0.w;
        g[4] = 0.25f * a[4] + g1.x; g[5] = 0.25f * a[5] + g1.y; g[6] = 0.25f * a[6] + g1.z; g[7] = 0.25f * a[7] + g1.w;
      } else {
#pragma unroll
        for (int c = 0; c < 8; c++) g[c] = 0.25f * a[c];
      }
      uint4 act = *(const uint4*)(e.act + k);
      *(uint4*)(y + k) = bf8_masked(g, act);
    }
  } else if (EPI == EPI_UPCAT) {
    if (co0 < OS) {
      const int mc = m >> 1;
      const size_t base = (size_t)z * mc * mc * OS;
      for (int q = threadIdx.x; q < (TH / 2) * (GT_W / 2) * C8; q += GT_NT) {
        int cp = q / C8, cg = q - cp * C8;
        int cy = cp / (GT_W / 2), cx = cp - cy * (GT_W / 2);
        const float* a00 = ofl + ((2 * cy) * GT_W + 2 * cx) * OFP + cg * 8;
        float g[8];
#pragma unroll
        for (int c = 0; c < 8; c++) g[c] = (a00[c] + a00[OFP + c]) + (a00[GT_W * OFP + c] + a00[(GT_W + 1) * OFP + c]);
        size_t k = base + ((size_t)((ty0 >> 1) + cy) * mc + (tx0 >> 1) + cx) * OS + co0 + cg * 8;
        uint4 act = *(const uint4*)(e.act + k);
        *(uint4*)(y + k) = bf8_masked(g, act);
      }
    } else {
      constexpr int CS = COUT - OS;
      const size_t base = (size_t)z * m * m * CS;
      for (int q = threadIdx.x; q < TH * GT_W * C8; q += GT_NT) {
        int pix = q / C8, cg = q - pix * C8;
        size_t k = base + ((size_t)(ty0 + pix / GT_W) * m + tx0 + pix % GT_W) * CS + (co0 - OS) + cg * 8;
        const float* a = ofl + pix * OFP + cg * 8;
        *(float4*)(e.gout + k) = float4{a[0], a[1], a[2], a[3]};
        *(float4*)(e.gout + k + 4) = float4{a[4], a[5], a[6], a[7]};
      }
    }
  }
}

// 3x3 periodic convolution, CIN -> COUT channels at an m x m level, over nimg images
// (tensors of consecutive images follow each other). Input modes as in conv_in_item.
// w: BF16 [COUT][9][CIN] (tap = 3 (dy + 1) + dx + 1); bias FP32 [COUT] (EPI_RELU).
// A workgroup computes TH rows x 16 columns of pixels and NBW blocks of 16 output
// channels (blockIdx.y selects which); each wavefront owns TH/4 rows (one MFMA
// M-block each). The workgroup's weight slice and its input tile (with the
// one-pixel halo) are staged in LDS, so every B fragment feeds TH/4 MFMAs from LDS.
template <int CIN, int COUT, int MODE, int C2, int TH, int NBW, int EPI = EPI_RELU, int OS = 0>
__global__ void __launch_bounds__(GT_NT)
k_conv3x3(const uint16_t* __restrict__ x1, const uint16_t* __restrict__ x2, const uint16_t* __restrict__ w,
          const float* __restrict__ bias, uint16_t* __restrict__ y, int m, int nimg, ConvEpiArgs e) {
  static_assert(CIN % 16 == 0 && COUT % (16 * NBW) == 0 && TH % 4 == 0, "tile shape");
  static_assert(EPI != EPI_UPCAT || (OS % (16 * NBW) == 0 && OS < COUT), "split on a channel block");
  constexpr int PS = CIN + 8;          // LDS pixel stride (bf16): 16-byte aligned, breaks A-read bank conflicts
  constexpr int WS = 9 * CIN + 4;      // LDS weight row stride (bf16): breaks B-read bank conflicts
  constexpr int CH8 = CIN / 8;
  constexpr int LH = TH + 2, LW = GT_W + 2;
  constexpr int MB = TH / 4;           // M-blocks (rows) per wavefront
  constexpr int OC = NBW * 16, OCP = OC + 8;  // output channels of the workgroup, staged stride
  constexpr int OFP = OC + 4;                 // FP32 staged stride (backward epilogues)
  __shared__ __align__(16) uint16_t tile[LH * LW * PS];
  __shared__ __align__(16) uint16_t wl[NBW * 16 * WS];
  __shared__ __align__(16) uint16_t otile[EPI == EPI_RELU ? TH * GT_W * OCP : 8];
  __shared__ __align__(16) float ofl[EPI == EPI_RELU ? 4 : TH * GT_W * OFP];
  const int tiles_x = m / GT_W, tpi = tiles_x * (m / TH), ntiles = tpi * nimg;
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
  for (int nb = 0; nb < NBW; nb++) bv[nb] = EPI == EPI_RELU ? bias[co0 + nb * 16 + li] : 0.f;
  // ---- input tiles (with halo), 8 channels (16 bytes) per item; the next tile's
  // items are loaded into registers while the current one is multiplied
  constexpr int TI = LH * LW * CH8, TK = (TI + GT_NT - 1) / GT_NT;
  uint4 pre[TK];
  auto prefetch = [&](int t) {
    const int z = t / tpi, tt = t - z * tpi;
    const int tx0 = (tt % tiles_x) * GT_W, ty0 = (tt / tiles_x) * TH;
    const uint16_t* x1z = x1 + z * conv_x1_image<CIN, MODE, C2>(m);
    const uint16_t* x2z = MODE == CONV_UPCAT ? x2 + z * conv_x2_image<C2>(m) : x2;
#pragma unroll
    for (int k = 0; k < TK; k++) {
      int q = threadIdx.x + k * GT_NT;
      if (q < TI) {
        int pix = q / CH8, cg = q - pix * CH8;
        int ly = pix / LW, lx = pix - ly * LW;
        pre[k] = conv_in_item<CIN, MODE, C2>(x1z, x2z, m, wrapm(ty0 - 1 + ly, m), wrapm(tx0 - 1 + lx, m), cg * 8);
      }
    }
  };
  if ((int)blockIdx.x < ntiles) prefetch(blockIdx.x);
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int z = t / tpi, tt = t - z * tpi;
    const int tx0 = (tt % tiles_x) * GT_W, ty0 = (tt / tiles_x) * TH;
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
    if (EPI == EPI_RELU) {
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
      uint16_t* yz = y + (size_t)z * m * m * COUT;
      constexpr int OI = TH * GT_W * (OC / 8);
      for (int q = threadIdx.x; q < OI; q += GT_NT) {
        int pix = q / (OC / 8), cg = q - pix * (OC / 8);
        int py = ty0 + pix / GT_W, px = tx0 + (pix % GT_W);
        *(uint4*)(yz + ((size_t)py * m + px) * COUT + co0 + cg * 8) = *(const uint4*)(otile + pix * OCP + cg * 8);
      }
    } else {
#pragma unroll
      for (int r = 0; r < MB; r++) {
        int row = wave * MB + r;
#pragma unroll
        for (int nb = 0; nb < NBW; nb++)
#pragma unroll
          for (int q = 0; q < 4; q++) ofl[(row * GT_W + lk + q) * OFP + nb * 16 + li] = acc[r][nb][q];
      }
      __syncthreads();
      conv_bwd_epilogue<COUT, TH, OC, OFP, EPI, OS>(ofl, y, e, m, z, tx0, ty0, co0);
    }
    __syncthreads();  // the next tile overwrites the staged input and outputs
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
