// fs_train.hpp -- training of the learned pressure guess on the device: the loss,
// the backward pass and Adam. The forward pass is the inference network's own
// kernels (fs_guess.hpp) run over a batch of images.
//
// Loss: mean over the batch and the pixels of (o - t)^2, o the network output and
// t = (p - p_prev) / (max|r| h^2) the increment the exact pressure p needs.
// Backward: the pre-activation gradient gz of every layer is stored in BF16 (NHWC,
// like the activations); the data gradient of a 3x3 layer is k_conv3x3 run on gz with
// the transposed, tap-flipped BF16 weights and a backward epilogue (ReLU mask,
// un-pooling, the split of a concatenation); the weight gradient is an MFMA
// reduction over pixels (k_conv3x3_wgrad) into per-workgroup FP32 partials summed
// in a fixed order. Parameters, gradients and the Adam moments are FP32, in the
// weights-blob order: per layer w[cout][k*k][cin] then b[cout].
#pragma once
#include "fs_guess.hpp"

namespace fsk {

// Weight and bias gradient of a 3x3 layer over nimg images at level m:
//   part[blockIdx.x][co][tap][ci] = sum over this workgroup's pixels p of
//                                   gz[p][co] * xin[p + off(tap)][ci]
//   part[blockIdx.x][COUT * 9 * CIN + co] = sum over the same pixels of gz[p][co]
// xin is the layer's input as the forward pass formed it (conv_in_item). A
// workgroup owns one block of 16 output channels (blockIdx.y) and loops over pixel
// tiles of TH rows x 16 columns; the K dimension of each MFMA is one row of 16
// pixels. The input tile (with halo) is staged as in the forward pass, the gradient
// tile transposed ([channel][pixel]) so an A fragment is one 8-byte read. The
// (tap, input-channel block) products are spread over the four wavefronts.
template <int CIN, int COUT, int MODE, int C2, int TH>
__global__ void __launch_bounds__(GT_NT)
k_conv3x3_wgrad(const uint16_t* __restrict__ x1, const uint16_t* __restrict__ x2, const uint16_t* __restrict__ gz,
                float* __restrict__ part, int m, int nimg) {
  static_assert(CIN % 16 == 0 && COUT % 16 == 0, "channel blocks");
  constexpr int PS = CIN + 8, CH8 = CIN / 8;
  constexpr int LH = TH + 2, LW = GT_W + 2;
  constexpr int NI = 9 * (CIN / 16), IPW = (NI + 3) / 4;
  constexpr int GS = TH * GT_W + 4;  // transposed gradient row stride (pixels)
  constexpr size_t PL = (size_t)COUT * 9 * CIN + COUT;
  __shared__ __align__(16) uint16_t tile[LH * LW * PS];
  __shared__ __align__(16) uint16_t gzt[16 * GS];
  const int tiles_x = m / GT_W, tpi = tiles_x * (m / TH), ntiles = tpi * nimg;
  const int co0 = blockIdx.y * 16;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = (lane >> 4) * 4;
  f32x4 acc[IPW];
#pragma unroll
  for (int j = 0; j < IPW; j++) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
  f32x4 accb = f32x4{0.f, 0.f, 0.f, 0.f};
  const bf16x4 ones = bf16x4{0x3f80, 0x3f80, 0x3f80, 0x3f80};
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int z = t / tpi, tt = t - z * tpi;
    const int tx0 = (tt % tiles_x) * GT_W, ty0 = (tt / tiles_x) * TH;
    const uint16_t* x1z = x1 + z * conv_x1_image<CIN, MODE, C2>(m);
    const uint16_t* x2z = MODE == CONV_UPCAT ? x2 + z * conv_x2_image<C2>(m) : x2;
    for (int q = threadIdx.x; q < LH * LW * CH8; q += GT_NT) {
      int pix = q / CH8, cg = q - pix * CH8;
      int ly = pix / LW, lx = pix - ly * LW;
      *(uint4*)(tile + pix * PS + cg * 8) =
          conv_in_item<CIN, MODE, C2>(x1z, x2z, m, wrapm(ty0 - 1 + ly, m), wrapm(tx0 - 1 + lx, m), cg * 8);
    }
    const uint16_t* gzz = gz + (size_t)z * m * m * COUT;
    for (int q = threadIdx.x; q < TH * GT_W * 2; q += GT_NT) {
      int pix = q >> 1, half = q & 1;
      uint4 v = *(const uint4*)(gzz + ((size_t)(ty0 + pix / GT_W) * m + tx0 + pix % GT_W) * COUT + co0 + half * 8);
      uint16_t* d = gzt + (half * 8) * GS + pix;
      d[0 * GS] = (uint16_t)(v.x & 0xffff);
      d[1 * GS] = (uint16_t)(v.x >> 16);
      d[2 * GS] = (uint16_t)(v.y & 0xffff);
      d[3 * GS] = (uint16_t)(v.y >> 16);
      d[4 * GS] = (uint16_t)(v.z & 0xffff);
      d[5 * GS] = (uint16_t)(v.z >> 16);
      d[6 * GS] = (uint16_t)(v.w & 0xffff);
      d[7 * GS] = (uint16_t)(v.w >> 16);
    }
    __syncthreads();
    for (int r = 0; r < TH; r++) {
      // A: 16 output channels x 16 pixels of row r; lane holds channel li, pixels lk..lk+3
      const bf16x4 afrag = *(const bf16x4*)(gzt + li * GS + r * GT_W + lk);
#pragma unroll
      for (int j = 0; j < IPW; j++) {
        const int it = wave + 4 * j;
        if (it < NI) {
          const int tap = it / (CIN / 16), cb = it - tap * (CIN / 16);
          const int dy = tap / 3, dx = tap - dy * 3;
          // B: 16 pixels x 16 input channels; lane holds channel cb*16 + li, pixels lk..lk+3
          const uint16_t* src = tile + ((r + dy) * LW + lk + dx) * PS + cb * 16 + li;
          bf16x4 bfrag = bf16x4{(short)src[0], (short)src[PS], (short)src[2 * PS], (short)src[3 * PS]};
          acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(afrag, bfrag, acc[j], 0, 0, 0);
        }
      }
      if (wave == 0) accb = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(afrag, ones, accb, 0, 0, 0);
    }
    __syncthreads();  // the next tile overwrites both staged tiles
  }
  // D[row][col]: lane holds rows (output channels) lk..lk+3 and column (input channel) li
  float* dst = part + (size_t)blockIdx.x * PL;
#pragma unroll
  for (int j = 0; j < IPW; j++) {
    const int it = wave + 4 * j;
    if (it < NI) {
      const int tap = it / (CIN / 16), cb = it - tap * (CIN / 16);
#pragma unroll
      for (int q = 0; q < 4; q++) dst[(size_t)(co0 + lk + q) * 9 * CIN + tap * CIN + cb * 16 + li] = acc[j][q];
    }
  }
  if (wave == 0 && li == 0) {
#pragma unroll
    for (int q = 0; q < 4; q++) dst[(size_t)COUT * 9 * CIN + co0 + lk + q] = accb[q];
  }
}

// dst[i] = sum over r < rows of part[r * len + i], in order of r.
__global__ void __launch_bounds__(256)
k_tr_reduce(const float* __restrict__ part, int rows, size_t len, float* __restrict__ dst) {
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= len) return;
  float a = 0.f;
  for (int r = 0; r < rows; r++) a += part[(size_t)r * len + i];
  dst[i] = a;
}

// First level of a two-level fixed-order sum: rows [y TR_RC, (y + 1) TR_RC) of part
// summed in order into out[y] (y = blockIdx.y); k_tr_reduce then sums the out rows.
constexpr int TR_RC = 32;
__global__ void __launch_bounds__(256)
k_tr_reduce_rows(const float* __restrict__ part, int rows, size_t len, float* __restrict__ out) {
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= len) return;
  const int r0 = blockIdx.y * TR_RC, r1 = min(rows, r0 + TR_RC);
  float a = 0.f;
  for (int r = r0; r < r1; r++) a += part[(size_t)r * len + i];
  out[(size_t)blockIdx.y * len + i] = a;
}

// Training pairs: image b is (p_prev, p) = (snap[idx[b]], snap[idx[b] + 1]), FP64 n x n.
// Writes p_prev to pp[b], r = L p - L p_prev (5-point periodic Laplacian) to r32[b] in
// FP32, and per-block maxima of |r| and |p_prev| (FP64) to mr[b][blk], mp[b][blk].
__global__ void __launch_bounds__(256)
k_tr_prep(const double* __restrict__ snap, const int* __restrict__ idx, int n, double inv_h2,
          double* __restrict__ pp, float* __restrict__ r32, double* __restrict__ mr, double* __restrict__ mp) {
  const int b = blockIdx.y;
  const size_t nn = (size_t)n * n;
  const double* P0 = snap + (size_t)idx[b] * nn;
  const double* P1 = P0 + nn;
  double ar = 0.0, ap = 0.0;
  for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < nn; k += (size_t)gridDim.x * 256) {
    int i = (int)(k % n), j = (int)(k / n);
    int ip = i + 1 == n ? 0 : i + 1, im = i == 0 ? n - 1 : i - 1;
    int jp = j + 1 == n ? 0 : j + 1, jm = j == 0 ? n - 1 : j - 1;
    auto lap = [&](const double* P) {
      return (((P[(size_t)j * n + ip] + P[(size_t)j * n + im]) + (P[(size_t)jp * n + i] + P[(size_t)jm * n + i])) -
              4.0 * P[k]) * inv_h2;
    };
    double r = lap(P1) - lap(P0);
    pp[(size_t)b * nn + k] = P0[k];
    r32[(size_t)b * nn + k] = (float)r;
    ar = fmax(ar, nan_to_inf(fabs(r)));
    ap = fmax(ap, nan_to_inf(fabs(P0[k])));
  }
  ar = block_max(ar);
  ap = block_max(ap);
  if (threadIdx.x == 0) {
    mr[(size_t)b * gridDim.x + blockIdx.x] = ar;
    mp[(size_t)b * gridDim.x + blockIdx.x] = ap;
  }
}

// scales[2b] = max|r|, scales[2b + 1] = max|p_prev| of image b (one block per image).
__global__ void __launch_bounds__(256)
k_tr_scales(const double* __restrict__ mr, const double* __restrict__ mp, int nblk, double* __restrict__ scales) {
  const int b = blockIdx.x;
  double a = 0.0, c = 0.0;
  for (int k = threadIdx.x; k < nblk; k += 256) {
    a = fmax(a, mr[(size_t)b * nblk + k]);
    c = fmax(c, mp[(size_t)b * nblk + k]);
  }
  a = block_max(a);
  c = block_max(c);
  if (threadIdx.x == 0) {
    scales[2 * b] = a;
    scales[2 * b + 1] = c;
  }
}

// Target of image b: t = (p - p_prev) / (max|r| h^2) in FP32 (0 where max|r| = 0).
__global__ void __launch_bounds__(256)
k_tr_target(const double* __restrict__ snap, const int* __restrict__ idx, int n, double h2,
            const double* __restrict__ scales, float* __restrict__ t) {
  const int b = blockIdx.y;
  const size_t nn = (size_t)n * n;
  const double* P0 = snap + (size_t)idx[b] * nn;
  const double* P1 = P0 + nn;
  const double sr = scales[2 * b];
  const double inv = sr > 0 ? 1.0 / (sr * h2) : 0.0;
  for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < nn; k += (size_t)gridDim.x * 256)
    t[(size_t)b * nn + k] = (float)((P1[k] - P0[k]) * inv);
}

// Output layer forward and backward over all B n^2 pixels: o = b + w . d0 (as the
// inference output), e = o - t, dO = 2 e / (B n^2); gz13 = BF16(dO w) where d0 > 0.
// Per-block partials: lpart[blk] = sum e^2 (FP64), opart[blk][c] = sum dO d0_c for
// c < 16 and opart[blk][16] = sum dO (FP32), each in a fixed order; err = e.
__global__ void __launch_bounds__(256)
k_tr_out(const uint16_t* __restrict__ d0, const float* __restrict__ w, const float* __restrict__ bias,
         const float* __restrict__ t, size_t count, uint16_t* __restrict__ gz, double* __restrict__ lpart,
         float* __restrict__ opart, float* __restrict__ err) {
  const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
  float vals[17];
  double e2 = 0.0;
#pragma unroll
  for (int c = 0; c < 17; c++) vals[c] = 0.f;
  if (k < count) {
    const uint4* src = (const uint4*)(d0 + k * 16);
    uint4 a = src[0], b = src[1];
    uint32_t v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    float x[16];
#pragma unroll
    for (int c = 0; c < 8; c++) {
      x[2 * c] = bf_val(v[c] & 0xffff);
      x[2 * c + 1] = bf_val(v[c] >> 16);
    }
    float o = bias[0];
#pragma unroll
    for (int c = 0; c < 16; c++) o = fmaf(w[c], x[c], o);
    const float e = o - t[k];
    err[k] = e;
    e2 = (double)e * (double)e;
    const float dO = e * (float)(2.0 / (double)count);
    uint32_t g[8];
#pragma unroll
    for (int c = 0; c < 8; c++) {
      uint32_t lo = x[2 * c] > 0.f ? (uint32_t)bf_bits(dO * w[2 * c]) : 0u;
      uint32_t hi = x[2 * c + 1] > 0.f ? (uint32_t)bf_bits(dO * w[2 * c + 1]) : 0u;
      g[c] = lo | (hi << 16);
    }
    uint4* dst = (uint4*)(gz + k * 16);
    dst[0] = uint4{g[0], g[1], g[2], g[3]};
    dst[1] = uint4{g[4], g[5], g[6], g[7]};
#pragma unroll
    for (int c = 0; c < 16; c++) vals[c] = dO * x[c];
    vals[16] = dO;
  }
  __shared__ float sv[4][17];
  __shared__ double se[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < 17; c++) {
    float a = vals[c];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    vals[c] = a;
  }
  e2 = wave_sum(e2);
  if (lane == 0) {
#pragma unroll
    for (int c = 0; c < 17; c++) sv[wave][c] = vals[c];
    se[wave] = e2;
  }
  __syncthreads();
  if (threadIdx.x < 17) opart[(size_t)blockIdx.x * 17 + threadIdx.x] =
      ((sv[0][threadIdx.x] + sv[1][threadIdx.x]) + (sv[2][threadIdx.x] + sv[3][threadIdx.x]));
  if (threadIdx.x == 0) lpart[blockIdx.x] = (se[0] + se[1]) + (se[2] + se[3]);
}

// loss = sum of lpart[0 .. nblk) / count (one block, fixed order).
__global__ void __launch_bounds__(256)
k_tr_loss(const double* __restrict__ lpart, int nblk, double count, double* __restrict__ loss) {
  double a = 0.0;
  for (int k = threadIdx.x; k < nblk; k += 256) a += lpart[k];
  __shared__ double sm[4];
  a = wave_sum(a);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) *loss = ((sm[0] + sm[1]) + (sm[2] + sm[3])) / count;
}

// Gradient of the first layer (2 -> 16, FP32) over a 16 x 16 pixel tile per step:
//   part[blk][co][tap][c] = sum gz0[p][co] x_c[p + off(tap)],  part[blk][288 + co] = sum gz0[p][co]
// with x formed exactly as k_guess_first forms it (r / max|r|, p_prev / max|p_prev|
// in FP32). Thread o < 304 owns output o; the tile is staged in LDS.
__global__ void __launch_bounds__(320)
k_tr_first_wgrad(const uint16_t* __restrict__ gz, const float* __restrict__ r32, const double* __restrict__ pp,
                 const double* __restrict__ scales, int n, int nimg, float* __restrict__ part) {
  constexpr int T = 16, L = T + 2;
  __shared__ float xs[L * L * 2];
  __shared__ float gs[T * T * 16];
  const int tiles_x = n / T, tpi = tiles_x * tiles_x, ntiles = tpi * nimg;
  const int o = threadIdx.x;
  const int co = o < 288 ? o / 18 : o - 288, tap = (o % 18) / 2, c = o & 1;
  const int dy = tap / 3, dx = tap % 3;
  float acc = 0.f;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int z = t / tpi, tt = t - z * tpi;
    const int tx0 = (tt % tiles_x) * T, ty0 = (tt / tiles_x) * T;
    const size_t nn = (size_t)n * n;
    const double sr = scales[2 * z], sp = scales[2 * z + 1];
    const float ir = sr > 0 ? (float)(1.0 / sr) : 0.f, ip = sp > 0 ? (float)(1.0 / sp) : 0.f;
    __syncthreads();
    for (int q = threadIdx.x; q < L * L; q += blockDim.x) {
      int ly = q / L, lx = q - ly * L;
      size_t k = z * nn + (size_t)wrapm(ty0 - 1 + ly, n) * n + wrapm(tx0 - 1 + lx, n);
      xs[2 * q] = r32[k] * ir;
      xs[2 * q + 1] = (float)pp[k] * ip;
    }
    for (int q = threadIdx.x; q < T * T * 2; q += blockDim.x) {
      int pix = q >> 1, half = q & 1;
      uint4 v = *(const uint4*)(gz + (z * nn + (size_t)(ty0 + pix / T) * n + tx0 + pix % T) * 16 + half * 8);
      float* d = gs + pix * 16 + half * 8;
      d[0] = bf_val(v.x & 0xffff); d[1] = bf_val(v.x >> 16); d[2] = bf_val(v.y & 0xffff); d[3] = bf_val(v.y >> 16);
      d[4] = bf_val(v.z & 0xffff); d[5] = bf_val(v.z >> 16); d[6] = bf_val(v.w & 0xffff); d[7] = bf_val(v.w >> 16);
    }
    __syncthreads();
    if (o < 288) {
      for (int py = 0; py < T; py++)
        for (int px = 0; px < T; px++) acc = fmaf(gs[(py * T + px) * 16 + co], xs[((py + dy) * L + px + dx) * 2 + c], acc);
    } else if (o < 304) {
      for (int p = 0; p < T * T; p++) acc += gs[p * 16 + co];
    }
  }
  if (o < 304) part[(size_t)blockIdx.x * 304 + o] = acc;
}

// Adam (the PyTorch form): m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2,
// w -= (lr / bc1) m / (sqrt(v) / sqrt(bc2) + eps), bc_i = 1 - b_i^step.
__global__ void __launch_bounds__(256)
k_tr_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1, float* __restrict__ m2,
          size_t count, float lr_bc1, float inv_sqrt_bc2, float b1, float b2, float eps) {
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  float gi = g[i];
  float a = b1 * m1[i] + (1.f - b1) * gi;
  float b = b2 * m2[i] + (1.f - b2) * (gi * gi);
  m1[i] = a;
  m2[i] = b;
  w[i] -= lr_bc1 * (a / (sqrtf(b) * inv_sqrt_bc2 + eps));
}

// BF16 weights of a 3x3 layer from its FP32 parameters w[cout][9][cin]: wf in the same
// layout (forward), wt[ci][tap][co] = w[co][8 - tap][ci] (backward data).
__global__ void __launch_bounds__(256)
k_tr_pack(const float* __restrict__ w, int cin, int cout, uint16_t* __restrict__ wf, uint16_t* __restrict__ wt) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= cout * 9 * cin) return;
  int ci = i % cin, tap = (i / cin) % 9, co = i / (9 * cin);
  uint16_t b = bf_bits(w[i]);
  wf[i] = b;
  wt[((size_t)ci * 9 + (8 - tap)) * cout + co] = b;
}

}  // namespace fsk
