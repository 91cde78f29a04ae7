// fs_mg_fused.hpp -- LDS-tiled multigrid kernels: one launch per level and direction.
//
// The V-cycle of src/gpu/multigrid.js:277-330 (red-black V(pre,post), full-weighting
// restriction, bilinear prolongation) is regrouped so that each grid-wide level
// costs two launches instead of 2*(pre+post)+2:
//
//   k_mgt_down<SW>   x = 0; SW red-black sweeps; residual; 2x2 restriction -> coarse rhs
//   k_mgt_down0<SW>  the same on level 0, fused with the FP64 refinement residual
//                    r = div u* - L(p + e_prev) that feeds it
//   k_mgt_up<SW>     x += prolong(coarse x); SW red-black sweeps
//   k_mg_coarse_ml   every level <= COARSE_N (down, coarsest solve, up) in one workgroup;
//                    levels of at most 8x8 cells run in one wavefront's registers
//
// A tile of TX x TY cells is loaded with a halo of 2*SW cells into LDS. Each
// half-sweep is valid on a region one cell smaller than the previous one, so after
// 2*SW half-sweeps the interior holds exactly what the same number of grid-wide
// colour passes produces: every cell sees the same operands in the same order, and
// the result is bitwise identical to the one-kernel-per-pass path.
// A correction that starts from zero makes the first red half-sweep neighbour-free
// (x = -h^2 rhs / 4), which is what lets it run on the whole loaded region.
#pragma once
#include "fs_kernels.hpp"

namespace fsk {

constexpr int COARSE_N = 64;   // levels at or below this size are solved by one workgroup

// Tile shape: TX x TY cells per workgroup of NT threads. Each grid-wide level picks
// the largest shape that still gives every CU a tile (256 tiles), so the small
// levels, which are latency-bound, are spread over the whole chip.
template <int TX_, int TY_, int NT_>
struct TileShape {
  static constexpr int TX = TX_, TY = TY_, NT = NT_;
};
using Tile64 = TileShape<64, 32, 512>;  // with 256 threads the level-0 down-sweep at 2048^2 took 9 % longer
using Tile32 = TileShape<32, 32, 256>;
using Tile16 = TileShape<16, 16, 128>;
using Tile8 = TileShape<8, 8, 64>;
using Tile64sq = TileShape<64, 64, 1024>;  // one tile per CU at 1024^2: a quarter of the
                                            // halo cells of 32 x 32 tiles recomputed
// the same tiles with more threads (fewer cells each): shorter passes on the small,
// latency-bound levels
using Tile32w = TileShape<32, 32, 1024>;
using Tile16w = TileShape<16, 16, 256>;
using Tile16x1024 = TileShape<16, 16, 1024>;  // the pair above an exactly solved 64 x 64 level, which
                                               // every workgroup solves (one output block per wavefront)
using Tile8x1024 = TileShape<8, 8, 1024>;  // k_mgt_down2 and k_mgt_up2 above the coarse solve: the
                                           // level-l region is ~25 times the tile, and these launches
                                           // are latency-bound: short passes over many threads

// Blocks are issued round-robin over the 8 XCDs; give each XCD a contiguous band of
// tiles so neighbouring tiles (which share halos) share an L2.
__device__ __forceinline__ int xcd_tile(int nblocks) {
  int b = blockIdx.x;
  if (nblocks & 7) return b;
  return (b & 7) * (nblocks >> 3) + (b >> 3);
}

__device__ __forceinline__ int wrapn(int i, int n) {
  // |i| < 2n for every caller (halo << n), so one correction suffices
  return i < 0 ? i + n : (i >= n ? i - n : i);
}

// Load a periodic RW x RH window with origin (ox, oy) of an n x n field into LDS.
// Every load a thread owns is issued before any of them is stored, so the
// window costs one memory latency instead of one per row of threads.
// BF16 windows with an even origin and width are read two cells per 4-byte load (a
// pair starts at an even index, so it never straddles the periodic seam).
template <int RW, int RH, int NT, typename T>
__device__ __forceinline__ void load_window(float* __restrict__ dst, const T* __restrict__ src, int n, int ox, int oy) {
  if constexpr (sizeof(T) == 2 && RW % 2 == 0) {
    if ((ox & 1) == 0) {
      constexpr int CNT2 = RW * RH / 2, K2 = (CNT2 + NT - 1) / NT, RW2 = RW / 2;
      uint32_t tw[K2];
#pragma unroll
      for (int k = 0; k < K2; k++) {
        int q = threadIdx.x + k * NT;
        if (q < CNT2) {
          int ly = q / RW2, lx = 2 * (q - ly * RW2);
          tw[k] = *(const uint32_t*)(src + (size_t)wrapn(oy + ly, n) * n + wrapn(ox + lx, n));
        }
      }
#pragma unroll
      for (int k = 0; k < K2; k++) {
        int q = threadIdx.x + k * NT;
        if (q < CNT2) {
          dst[2 * q] = __uint_as_float(tw[k] << 16);
          dst[2 * q + 1] = __uint_as_float(tw[k] & 0xffff0000u);
        }
      }
      return;
    }
  }
  constexpr int CNT = RW * RH, K = (CNT + NT - 1) / NT;
  float tmp[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    int q = threadIdx.x + k * NT;
    if (q < CNT) {
      int ly = q / RW, lx = q - ly * RW;
      tmp[k] = ld(src[(size_t)wrapn(oy + ly, n) * n + wrapn(ox + lx, n)]);
    }
  }
#pragma unroll
  for (int k = 0; k < K; k++) {
    int q = threadIdx.x + k * NT;
    if (q < CNT) dst[q] = tmp[k];
  }
}

// A window's loads held in registers: issue() starts them, commit() writes them to LDS
// as FP32 (the values load_window stores). A kernel that issues every window it needs
// before committing any pays one memory latency for all of them.
template <int RW, int RH, int NT, typename T>
struct Window {
  static constexpr int CNT = RW * RH, K = (CNT + NT - 1) / NT;
  static constexpr bool PAIRS = sizeof(T) == 2 && RW % 2 == 0;
  static constexpr int RW2 = RW / 2, CNT2 = CNT / 2, K2 = (CNT2 + NT - 1) / NT;
  uint32_t w[PAIRS ? K : 1];  // BF16: two cells per word (even origin) or one cell's bits
  float f[PAIRS ? 1 : K];
  bool pairs;
  __device__ __forceinline__ void issue(const T* __restrict__ src, int n, int ox, int oy) {
    if constexpr (PAIRS) {
      pairs = (ox & 1) == 0;
      if (pairs) {
#pragma unroll
        for (int k = 0; k < K2; k++) {
          const int q = threadIdx.x + k * NT;
          if (q < CNT2) {
            const int ly = q / RW2, lx = 2 * (q - ly * RW2);
            w[k] = *(const uint32_t*)(src + (size_t)wrapn(oy + ly, n) * n + wrapn(ox + lx, n));
          }
        }
      } else {
#pragma unroll
        for (int k = 0; k < K; k++) {
          const int q = threadIdx.x + k * NT;
          if (q < CNT) {
            const int ly = q / RW, lx = q - ly * RW;
            w[k] = *(const uint16_t*)(src + (size_t)wrapn(oy + ly, n) * n + wrapn(ox + lx, n));
          }
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int q = threadIdx.x + k * NT;
        if (q < CNT) {
          const int ly = q / RW, lx = q - ly * RW;
          f[k] = ld(src[(size_t)wrapn(oy + ly, n) * n + wrapn(ox + lx, n)]);
        }
      }
    }
  }
  __device__ __forceinline__ void commit(float* __restrict__ dst) const {
    if constexpr (PAIRS) {
      if (pairs) {
#pragma unroll
        for (int k = 0; k < K2; k++) {
          const int q = threadIdx.x + k * NT;
          if (q < CNT2) {
            dst[2 * q] = __uint_as_float(w[k] << 16);
            dst[2 * q + 1] = __uint_as_float(w[k] & 0xffff0000u);
          }
        }
      } else {
#pragma unroll
        for (int k = 0; k < K; k++) {
          const int q = threadIdx.x + k * NT;
          if (q < CNT) dst[q] = __uint_as_float(w[k] << 16);
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int q = threadIdx.x + k * NT;
        if (q < CNT) dst[q] = f[k];
      }
    }
  }
};

// One red-black half-sweep on the LDS region [LO, RW-LO) x [LO, RH-LO), visiting
// only the cells of `colour` (par0 = colour parity of LDS cell (0,0)). A colour pass
// reads only the other colour, so all of a thread's reads are issued before its writes.
template <int RW, int RH, int LO, int NT>
__device__ __forceinline__ void lds_half_sweep(float* __restrict__ sx, const float* __restrict__ sr, int colour,
                                               int par0, float h2) {
  constexpr int w = RW - 2 * LO, hgt = RH - 2 * LO, hw = w / 2;  // w is even
  constexpr int CNT = hw * hgt, K = (CNT + NT - 1) / NT;
  float acc[K];
  int kk[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    int q = threadIdx.x + k * NT;
    if (q < CNT) {
      int ry = q / hw, rx = q - ry * hw;
      int ly = LO + ry;
      int lx = LO + 2 * rx + ((colour + par0 + LO + ly) & 1);
      int c = ly * RW + lx;
      kk[k] = c;
      float sum = sx[c - 1] + sx[c + 1] + sx[c - RW] + sx[c + RW];
      acc[k] = (sum - h2 * sr[c]) * 0.25f;
    }
  }
#pragma unroll
  for (int k = 0; k < K; k++) {
    int q = threadIdx.x + k * NT;
    if (q < CNT) sx[kk[k]] = acc[k];
  }
}

// Half-sweeps S .. SEND-1 on regions LO = S + LOOFF (one cell smaller each time).
template <int RW, int RH, int S, int SEND, int LOOFF, int NT>
struct HalfSweeps {
  static __device__ __forceinline__ void run(float* sx, const float* sr, int par0, float h2) {
    lds_half_sweep<RW, RH, S + LOOFF, NT>(sx, sr, S & 1, par0, h2);
    __syncthreads();
    HalfSweeps<RW, RH, S + 1, SEND, LOOFF, NT>::run(sx, sr, par0, h2);
  }
};
template <int RW, int RH, int SEND, int LOOFF, int NT>
struct HalfSweeps<RW, RH, SEND, SEND, LOOFF, NT> {
  static __device__ __forceinline__ void run(float*, const float*, int, float) {}
};

// Pre-smoothing from zero: red half-sweep on the whole region, then 2*SW-1
// half-sweeps on shrinking regions.
template <int SW, int RW, int RH, int NT>
__device__ __forceinline__ void tile_smooth_from_zero(float* __restrict__ sx, const float* __restrict__ sr, int par0,
                                                      float h2) {
  for (int q = threadIdx.x; q < RW * RH; q += NT) {
    int ly = q / RW, lx = q - ly * RW;
    bool red = ((lx + ly + par0) & 1) == 0;
    sx[q] = red ? (0.f - h2 * sr[q]) * 0.25f : 0.f;
  }
  __syncthreads();
  HalfSweeps<RW, RH, 1, 2 * SW, 0, NT>::run(sx, sr, par0, h2);
}

// Store the tile interior and the restricted residual of each 2x2 block of it.
template <int H, int RW, class TS, typename S>
__device__ __forceinline__ void tile_store_and_restrict(const float* __restrict__ sx, const float* __restrict__ sr,
                                                        S* __restrict__ x, S* __restrict__ crhs, int n,
                                                        int tx0, int ty0, float ih2) {
  constexpr int TX = TS::TX, TY = TS::TY, NT = TS::NT;
  for (int q = threadIdx.x; q < TX * TY; q += NT) {
    int iy = q / TX, ix = q - iy * TX;
    int gi = tx0 + ix, gj = ty0 + iy;
    if (gi < n && gj < n) x[(size_t)gj * n + gi] = st<S>(sx[(iy + H) * RW + ix + H]);
  }
  constexpr int CX = TX / 2, CY = TY / 2;
  int cn = n >> 1;
  for (int q = threadIdx.x; q < CX * CY; q += NT) {
    int cy = q / CX, cx = q - cy * CX;
    int I = (tx0 >> 1) + cx, J = (ty0 >> 1) + cy;
    if (I >= cn || J >= cn) continue;
    float s = 0.f;
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int a = 0; a < 2; a++) {
        int k = (2 * cy + b + H) * RW + (2 * cx + a + H);
        float lap = (sx[k - 1] + sx[k + 1] + sx[k - RW] + sx[k + RW] - 4.f * sx[k]) * ih2;
        s += sr[k] - lap;
      }
    crhs[(size_t)J * cn + I] = st<S>(s * 0.25f);
  }
}

// -------------------------------------------------------------- level >= 1 down

template <int SW, class TS, typename S>
__global__ void __launch_bounds__(TS::NT)
k_mgt_down(S* __restrict__ x, const S* __restrict__ rhs, int n, float h2, float ih2,
           S* __restrict__ crhs, int tiles_x, int nblocks, Gate gate) {
  if (gate_closed(gate)) return;
  constexpr int TX = TS::TX, TY = TS::TY, NT = TS::NT;
  constexpr int H = 2 * SW;
  constexpr int RW = TX + 2 * H, RH = TY + 2 * H;
  __shared__ float sx[RH * RW];
  __shared__ float sr[RH * RW];
  int t = xcd_tile(nblocks);
  int tx0 = (t % tiles_x) * TX, ty0 = (t / tiles_x) * TY;
  int ox = tx0 - H, oy = ty0 - H;  // global coordinate of LDS (0,0)
  load_window<RW, RH, NT>(sr, rhs, n, ox, oy);
  __syncthreads();
  tile_smooth_from_zero<SW, RW, RH, NT>(sx, sr, (ox + oy) & 1, h2);
  tile_store_and_restrict<H, RW, TS>(sx, sr, x, crhs, n, tx0, ty0, ih2);
}

// Two grid-wide levels down in one launch, level l (n0) and level l+1 (n0 / 2),
// for the small levels whose single launches are latency-bound. A workgroup owns a
// TX x TY tile of level l+1 and the 2TX x 2TY cells of level l under it. It smooths
// level l from zero on the region whose residual covers the level-(l+1) tile and its
// smoothing halo (neighbouring workgroups recompute the overlap), restricts that
// residual into LDS at the stored precision, then runs the level-(l+1) down-sweep on
// it exactly as k_mgt_down does. Every stored value (x at both levels, the rhs of
// level l+1 and of level l+2) equals the two separate launches' bit for bit.
template <int SW, class TS, typename S>
__global__ void __launch_bounds__(TS::NT)
k_mgt_down2(S* __restrict__ x0, const S* __restrict__ rhs0, int n0, float h20, float ih20, S* __restrict__ x1,
            S* __restrict__ rhs1, float h21, float ih21, S* __restrict__ crhs, int tiles_x, int nblocks,
            Gate gate) {
  if (gate_closed(gate)) return;
  constexpr int TX = TS::TX, TY = TS::TY, NT = TS::NT;
  constexpr int H = 2 * SW;
  constexpr int RW1 = TX + 2 * H, RH1 = TY + 2 * H;            // level l+1 region
  constexpr int RW0 = 2 * RW1 + 2 * H, RH0 = 2 * RH1 + 2 * H;  // level l region: residual on 2 RW1 x 2 RH1
  __shared__ float sx0[RH0 * RW0];
  __shared__ float sr0[RH0 * RW0];
  __shared__ float sx1[RH1 * RW1];
  __shared__ float sr1[RH1 * RW1];
  const int n1 = n0 >> 1;
  const int t = xcd_tile(nblocks);
  const int tx1 = (t % tiles_x) * TX, ty1 = (t / tiles_x) * TY;
  const int ox1 = tx1 - H, oy1 = ty1 - H;          // level-(l+1) coordinate of sx1 (0, 0)
  const int ox0 = 2 * ox1 - H, oy0 = 2 * oy1 - H;  // level-l coordinate of sx0 (0, 0)
  load_window<RW0, RH0, NT>(sr0, rhs0, n0, ox0, oy0);
  __syncthreads();
  // after 2 SW half-sweeps from zero, x is exact on [H - 1, RW0 - H + 1) and the
  // residual on [H, RW0 - H), the 2 RW1 x 2 RH1 block under the level-(l+1) region
  tile_smooth_from_zero<SW, RW0, RH0, NT>(sx0, sr0, (ox0 + oy0) & 1, h20);
  for (int q = threadIdx.x; q < 4 * TX * TY; q += NT) {
    const int iy = q / (2 * TX), ix = q - iy * (2 * TX);
    x0[(size_t)(2 * ty1 + iy) * n0 + 2 * tx1 + ix] = st<S>(sx0[(3 * H + iy) * RW0 + 3 * H + ix]);
  }
  for (int q = threadIdx.x; q < RW1 * RH1; q += NT) {
    const int cy = q / RW1, cx = q - cy * RW1;
    float s = 0.f;
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int a = 0; a < 2; a++) {
        const int k = (2 * cy + b + H) * RW0 + (2 * cx + a + H);
        const float lap = (sx0[k - 1] + sx0[k + 1] + sx0[k - RW0] + sx0[k + RW0] - 4.f * sx0[k]) * ih20;
        s += sr0[k] - lap;
      }
    sr1[q] = rnd<S>(s * 0.25f);  // as the level-(l+1) sweep would load it
    if (cx >= H && cx < H + TX && cy >= H && cy < H + TY)
      rhs1[(size_t)(oy1 + cy) * n1 + ox1 + cx] = st<S>(s * 0.25f);
  }
  __syncthreads();
  tile_smooth_from_zero<SW, RW1, RH1, NT>(sx1, sr1, (ox1 + oy1) & 1, h21);
  tile_store_and_restrict<H, RW1, TS>(sx1, sr1, x1, crhs, n1, tx1, ty1, ih21);
}

// -------------------------------------------------------------- level 0 down

// FP64 refinement residual fused with the level-0 pre-smoothing and restriction.
// p (FP64, may be null = zero) and e_prev (FP32, may be null = zero) give p' = p + e_prev;
// r = div u* - L p' (src/core/grid.js:191-206, src/cpu/projection.js:48-62) on the
// tile + halo 2*SW; p' and r are stored for the interior and the tile's max|r| goes
// to rpart[tile]. residual_only stops after that.
template <int SW, class TS, typename S>
__global__ void __launch_bounds__(TS::NT)
k_mgt_down0(const float* __restrict__ u, const float* __restrict__ v, const double* __restrict__ p,
            const float* __restrict__ e_prev, double* __restrict__ pout, S* __restrict__ r32,
            S* __restrict__ x, S* __restrict__ crhs, int n, double idx, double idx2, float h2, float ih2,
            int tiles_x, int nblocks, double* __restrict__ rpart, int residual_only, Gate gate) {
  if (gate_closed(gate)) return;
  constexpr int TX = TS::TX, TY = TS::TY, NT = TS::NT;
  constexpr int H = 2 * SW;
  constexpr int RW = TX + 2 * H, RH = TY + 2 * H;
  constexpr int PW = RW + 2, PH = RH + 2;  // p' needs one more ring
  __shared__ double sp[PH * PW];
  __shared__ float sx[RH * RW];
  __shared__ float sr[RH * RW];
  __shared__ double wm[NT / 64];
  int t = xcd_tile(nblocks);
  int tx0 = (t % tiles_x) * TX, ty0 = (t / tiles_x) * TY;
  int ox = tx0 - H, oy = ty0 - H;
  // every load (p, e, the face velocities) in flight before the first LDS store
  constexpr int CNTV = RW * RH, KV = (CNTV + NT - 1) / NT;
  float uL[KV], uR[KV], vB[KV], vT[KV];
#pragma unroll
  for (int k = 0; k < KV; k++) {
    int q = threadIdx.x + k * NT;
    if (q < CNTV) {
      int ly = q / RW, lx = q - ly * RW;
      int gi = wrapn(ox + lx, n), gj = wrapn(oy + ly, n);
      int gip = gi + 1 == n ? 0 : gi + 1, gjp = gj + 1 == n ? 0 : gj + 1;
      uL[k] = u[(size_t)gj * n + gi];
      uR[k] = u[(size_t)gj * n + gip];
      vB[k] = v[(size_t)gj * n + gi];
      vT[k] = v[(size_t)gjp * n + gi];
    }
  }
  {
    constexpr int CNT = PW * PH, K = (CNT + NT - 1) / NT;
    double tp[K];
    float te[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
      int q = threadIdx.x + k * NT;
      if (q < CNT) {
        int ly = q / PW, lx = q - ly * PW;
        size_t g = (size_t)wrapn(oy - 1 + ly, n) * n + wrapn(ox - 1 + lx, n);
        tp[k] = p ? p[g] : 0.0;
        te[k] = e_prev ? e_prev[g] : 0.f;
      }
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
      int q = threadIdx.x + k * NT;
      if (q < CNT) sp[q] = e_prev ? tp[k] + (double)te[k] : tp[k];
    }
  }
  __syncthreads();
  double amax = 0.0;
  {
    constexpr int CNT = CNTV, K = KV;
#pragma unroll
    for (int k = 0; k < K; k++) {
      int q = threadIdx.x + k * NT;
      if (q < CNT) {
        int ly = q / RW, lx = q - ly * RW;
        int kp = (ly + 1) * PW + lx + 1;
        double pc = sp[kp];
        double lap = (sp[kp - 1] - 2 * pc + sp[kp + 1]) * idx2 + (sp[kp - PW] - 2 * pc + sp[kp + PW]) * idx2;
        double dv = ((double)uR[k] - (double)uL[k]) * idx + ((double)vT[k] - (double)vB[k]) * idx;
        double r = dv - lap;
        sr[q] = rnd<S>((float)r);  // the smoother sees the stored right-hand side
        bool inner = lx >= H && lx < H + TX && ly >= H && ly < H + TY && ox + lx < n && oy + ly < n;
        if (inner) {
          size_t g = (size_t)(oy + ly) * n + (ox + lx);
          pout[g] = pc;
          r32[g] = st<S>((float)r);
          amax = fmax(amax, nan_to_inf(fabs(r)));
        }
      }
    }
  }
  {
    double m = wave_max(amax);
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      double b = wm[0];
      for (int w = 1; w < NT / 64; w++) b = fmax(b, wm[w]);
      rpart[t] = b;
      gate_record(gate, b);
    }
  }
  if (residual_only) return;
  tile_smooth_from_zero<SW, RW, RH, NT>(sx, sr, (ox + oy) & 1, h2);
  tile_store_and_restrict<H, RW, TS>(sx, sr, x, crhs, n, tx0, ty0, ih2);
}

// ------------------------------------------- fixed thread -> cell-pair map passes

// Thread t owns the cell pairs (2p, 2p + 1) of rows r, r + RG, r + 2 RG, ... of an
// RW x RH LDS region, p = t % PC, r = t / PC. The red-black passes below update every
// interior cell of one colour on the whole region instead of shrinking it by a cell per
// half-sweep: a cell at distance d from the region's edge is exact after k half-sweeps
// whenever d >= k, as in the shrinking passes, and the cells nearer the edge are never
// read by a consumer. Each cell sees the same operands in the same order as
// lds_half_sweep, so the results are the same bits; the indices are computed once per
// launch instead of once per pass.
template <int RW, int RH, int NT>
struct PairMap {
  static constexpr int PC = RW / 2, RG = NT / PC, ROWS = (RH + RG - 1) / RG;
  static_assert(RW % 2 == 0 && RG >= 1, "pair map needs an even width and a row of pairs per group");
  int base, r, p;
  bool on;
  __device__ __forceinline__ PairMap() {
    const int t = threadIdx.x;
    r = t / PC;
    p = t - r * PC;
    on = r < RG;
    base = r * RW + 2 * p;
  }
  __device__ __forceinline__ bool row_ok(int k) const { return on && r + k * RG < RH; }
};

// Red pass from zero on the whole region: x = -h2 rhs / 4 on red cells, 0 on black.
template <int RW, int RH, int NT>
__device__ __forceinline__ void fm_from_zero(float* __restrict__ sx, const float* __restrict__ sr,
                                             const PairMap<RW, RH, NT>& m, int par0, float h2) {
  using M = PairMap<RW, RH, NT>;
#pragma unroll
  for (int k = 0; k < M::ROWS; k++) {
    if (!m.row_ok(k)) continue;
    const int c = m.base + k * M::RG * RW;
    const bool red0 = ((m.r + k * M::RG + par0) & 1) == 0;  // is cell 2p red
    const float2 r2 = *(const float2*)(sr + c);
    float2 x2;
    x2.x = red0 ? (0.f - h2 * r2.x) * 0.25f : 0.f;
    x2.y = red0 ? 0.f : (0.f - h2 * r2.y) * 0.25f;
    *(float2*)(sx + c) = x2;
  }
}

// One colour pass on every interior cell of the colour (rows 1 .. RH-2, columns 1 .. RW-2).
template <int RW, int RH, int NT>
__device__ __forceinline__ void fm_half_sweep(float* __restrict__ sx, const float* __restrict__ sr,
                                              const PairMap<RW, RH, NT>& m, int colour, int par0, float h2) {
  using M = PairMap<RW, RH, NT>;
  float acc[M::ROWS];
  int cc[M::ROWS];
  bool act[M::ROWS];
#pragma unroll
  for (int k = 0; k < M::ROWS; k++) {
    const int ly = m.r + k * M::RG;
    const int off = (colour + par0 + ly) & 1;  // the colour's cell of the pair
    act[k] = m.row_ok(k) && ly >= 1 && ly <= RH - 2 && !(off == 0 && m.p == 0) && !(off == 1 && m.p == M::PC - 1);
    const int c = m.base + k * M::RG * RW + off;
    cc[k] = c;
    if (act[k]) {
      const float sum = sx[c - 1] + sx[c + 1] + sx[c - RW] + sx[c + RW];
      acc[k] = (sum - h2 * sr[c]) * 0.25f;
    }
  }
#pragma unroll
  for (int k = 0; k < M::ROWS; k++)
    if (act[k]) sx[cc[k]] = acc[k];
}

// Pre-smoothing from zero with SW sweeps (2 SW half-sweeps: red from zero, then black,
// red, ...), the last one followed by a barrier.
template <int SW, int RW, int RH, int NT>
__device__ __forceinline__ void fm_smooth_from_zero(float* __restrict__ sx, const float* __restrict__ sr, int par0,
                                                    float h2) {
  const PairMap<RW, RH, NT> m;
  fm_from_zero<RW, RH, NT>(sx, sr, m, par0, h2);
  __syncthreads();
#pragma unroll
  for (int s = 1; s < 2 * SW; s++) {
    fm_half_sweep<RW, RH, NT>(sx, sr, m, s & 1, par0, h2);
    __syncthreads();
  }
}

template <typename S>
__device__ __forceinline__ void store4(S* __restrict__ dst, float a, float b, float c, float d);
template <>
__device__ __forceinline__ void store4<float>(float* __restrict__ dst, float a, float b, float c, float d) {
  *(float4*)dst = make_float4(a, b, c, d);
}
template <>
__device__ __forceinline__ void store4<bf16>(bf16* __restrict__ dst, float a, float b, float c, float d) {
  const uint32_t lo = (uint32_t)st<bf16>(a).b | ((uint32_t)st<bf16>(b).b << 16);
  const uint32_t hi = (uint32_t)st<bf16>(c).b | ((uint32_t)st<bf16>(d).b << 16);
  *(uint2*)dst = make_uint2(lo, hi);
}

// k_mgt_down0 with every global access 8 or 16 bytes wide and fixed thread maps: u, v, p
// and the previous correction are staged in LDS (p + e in FP64), the FP64 residual is
// evaluated four cells per thread from LDS, the passes use PairMap, and x leaves four
// cells per store. Same operands in the same order as k_mgt_down0, so the same bits.
// The host uses it when H = 2 SW is a multiple of 4 and n is a multiple of 4, TX and TY
// (tiles are whole, quads and pairs never straddle the periodic seam, and every
// 16-byte access is aligned); n * n < 2^31 (32-bit offsets).
template <int SW, class TS, typename S>
__global__ void __launch_bounds__(TS::NT)
k_mgt_down0q(const float* __restrict__ u, const float* __restrict__ v, const double* __restrict__ p,
             const float* __restrict__ e_prev, double* __restrict__ pout, S* __restrict__ r32, S* __restrict__ x,
             S* __restrict__ crhs, int n, double idx, double idx2, float h2, float ih2, int tiles_x, int nblocks,
             double* __restrict__ rpart, int residual_only, Gate gate) {
  if (gate_closed(gate)) return;
  constexpr int TX = TS::TX, TY = TS::TY, NT = TS::NT;
  constexpr int H = 2 * SW;
  static_assert(H % 4 == 0 && TX % 4 == 0, "quad accesses need H and TX multiples of 4");
  constexpr int RW = TX + 2 * H, RH = TY + 2 * H;
  constexpr int UQ = RW / 4 + 1, UW = 4 * UQ;  // u: columns ox .. ox + RW (one past the region for uR)
  constexpr int VQ = RW / 4;                   // v: rows oy .. oy + RH (one past for vT)
  constexpr int PP = RW / 2 + 2, PW = 2 * PP;  // p, e: columns ox - 2 .. ox + RW + 1 in pairs
  constexpr int PH = RH + 2;                   // rows oy - 1 .. oy + RH
  __shared__ __align__(16) double sp[PH * PW];
  __shared__ __align__(16) float sxu[RH * UW];  // u window, then the correction x (row stride RW)
  // AL: the residual takes the v window's LDS once every thread has read it, so that
  // tiles of at most 512 threads, several to a CU, fit one more per CU (at 2048^2 the
  // 32x32 tile: 34.8 -> 28.4 KB, 35.7 -> 33.2 us); a 1024-thread tile, alone on its CU,
  // writes it directly and skips the extra barrier
  constexpr bool AL = NT <= 512;
  __shared__ __align__(16) float svr[(RH + 1) * RW + (AL ? 0 : RH * RW)];
  __shared__ double wm[NT / 64];
  float* __restrict__ sv = svr;
  float* __restrict__ sr = AL ? svr : svr + (RH + 1) * RW;
  float* __restrict__ su = sxu;
  float* __restrict__ sx = sxu;
  const int t = xcd_tile(nblocks);
  const int tx0 = (t % tiles_x) * TX, ty0 = (t / tiles_x) * TY;
  const int ox = tx0 - H, oy = ty0 - H;
  const int tid = threadIdx.x;
  // ---- every load in flight before the first LDS store
  constexpr int CU = UQ * RH, KU = (CU + NT - 1) / NT;
  constexpr int CV = VQ * (RH + 1), KV = (CV + NT - 1) / NT;
  constexpr int CP = PP * PH, KP = (CP + NT - 1) / NT;
  float tu[KU][4], tv[KV][4];
  double tp[KP][2];
  float te[KP][2];
#pragma unroll
  for (int k = 0; k < KU; k++) {
    const int q = tid + k * NT;
    if (q < CU) {
      const int ly = q / UQ, qx = q - ly * UQ;
      const float4 w = *(const float4*)(u + wrapn(oy + ly, n) * n + wrapn(ox + 4 * qx, n));
      tu[k][0] = w.x, tu[k][1] = w.y, tu[k][2] = w.z, tu[k][3] = w.w;
    }
  }
#pragma unroll
  for (int k = 0; k < KV; k++) {
    const int q = tid + k * NT;
    if (q < CV) {
      const int ly = q / VQ, qx = q - ly * VQ;
      const float4 w = *(const float4*)(v + wrapn(oy + ly, n) * n + wrapn(ox + 4 * qx, n));
      tv[k][0] = w.x, tv[k][1] = w.y, tv[k][2] = w.z, tv[k][3] = w.w;
    }
  }
#pragma unroll
  for (int k = 0; k < KP; k++) {
    const int q = tid + k * NT;
    tp[k][0] = tp[k][1] = 0.0;
    te[k][0] = te[k][1] = 0.f;
    if (q < CP) {
      const int ly = q / PP, qx = q - ly * PP;
      const int g = wrapn(oy - 1 + ly, n) * n + wrapn(ox - 2 + 2 * qx, n);
      if (p) {
        const double2 w = *(const double2*)(p + g);
        tp[k][0] = w.x, tp[k][1] = w.y;
      }
      if (e_prev) {
        const float2 w = *(const float2*)(e_prev + g);
        te[k][0] = w.x, te[k][1] = w.y;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < KU; k++) {
    const int q = tid + k * NT;
    if (q < CU) *(float4*)(su + 4 * q) = make_float4(tu[k][0], tu[k][1], tu[k][2], tu[k][3]);
  }
#pragma unroll
  for (int k = 0; k < KV; k++) {
    const int q = tid + k * NT;
    if (q < CV) *(float4*)(sv + 4 * q) = make_float4(tv[k][0], tv[k][1], tv[k][2], tv[k][3]);
  }
#pragma unroll
  for (int k = 0; k < KP; k++) {
    const int q = tid + k * NT;
    if (q < CP) {
      double2 s;
      s.x = e_prev ? tp[k][0] + (double)te[k][0] : tp[k][0];
      s.y = e_prev ? tp[k][1] + (double)te[k][1] : tp[k][1];
      *(double2*)(sp + 2 * q) = s;
    }
  }
  __syncthreads();
  // ---- FP64 residual r = div u* - L p' on the region, four cells per thread
  double amax = 0.0;
  {
    // AL: the residual stays in registers until every thread has read its v window
    constexpr int CQ = VQ * RH, KQ = (CQ + NT - 1) / NT;
    float4 rq[AL ? KQ : 1];
#pragma unroll
    for (int k = 0; k < KQ; k++) {
      const int q = tid + k * NT;
      if (q < CQ) {
        const int ly = q / VQ, qx = q - ly * VQ, lx0 = 4 * qx;
        const float4 ul = *(const float4*)(su + ly * UW + lx0);
        const float ur4 = su[ly * UW + lx0 + 4];
        const float4 vb = *(const float4*)(sv + ly * RW + lx0);
        const float4 vt = *(const float4*)(sv + (ly + 1) * RW + lx0);
        const double* pc = sp + (ly + 1) * PW + lx0 + 2;  // p' of cell (lx0, ly)
        const double p0 = pc[0], p1 = pc[1], p2 = pc[2], p3 = pc[3];
        auto res = [&](double pl, double c, double pr, double pb, double pt, float uL, float uR, float vB,
                       float vT) {
          const double lap = (pl - 2 * c + pr) * idx2 + (pb - 2 * c + pt) * idx2;
          const double dv = ((double)uR - (double)uL) * idx + ((double)vT - (double)vB) * idx;
          return dv - lap;
        };
        const double r0 = res(pc[-1], p0, p1, pc[-PW], pc[PW], ul.x, ul.y, vb.x, vt.x);
        const double r1 = res(p0, p1, p2, pc[1 - PW], pc[1 + PW], ul.y, ul.z, vb.y, vt.y);
        const double r2 = res(p1, p2, p3, pc[2 - PW], pc[2 + PW], ul.z, ul.w, vb.z, vt.z);
        const double r3 = res(p2, p3, pc[4], pc[3 - PW], pc[3 + PW], ul.w, ur4, vb.w, vt.w);
        const float4 rv = make_float4(rnd<S>((float)r0), rnd<S>((float)r1), rnd<S>((float)r2), rnd<S>((float)r3));
        if constexpr (AL) rq[k] = rv;
        else *(float4*)(sr + ly * RW + lx0) = rv;
        if (lx0 >= H && lx0 < H + TX && ly >= H && ly < H + TY) {
          const int g = (oy + ly) * n + ox + lx0;
          *(double2*)(pout + g) = make_double2(p0, p1);
          *(double2*)(pout + g + 2) = make_double2(p2, p3);
          store4<S>(r32 + g, (float)r0, (float)r1, (float)r2, (float)r3);
          amax = fmax(amax, fmax(fmax(nan_to_inf(fabs(r0)), nan_to_inf(fabs(r1))),
                                 fmax(nan_to_inf(fabs(r2)), nan_to_inf(fabs(r3)))));
        }
      }
    }
    if constexpr (AL) {
      __syncthreads();
#pragma unroll
      for (int k = 0; k < KQ; k++) {
        const int q = tid + k * NT;
        if (q < CQ) {
          const int ly = q / VQ, lx0 = 4 * (q - ly * VQ);
          *(float4*)(sr + ly * RW + lx0) = rq[k];
        }
      }
    }
  }
  {
    double m = wave_max(amax);
    if ((tid & 63) == 0) wm[tid >> 6] = m;
    __syncthreads();
    if (tid == 0) {
      double b = wm[0];
      for (int w = 1; w < NT / 64; w++) b = fmax(b, wm[w]);
      rpart[t] = b;
      gate_record(gate, b);
    }
  }
  if (residual_only) return;
  // ---- S sweeps from zero (the u window is no longer read: sx reuses its LDS)
  fm_smooth_from_zero<SW, RW, RH, NT>(sx, sr, (ox + oy) & 1, h2);
  // ---- x on the tile, four cells per store; the restricted residual of each 2x2 block
  constexpr int XQ = TX / 4, CXQ = XQ * TY;
  for (int q = tid; q < CXQ; q += NT) {
    const int iy = q / XQ, m4 = 4 * (q - iy * XQ);
    const float4 xv = *(const float4*)(sx + (iy + H) * RW + H + m4);
    store4<S>(x + (ty0 + iy) * n + tx0 + m4, xv.x, xv.y, xv.z, xv.w);
  }
  constexpr int CX = TX / 2, CY = TY / 2;
  const int cn = n >> 1;
  for (int q = tid; q < CX * CY; q += NT) {
    const int cy = q / CX, cx = q - cy * CX;
    float s = 0.f;
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int a = 0; a < 2; a++) {
        const int k = (2 * cy + b + H) * RW + (2 * cx + a + H);
        const float lap = (sx[k - 1] + sx[k + 1] + sx[k - RW] + sx[k + RW] - 4.f * sx[k]) * ih2;
        s += sr[k] - lap;
      }
    crhs[((ty0 >> 1) + cy) * cn + (tx0 >> 1) + cx] = st<S>(s * 0.25f);
  }
}

// ------------------------------------------------- exact coarsest-level solve
// The 5-point Laplacian of an n x n level (n <= COARSE_N) is T (x) I + I (x) T with T
// the 1D second difference, and T = Q diag(lam) Q^T for an orthonormal eigenbasis Q:
// the real Fourier basis on a periodic level, the DCT-II basis on a walled (zero-flux)
// one; the host builds Q and W for either. So
// L x = r - mean(r) is solved by x = Q ((Q^T r Q) * W) Q^T with W[a][b] = 1 / (lam_a +
// lam_b) and W[0][0] = 0 (the zero-mean solution, as a pseudo-inverse gives). The four
// 64 x 64 x 64 products run on the BF16 matrix cores (v_mfma_f32_16x16x32_bf16) with
// every operand split into a BF16 head and tail, x = h + l with h = bf16(x) and
// l = bf16(x - h): a b ~ ha hb + ha lb + la hb accumulated in FP32 (the dropped la lb
// and the tails' own rounding leave ~2^-16 relative per product). One 16 x 16 output
// block per wavefront; Q and W come from the host (built in FP64), zero-padded to
// 64 x 64 and laid out in the order the lanes read them, so n < 64 needs no special case.
//
// v_mfma_f32_16x16x32_bf16: lane l supplies A[l % 16][8 (l / 16) + t] and
// B[8 (l / 16) + t][l % 16], t = 0..7, and receives D[4 (l / 16) + i][l % 16] in its
// i-th accumulator.
constexpr int EXN = 64;   // padded level size
constexpr int EXB = 72;   // LDS row pitch (BF16): 16-byte reads of 16 rows hit distinct banks
constexpr int EXT = 1024; // 16 wavefronts, one 16 x 16 block of the 64 x 64 result each
// Host-built operand table, in 16-byte units: Q[k][16 T + i] (T the block, i = l % 16,
// k = 32 kk + 8 (l / 16) + t) heads then tails, Q[16 T + i][k] heads then tails (each
// [T 4][lane 64][kk 2] x 8 BF16), then W at each lane's four outputs ([wave 16][lane 64]
// x 4 FP32).
constexpr int EX_TAB = 4 * 512 + 1024;
typedef __bf16 ex_bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 ex_bf16x4 __attribute__((ext_vector_type(4)));
typedef float ex_f32x4 __attribute__((ext_vector_type(4)));

// x into LDS as head (at h[o]) and tail (at l[o])
__device__ __forceinline__ void ex_put(__bf16* h, __bf16* l, int o, float x) {
  const __bf16 hx = (__bf16)x;
  h[o] = hx;
  l[o] = (__bf16)(x - (float)hx);
}

// acc over the 64-long k range: heads and tails of the two operands of k-chunk kk
__device__ __forceinline__ ex_f32x4 ex_mma3(ex_f32x4 acc, ex_bf16x8 ah, ex_bf16x8 al, ex_bf16x8 bh, ex_bf16x8 bl) {
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
}

// The exact solve by the NT / 64 wavefronts of a workgroup, for any 16 x 16 output
// blocks: products 1 and 2 on all 16 blocks, products 3 and 4 on the blocks in the row
// blocks of rbm (bit tm) and, for the last product, the column blocks of cbm (bit tn);
// put(row, col, x) receives every element of the last product's blocks (row, col < n).
// Sixteen wavefronts compute block (w / 4, w % 4) of every product; four compute the
// blocks (w, j) of product 1 and (j, w) of the others, j = 0..3. Each block is the same
// MFMA sequence on the same operands either way, so the same bits.
// LDS: the four 64 x EXB BF16 matrices sah, sal, sbh, sbl (A and B, heads and tails).
template <int NT, int NS, typename S, class Put>
__device__ __forceinline__ void ex_solve(const S* __restrict__ rg, int n, const uint4* __restrict__ tab,
                                         __bf16* __restrict__ sah, __bf16* __restrict__ sal, __bf16* __restrict__ sbh,
                                         __bf16* __restrict__ sbl, unsigned rbm, unsigned cbm, Put put) {
  constexpr int NW = NT / 64, NJ = 16 / NW;
  static_assert(NW == 16 || NW == 4, "the exact solve runs on 4 or 16 wavefronts");
  static_assert(NS >= 1 && NS <= NJ, "row-block slots of products 3 and 4");
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lg = lane >> 4;
  // block j of wavefront w: product 1 (tm, tn) = P1, the others P2
  auto tm1 = [&](int j) { return NW == 16 ? w >> 2 : w; };
  auto tn1 = [&](int j) { return NW == 16 ? w & 3 : j; };
  auto tm2 = [&](int j) { return NW == 16 ? w >> 2 : j; };
  auto tn2 = [&](int j) { return NW == 16 ? w & 3 : w; };
  // the right-hand side first (four consecutive cells of one row per thread), then the
  // Q operands in the order the products use them
  constexpr int KR = 1024 / NT;
  float rv[KR][4];
#pragma unroll
  for (int k = 0; k < KR; k++) {
    const int q = threadIdx.x + k * NT, ri = q >> 4, rj = 4 * (q & 15);
    if (n == EXN) {
      if constexpr (sizeof(S) == 4) {
        const float4 t = *(const float4*)((const float*)rg + ri * EXN + rj);
        rv[k][0] = t.x, rv[k][1] = t.y, rv[k][2] = t.z, rv[k][3] = t.w;
      } else {
        const uint2 t = *(const uint2*)((const uint16_t*)rg + ri * EXN + rj);
        rv[k][0] = __uint_as_float(t.x << 16), rv[k][1] = __uint_as_float(t.x & 0xffff0000u);
        rv[k][2] = __uint_as_float(t.y << 16), rv[k][3] = __uint_as_float(t.y & 0xffff0000u);
      }
    } else {
#pragma unroll
      for (int u = 0; u < 4; u++) rv[k][u] = (ri < n && rj + u < n) ? ld(rg[ri * n + rj + u]) : 0.f;
    }
  }
  auto q8 = [&](int part, int T, int kk) { return __builtin_bit_cast(ex_bf16x8, tab[part * 512 + (T * 64 + lane) * 2 + kk]); };
  // operands of products 1 and 2 (fixed per wavefront), W at the lane's outputs
  ex_bf16x8 a1h[2], a1l[2], b2h[2], b2l[2];
  float4 wv[NJ];
#pragma unroll
  for (int kk = 0; kk < 2; kk++) a1h[kk] = q8(0, tm1(0), kk), a1l[kk] = q8(1, tm1(0), kk);
#pragma unroll
  for (int kk = 0; kk < 2; kk++) b2h[kk] = q8(0, tn2(0), kk), b2l[kk] = q8(1, tn2(0), kk);
#pragma unroll
  for (int j = 0; j < NJ; j++) wv[j] = ((const float4*)(tab + 2048))[(4 * tm2(j) + tn2(j)) * 64 + lane];
  // r transposed into A: RT[j][i] = r[i][j]
#pragma unroll
  for (int k = 0; k < KR; k++) {
    const int q = threadIdx.x + k * NT, ri = q >> 4, rj = 4 * (q & 15);
#pragma unroll
    for (int u = 0; u < 4; u++) ex_put(sah, sal, (rj + u) * EXB + ri, rv[k][u]);
  }
  __syncthreads();
  // C1 = Q^T r (B operand RT rows), row-major into B
#pragma unroll
  for (int j = 0; j < NJ; j++) {
    const int m0 = 16 * tm1(j), n0 = 16 * tn1(j);
    ex_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < 2; kk++) {
      const int o = (n0 + li) * EXB + 32 * kk + 8 * lg;
      acc = ex_mma3(acc, a1h[kk], a1l[kk], *(const ex_bf16x8*)(sah + o), *(const ex_bf16x8*)(sal + o));
    }
#pragma unroll
    for (int i = 0; i < 4; i++) ex_put(sbh, sbl, (m0 + 4 * lg + i) * EXB + n0 + li, acc[i]);
  }
  // the operands of products 3 and 4 load while the products run: the row blocks of
  // rbm this wavefront computes, at most NS of them (slot tm -1: none)
  int stm[NS];
  {
    int k = 0;
#pragma unroll
    for (int j = 0; j < NJ; j++) {
      const bool on = (rbm >> tm2(j)) & 1;
#pragma unroll
      for (int q = 0; q < NS; q++)
        if (q == k) stm[q] = on ? tm2(j) : -1;
      k += on ? 1 : 0;
    }
#pragma unroll
    for (int q = 0; q < NS; q++)
      if (q >= k) stm[q] = -1;
  }
  ex_bf16x8 a3h[NS][2], a3l[NS][2], b4h[2], b4l[2];
#pragma unroll
  for (int q = 0; q < NS; q++)
    if (stm[q] >= 0)
#pragma unroll
      for (int kk = 0; kk < 2; kk++) a3h[q][kk] = q8(2, stm[q], kk), a3l[q][kk] = q8(3, stm[q], kk);
  if ((cbm >> tn2(0)) & 1)
#pragma unroll
    for (int kk = 0; kk < 2; kk++) b4h[kk] = q8(2, tn2(0), kk), b4l[kk] = q8(3, tn2(0), kk);
  __syncthreads();
  // C2 = (C1 Q) * W (A operand C1 rows), transposed into A: C2T[n][m]
#pragma unroll
  for (int j = 0; j < NJ; j++) {
    const int m0 = 16 * tm2(j), n0 = 16 * tn2(j);
    ex_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < 2; kk++) {
      const int o = (m0 + li) * EXB + 32 * kk + 8 * lg;
      acc = ex_mma3(acc, *(const ex_bf16x8*)(sbh + o), *(const ex_bf16x8*)(sbl + o), b2h[kk], b2l[kk]);
    }
    const float c[4] = {acc[0] * wv[j].x, acc[1] * wv[j].y, acc[2] * wv[j].z, acc[3] * wv[j].w};
    ex_bf16x4 h4, l4;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const __bf16 hx = (__bf16)c[i];
      h4[i] = hx;
      l4[i] = (__bf16)(c[i] - (float)hx);
    }
    const int o = (n0 + li) * EXB + m0 + 4 * lg;
    *(ex_bf16x4*)(sah + o) = h4;
    *(ex_bf16x4*)(sal + o) = l4;
  }
  __syncthreads();
  // C3 = Q C2 (B operand C2T rows), row-major into B, on the row blocks of rbm
#pragma unroll
  for (int q = 0; q < NS; q++) {
    if (stm[q] < 0) continue;
    const int m0 = 16 * stm[q], n0 = 16 * tn2(0);
    ex_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < 2; kk++) {
      const int o = (n0 + li) * EXB + 32 * kk + 8 * lg;
      acc = ex_mma3(acc, a3h[q][kk], a3l[q][kk], *(const ex_bf16x8*)(sah + o), *(const ex_bf16x8*)(sal + o));
    }
#pragma unroll
    for (int i = 0; i < 4; i++) ex_put(sbh, sbl, (m0 + 4 * lg + i) * EXB + n0 + li, acc[i]);
  }
  __syncthreads();
  // x = C3 Q^T (A operand C3 rows) on the blocks of rbm x cbm
  if (!((cbm >> tn2(0)) & 1)) return;
#pragma unroll
  for (int q = 0; q < NS; q++) {
    if (stm[q] < 0) continue;
    const int m0 = 16 * stm[q], n0 = 16 * tn2(0);
    ex_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < 2; kk++) {
      const int o = (m0 + li) * EXB + 32 * kk + 8 * lg;
      acc = ex_mma3(acc, *(const ex_bf16x8*)(sbh + o), *(const ex_bf16x8*)(sbl + o), b4h[kk], b4l[kk]);
    }
    const int col = n0 + li;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int row = m0 + 4 * lg + i;
      if (row < n && col < n) put(row, col, acc[i]);
    }
  }
}

// -------------------------------------------------------------------- up

// x += bilinear prolongation of the coarse correction (PROLONG, src/gpu/multigrid.js:99-131),
// then SW red-black sweeps. Reads xin and writes xout: a tile's halo is another
// tile's interior, so the update cannot be in place across blocks.
template <int SW, class TS, typename S, typename SO = S>
__global__ void __launch_bounds__(TS::NT)
k_mgt_up(const S* __restrict__ xin, SO* __restrict__ xout, const S* __restrict__ rhs,
         const SO* __restrict__ xc, int n, float h2, int tiles_x, int nblocks, Gate gate) {
  if (gate_closed(gate)) return;
  constexpr int TX = TS::TX, TY = TS::TY, NT = TS::NT;
  constexpr int H = 2 * SW;
  constexpr int RW = TX + 2 * H, RH = TY + 2 * H;
  constexpr int CW = RW / 2 + 3, CH = RH / 2 + 3;  // coarse footprint of the region
  __shared__ float sx[RH * RW];
  __shared__ float sr[RH * RW];
  __shared__ float sc[CH * CW];
  int t = xcd_tile(nblocks);
  int tx0 = (t % tiles_x) * TX, ty0 = (t / tiles_x) * TY;
  int ox = tx0 - H, oy = ty0 - H;
  int cn = n >> 1;
  // lowest coarse node any region cell interpolates from (ox, oy are even)
  int cox = (ox >> 1) - 1, coy = (oy >> 1) - 1;
  constexpr int CNT = RW * RH, K = (CNT + NT - 1) / NT;
  // x, the coarse correction and rhs all in flight at once; FP32 x stays in registers,
  // BF16 x is staged through LDS two cells per load
  using WX = Window<RW, RH, NT, S>;
  constexpr bool PAIRS = WX::PAIRS;
  WX wx;
  Window<CW, CH, NT, SO> wc;
  Window<RW, RH, NT, S> wr;
  wx.issue(xin, n, ox, oy);
  wc.issue(xc, cn, cox, coy);
  wr.issue(rhs, n, ox, oy);
  if constexpr (PAIRS) wx.commit(sx);
  wc.commit(sc);
  wr.commit(sr);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; k++) {
    int q = threadIdx.x + k * NT;
    if (q < CNT) {
      int ly = q / RW, lx = q - ly * RW;
      int i = ox + lx, j = oy + ly;  // unwrapped fine coordinates (parity is what matters)
      int i0 = (i & 1) ? (i >> 1) : (i >> 1) - 1;
      int j0 = (j & 1) ? (j >> 1) : (j >> 1) - 1;
      float fx = (i & 1) ? 0.25f : 0.75f, fy = (j & 1) ? 0.25f : 0.75f;
      int cx = i0 - cox, cy = j0 - coy;
      float a = sc[cy * CW + cx], b = sc[cy * CW + cx + 1];
      float c = sc[(cy + 1) * CW + cx], d = sc[(cy + 1) * CW + cx + 1];
      float top = a * (1.f - fx) + b * fx, bot = c * (1.f - fx) + d * fx;
      sx[q] = (PAIRS ? sx[q] : wx.f[PAIRS ? 0 : k]) + (top * (1.f - fy) + bot * fy);
    }
  }
  __syncthreads();
  HalfSweeps<RW, RH, 0, 2 * SW, 1, NT>::run(sx, sr, (ox + oy) & 1, h2);
  for (int q = threadIdx.x; q < TX * TY; q += NT) {
    int iy = q / TX, ix = q - iy * TX;
    int gi = tx0 + ix, gj = ty0 + iy;
    if (gi < n && gj < n) xout[(size_t)gj * n + gi] = st<SO>(sx[(iy + H) * RW + ix + H]);
  }
}

// Two grid-wide levels up in one launch, level l + 1 (n0 / 2) then level l (n0), the
// mirror of k_mgt_down2. A workgroup owns a TX x TY tile of level l. It runs the
// level-(l+1) up-sweep (stored x + prolonged correction of level l + 2, SW sweeps) on
// a region whose valid interior covers the level-(l+1) nodes the tile's level-l region
// interpolates from, then the level-l up-sweep with those values read from LDS. The
// level-(l+1) correction is used as the separate launch would store it (FP32) and is
// not stored; the level-l correction equals the separate launches' bit for bit.
//
// FOLD: level l + 2 is the 64 x 64 level solved exactly, and every workgroup solves it
// itself (ex_solve on the level's rhs, rhs2) instead of reading the correction a
// k_mg_coarse_exact launch would store: products 3 and 4 only on the output blocks that
// cover the workgroup's footprint window, whose values go straight to LDS. The same
// MFMA sequence per block, so the same bits, and one launch fewer per V-cycle.
template <int SW, class TS, typename S, typename SO, bool FOLD = false>
__global__ void __launch_bounds__(TS::NT)
k_mgt_up2(const S* __restrict__ x1in, const S* __restrict__ rhs1, const SO* __restrict__ xc2,
          const S* __restrict__ x0in, SO* __restrict__ x0out, const S* __restrict__ rhs0, int n0, float h20,
          float h21, int tiles_x, int nblocks, Gate gate, const S* __restrict__ rhs2, const uint4* __restrict__ tab) {
  if (gate_closed(gate)) return;
  constexpr int TX = TS::TX, TY = TS::TY, NT = TS::NT;
  constexpr int H = 2 * SW;
  constexpr int RW0 = TX + 2 * H, RH0 = TY + 2 * H;         // level-l region
  constexpr int CW0 = RW0 / 2 + 3, CH0 = RH0 / 2 + 3;       // its level-(l+1) footprint
  constexpr int D = H / 2 + 1;                              // footprint origin = tile origin / 2 - D
  constexpr int E = (D + H + 1) & ~1;                       // level-(l+1) region origin = tile origin / 2 - E (even)
  constexpr int RW1 = (E - D + CW0 + H + 1) & ~1, RH1 = (E - D + CH0 + H + 1) & ~1;
  constexpr int CW1 = RW1 / 2 + 3, CH1 = RH1 / 2 + 3;       // level-(l+2) footprint of the level-(l+1) region
  static_assert(TX % 8 == 0 && TY % 8 == 0, "level-(l+1) region origin must be even");
  __shared__ float sx1[RH1 * RW1];
  __shared__ float sr1[RH1 * RW1];
  __shared__ float sc1[CH1 * CW1];
  __shared__ float sx0[RH0 * RW0];
  __shared__ float sr0[RH0 * RW0];
  constexpr int EXM = FOLD ? EXN * EXB : 1;
  __shared__ __align__(16) __bf16 sah[EXM], sal[EXM], sbh[EXM], sbl[EXM];
  const int n1 = n0 >> 1, n2 = n1 >> 1;
  const int t = xcd_tile(nblocks);
  const int tx0 = (t % tiles_x) * TX, ty0 = (t / tiles_x) * TY;
  const int ox0 = tx0 - H, oy0 = ty0 - H;
  const int ox1 = (tx0 >> 1) - E, oy1 = (ty0 >> 1) - E;
  const int cox1 = (ox1 >> 1) - 1, coy1 = (oy1 >> 1) - 1;
  // every window of both levels in flight at once; FP32 x stays in registers, BF16 x
  // is staged through LDS two cells per load
  using WX1 = Window<RW1, RH1, NT, S>;
  using WX0 = Window<RW0, RH0, NT, S>;
  WX1 wx1;
  Window<CW1, CH1, NT, SO> wc1;
  Window<RW1, RH1, NT, S> wr1;
  WX0 wx0;
  Window<RW0, RH0, NT, S> wr0;
  wx1.issue(x1in, n1, ox1, oy1);
  if constexpr (!FOLD) wc1.issue(xc2, n2, cox1, coy1);
  wr1.issue(rhs1, n1, ox1, oy1);
  wx0.issue(x0in, n0, ox0, oy0);
  wr0.issue(rhs0, n0, ox0, oy0);
  if constexpr (FOLD) {
    // the footprint window of level l + 2 (n2 = 64): its row and column blocks, then
    // the solve with the window's values put in sc1
    const int wr0_ = coy1 & (EXN - 1), wc0_ = cox1 & (EXN - 1);
    unsigned rbm = 0, cbm = 0;
    for (int r = 0; r < CH1; r++) rbm |= 1u << (((wr0_ + r) & (EXN - 1)) >> 4);
    for (int c = 0; c < CW1; c++) cbm |= 1u << (((wc0_ + c) & (EXN - 1)) >> 4);
    // a window of CH1 rows touches at most CH1 / 16 + 2 of the four row blocks
    constexpr int NS = NT == 1024 ? 1 : ((CH1 + 15) / 16 + 1 < 4 ? (CH1 + 15) / 16 + 1 : 4);
    ex_solve<NT, NS, S>(rhs2, EXN, tab, sah, sal, sbh, sbl, rbm, cbm, [&](int row, int col, float x) {
      const int lr = (row - wr0_) & (EXN - 1), lc = (col - wc0_) & (EXN - 1);
      if (lr < CH1 && lc < CW1) sc1[lr * CW1 + lc] = x;
    });
  }
  // ---- level l + 1: x + prolong(correction of level l + 2), then the post-sweeps
  {
    constexpr int CNT = RW1 * RH1, K = (CNT + NT - 1) / NT;
    constexpr bool PAIRS = WX1::PAIRS;
    if constexpr (PAIRS) wx1.commit(sx1);
    if constexpr (!FOLD) wc1.commit(sc1);
    wr1.commit(sr1);
    if constexpr (WX0::PAIRS) wx0.commit(sx0);
    wr0.commit(sr0);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int q = threadIdx.x + k * NT;
      if (q < CNT) {
        const int ly = q / RW1, lx = q - ly * RW1;
        const int i = ox1 + lx, j = oy1 + ly;
        const int i0 = (i & 1) ? (i >> 1) : (i >> 1) - 1;
        const int j0 = (j & 1) ? (j >> 1) : (j >> 1) - 1;
        const float fx = (i & 1) ? 0.25f : 0.75f, fy = (j & 1) ? 0.25f : 0.75f;
        const int cx = i0 - cox1, cy = j0 - coy1;
        const float a = sc1[cy * CW1 + cx], b = sc1[cy * CW1 + cx + 1];
        const float c = sc1[(cy + 1) * CW1 + cx], d = sc1[(cy + 1) * CW1 + cx + 1];
        const float top = a * (1.f - fx) + b * fx, bot = c * (1.f - fx) + d * fx;
        sx1[q] = (PAIRS ? sx1[q] : wx1.f[PAIRS ? 0 : k]) + (top * (1.f - fy) + bot * fy);
      }
    }
    __syncthreads();
    HalfSweeps<RW1, RH1, 0, 2 * SW, 1, NT>::run(sx1, sr1, (ox1 + oy1) & 1, h21);
  }
  // ---- level l: x + prolong(the level-(l+1) correction, from LDS), then the post-sweeps
  {
    constexpr int CNT = RW0 * RH0, K = (CNT + NT - 1) / NT;
    constexpr bool PAIRS = WX0::PAIRS;
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int q = threadIdx.x + k * NT;
      if (q < CNT) {
        const int ly = q / RW0, lx = q - ly * RW0;
        const int i = ox0 + lx, j = oy0 + ly;
        const int i0 = (i & 1) ? (i >> 1) : (i >> 1) - 1;
        const int j0 = (j & 1) ? (j >> 1) : (j >> 1) - 1;
        const float fx = (i & 1) ? 0.25f : 0.75f, fy = (j & 1) ? 0.25f : 0.75f;
        // the level-(l+1) node i0 sits at sx1 column i0 - ox1, inside its valid interior
        const int cx = i0 - ox1, cy = j0 - oy1;
        const float a = rnd<SO>(sx1[cy * RW1 + cx]), b = rnd<SO>(sx1[cy * RW1 + cx + 1]);
        const float c = rnd<SO>(sx1[(cy + 1) * RW1 + cx]), d = rnd<SO>(sx1[(cy + 1) * RW1 + cx + 1]);
        const float top = a * (1.f - fx) + b * fx, bot = c * (1.f - fx) + d * fx;
        sx0[q] = (PAIRS ? sx0[q] : wx0.f[PAIRS ? 0 : k]) + (top * (1.f - fy) + bot * fy);
      }
    }
    __syncthreads();
    HalfSweeps<RW0, RH0, 0, 2 * SW, 1, NT>::run(sx0, sr0, (ox0 + oy0) & 1, h20);
    for (int q = threadIdx.x; q < TX * TY; q += NT) {
      const int iy = q / TX, ix = q - iy * TX;
      x0out[(size_t)(ty0 + iy) * n0 + tx0 + ix] = st<SO>(sx0[(iy + H) * RW0 + ix + H]);
    }
  }
}

// A window of whole quads, RW / 4 per row (RW and the origin ox multiples of 4 and n a
// multiple of 4, so a quad never straddles the periodic seam): 16-byte (FP32) or 8-byte
// (BF16) loads issued into registers, committed to LDS as FP32, the values Window holds.
template <int RW, int RH, int NT, typename T>
struct QWindow {
  static constexpr int RQ = RW / 4, CNT = RQ * RH, K = (CNT + NT - 1) / NT;
  float f[K][4];
  __device__ __forceinline__ void issue(const T* __restrict__ src, int n, int ox, int oy) {
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int q = threadIdx.x + k * NT;
      if (q < CNT) {
        const int ly = q / RQ, qx = q - ly * RQ;
        const T* a = src + wrapn(oy + ly, n) * n + wrapn(ox + 4 * qx, n);
        if constexpr (sizeof(T) == 4) {
          const float4 w = *(const float4*)a;
          f[k][0] = w.x, f[k][1] = w.y, f[k][2] = w.z, f[k][3] = w.w;
        } else {
          const uint2 w = *(const uint2*)a;
          f[k][0] = __uint_as_float(w.x << 16), f[k][1] = __uint_as_float(w.x & 0xffff0000u);
          f[k][2] = __uint_as_float(w.y << 16), f[k][3] = __uint_as_float(w.y & 0xffff0000u);
        }
      }
    }
  }
  __device__ __forceinline__ void commit(float* __restrict__ dst) const {
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int q = threadIdx.x + k * NT;
      if (q < CNT) *(float4*)(dst + 4 * q) = make_float4(f[k][0], f[k][1], f[k][2], f[k][3]);
    }
  }
};

// x + the bilinear prolongation (PROLONG) of a coarse LDS array cw (row stride CWS,
// origin (cox, coy) in coarse cells) for the four cells (i .. i + 3, j), i even: they
// interpolate from the coarse columns i/2 - 1 .. i/2 + 2 with fx = 3/4, 1/4, 3/4, 1/4.
// R rounds a coarse value as stored (the level above keeps its correction in SO). Each
// cell's arithmetic is that of the one-cell loop, in the same order.
template <int CWS, bool R, typename SO>
__device__ __forceinline__ float4 prolong_quad(const float* __restrict__ cw, int cox, int coy, int i, int j,
                                               float4 x) {
  const int j0 = (j & 1) ? (j >> 1) : (j >> 1) - 1;
  const float fy = (j & 1) ? 0.25f : 0.75f;
  const float* r0 = cw + (j0 - coy) * CWS + (i >> 1) - 1 - cox;
  const float* r1 = r0 + CWS;
  auto g = [&](float v) { return R ? rnd<SO>(v) : v; };
  const float a0 = g(r0[0]), a1 = g(r0[1]), a2 = g(r0[2]), a3 = g(r0[3]);
  const float b0 = g(r1[0]), b1 = g(r1[1]), b2 = g(r1[2]), b3 = g(r1[3]);
  auto bl = [&](float a, float b, float c, float d, float fx) {
    const float top = a * (1.f - fx) + b * fx, bot = c * (1.f - fx) + d * fx;
    return top * (1.f - fy) + bot * fy;
  };
  return make_float4(x.x + bl(a0, a1, b0, b1, 0.75f), x.y + bl(a1, a2, b1, b2, 0.25f),
                     x.z + bl(a1, a2, b1, b2, 0.75f), x.w + bl(a2, a3, b2, b3, 0.25f));
}

// Post-smoothing: 2 SW half-sweeps (red, black, ...) on the whole region with PairMap,
// each followed by a barrier.
template <int SW, int RW, int RH, int NT>
__device__ __forceinline__ void fm_post_sweeps(float* __restrict__ sx, const float* __restrict__ sr, int par0,
                                               float h2) {
  const PairMap<RW, RH, NT> m;
#pragma unroll
  for (int s = 0; s < 2 * SW; s++) {
    fm_half_sweep<RW, RH, NT>(sx, sr, m, s & 1, par0, h2);
    __syncthreads();
  }
}

// k_mgt_up2 with quad accesses (H = 2 SW a multiple of 4; the host launches k_mgt_up2
// only on grids of whole tiles whose sizes are multiples of 16): x and rhs of both levels
// staged through QWindow, the prolongation four cells at a time from the eight coarse
// nodes they share, PairMap passes, x leaving four cells per store. Same bits.
template <int SW, class TS, typename S, typename SO, bool FOLD = false>
__global__ void __launch_bounds__(TS::NT)
k_mgt_up2q(const S* __restrict__ x1in, const S* __restrict__ rhs1, const SO* __restrict__ xc2,
           const S* __restrict__ x0in, SO* __restrict__ x0out, const S* __restrict__ rhs0, int n0, float h20,
           float h21, int tiles_x, int nblocks, Gate gate, const S* __restrict__ rhs2, const uint4* __restrict__ tab) {
  if (gate_closed(gate)) return;
  constexpr int TX = TS::TX, TY = TS::TY, NT = TS::NT;
  constexpr int H = 2 * SW;
  constexpr int RW0 = TX + 2 * H, RH0 = TY + 2 * H;         // level-l region
  constexpr int CW0 = RW0 / 2 + 3, CH0 = RH0 / 2 + 3;       // its level-(l+1) footprint
  constexpr int D = H / 2 + 1;                              // footprint origin = tile origin / 2 - D
  constexpr int E = (D + H + 1) & ~1;                       // level-(l+1) region origin = tile origin / 2 - E (even)
  constexpr int RW1 = (E - D + CW0 + H + 1) & ~1, RH1 = (E - D + CH0 + H + 1) & ~1;
  constexpr int CW1 = RW1 / 2 + 3, CH1 = RH1 / 2 + 3;       // level-(l+2) footprint of the level-(l+1) region
  static_assert(TX % 8 == 0 && TY % 8 == 0 && H % 4 == 0 && E % 4 == 0 && RW1 % 4 == 0 && RW0 % 4 == 0,
                "quad windows");
  __shared__ __align__(16) float sx1[RH1 * RW1];
  __shared__ __align__(16) float sr1[RH1 * RW1];
  __shared__ float sc1[CH1 * CW1];
  __shared__ __align__(16) float sx0[RH0 * RW0];
  __shared__ __align__(16) float sr0[RH0 * RW0];
  constexpr int EXM = FOLD ? EXN * EXB : 1;
  __shared__ __align__(16) __bf16 sah[EXM], sal[EXM], sbh[EXM], sbl[EXM];
  const int n1 = n0 >> 1, n2 = n1 >> 1;
  const int t = xcd_tile(nblocks);
  const int tx0 = (t % tiles_x) * TX, ty0 = (t / tiles_x) * TY;
  const int ox0 = tx0 - H, oy0 = ty0 - H;
  const int ox1 = (tx0 >> 1) - E, oy1 = (ty0 >> 1) - E;
  const int cox1 = (ox1 >> 1) - 1, coy1 = (oy1 >> 1) - 1;
  // every window of both levels in flight at once
  QWindow<RW1, RH1, NT, S> wx1, wr1;
  QWindow<RW0, RH0, NT, S> wx0, wr0;
  Window<CW1, CH1, NT, SO> wc1;
  wx1.issue(x1in, n1, ox1, oy1);
  if constexpr (!FOLD) wc1.issue(xc2, n2, cox1, coy1);
  wr1.issue(rhs1, n1, ox1, oy1);
  wx0.issue(x0in, n0, ox0, oy0);
  wr0.issue(rhs0, n0, ox0, oy0);
  if constexpr (FOLD) {
    // the footprint window of level l + 2 (n2 = 64): its row and column blocks, then
    // the solve with the window's values put in sc1
    const int wr0_ = coy1 & (EXN - 1), wc0_ = cox1 & (EXN - 1);
    unsigned rbm = 0, cbm = 0;
    for (int r = 0; r < CH1; r++) rbm |= 1u << (((wr0_ + r) & (EXN - 1)) >> 4);
    for (int c = 0; c < CW1; c++) cbm |= 1u << (((wc0_ + c) & (EXN - 1)) >> 4);
    // a window of CH1 rows touches at most CH1 / 16 + 2 of the four row blocks
    constexpr int NS = NT == 1024 ? 1 : ((CH1 + 15) / 16 + 1 < 4 ? (CH1 + 15) / 16 + 1 : 4);
    ex_solve<NT, NS, S>(rhs2, EXN, tab, sah, sal, sbh, sbl, rbm, cbm, [&](int row, int col, float x) {
      const int lr = (row - wr0_) & (EXN - 1), lc = (col - wc0_) & (EXN - 1);
      if (lr < CH1 && lc < CW1) sc1[lr * CW1 + lc] = x;
    });
  }
  if constexpr (!FOLD) wc1.commit(sc1);
  wr1.commit(sr1);
  wr0.commit(sr0);
  __syncthreads();
  // ---- level l + 1: x + prolong(correction of level l + 2), then the post-sweeps
  {
    using W = QWindow<RW1, RH1, NT, S>;
#pragma unroll
    for (int k = 0; k < W::K; k++) {
      const int q = threadIdx.x + k * NT;
      if (q < W::CNT) {
        const int ly = q / W::RQ, lx0 = 4 * (q - ly * W::RQ);
        const float4 xv = make_float4(wx1.f[k][0], wx1.f[k][1], wx1.f[k][2], wx1.f[k][3]);
        *(float4*)(sx1 + ly * RW1 + lx0) = prolong_quad<CW1, false, SO>(sc1, cox1, coy1, ox1 + lx0, oy1 + ly, xv);
      }
    }
    __syncthreads();
    fm_post_sweeps<SW, RW1, RH1, NT>(sx1, sr1, (ox1 + oy1) & 1, h21);
  }
  // ---- level l: x + prolong(the level-(l+1) correction, from LDS), then the post-sweeps
  {
    using W = QWindow<RW0, RH0, NT, S>;
#pragma unroll
    for (int k = 0; k < W::K; k++) {
      const int q = threadIdx.x + k * NT;
      if (q < W::CNT) {
        const int ly = q / W::RQ, lx0 = 4 * (q - ly * W::RQ);
        const float4 xv = make_float4(wx0.f[k][0], wx0.f[k][1], wx0.f[k][2], wx0.f[k][3]);
        // the level-(l+1) nodes sit at sx1 column i0 - ox1, inside its valid interior
        *(float4*)(sx0 + ly * RW0 + lx0) = prolong_quad<RW1, true, SO>(sx1, ox1, oy1, ox0 + lx0, oy0 + ly, xv);
      }
    }
    __syncthreads();
    fm_post_sweeps<SW, RW0, RH0, NT>(sx0, sr0, (ox0 + oy0) & 1, h20);
    constexpr int XQ = TX / 4;
    for (int q = threadIdx.x; q < XQ * TY; q += NT) {
      const int iy = q / XQ, m4 = 4 * (q - iy * XQ);
      const float4 xv = *(const float4*)(sx0 + (iy + H) * RW0 + H + m4);
      store4<SO>(x0out + (ty0 + iy) * n0 + tx0 + m4, xv.x, xv.y, xv.z, xv.w);
    }
  }
}

// Quad-access, fixed-map forms of the single-level down and up and of the two-level
// down (H = 2 SW a multiple of 4; the host uses them on whole tiles with n a multiple of
// 4): the same operands in the same order as k_mgt_down, k_mgt_up and k_mgt_down2.

template <int SW, class TS, typename S>
__global__ void __launch_bounds__(TS::NT)
k_mgt_downq(S* __restrict__ x, const S* __restrict__ rhs, int n, float h2, float ih2, S* __restrict__ crhs,
            int tiles_x, int nblocks, Gate gate) {
  if (gate_closed(gate)) return;
  constexpr int TX = TS::TX, TY = TS::TY, NT = TS::NT;
  constexpr int H = 2 * SW;
  constexpr int RW = TX + 2 * H, RH = TY + 2 * H;
  static_assert(H % 4 == 0 && TX % 4 == 0, "quad windows");
  __shared__ __align__(16) float sx[RH * RW];
  __shared__ __align__(16) float sr[RH * RW];
  const int t = xcd_tile(nblocks);
  const int tx0 = (t % tiles_x) * TX, ty0 = (t / tiles_x) * TY;
  const int ox = tx0 - H, oy = ty0 - H;
  {
    QWindow<RW, RH, NT, S> wr;
    wr.issue(rhs, n, ox, oy);
    wr.commit(sr);
  }
  __syncthreads();
  fm_smooth_from_zero<SW, RW, RH, NT>(sx, sr, (ox + oy) & 1, h2);
  constexpr int XQ = TX / 4;
  for (int q = threadIdx.x; q < XQ * TY; q += NT) {
    const int iy = q / XQ, m4 = 4 * (q - iy * XQ);
    const float4 xv = *(const float4*)(sx + (iy + H) * RW + H + m4);
    store4<S>(x + (ty0 + iy) * n + tx0 + m4, xv.x, xv.y, xv.z, xv.w);
  }
  constexpr int CX = TX / 2, CY = TY / 2;
  const int cn = n >> 1;
  for (int q = threadIdx.x; q < CX * CY; q += NT) {
    const int cy = q / CX, cx = q - cy * CX;
    float s = 0.f;
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int a = 0; a < 2; a++) {
        const int k = (2 * cy + b + H) * RW + (2 * cx + a + H);
        const float lap = (sx[k - 1] + sx[k + 1] + sx[k - RW] + sx[k + RW] - 4.f * sx[k]) * ih2;
        s += sr[k] - lap;
      }
    crhs[((ty0 >> 1) + cy) * cn + (tx0 >> 1) + cx] = st<S>(s * 0.25f);
  }
}

template <int SW, class TS, typename S, typename SO = S>
__global__ void __launch_bounds__(TS::NT)
k_mgt_upq(const S* __restrict__ xin, SO* __restrict__ xout, const S* __restrict__ rhs,
          const SO* __restrict__ xc, int n, float h2, int tiles_x, int nblocks, Gate gate) {
  if (gate_closed(gate)) return;
  constexpr int TX = TS::TX, TY = TS::TY, NT = TS::NT;
  constexpr int H = 2 * SW;
  constexpr int RW = TX + 2 * H, RH = TY + 2 * H;
  constexpr int CW = RW / 2 + 3, CH = RH / 2 + 3;  // coarse footprint of the region
  static_assert(H % 4 == 0 && TX % 4 == 0, "quad windows");
  __shared__ __align__(16) float sx[RH * RW];
  __shared__ __align__(16) float sr[RH * RW];
  __shared__ float sc[CH * CW];
  const int t = xcd_tile(nblocks);
  const int tx0 = (t % tiles_x) * TX, ty0 = (t / tiles_x) * TY;
  const int ox = tx0 - H, oy = ty0 - H;
  const int cn = n >> 1;
  const int cox = (ox >> 1) - 1, coy = (oy >> 1) - 1;
  QWindow<RW, RH, NT, S> wx, wr;
  Window<CW, CH, NT, SO> wc;
  wx.issue(xin, n, ox, oy);
  wc.issue(xc, cn, cox, coy);
  wr.issue(rhs, n, ox, oy);
  wc.commit(sc);
  wr.commit(sr);
  __syncthreads();
  {
    using W = QWindow<RW, RH, NT, S>;
#pragma unroll
    for (int k = 0; k < W::K; k++) {
      const int q = threadIdx.x + k * NT;
      if (q < W::CNT) {
        const int ly = q / W::RQ, lx0 = 4 * (q - ly * W::RQ);
        const float4 xv = make_float4(wx.f[k][0], wx.f[k][1], wx.f[k][2], wx.f[k][3]);
        *(float4*)(sx + ly * RW + lx0) = prolong_quad<CW, false, SO>(sc, cox, coy, ox + lx0, oy + ly, xv);
      }
    }
  }
  __syncthreads();
  fm_post_sweeps<SW, RW, RH, NT>(sx, sr, (ox + oy) & 1, h2);
  constexpr int XQ = TX / 4;
  for (int q = threadIdx.x; q < XQ * TY; q += NT) {
    const int iy = q / XQ, m4 = 4 * (q - iy * XQ);
    const float4 xv = *(const float4*)(sx + (iy + H) * RW + H + m4);
    store4<SO>(xout + (ty0 + iy) * n + tx0 + m4, xv.x, xv.y, xv.z, xv.w);
  }
}

template <int SW, class TS, typename S>
__global__ void __launch_bounds__(TS::NT)
k_mgt_down2q(S* __restrict__ x0, const S* __restrict__ rhs0, int n0, float h20, float ih20, S* __restrict__ x1,
             S* __restrict__ rhs1, float h21, float ih21, S* __restrict__ crhs, int tiles_x, int nblocks,
             Gate gate) {
  if (gate_closed(gate)) return;
  constexpr int TX = TS::TX, TY = TS::TY, NT = TS::NT;
  constexpr int H = 2 * SW;
  constexpr int RW1 = TX + 2 * H, RH1 = TY + 2 * H;            // level l+1 region
  constexpr int RW0 = 2 * RW1 + 2 * H, RH0 = 2 * RH1 + 2 * H;  // level l region: residual on 2 RW1 x 2 RH1
  static_assert(H % 4 == 0 && TX % 4 == 0, "quad windows");
  __shared__ __align__(16) float sx0[RH0 * RW0];
  __shared__ __align__(16) float sr0[RH0 * RW0];
  __shared__ __align__(16) float sx1[RH1 * RW1];
  __shared__ __align__(16) float sr1[RH1 * RW1];
  const int n1 = n0 >> 1;
  const int t = xcd_tile(nblocks);
  const int tx1 = (t % tiles_x) * TX, ty1 = (t / tiles_x) * TY;
  const int ox1 = tx1 - H, oy1 = ty1 - H;          // level-(l+1) coordinate of sx1 (0, 0)
  const int ox0 = 2 * ox1 - H, oy0 = 2 * oy1 - H;  // level-l coordinate of sx0 (0, 0)
  {
    QWindow<RW0, RH0, NT, S> wr;
    wr.issue(rhs0, n0, ox0, oy0);
    wr.commit(sr0);
  }
  __syncthreads();
  // after 2 SW half-sweeps from zero, x is exact on [H - 1, RW0 - H + 1) and the
  // residual on [H, RW0 - H), the 2 RW1 x 2 RH1 block under the level-(l+1) region
  fm_smooth_from_zero<SW, RW0, RH0, NT>(sx0, sr0, (ox0 + oy0) & 1, h20);
  constexpr int XQ0 = 2 * TX / 4;
  for (int q = threadIdx.x; q < XQ0 * 2 * TY; q += NT) {
    const int iy = q / XQ0, m4 = 4 * (q - iy * XQ0);
    const float4 xv = *(const float4*)(sx0 + (3 * H + iy) * RW0 + 3 * H + m4);
    store4<S>(x0 + (2 * ty1 + iy) * n0 + 2 * tx1 + m4, xv.x, xv.y, xv.z, xv.w);
  }
  for (int q = threadIdx.x; q < RW1 * RH1; q += NT) {
    const int cy = q / RW1, cx = q - cy * RW1;
    float s = 0.f;
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int a = 0; a < 2; a++) {
        const int k = (2 * cy + b + H) * RW0 + (2 * cx + a + H);
        const float lap = (sx0[k - 1] + sx0[k + 1] + sx0[k - RW0] + sx0[k + RW0] - 4.f * sx0[k]) * ih20;
        s += sr0[k] - lap;
      }
    sr1[q] = rnd<S>(s * 0.25f);  // as the level-(l+1) sweep would load it
    if (cx >= H && cx < H + TX && cy >= H && cy < H + TY)
      rhs1[(size_t)(oy1 + cy) * n1 + ox1 + cx] = st<S>(s * 0.25f);
  }
  __syncthreads();
  fm_smooth_from_zero<SW, RW1, RH1, NT>(sx1, sr1, (ox1 + oy1) & 1, h21);
  constexpr int XQ1 = TX / 4;
  for (int q = threadIdx.x; q < XQ1 * TY; q += NT) {
    const int iy = q / XQ1, m4 = 4 * (q - iy * XQ1);
    const float4 xv = *(const float4*)(sx1 + (iy + H) * RW1 + H + m4);
    store4<S>(x1 + (ty1 + iy) * n1 + tx1 + m4, xv.x, xv.y, xv.z, xv.w);
  }
  constexpr int CX = TX / 2, CY = TY / 2;
  const int cn = n1 >> 1;
  for (int q = threadIdx.x; q < CX * CY; q += NT) {
    const int cy = q / CX, cx = q - cy * CX;
    float s = 0.f;
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int a = 0; a < 2; a++) {
        const int k = (2 * cy + b + H) * RW1 + (2 * cx + a + H);
        const float lap = (sx1[k - 1] + sx1[k + 1] + sx1[k - RW1] + sx1[k + RW1] - 4.f * sx1[k]) * ih21;
        s += sr1[k] - lap;
      }
    crhs[((ty1 >> 1) + cy) * cn + (tx1 >> 1) + cx] = st<S>(s * 0.25f);
  }
}

// ----------------------------------------------------------- coarse levels

// All levels from n0 (<= COARSE_N) down to the coarsest, in one workgroup:
// pre-smoothing from zero, residual + restriction, `coarse` sweeps on the coarsest
// level after removing its rhs mean (the periodic compatibility condition,
// src/cpu/projection.js:91-102), then prolongation + post-smoothing back up.
// Levels larger than 8x8 live in LDS and are swept by the whole workgroup; the
// tail of levels with at most 64 cells runs in wavefront 0, one cell per lane,
// neighbours exchanged with cross-lane shuffles and no barriers.
// Reads the level-n0 rhs and writes the level-n0 correction.
constexpr int CML_CELLS = 4096 + 1024 + 256 + 64 + 16 + 4 + 1 + 1;
constexpr int CTH = 1024;
constexpr int CK = (COARSE_N * COARSE_N / 4 + CTH - 1) / CTH;  // coarse cells per thread (restriction)
constexpr int CKF = (COARSE_N * COARSE_N + CTH - 1) / CTH;     // fine cells per thread (prolongation)
constexpr int CML_MAXL = 12;
constexpr int REG_N = 8;       // levels with n <= REG_N run in registers
constexpr int REG_MAXL = 4;    // 8, 4, 2, 1
// Per-level table passed by value (kernel arguments are uniform and indexable).
struct LevelTable {
  float h2[CML_MAXL];   // (float)(dx_l * dx_l)
  float ih2[CML_MAXL];  // (float)(1 / (dx_l * dx_l))
  int n[CML_MAXL];      // level size
  int off[CML_MAXL];    // offset of the level in the LDS arrays
  int mn[CML_MAXL];     // qdiv magic for division by n
  int mh[CML_MAXL];     // qdiv magic for division by n / 2
  int levels;           // levels in this kernel
  int tail;             // first level run in registers (== levels: none)
  int wave;             // first level run as register blocks in wavefront 0 (== tail: none); the
                        // levels wave .. tail-1 are 32 x 32 then 16 x 16, or 16 x 16 alone
};

// q / d for 0 <= q < 4096 and 1 <= d <= 64 as one multiply and shift, with
// m = ceil(2^18 / d): q (m d - 2^18) < 2^18 over that range makes the floor exact.
// Runtime integer division costs tens of instructions and the coarse levels run
// on one CU, where that index arithmetic is on the critical path.
__host__ __device__ inline int qdiv_magic(int d) { return (int)(((1u << 18) + (unsigned)d - 1) / (unsigned)d); }
__device__ __forceinline__ int qdiv(int q, int m) { return (int)(((unsigned)q * (unsigned)m) >> 18); }

// One colour pass on an LDS level (n even, n <= 64). Threads map to the cells of
// the colour only (row j, pair m, i = 2m + ((j + colour) & 1)), so every active
// lane works and small levels occupy few wavefronts; the indices come from one
// multiply-shift per cell. A pass reads only the other colour, so all reads precede
// the writes. Levels handed to this kernel are even (the hierarchy stops before an
// odd level).
constexpr int CKP = (COARSE_N * COARSE_N / 2 + CTH - 1) / CTH;  // cells of one colour per thread
__device__ __forceinline__ void cml_pass(float* __restrict__ x, const float* __restrict__ r, int n, int mh,
                                         int colour, float h2, bool from_zero) {
  const int hn = n >> 1, cells = hn * n;
  float acc[CKP];
  int cc[CKP];
#pragma unroll
  for (int k = 0; k < CKP; k++) {
    int q = threadIdx.x + k * CTH;
    if (q < cells) {
      int j = qdiv(q, mh), m = q - j * hn;
      int i = 2 * m + ((j + colour) & 1);
      int row = j * n, c = row + i;
      cc[k] = c;
      float sum = 0.f;
      if (!from_zero) {
        int il = i == 0 ? n - 1 : i - 1, ir = i + 1 == n ? 0 : i + 1;
        int jb = j == 0 ? n - 1 : j - 1, jt = j + 1 == n ? 0 : j + 1;
        sum = x[row + il] + x[row + ir] + x[jb * n + i] + x[jt * n + i];
      }
      acc[k] = (sum - h2 * r[c]) * 0.25f;
    }
  }
#pragma unroll
  for (int k = 0; k < CKP; k++) {
    int q = threadIdx.x + k * CTH;
    if (q < cells) x[cc[k]] = acc[k];
  }
}

// register-level red-black pass: lane = j*n + i holds x of its cell
__device__ __forceinline__ float reg_colour(float x, float r, int n, int i, int j, bool active, float h2, int colour,
                                            bool from_zero) {
  int im = i == 0 ? n - 1 : i - 1, ip = i + 1 == n ? 0 : i + 1;
  int jm = j == 0 ? n - 1 : j - 1, jp = j + 1 == n ? 0 : j + 1;
  float xl = __shfl(x, j * n + im, 64), xr = __shfl(x, j * n + ip, 64);
  float xb = __shfl(x, jm * n + i, 64), xt = __shfl(x, jp * n + i, 64);
  float sum = from_zero ? 0.f : xl + xr + xb + xt;
  float nv = (sum - h2 * r) * 0.25f;
  return (active && ((i + j) & 1) == colour) ? nv : x;
}

__device__ __forceinline__ float reg_residual(float x, float r, int n, int i, int j, float ih2) {
  int im = i == 0 ? n - 1 : i - 1, ip = i + 1 == n ? 0 : i + 1;
  int jm = j == 0 ? n - 1 : j - 1, jp = j + 1 == n ? 0 : j + 1;
  float xl = __shfl(x, j * n + im, 64), xr = __shfl(x, j * n + ip, 64);
  float xb = __shfl(x, jm * n + i, 64), xt = __shfl(x, jp * n + i, 64);
  float lap = (xl + xr + xb + xt - 4.f * x) * ih2;
  return r - lap;
}

// ---- register blocks: a level of n = 8 B cells (B = 4 or 2) in wavefront 0, lane
// (bx, by) = (lane & 7, lane >> 3) holding the B x B cells (B bx + a, B by + b) at
// x[b B + a]. Block origins are even, so the red-black colour of a slot is the same in
// every lane and each neighbour a block edge needs is one uniform cross-lane shuffle.
// The arithmetic and its order are those of the LDS passes (same bits).
__device__ __forceinline__ int wb_lane(int bx, int by) { return ((by & 7) << 3) | (bx & 7); }

template <int B>
struct WbNbr {  // the lanes of the four neighbouring blocks (periodic over the 8 x 8 blocks)
  int l, r, d, u;
  __device__ __forceinline__ WbNbr(int bx, int by)
      : l(wb_lane(bx + 7, by)), r(wb_lane(bx + 1, by)), d(wb_lane(bx, by + 7)), u(wb_lane(bx, by + 1)) {}
};

// the four neighbours of slot (a, b): left, right, below (j - 1), above (j + 1)
template <int B>
__device__ __forceinline__ void wb_nbrs(const float (&x)[B * B], const WbNbr<B>& nb, int a, int b, float& xl,
                                        float& xr, float& xb, float& xt) {
  xl = a > 0 ? x[b * B + a - 1] : __shfl(x[b * B + B - 1], nb.l, 64);
  xr = a < B - 1 ? x[b * B + a + 1] : __shfl(x[b * B], nb.r, 64);
  xb = b > 0 ? x[(b - 1) * B + a] : __shfl(x[(B - 1) * B + a], nb.d, 64);
  xt = b < B - 1 ? x[(b + 1) * B + a] : __shfl(x[a], nb.u, 64);
}

// one colour pass (cml_pass / reg_colour): cells of colour C read only the other colour
template <int B, int C, bool FZ>
__device__ __forceinline__ void wb_pass(float (&x)[B * B], const float (&r)[B * B], const WbNbr<B>& nb, float h2) {
#pragma unroll
  for (int b = 0; b < B; b++)
#pragma unroll
    for (int a = 0; a < B; a++) {
      if (((a + b) & 1) != C) continue;
      float sum = 0.f;
      if (!FZ) {
        float xl, xr, xb, xt;
        wb_nbrs<B>(x, nb, a, b, xl, xr, xb, xt);
        sum = xl + xr + xb + xt;
      }
      x[b * B + a] = (sum - h2 * r[b * B + a]) * 0.25f;
    }
}

// pre (or post) red-black sweeps; from_zero: x is a zero start (its first red pass is
// neighbour-free)
template <int B>
__device__ __forceinline__ void wb_sweeps(float (&x)[B * B], const float (&r)[B * B], const WbNbr<B>& nb, float h2,
                                          int sweeps, bool from_zero) {
  for (int s = 0; s < sweeps; s++) {
    if (s == 0 && from_zero) wb_pass<B, 0, true>(x, r, nb, h2);
    else wb_pass<B, 0, false>(x, r, nb, h2);
    wb_pass<B, 1, false>(x, r, nb, h2);
  }
}

// residual of the four children of each coarse cell, summed in order and quartered
// (the LDS restriction); the coarse block of B/2 x B/2 cells is this lane's own
template <int B>
__device__ __forceinline__ void wb_restrict(const float (&x)[B * B], const float (&r)[B * B], const WbNbr<B>& nb,
                                            float ih2, float (&rc)[(B / 2) * (B / 2)]) {
  constexpr int BC = B / 2;
  float res[B * B];
#pragma unroll
  for (int b = 0; b < B; b++)
#pragma unroll
    for (int a = 0; a < B; a++) {
      float xl, xr, xb, xt;
      wb_nbrs<B>(x, nb, a, b, xl, xr, xb, xt);
      float lap = (xl + xr + xb + xt - 4.f * x[b * B + a]) * ih2;
      res[b * B + a] = r[b * B + a] - lap;
    }
#pragma unroll
  for (int cb = 0; cb < BC; cb++)
#pragma unroll
    for (int ca = 0; ca < BC; ca++) {
      float sum = 0.f;
      sum += res[(2 * cb) * B + 2 * ca];
      sum += res[(2 * cb) * B + 2 * ca + 1];
      sum += res[(2 * cb + 1) * B + 2 * ca];
      sum += res[(2 * cb + 1) * B + 2 * ca + 1];
      rc[cb * BC + ca] = sum * 0.25f;
    }
}

// x += bilinear cell-centred prolongation of the coarse blocks xc (B/2 x B/2 per lane;
// PROLONG, src/gpu/multigrid.js:99-131): fine slot a reads coarse offsets a/2 - 1 + (a & 1)
// and one more, relative to this lane's coarse origin, so from the lane and its eight
// neighbours
template <int B>
__device__ __forceinline__ void wb_prolong(float (&x)[B * B], const float (&xc)[(B / 2) * (B / 2)], int bx, int by) {
  constexpr int BC = B / 2, W = BC + 2;
  float ring[W * W];  // coarse values at offsets -1 .. BC, row-major
#pragma unroll
  for (int oy = -1; oy <= BC; oy++)
#pragma unroll
    for (int ox = -1; ox <= BC; ox++) {
      const int dx = ox < 0 ? -1 : (ox >= BC ? 1 : 0), dy = oy < 0 ? -1 : (oy >= BC ? 1 : 0);
      const int sx = ox - dx * BC, sy = oy - dy * BC;
      const float v = xc[sy * BC + sx];
      ring[(oy + 1) * W + ox + 1] = (dx == 0 && dy == 0) ? v : __shfl(v, wb_lane(bx + dx + 8, by + dy + 8), 64);
    }
#pragma unroll
  for (int b = 0; b < B; b++)
#pragma unroll
    for (int a = 0; a < B; a++) {
      const int ox0 = (a & 1) ? (a >> 1) : (a >> 1) - 1, oy0 = (b & 1) ? (b >> 1) : (b >> 1) - 1;
      const float fx = (a & 1) ? 0.25f : 0.75f, fy = (b & 1) ? 0.25f : 0.75f;
      const float va = ring[(oy0 + 1) * W + ox0 + 1], vb = ring[(oy0 + 1) * W + ox0 + 2];
      const float vc = ring[(oy0 + 2) * W + ox0 + 1], vd = ring[(oy0 + 2) * W + ox0 + 2];
      const float top = va * (1.f - fx) + vb * fx, bot = vc * (1.f - fx) + vd * fx;
      x[b * B + a] = x[b * B + a] + (top * (1.f - fy) + bot * fy);
    }
}

template <int B>
__device__ __forceinline__ void wb_load(float (&v)[B * B], const float* __restrict__ lds, int bx, int by) {
  constexpr int n = 8 * B;
#pragma unroll
  for (int b = 0; b < B; b++)
#pragma unroll
    for (int a = 0; a < B; a++) v[b * B + a] = lds[(B * by + b) * n + B * bx + a];
}
template <int B>
__device__ __forceinline__ void wb_store(const float (&v)[B * B], float* __restrict__ lds, int bx, int by) {
  constexpr int n = 8 * B;
#pragma unroll
  for (int b = 0; b < B; b++)
#pragma unroll
    for (int a = 0; a < B; a++) lds[(B * by + b) * n + B * bx + a] = v[b * B + a];
}

template <typename S, typename SO = S>
__global__ void __launch_bounds__(CTH)
k_mg_coarse_ml(SO* __restrict__ xg, const S* __restrict__ rg, const LevelTable tb, int pre, int post,
               int coarse, const float* __restrict__ pinv, Gate gate) {
  if (gate_closed(gate)) return;
  __shared__ float bx[CML_CELLS];
  __shared__ float br[CML_CELLS];
  __shared__ float spv[REG_N * REG_N * REG_N * REG_N];  // pseudo-inverse of the coarsest register level
  __shared__ float srr[REG_N * REG_N];                  // its right-hand side, read back as broadcasts
  __shared__ double red[CTH / 64];
  const int levels = tb.levels, tail = tb.tail, last = levels - 1;
  const int n0 = tb.n[0];
  const int ncells = tb.n[last] * tb.n[last];
  const bool use_pinv = pinv != nullptr && tail < levels;
  {
    constexpr int K = CKF, KP = (REG_N * REG_N * REG_N * REG_N + CTH - 1) / CTH;
    float tmp[K], tp[KP];
#pragma unroll
    for (int k = 0; k < K; k++) {
      int q = threadIdx.x + k * CTH;
      if (q < n0 * n0) tmp[k] = ld(rg[q]);
    }
#pragma unroll
    for (int k = 0; k < KP; k++) {
      int q = threadIdx.x + k * CTH;
      if (use_pinv && q < ncells * ncells) tp[k] = pinv[q];
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
      int q = threadIdx.x + k * CTH;
      if (q < n0 * n0) br[q] = tmp[k];
    }
#pragma unroll
    for (int k = 0; k < KP; k++) {
      int q = threadIdx.x + k * CTH;
      if (use_pinv && q < ncells * ncells) spv[q] = tp[k];
    }
  }
  __syncthreads();
  // ---- LDS levels above the register levels (or above the LDS coarsest): down
  const int wave = tb.wave;
  const int ldsDown = tail < levels ? wave : last;
  for (int l = 0; l < ldsDown; l++) {
    int n = tb.n[l];
    float h2 = tb.h2[l], ih2 = tb.ih2[l];
    float* x = bx + tb.off[l];
    float* r = br + tb.off[l];
    if (pre == 0) {
      for (int q = threadIdx.x; q < n * n; q += CTH) x[q] = 0.f;
      __syncthreads();
    }
    if (pre > 0) {
      for (int s = 0; s < pre; s++) {
        cml_pass(x, r, n, tb.mh[l], 0, h2, s == 0);  // from zero: black cells are stale but never read
        __syncthreads();
        cml_pass(x, r, n, tb.mh[l], 1, h2, false);
        __syncthreads();
      }
    }
    int cn = n >> 1;
    float* rc = br + tb.off[l + 1];
    {
      float acc[CK];
#pragma unroll
      for (int k = 0; k < CK; k++) {
        int q = threadIdx.x + k * CTH;
        if (q < cn * cn) {
          int J = qdiv(q, tb.mn[l + 1]), I = q - J * cn;
          float s = 0.f;
#pragma unroll
          for (int b = 0; b < 2; b++)
#pragma unroll
            for (int a = 0; a < 2; a++) {
              int i = 2 * I + a, j = 2 * J + b;
              int c = j * n + i;
              float lap = (x[j * n + (i == 0 ? n - 1 : i - 1)] + x[j * n + (i + 1 == n ? 0 : i + 1)] +
                           x[(j == 0 ? n - 1 : j - 1) * n + i] + x[(j + 1 == n ? 0 : j + 1) * n + i] - 4.f * x[c]) *
                          ih2;
              s += r[c] - lap;
            }
          acc[k] = s * 0.25f;
        }
      }
#pragma unroll
      for (int k = 0; k < CK; k++) {
        int q = threadIdx.x + k * CTH;
        if (q < cn * cn) rc[q] = acc[k];
      }
    }
    __syncthreads();
  }
  if (tail == levels) {
    // the coarsest level is an LDS level: remove the rhs mean, then sweep
    int n = tb.n[last];
    float h2 = tb.h2[last];
    float* x = bx + tb.off[last];
    float* r = br + tb.off[last];
    double s = 0.0;
    for (int q = threadIdx.x; q < n * n; q += CTH) s += r[q];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    double tot = 0.0;
    for (int w = 0; w < CTH / 64; w++) tot += red[w];
    float mean = (float)(tot / (n * n));
    __syncthreads();
    for (int q = threadIdx.x; q < n * n; q += CTH) {
      r[q] -= mean;
      x[q] = 0.f;
    }
    __syncthreads();
    for (int sw = 0; sw < coarse; sw++) {
      cml_pass(x, r, n, tb.mh[last], 0, h2, false);
      __syncthreads();
      cml_pass(x, r, n, tb.mh[last], 1, h2, false);
      __syncthreads();
    }
  } else if (threadIdx.x < 64) {
    // ---- register tail: levels tail..last in wavefront 0
    int lane = threadIdx.x;
    float xr[REG_MAXL], rr[REG_MAXL];
#pragma unroll
    for (int k = 0; k < REG_MAXL; k++) xr[k] = rr[k] = 0.f;
    // register blocks above the tail: 32 x 32 (4 x 4 per lane) and 16 x 16 (2 x 2 per lane)
    const int bxl = lane & 7, byl = lane >> 3;
    const int nblk = tail - wave;  // 0, 1 (16) or 2 (32, 16)
    float x32[16], r32[16], x16[4], r16[4];
    if (nblk == 2) {
      const WbNbr<4> nb4(bxl, byl);
      wb_load<4>(r32, br + tb.off[wave], bxl, byl);
#pragma unroll
      for (int q = 0; q < 16; q++) x32[q] = 0.f;
      wb_sweeps<4>(x32, r32, nb4, tb.h2[wave], pre, true);
      wb_restrict<4>(x32, r32, nb4, tb.ih2[wave], r16);
    } else if (nblk == 1) {
      wb_load<2>(r16, br + tb.off[wave], bxl, byl);
    }
    if (nblk >= 1) {
      const WbNbr<2> nb2(bxl, byl);
      const int l16 = tail - 1;
#pragma unroll
      for (int q = 0; q < 4; q++) x16[q] = 0.f;
      wb_sweeps<2>(x16, r16, nb2, tb.h2[l16], pre, true);
      float rc1[1];
      wb_restrict<2>(x16, r16, nb2, tb.ih2[l16], rc1);
      rr[0] = rc1[0];  // lane = J * 8 + I of the 8 x 8 tail level
    } else {
      int n = tb.n[tail];
      rr[0] = lane < n * n ? br[tb.off[tail] + lane] : 0.f;
    }
    const int nt = last - tail + 1;  // register levels (<= REG_MAXL)
#pragma unroll
    for (int k = 0; k < REG_MAXL; k++) {
      if (k < nt) {
        int l = tail + k, n = tb.n[l];
        bool act = lane < n * n;
        int j = act ? qdiv(lane, tb.mn[l]) : 0, i = act ? lane - j * n : 0;
        float h2 = tb.h2[l], ih2 = tb.ih2[l];
        if (k == nt - 1) {
          double s = act ? (double)rr[k] : 0.0;
          s = wave_sum(s);
          float mean = (float)(s / (n * n));
          rr[k] = rr[k] - mean;
          float x = 0.f;
          if (use_pinv) {
            // the sweeps would converge below FP32 rounding: apply the pseudo-inverse
            // of the periodic Laplacian instead (spv[c*cells + lane]); the rhs goes
            // through LDS so that every column weight is an independent broadcast read.
            // Only wavefront 0 runs here, and one wavefront's LDS accesses complete in
            // issue order, so its own writes are visible to its following reads.
            int cells = n * n;
            if (act) srr[lane] = rr[k];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll 8
            for (int c = 0; c < cells; c++) x = fmaf(act ? spv[c * cells + lane] : 0.f, srr[c], x);
          } else {
            for (int sw = 0; sw < coarse; sw++) {
              x = reg_colour(x, rr[k], n, i, j, act, h2, 0, false);
              x = reg_colour(x, rr[k], n, i, j, act, h2, 1, false);
            }
          }
          xr[k] = x;
        } else {
          float x = 0.f;
          if (pre > 0) {
            x = reg_colour(x, rr[k], n, i, j, act, h2, 0, true);
            x = reg_colour(x, rr[k], n, i, j, act, h2, 1, false);
            for (int s = 1; s < pre; s++) {
              x = reg_colour(x, rr[k], n, i, j, act, h2, 0, false);
              x = reg_colour(x, rr[k], n, i, j, act, h2, 1, false);
            }
          }
          xr[k] = x;
          // restriction: coarse lane (I,J) sums the residuals of its four children in order
          float res = reg_residual(x, rr[k], n, i, j, ih2);
          int cn = n >> 1;
          bool cact = lane < cn * cn;
          int J = cact ? qdiv(lane, tb.mn[l + 1]) : 0, I = cact ? lane - J * cn : 0;
          float s = 0.f;
          s += __shfl(res, (2 * J) * n + 2 * I, 64);
          s += __shfl(res, (2 * J) * n + 2 * I + 1, 64);
          s += __shfl(res, (2 * J + 1) * n + 2 * I, 64);
          s += __shfl(res, (2 * J + 1) * n + 2 * I + 1, 64);
          if (k + 1 < REG_MAXL) rr[k + 1] = cact ? s * 0.25f : 0.f;
        }
      }
    }
#pragma unroll
    for (int k = REG_MAXL - 2; k >= 0; k--) {
      if (k < nt - 1) {
        int l = tail + k, n = tb.n[l], cn = n >> 1;
        bool act = lane < n * n;
        int j = act ? qdiv(lane, tb.mn[l]) : 0, i = act ? lane - j * n : 0;
        float h2 = tb.h2[l];
        int i0 = (i & 1) ? (i >> 1) : (i >> 1) - 1;
        int j0 = (j & 1) ? (j >> 1) : (j >> 1) - 1;
        float fx = (i & 1) ? 0.25f : 0.75f, fy = (j & 1) ? 0.25f : 0.75f;
        int x0 = i0 < 0 ? i0 + cn : i0, y0 = j0 < 0 ? j0 + cn : j0;
        int x1 = x0 + 1 == cn ? 0 : x0 + 1, y1 = y0 + 1 == cn ? 0 : y0 + 1;
        float xc = xr[k + 1];
        float a = __shfl(xc, y0 * cn + x0, 64), b = __shfl(xc, y0 * cn + x1, 64);
        float c = __shfl(xc, y1 * cn + x0, 64), d = __shfl(xc, y1 * cn + x1, 64);
        float top = a * (1.f - fx) + b * fx, bot = c * (1.f - fx) + d * fx;
        float x = xr[k] + (top * (1.f - fy) + bot * fy);
        for (int s = 0; s < post; s++) {
          x = reg_colour(x, rr[k], n, i, j, act, h2, 0, false);
          x = reg_colour(x, rr[k], n, i, j, act, h2, 1, false);
        }
        xr[k] = x;
      }
    }
    if (nblk >= 1) {
      const WbNbr<2> nb2(bxl, byl);
      const float xc1[1] = {xr[0]};
      wb_prolong<2>(x16, xc1, bxl, byl);
      wb_sweeps<2>(x16, r16, nb2, tb.h2[tail - 1], post, false);
      if (nblk == 2) {
        const WbNbr<4> nb4(bxl, byl);
        wb_prolong<4>(x32, x16, bxl, byl);
        wb_sweeps<4>(x32, r32, nb4, tb.h2[wave], post, false);
        wb_store<4>(x32, bx + tb.off[wave], bxl, byl);
      } else {
        wb_store<2>(x16, bx + tb.off[wave], bxl, byl);
      }
    } else {
      int n = tb.n[tail];
      if (lane < n * n) bx[tb.off[tail] + lane] = xr[0];
    }
  }
  __syncthreads();
  // ---- LDS levels: up
  for (int l = ldsDown - 1; l >= 0; l--) {
    int n = tb.n[l], cn = n >> 1;
    float h2 = tb.h2[l];
    float* x = bx + tb.off[l];
    float* r = br + tb.off[l];
    const float* xc = bx + tb.off[l + 1];
    {
      float acc[CKF];
#pragma unroll
      for (int k = 0; k < CKF; k++) {
        int q = threadIdx.x + k * CTH;
        if (q < n * n) {
          int j = qdiv(q, tb.mn[l]), i = q - j * n;
          int i0 = (i & 1) ? (i >> 1) : (i >> 1) - 1;
          int j0 = (j & 1) ? (j >> 1) : (j >> 1) - 1;
          float fx = (i & 1) ? 0.25f : 0.75f, fy = (j & 1) ? 0.25f : 0.75f;
          int x0 = i0 < 0 ? i0 + cn : i0, y0 = j0 < 0 ? j0 + cn : j0;
          int x1 = x0 + 1 == cn ? 0 : x0 + 1, y1 = y0 + 1 == cn ? 0 : y0 + 1;
          float a = xc[y0 * cn + x0], b = xc[y0 * cn + x1], c = xc[y1 * cn + x0], d = xc[y1 * cn + x1];
          float top = a * (1.f - fx) + b * fx, bot = c * (1.f - fx) + d * fx;
          acc[k] = x[q] + (top * (1.f - fy) + bot * fy);
        }
      }
#pragma unroll
      for (int k = 0; k < CKF; k++) {
        int q = threadIdx.x + k * CTH;
        if (q < n * n) x[q] = acc[k];
      }
    }
    __syncthreads();
    if (post > 0) {
      for (int s = 0; s < post; s++) {
        cml_pass(x, r, n, tb.mh[l], 0, h2, false);
        __syncthreads();
        cml_pass(x, r, n, tb.mh[l], 1, h2, false);
        __syncthreads();
      }
    }
  }
  for (int q = threadIdx.x; q < n0 * n0; q += CTH) xg[q] = st<SO>(bx[q]);
}

// base: null for one grid, else one grid per workgroup at element offset base[blockIdx.x]
// of xg and rg (a batch of cavities, each with its own level of n x n cells)
template <typename S>
__global__ void __launch_bounds__(EXT)
k_mg_coarse_exact(float* __restrict__ xg, const S* __restrict__ rg, int n, const long* __restrict__ base,
                  const uint4* __restrict__ tab, Gate gate) {
  if (gate_closed(gate)) return;
  if (base) {
    xg += base[blockIdx.x];
    rg += base[blockIdx.x];
  }
  // two matrices, heads and tails; A/B alternate as input and output of the products
  __shared__ __align__(16) __bf16 sah[EXN * EXB], sal[EXN * EXB], sbh[EXN * EXB], sbl[EXN * EXB];
  ex_solve<EXT, 1, S>(rg, n, tab, sah, sal, sbh, sbl, 0xfu, 0xfu,
                   [&](int row, int col, float x) { xg[row * n + col] = x; });
}

}  // namespace fsk
