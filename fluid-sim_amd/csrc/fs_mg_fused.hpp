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

constexpr int TX = 64;
constexpr int TY = 32;
constexpr int NT = 256;
constexpr int COARSE_N = 64;   // levels at or below this size are solved by one workgroup

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

// One red-black half-sweep on the LDS region [lo, RW-lo) x [lo, RH-lo), visiting
// only the cells of `colour`. par0 is the colour parity of LDS cell (0,0).
template <int RW, int RH>
__device__ __forceinline__ void lds_half_sweep(float* __restrict__ sx, const float* __restrict__ sr, int lo,
                                               int colour, int par0, float h2) {
  const int w = RW - 2 * lo, hgt = RH - 2 * lo, hw = w >> 1;  // w is even
  for (int q = threadIdx.x; q < hw * hgt; q += NT) {
    int ry = q / hw, rx = q - ry * hw;
    int ly = lo + ry;
    int lx = lo + 2 * rx + ((colour + par0 + lo + ly) & 1);
    int k = ly * RW + lx;
    float sum = sx[k - 1] + sx[k + 1] + sx[k - RW] + sx[k + RW];
    sx[k] = (sum - h2 * sr[k]) * 0.25f;
  }
}

// Pre-smoothing from zero: red half-sweep on the whole region, then 2*SW-1
// half-sweeps on shrinking regions.
template <int SW, int RW, int RH>
__device__ __forceinline__ void tile_smooth_from_zero(float* __restrict__ sx, const float* __restrict__ sr, int par0,
                                                      float h2) {
  for (int q = threadIdx.x; q < RW * RH; q += NT) {
    int ly = q / RW, lx = q - ly * RW;
    bool red = ((lx + ly + par0) & 1) == 0;
    sx[q] = red ? (0.f - h2 * sr[q]) * 0.25f : 0.f;
  }
  __syncthreads();
#pragma unroll
  for (int s = 1; s < 2 * SW; s++) {
    lds_half_sweep<RW, RH>(sx, sr, s, s & 1, par0, h2);
    __syncthreads();
  }
}

// Store the tile interior and the restricted residual of each 2x2 block of it.
template <int H, int RW>
__device__ __forceinline__ void tile_store_and_restrict(const float* __restrict__ sx, const float* __restrict__ sr,
                                                        float* __restrict__ x, float* __restrict__ crhs, int n,
                                                        int tx0, int ty0, float ih2) {
  for (int q = threadIdx.x; q < TX * TY; q += NT) {
    int iy = q / TX, ix = q - iy * TX;
    int gi = tx0 + ix, gj = ty0 + iy;
    if (gi < n && gj < n) x[(size_t)gj * n + gi] = sx[(iy + H) * RW + ix + H];
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
    crhs[(size_t)J * cn + I] = s * 0.25f;
  }
}

// -------------------------------------------------------------- level >= 1 down

template <int SW>
__global__ void __launch_bounds__(NT)
k_mgt_down(float* __restrict__ x, const float* __restrict__ rhs, int n, float h2, float ih2,
           float* __restrict__ crhs, int tiles_x, int nblocks) {
  constexpr int H = 2 * SW;
  constexpr int RW = TX + 2 * H, RH = TY + 2 * H;
  __shared__ float sx[RH * RW];
  __shared__ float sr[RH * RW];
  int t = xcd_tile(nblocks);
  int tx0 = (t % tiles_x) * TX, ty0 = (t / tiles_x) * TY;
  int ox = tx0 - H, oy = ty0 - H;  // global coordinate of LDS (0,0)
  for (int q = threadIdx.x; q < RW * RH; q += NT) {
    int ly = q / RW, lx = q - ly * RW;
    sr[q] = rhs[(size_t)wrapn(oy + ly, n) * n + wrapn(ox + lx, n)];
  }
  __syncthreads();
  tile_smooth_from_zero<SW, RW, RH>(sx, sr, (ox + oy) & 1, h2);
  tile_store_and_restrict<H, RW>(sx, sr, x, crhs, n, tx0, ty0, ih2);
}

// -------------------------------------------------------------- level 0 down

// FP64 refinement residual fused with the level-0 pre-smoothing and restriction.
// p (FP64, may be null = zero) and e_prev (FP32, may be null = zero) give p' = p + e_prev;
// r = div u* - L p' (src/core/grid.js:191-206, src/cpu/projection.js:48-62) on the
// tile + halo 2*SW; p' and r are stored for the interior and the tile's max|r| goes
// to rpart[tile]. residual_only stops after that.
template <int SW>
__global__ void __launch_bounds__(NT)
k_mgt_down0(const float* __restrict__ u, const float* __restrict__ v, const double* __restrict__ p,
            const float* __restrict__ e_prev, double* __restrict__ pout, float* __restrict__ r32,
            float* __restrict__ x, float* __restrict__ crhs, int n, double idx, double idx2, float h2, float ih2,
            int tiles_x, int nblocks, double* __restrict__ rpart, int residual_only) {
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
  for (int q = threadIdx.x; q < PW * PH; q += NT) {
    int ly = q / PW, lx = q - ly * PW;
    size_t g = (size_t)wrapn(oy - 1 + ly, n) * n + wrapn(ox - 1 + lx, n);
    double pv = p ? p[g] : 0.0;
    if (e_prev) pv += (double)e_prev[g];
    sp[q] = pv;
  }
  __syncthreads();
  double amax = 0.0;
  for (int q = threadIdx.x; q < RW * RH; q += NT) {
    int ly = q / RW, lx = q - ly * RW;
    int gi = wrapn(ox + lx, n), gj = wrapn(oy + ly, n);
    int gip = gi + 1 == n ? 0 : gi + 1, gjp = gj + 1 == n ? 0 : gj + 1;
    int kp = (ly + 1) * PW + lx + 1;
    double pc = sp[kp];
    double lap = (sp[kp - 1] - 2 * pc + sp[kp + 1]) * idx2 + (sp[kp - PW] - 2 * pc + sp[kp + PW]) * idx2;
    double dv = ((double)u[(size_t)gj * n + gip] - (double)u[(size_t)gj * n + gi]) * idx +
                ((double)v[(size_t)gjp * n + gi] - (double)v[(size_t)gj * n + gi]) * idx;
    double r = dv - lap;
    sr[q] = (float)r;
    bool inner = lx >= H && lx < H + TX && ly >= H && ly < H + TY && ox + lx < n && oy + ly < n;
    if (inner) {
      size_t g = (size_t)gj * n + gi;
      pout[g] = pc;
      r32[g] = (float)r;
      amax = fmax(amax, nan_to_inf(fabs(r)));
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
    }
  }
  if (residual_only) return;
  tile_smooth_from_zero<SW, RW, RH>(sx, sr, (ox + oy) & 1, h2);
  tile_store_and_restrict<H, RW>(sx, sr, x, crhs, n, tx0, ty0, ih2);
}

// -------------------------------------------------------------------- up

// x += bilinear prolongation of the coarse correction (PROLONG, src/gpu/multigrid.js:99-131),
// then SW red-black sweeps. Reads xin and writes xout: a tile's halo is another
// tile's interior, so the update cannot be in place across blocks.
template <int SW>
__global__ void __launch_bounds__(NT)
k_mgt_up(const float* __restrict__ xin, float* __restrict__ xout, const float* __restrict__ rhs,
         const float* __restrict__ xc, int n, float h2, int tiles_x, int nblocks) {
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
  for (int q = threadIdx.x; q < CW * CH; q += NT) {
    int ly = q / CW, lx = q - ly * CW;
    sc[q] = xc[(size_t)wrapn(coy + ly, cn) * cn + wrapn(cox + lx, cn)];
  }
  for (int q = threadIdx.x; q < RW * RH; q += NT) {
    int ly = q / RW, lx = q - ly * RW;
    size_t g = (size_t)wrapn(oy + ly, n) * n + wrapn(ox + lx, n);
    sr[q] = rhs[g];
    sx[q] = xin[g];
  }
  __syncthreads();
  for (int q = threadIdx.x; q < RW * RH; q += NT) {
    int ly = q / RW, lx = q - ly * RW;
    int i = ox + lx, j = oy + ly;  // unwrapped fine coordinates (parity is what matters)
    int i0 = (i & 1) ? (i >> 1) : (i >> 1) - 1;
    int j0 = (j & 1) ? (j >> 1) : (j >> 1) - 1;
    float fx = (i & 1) ? 0.25f : 0.75f, fy = (j & 1) ? 0.25f : 0.75f;
    int cx = i0 - cox, cy = j0 - coy;
    float a = sc[cy * CW + cx], b = sc[cy * CW + cx + 1];
    float c = sc[(cy + 1) * CW + cx], d = sc[(cy + 1) * CW + cx + 1];
    float top = a * (1.f - fx) + b * fx, bot = c * (1.f - fx) + d * fx;
    sx[q] = sx[q] + (top * (1.f - fy) + bot * fy);
  }
  __syncthreads();
  int par0 = (ox + oy) & 1;
#pragma unroll
  for (int s = 0; s < 2 * SW; s++) {
    lds_half_sweep<RW, RH>(sx, sr, s + 1, s & 1, par0, h2);
    __syncthreads();
  }
  for (int q = threadIdx.x; q < TX * TY; q += NT) {
    int iy = q / TX, ix = q - iy * TX;
    int gi = tx0 + ix, gj = ty0 + iy;
    if (gi < n && gj < n) xout[(size_t)gj * n + gi] = sx[(iy + H) * RW + ix + H];
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
constexpr int CML_MAXL = 12;
constexpr int REG_N = 8;       // levels with n <= REG_N run in registers
constexpr int REG_MAXL = 4;    // 8, 4, 2, 1
struct LevelSpacing {
  float h2[CML_MAXL];   // (float)(dx_l * dx_l)
  float ih2[CML_MAXL];  // (float)(1 / (dx_l * dx_l))
};

// one colour pass on an LDS level (n <= 64): lane -> column, 16 rows per pass
__device__ __forceinline__ void cml_colour(float* __restrict__ x, const float* __restrict__ r, int n, float h2,
                                           int colour, bool from_zero) {
  int i = threadIdx.x & 63;
  if (i >= n) return;
  int im = i == 0 ? n - 1 : i - 1, ip = i + 1 == n ? 0 : i + 1;
  for (int j = threadIdx.x >> 6; j < n; j += CTH / 64) {
    if (((i + j) & 1) != colour) continue;
    float sum = 0.f;
    if (!from_zero)
      sum = x[j * n + im] + x[j * n + ip] + x[(j == 0 ? n - 1 : j - 1) * n + i] + x[(j + 1 == n ? 0 : j + 1) * n + i];
    x[j * n + i] = (sum - h2 * r[j * n + i]) * 0.25f;
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

__global__ void __launch_bounds__(CTH)
k_mg_coarse_ml(float* __restrict__ xg, const float* __restrict__ rg, int n0, int levels, const LevelSpacing sp,
               int pre, int post, int coarse) {
  __shared__ float bx[CML_CELLS];
  __shared__ float br[CML_CELLS];
  __shared__ double red[CTH / 64];
  int off[CML_MAXL];
  int nl[CML_MAXL];
  int tail = levels;  // first level handled in registers
  {
    int o = 0, m = n0;
    for (int l = 0; l < levels; l++) {
      off[l] = o;
      nl[l] = m;
      if (tail == levels && m <= REG_N) tail = l;
      o += m * m;
      m >>= 1;
    }
  }
  for (int q = threadIdx.x; q < n0 * n0; q += CTH) br[q] = rg[q];
  __syncthreads();
  const int last = levels - 1;
  // ---- LDS levels above the register tail (or above the LDS coarsest): down
  const int ldsDown = tail < levels ? tail : last;
  for (int l = 0; l < ldsDown; l++) {
    int n = nl[l];
    float h2 = sp.h2[l], ih2 = sp.ih2[l];
    float* x = bx + off[l];
    float* r = br + off[l];
    if (pre == 0) {
      for (int q = threadIdx.x; q < n * n; q += CTH) x[q] = 0.f;
      __syncthreads();
    }
    for (int s = 0; s < pre; s++) {
      cml_colour(x, r, n, h2, 0, s == 0);  // from zero: black cells are stale but never read
      __syncthreads();
      cml_colour(x, r, n, h2, 1, false);
      __syncthreads();
    }
    int cn = n >> 1;
    float* rc = br + off[l + 1];
    int I = threadIdx.x & 63;
    if (I < cn) {
      for (int J = threadIdx.x >> 6; J < cn; J += CTH / 64) {
        float s = 0.f;
        for (int b = 0; b < 2; b++)
          for (int a = 0; a < 2; a++) {
            int i = 2 * I + a, j = 2 * J + b;
            int k = j * n + i;
            float lap = (x[j * n + (i == 0 ? n - 1 : i - 1)] + x[j * n + (i + 1 == n ? 0 : i + 1)] +
                         x[(j == 0 ? n - 1 : j - 1) * n + i] + x[(j + 1 == n ? 0 : j + 1) * n + i] - 4.f * x[k]) *
                        ih2;
            s += r[k] - lap;
          }
        rc[J * cn + I] = s * 0.25f;
      }
    }
    __syncthreads();
  }
  if (tail == levels) {
    // the coarsest level is an LDS level: remove the rhs mean, then sweep
    int n = nl[last];
    float h2 = sp.h2[last];
    float* x = bx + off[last];
    float* r = br + off[last];
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
      cml_colour(x, r, n, h2, 0, false);
      __syncthreads();
      cml_colour(x, r, n, h2, 1, false);
      __syncthreads();
    }
  } else if (threadIdx.x < 64) {
    // ---- register tail: levels tail..last in wavefront 0
    int lane = threadIdx.x;
    float xr[REG_MAXL], rr[REG_MAXL];
#pragma unroll
    for (int k = 0; k < REG_MAXL; k++) xr[k] = rr[k] = 0.f;
    {
      int n = nl[tail];
      rr[0] = lane < n * n ? br[off[tail] + lane] : 0.f;
    }
    const int nt = last - tail + 1;  // register levels (<= REG_MAXL)
#pragma unroll
    for (int k = 0; k < REG_MAXL; k++) {
      if (k < nt) {
        int l = tail + k, n = nl[l];
        bool act = lane < n * n;
        int j = act ? lane / n : 0, i = act ? lane - j * n : 0;
        float h2 = sp.h2[l], ih2 = sp.ih2[l];
        if (k == nt - 1) {
          double s = act ? (double)rr[k] : 0.0;
          s = wave_sum(s);
          float mean = (float)(s / (n * n));
          rr[k] = rr[k] - mean;
          float x = 0.f;
          for (int sw = 0; sw < coarse; sw++) {
            x = reg_colour(x, rr[k], n, i, j, act, h2, 0, false);
            x = reg_colour(x, rr[k], n, i, j, act, h2, 1, false);
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
          int J = cact ? lane / cn : 0, I = cact ? lane - J * cn : 0;
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
        int l = tail + k, n = nl[l], cn = n >> 1;
        bool act = lane < n * n;
        int j = act ? lane / n : 0, i = act ? lane - j * n : 0;
        float h2 = sp.h2[l];
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
    int n = nl[tail];
    if (lane < n * n) bx[off[tail] + lane] = xr[0];
  }
  __syncthreads();
  // ---- LDS levels: up
  for (int l = ldsDown - 1; l >= 0; l--) {
    int n = nl[l], cn = n >> 1;
    float h2 = sp.h2[l];
    float* x = bx + off[l];
    float* r = br + off[l];
    const float* xc = bx + off[l + 1];
    int i = threadIdx.x & 63;
    if (i < n) {
      int i0 = (i & 1) ? (i >> 1) : (i >> 1) - 1;
      float fx = (i & 1) ? 0.25f : 0.75f;
      int x0 = i0 < 0 ? i0 + cn : i0;
      int x1 = x0 + 1 == cn ? 0 : x0 + 1;
      for (int j = threadIdx.x >> 6; j < n; j += CTH / 64) {
        int j0 = (j & 1) ? (j >> 1) : (j >> 1) - 1;
        float fy = (j & 1) ? 0.25f : 0.75f;
        int y0 = j0 < 0 ? j0 + cn : j0;
        int y1 = y0 + 1 == cn ? 0 : y0 + 1;
        float a = xc[y0 * cn + x0], b = xc[y0 * cn + x1], c = xc[y1 * cn + x0], d = xc[y1 * cn + x1];
        float top = a * (1.f - fx) + b * fx, bot = c * (1.f - fx) + d * fx;
        x[j * n + i] = x[j * n + i] + (top * (1.f - fy) + bot * fy);
      }
    }
    __syncthreads();
    for (int s = 0; s < post; s++) {
      cml_colour(x, r, n, h2, 0, false);
      __syncthreads();
      cml_colour(x, r, n, h2, 1, false);
      __syncthreads();
    }
  }
  for (int q = threadIdx.x; q < n0 * n0; q += CTH) xg[q] = bx[q];
}

}  // namespace fsk
