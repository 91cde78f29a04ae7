// fs_batch.hpp -- batched lid-driven cavities: many independent unit-box
// simulations of different power-of-two sizes stepped together, one launch per
// stage over all of them.
//
// Walls (no counterpart in the reference, whose domain is periodic): the MAC layout
// of the periodic solver with u[j n + 0] and v[0 n + i] as wall faces that stay 0
// and the right/top wall faces implied 0; no slip on every wall, the top wall (lid)
// moving with velocity U. Beyond a wall, ghosts put the wall value on the wall:
// u below the bottom -u, above the lid 2U - u; v beyond a side wall -v. Departure
// points are clamped into the box. Pressure: zero normal derivative at the walls.
//
// Layout: the simulations are ordered by size, largest first. The velocity and
// pressure of all of them live in one arena (simulation k at cell offset off_k);
// since every size is a multiple of 16 x 16, a 256-thread block never straddles two
// simulations and finds its own by a binary search over block offsets. The
// multigrid levels of one size s (all unit boxes, so the same spacing 1/s) form one
// contiguous class holding the simulations 0..count_s-1 that are at least that
// large, so a level operator is one flat launch over a class.
#pragma once
#include "fs_kernels.hpp"
#include "fs_mg_fused.hpp"

namespace fsk {

struct CavSim {
  int n, lg;            // cells per side, log2(n)
  long off;             // first cell in the field arena
  long lvl;             // first cell of its finest multigrid level in the level buffers
  float lid, nu;        // lid velocity, viscosity
  int disk;             // 1: a pinned disk
  float dx_, dy_, dr;   // disk centre and radius
  double dt;            // this step's dt (written on the device)
};

struct CavCell {
  int k, n, i, j;
  long c;   // cell index in the field arena
  int blk;  // the 256-cell block this workgroup covers
};

// A workgroup covers the 256-cell block xcd_tile(gridDim.x): consecutive blocks (and
// so the rows above and below a block of a large cavity) go to the same XCD's L2.
// blk0 holds each cavity's first block, then (from index nsim) each block's cavity.
__device__ __forceinline__ CavCell cav_cell(const CavSim* __restrict__ sims, const int* __restrict__ blk0, int nsim) {
  CavCell r;
  r.blk = xcd_tile((int)gridDim.x);
  r.k = blk0[nsim + r.blk];
  const CavSim& S = sims[r.k];
  int local = (r.blk - blk0[r.k]) * 256 + threadIdx.x;
  r.n = S.n;
  r.j = local >> S.lg;
  r.i = local & (S.n - 1);
  r.c = S.off + local;
  return r;
}

// wall-aware face values (ghosts beyond the walls)
__device__ __forceinline__ float cu_at(const float* __restrict__ u, int n, float U, int i, int j) {
  if (i <= 0 || i >= n) return 0.f;
  if (j < 0) return -u[i];
  if (j >= n) return 2.f * U - u[(size_t)(n - 1) * n + i];
  return u[(size_t)j * n + i];
}
__device__ __forceinline__ float cv_at(const float* __restrict__ v, int n, int i, int j) {
  if (j <= 0 || j >= n) return 0.f;
  if (i < 0) return -v[(size_t)j * n];
  if (i >= n) return -v[(size_t)j * n + n - 1];
  return v[(size_t)j * n + i];
}

// a[idx] with a 32-bit byte offset (idx < 2^30), so a block-uniform base stays in
// scalar registers and the load takes a scalar base plus a 32-bit lane offset
// x / k for the neighbour count k of a Neumann cell (2, 3 or 4): an interior cell's
// x / 4 is x * 0.25 exactly (a power-of-two scaling rounds once, either way), which
// saves the division sequence on all but the boundary cells; same bits.
__device__ __forceinline__ float nm_div(float x, int k) { return k == 4 ? x * 0.25f : x / (float)k; }

__device__ __forceinline__ float ld_u32(const float* __restrict__ a, int idx) {
  return *(const float*)((const char*)a + (unsigned)((unsigned)idx * 4u));
}

// bilinear sample at (x, y) in the unit box of component comp (0: u, 1: v); lo/hi get
// the range of the four source values. Branch-free: the four loads use indices
// clamped into the simulation, then selects apply what cu_at / cv_at return there
// (the ghost beyond a wall from the clamped row or column, 0 on a wall face).
template <int COMP>
__device__ __forceinline__ float cav_sample(const float* __restrict__ a, int n, float U, float x, float y, float* lo,
                                            float* hi) {
  float fn = (float)n;
  x = fminf(fmaxf(x, 0.f), 1.f);
  y = fminf(fmaxf(y, 0.f), 1.f);
  float gx = COMP == 0 ? x * fn : x * fn - 0.5f;
  float gy = COMP == 0 ? y * fn - 0.5f : y * fn;
  int i0 = (int)floorf(gx), j0 = (int)floorf(gy);
  i0 = min(i0, n - 1);
  j0 = min(j0, n - 1);
  float fx = gx - (float)i0, fy = gy - (float)j0;
  // u: i0 in [0, n-1], j0 in [-1, n-1]; v: i0 in [-1, n-1], j0 in [0, n-1]
  const int i1 = i0 + 1, j1 = j0 + 1;
  const int r0 = max(j0, 0) * n, r1 = min(j1, n - 1) * n;
  const int c0 = max(i0, 0), c1 = min(i1, n - 1);
  float s00 = ld_u32(a, r0 + c0), s10 = ld_u32(a, r0 + c1), s01 = ld_u32(a, r1 + c0), s11 = ld_u32(a, r1 + c1);
  if (COMP == 0) {
    // below the bottom wall -u, above the lid 2U - u; the side wall faces are 0
    if (j0 < 0) { s00 = -s00; s10 = -s10; }
    if (j1 >= n) { s01 = 2.f * U - s01; s11 = 2.f * U - s11; }
    if (i0 <= 0) { s00 = 0.f; s01 = 0.f; }
    if (i1 >= n) { s10 = 0.f; s11 = 0.f; }
  } else {
    // beyond a side wall -v; the bottom and top wall faces are 0
    if (i0 < 0) { s00 = -s00; s01 = -s01; }
    if (i1 >= n) { s10 = -s10; s11 = -s11; }
    if (j0 <= 0) { s00 = 0.f; s10 = 0.f; }
    if (j1 >= n) { s01 = 0.f; s11 = 0.f; }
  }
  if (lo) {
    *lo = fminf(fminf(s00, s10), fminf(s01, s11));
    *hi = fmaxf(fmaxf(s00, s10), fmaxf(s01, s11));
  }
  // fused multiply-adds: the walled batch has no bit-for-bit counterpart to match, only
  // the f64 oracle within its tolerance
  float bot = fmaf(s10 - s00, fx, s00), top = fmaf(s11 - s01, fx, s01);
  return fmaf(top - bot, fy, bot);
}

// RK2 departure point from (x, y) over dt, given the velocity (u1, v1) at (x, y).
__device__ __forceinline__ void cav_trace(const float* __restrict__ u, const float* __restrict__ v, int n, float U,
                                          float x, float y, float u1, float v1, float dt, float* xo, float* yo) {
  float xm = fmaf(-0.5f * dt, u1, x), ym = fmaf(-0.5f * dt, v1, y);
  float u2 = cav_sample<0>(u, n, U, xm, ym, nullptr, nullptr), v2 = cav_sample<1>(v, n, 0.f, xm, ym, nullptr, nullptr);
  *xo = fmaf(-dt, u2, x);
  *yo = fmaf(-dt, v2, y);
}

// The velocity at a face's own position: the face's value itself (the bilinear
// weights there are exactly 0, and x, y > 0 make the sign of a zero product
// irrelevant to the trace) and the other component interpolated.
__device__ __forceinline__ void cav_vel_u_face(const float* __restrict__ u, const float* __restrict__ v, int n, int i,
                                               int j, float x, float y, float* u1, float* v1) {
  *u1 = ld_u32(u, j * n + i);
  *v1 = cav_sample<1>(v, n, 0.f, x, y, nullptr, nullptr);
}
__device__ __forceinline__ void cav_vel_v_face(const float* __restrict__ u, const float* __restrict__ v, int n, float U,
                                               int i, int j, float x, float y, float* u1, float* v1) {
  *u1 = cav_sample<0>(u, n, U, x, y, nullptr, nullptr);
  *v1 = ld_u32(v, j * n + i);
}

// Semi-Lagrangian forward pass of the interior faces (the MacCormack predictor).
__global__ void __launch_bounds__(256)
k_cav_advect_sl(const float* __restrict__ u, const float* __restrict__ v, float* __restrict__ ou,
                float* __restrict__ ov, const CavSim* __restrict__ sims, const int* __restrict__ blk0, int nsim) {
  CavCell q = cav_cell(sims, blk0, nsim);
  const CavSim& S = sims[q.k];
  const float* us = u + S.off;
  const float* vs = v + S.off;
  const int n = q.n;
  float h = 1.f / (float)n, dt = (float)S.dt, U = S.lid;
  float xd, yd, u1, v1, a = 0.f, b = 0.f;
  if (q.i > 0) {
    float x = q.i * h, y = (q.j + 0.5f) * h;
    cav_vel_u_face(us, vs, n, q.i, q.j, x, y, &u1, &v1);
    cav_trace(us, vs, n, U, x, y, u1, v1, dt, &xd, &yd);
    a = cav_sample<0>(us, n, U, xd, yd, nullptr, nullptr);
  }
  if (q.j > 0) {
    float x = (q.i + 0.5f) * h, y = q.j * h;
    cav_vel_v_face(us, vs, n, U, q.i, q.j, x, y, &u1, &v1);
    cav_trace(us, vs, n, U, x, y, u1, v1, dt, &xd, &yd);
    b = cav_sample<1>(vs, n, 0.f, xd, yd, nullptr, nullptr);
  }
  ou[q.c] = a;
  ov[q.c] = b;
}

// MacCormack corrector: back-trace of the forward state through the original velocity,
// fwd + (u - back) / 2, clamped to the range of the forward trace's four sources.
__global__ void __launch_bounds__(256)
k_cav_advect_mc(const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ fu,
                const float* __restrict__ fv, float* __restrict__ ou, float* __restrict__ ov,
                const CavSim* __restrict__ sims, const int* __restrict__ blk0, int nsim) {
  CavCell q = cav_cell(sims, blk0, nsim);
  const CavSim& S = sims[q.k];
  const float* us = u + S.off;
  const float* vs = v + S.off;
  const int n = q.n, own = q.j * n + q.i;
  float h = 1.f / (float)n, dt = (float)S.dt, U = S.lid;
  float xd, yd, u1, v1, lo, hi, a = 0.f, b = 0.f;
  // the forward and the backward trace start from the same velocity at the face
  if (q.i > 0) {
    float x = q.i * h, y = (q.j + 0.5f) * h;
    cav_vel_u_face(us, vs, n, q.i, q.j, x, y, &u1, &v1);
    cav_trace(us, vs, n, U, x, y, u1, v1, dt, &xd, &yd);
    cav_sample<0>(us, n, U, xd, yd, &lo, &hi);
    cav_trace(us, vs, n, U, x, y, u1, v1, -dt, &xd, &yd);
    float back = cav_sample<0>(fu + S.off, n, U, xd, yd, nullptr, nullptr);
    float c = fu[q.c] + 0.5f * (ld_u32(us, own) - back);
    a = fminf(fmaxf(c, lo), hi);
  }
  if (q.j > 0) {
    float x = (q.i + 0.5f) * h, y = q.j * h;
    cav_vel_v_face(us, vs, n, U, q.i, q.j, x, y, &u1, &v1);
    cav_trace(us, vs, n, U, x, y, u1, v1, dt, &xd, &yd);
    cav_sample<1>(vs, n, 0.f, xd, yd, &lo, &hi);
    cav_trace(us, vs, n, U, x, y, u1, v1, -dt, &xd, &yd);
    float back = cav_sample<1>(fv + S.off, n, 0.f, xd, yd, nullptr, nullptr);
    float c = fv[q.c] + 0.5f * (ld_u32(vs, own) - back);
    b = fminf(fmaxf(c, lo), hi);
  }
  ou[q.c] = a;
  ov[q.c] = b;
}

// Explicit diffusion of the interior faces with the wall ghosts, then the pinned
// disk's direct forcing u += chi (0 - u), `passes` times (src/cpu/solid.js:165-233
// with plain separations; the wall faces are never forced).
__global__ void __launch_bounds__(256)
k_cav_diffuse_solid(const float* __restrict__ u, const float* __restrict__ v, float* __restrict__ ou,
                    float* __restrict__ ov, const CavSim* __restrict__ sims, const int* __restrict__ blk0, int nsim,
                    float mask_width, int passes) {
  CavCell q = cav_cell(sims, blk0, nsim);
  const CavSim& S = sims[q.k];
  const float* us = u + S.off;
  const float* vs = v + S.off;
  int n = q.n, i = q.i, j = q.j;
  float h = 1.f / (float)n;
  float a = S.nu > 0.f ? (float)((double)S.nu * S.dt * (double)n * (double)n) : 0.f;
  float U = S.lid;
  float un = 0.f, vn = 0.f;
  // the five values cu_at / cv_at return, from loads at indices clamped into the
  // simulation (32-bit offsets) and selects for the wall faces and ghosts
  const int own = j * n + i;
  const int il = max(i - 1, 0), ir = min(i + 1, n - 1), jb = max(j - 1, 0), jt = min(j + 1, n - 1);
  if (i > 0) {
    const float c = ld_u32(us, own);
    float l = ld_u32(us, j * n + il), r = ld_u32(us, j * n + ir), b = ld_u32(us, jb * n + i), t = ld_u32(us, jt * n + i);
    if (i - 1 == 0) l = 0.f;                // the left wall face
    if (i + 1 == n) r = 0.f;                // the right wall face
    if (j == 0) b = -c;                     // below the bottom wall: -u
    if (j + 1 == n) t = 2.f * U - c;        // above the lid: 2U - u
    un = c + a * (l + r + b + t - 4.f * c);
  }
  if (j > 0) {
    const float c = ld_u32(vs, own);
    float l = ld_u32(vs, j * n + il), r = ld_u32(vs, j * n + ir), b = ld_u32(vs, jb * n + i), t = ld_u32(vs, jt * n + i);
    if (i == 0) l = -c;                     // beyond the left wall: -v
    if (i + 1 == n) r = -c;                 // beyond the right wall: -v
    if (j - 1 == 0) b = 0.f;                // the bottom wall face
    if (j + 1 == n) t = 0.f;                // the top wall face
    vn = c + a * (l + r + b + t - 4.f * c);
  }
  if (S.disk) {
    float w = mask_width * h;
    if (i > 0) {
      float dx = i * h - S.dx_, dy = (j + 0.5f) * h - S.dy_;
      float t = (sqrtf(dx * dx + dy * dy) - S.dr) / w;
      float chi = t > 8.f ? 0.f : 0.5f * (1.f - tanhf(t));
      if (chi >= 1e-6f)
        for (int p = 0; p < passes; p++) un += chi * (0.f - un);
    }
    if (j > 0) {
      float dx = (i + 0.5f) * h - S.dx_, dy = j * h - S.dy_;
      float t = (sqrtf(dx * dx + dy * dy) - S.dr) / w;
      float chi = t > 8.f ? 0.f : 0.5f * (1.f - tanhf(t));
      if (chi >= 1e-6f)
        for (int p = 0; p < passes; p++) vn += chi * (0.f - vn);
    }
  }
  ou[q.c] = un;
  ov[q.c] = vn;
}

// FP64 refinement residual r = div u - L_N (p + e) of every simulation; p' = p + e is
// stored, r goes (FP32) to the simulation's finest multigrid rhs, block max |r| to part.
// One wavefront per 256-cell block: lane l takes the 4 consecutive cells 4l .. 4l + 3 of
// it (one row; sizes are multiples of 64, arena offsets multiples of 4096 cells), with
// 16-byte loads and stores; each cell's arithmetic is the one-cell formula's. A
// workgroup holds CAV_WPB consecutive blocks, one per wavefront (a launch of one
// wavefront per workgroup over 20 M cells is bound by dispatching its workgroups).
constexpr int CAV_WPB = 4;
__global__ void __launch_bounds__(64 * CAV_WPB)
k_cav_residual64(const float* __restrict__ u, const float* __restrict__ v, const double* __restrict__ p,
                 const float* __restrict__ xe, double* __restrict__ pout, float* __restrict__ rhs,
                 const CavSim* __restrict__ sims, const int* __restrict__ blk0, int nsim, double* __restrict__ part,
                 int nblocks) {
  const int blk = xcd_tile((int)gridDim.x) * CAV_WPB + (int)(threadIdx.x >> 6);
  if (blk >= nblocks) return;
  const int lane = (int)threadIdx.x & 63;
  const int k = blk0[nsim + blk];
  const CavSim& S = sims[k];
  const int n = S.n;
  const int local = (blk - blk0[k]) * 256 + 4 * lane;
  const int j = local >> S.lg, i = local & (n - 1);
  const long c = S.off + local, cl = S.lvl + local;
  // p + e at 4 cells of row jj (0.0 + e with no p, as the one-cell formula)
  auto pv4 = [&](long cc, long lv, double out[4]) {
    double2 a01{0.0, 0.0}, a23{0.0, 0.0};
    if (p) {
      a01 = *(const double2*)(p + cc);
      a23 = *(const double2*)(p + cc + 2);
    }
    out[0] = a01.x;
    out[1] = a01.y;
    out[2] = a23.x;
    out[3] = a23.y;
    if (xe) {
      const float4 e = *(const float4*)(xe + lv);
      out[0] += (double)e.x;
      out[1] += (double)e.y;
      out[2] += (double)e.z;
      out[3] += (double)e.w;
    }
  };
  auto pv1 = [&](long cc, long lv) -> double {
    double a = p ? p[cc] : 0.0;
    if (xe) a += (double)xe[lv];
    return a;
  };
  double pc[4], pb[4], pt[4];
  pv4(c, cl, pc);
  if (j > 0) pv4(c - n, cl - n, pb);
  if (j < n - 1) pv4(c + n, cl + n, pt);
  const double pl = i > 0 ? pv1(c - 1, cl - 1) : 0.0;
  const double pr = i + 4 < n ? pv1(c + 4, cl + 4) : 0.0;
  const float* us = u + S.off + (long)j * n;
  const float* vs = v + S.off;
  const float4 u4 = *(const float4*)(us + i);
  const float un = i + 4 < n ? us[i + 4] : 0.f;
  const float4 vb4 = j > 0 ? *(const float4*)(vs + (long)j * n + i) : float4{0.f, 0.f, 0.f, 0.f};
  const float4 vt4 = j + 1 < n ? *(const float4*)(vs + (long)(j + 1) * n + i) : float4{0.f, 0.f, 0.f, 0.f};
  const float uf[5] = {i > 0 ? u4.x : 0.f, u4.y, u4.z, u4.w, i + 4 < n ? un : 0.f};  // faces i .. i + 4
  const float vb[4] = {vb4.x, vb4.y, vb4.z, vb4.w}, vt[4] = {vt4.x, vt4.y, vt4.z, vt4.w};
  const double dn = (double)n;
  double out[4];
  double amax = 0.0;
  float r4[4];
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const int im = i + m;
    double s = 0.0;
    if (im > 0) s += (m > 0 ? pc[m - 1] : pl) - pc[m];
    if (im < n - 1) s += (m < 3 ? pc[m + 1] : pr) - pc[m];
    if (j > 0) s += pb[m] - pc[m];
    if (j < n - 1) s += pt[m] - pc[m];
    const double lap = s * dn * dn;
    const double div = ((double)uf[m + 1] - (double)uf[m]) * dn + ((double)vt[m] - (double)vb[m]) * dn;
    const double r = div - lap;
    out[m] = pc[m];
    r4[m] = (float)r;
    amax = fmax(amax, nan_to_inf(fabs(r)));
  }
  *(double2*)(pout + c) = double2{out[0], out[1]};
  *(double2*)(pout + c + 2) = double2{out[2], out[3]};
  *(float4*)(rhs + cl) = float4{r4[0], r4[1], r4[2], r4[3]};
  amax = wave_max(amax);
  if (lane == 0) part[blk] = amax;
}

// u -= grad p on the interior faces; block max |u|, |v| for the next step's dt.
// One wavefront per 256-cell block, 4 cells of a row per lane, CAV_WPB blocks per
// workgroup (as k_cav_residual64).
__global__ void __launch_bounds__(64 * CAV_WPB)
k_cav_gradient64(float* __restrict__ u, float* __restrict__ v, const double* __restrict__ p,
                 const CavSim* __restrict__ sims, const int* __restrict__ blk0, int nsim, double* __restrict__ part,
                 int nblocks) {
  const int blk = xcd_tile((int)gridDim.x) * CAV_WPB + (int)(threadIdx.x >> 6);
  if (blk >= nblocks) return;
  const int lane = (int)threadIdx.x & 63;
  const int k = blk0[nsim + blk];
  const CavSim& S = sims[k];
  const int n = S.n;
  const int local = (blk - blk0[k]) * 256 + 4 * lane;
  const int j = local >> S.lg, i = local & (n - 1);
  const long c = S.off + local;
  const double dn = (double)n;
  const double2 c01 = *(const double2*)(p + c), c23 = *(const double2*)(p + c + 2);
  const double pc[4] = {c01.x, c01.y, c23.x, c23.y};
  const double pl[4] = {i > 0 ? p[c - 1] : 0.0, c01.x, c01.y, c23.x};  // the cell to the left of each
  float4 u4 = *(const float4*)(u + c);
  float uu[4] = {u4.x, u4.y, u4.z, u4.w}, un[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int m = 0; m < 4; m++)
    if (i + m > 0) {
      un[m] = (float)((double)uu[m] - (pc[m] - pl[m]) * dn);
      uu[m] = un[m];
    }
  *(float4*)(u + c) = float4{uu[0], uu[1], uu[2], uu[3]};
  float vn[4] = {0.f, 0.f, 0.f, 0.f};
  if (j > 0) {
    const double2 b01 = *(const double2*)(p + c - n), b23 = *(const double2*)(p + c - n + 2);
    const double pb[4] = {b01.x, b01.y, b23.x, b23.y};
    const float4 v4 = *(const float4*)(v + c);
    const float vv[4] = {v4.x, v4.y, v4.z, v4.w};
#pragma unroll
    for (int m = 0; m < 4; m++) vn[m] = (float)((double)vv[m] - (pc[m] - pb[m]) * dn);
    *(float4*)(v + c) = float4{vn[0], vn[1], vn[2], vn[3]};
  }
  double a = 0.0, b = 0.0;
#pragma unroll
  for (int m = 0; m < 4; m++) {
    a = fmax(a, nan_to_inf(fabs((double)un[m])));
    b = fmax(b, nan_to_inf(fabs((double)vn[m])));
  }
  a = wave_max(a);
  b = wave_max(b);
  if (lane == 0) {
    part[2 * blk] = a;
    part[2 * blk + 1] = b;
  }
}

// dt of every simulation (one workgroup each): cfl h / (max(max|u|, |U|) + max|v|),
// from the block maxima of the last gradient pass; NaN if a velocity is not finite.
__global__ void __launch_bounds__(256)
k_cav_dt(const double* __restrict__ part, CavSim* __restrict__ sims, const int* __restrict__ blk0, int nsim,
         double cfl, double dt_max) {
  int k = blockIdx.x;
  int b0 = blk0[k], b1 = k + 1 < nsim ? blk0[k + 1] : b0 + ((sims[k].n * sims[k].n) >> 8);
  double mu = 0.0, mv = 0.0;
  for (int b = b0 + threadIdx.x; b < b1; b += 256) {
    mu = fmax(mu, part[2 * b]);
    mv = fmax(mv, part[2 * b + 1]);
  }
  mu = block_max(mu);
  mv = block_max(mv);
  if (threadIdx.x == 0) {
    double U = fabs((double)sims[k].lid);
    double rate = (fmax(mu, U) + mv) * sims[k].n;
    double dt = rate > 0 ? cfl / rate : dt_max;
    if (isinf(mu) || isinf(mv)) dt = __longlong_as_double(0x7ff8000000000000LL);
    sims[k].dt = fmin(dt, dt_max);
  }
}

// ------------------------------------------------ multigrid, zero-gradient walls
// One size class: `count` grids of s x s cells (spacing 1/s), one after another.

// Red-black colour pass: x = (sum of the neighbours inside the box - h^2 rhs) / their count.
__global__ void __launch_bounds__(256)
k_nm_rb(float* __restrict__ x, const float* __restrict__ rhs, int lg, long count, float h2, int colour, int x_zero) {
  long q = (long)blockIdx.x * 256 + threadIdx.x;
  const int s = 1 << lg, hs = s >> 1;
  if (q >= count * s * hs) return;
  long g = q >> (2 * lg - 1);
  int rem = (int)(q - (g << (2 * lg - 1)));
  int j = rem / hs, m = rem - j * hs;
  int i = 2 * m + ((j + colour) & 1);
  float* xg = x + (g << (2 * lg));
  long c = (long)j * s + i;
  float sum = 0.f;
  int k = 0;
  if (i > 0) { sum += x_zero ? 0.f : xg[c - 1]; k++; }
  if (i < s - 1) { sum += x_zero ? 0.f : xg[c + 1]; k++; }
  if (j > 0) { sum += x_zero ? 0.f : xg[c - s]; k++; }
  if (j < s - 1) { sum += x_zero ? 0.f : xg[c + s]; k++; }
  xg[c] = nm_div(sum - h2 * rhs[(g << (2 * lg)) + c], k);
}

// The black half of a from-zero first sweep reads red values only; when x_zero is set
// for the red pass, the black cells hold stale values that are never read.

// Residual of the four children of each coarse cell, averaged (restriction).
__global__ void __launch_bounds__(256)
k_nm_restrict(const float* __restrict__ x, const float* __restrict__ rhs, int lg, long count, float ih2,
              float* __restrict__ crhs) {
  long q = (long)blockIdx.x * 256 + threadIdx.x;
  const int s = 1 << lg, cs = s >> 1, clg = lg - 1;
  if (q >= count * cs * cs) return;
  long g = q >> (2 * clg);
  int rem = (int)(q - (g << (2 * clg)));
  int J = rem >> clg, I = rem & (cs - 1);
  const float* xg = x + (g << (2 * lg));
  const float* rg = rhs + (g << (2 * lg));
  float acc = 0.f;
#pragma unroll
  for (int b = 0; b < 2; b++)
#pragma unroll
    for (int a = 0; a < 2; a++) {
      int i = 2 * I + a, j = 2 * J + b;
      long c = (long)j * s + i;
      float xc = xg[c], sum = 0.f;
      if (i > 0) sum += xg[c - 1] - xc;
      if (i < s - 1) sum += xg[c + 1] - xc;
      if (j > 0) sum += xg[c - s] - xc;
      if (j < s - 1) sum += xg[c + s] - xc;
      acc += rg[c] - sum * ih2;
    }
  crhs[q] = acc * 0.25f;
}

// x = xin + bilinear cell-centred prolongation of the coarse correction, the coarse
// indices clamped into the box (constant extrapolation, zero gradient at the walls).
__global__ void __launch_bounds__(256)
k_nm_prolong(const float* __restrict__ xc, const float* __restrict__ xin, float* __restrict__ x, int lg, long count) {
  long q = (long)blockIdx.x * 256 + threadIdx.x;
  const int s = 1 << lg, cs = s >> 1, clg = lg - 1;
  if (q >= count * s * s) return;
  long g = q >> (2 * lg);
  int rem = (int)(q - (g << (2 * lg)));
  int j = rem >> lg, i = rem & (s - 1);
  int i0 = (i & 1) ? (i >> 1) : (i >> 1) - 1, j0 = (j & 1) ? (j >> 1) : (j >> 1) - 1;
  float fx = (i & 1) ? 0.25f : 0.75f, fy = (j & 1) ? 0.25f : 0.75f;
  int x0 = max(i0, 0), x1 = min(i0 + 1, cs - 1), y0 = max(j0, 0), y1 = min(j0 + 1, cs - 1);
  const float* cg = xc + (g << (2 * clg));
  float a = cg[y0 * cs + x0], b = cg[y0 * cs + x1], c = cg[y1 * cs + x0], d = cg[y1 * cs + x1];
  float top = a * (1.f - fx) + b * fx, bot = c * (1.f - fx) + d * fx;
  x[q] = xin[q] + (top * (1.f - fy) + bot * fy);
}

// ---- fused level kernels (LDS tiles, as on the periodic grid): a tile with a halo of
// 2 SW cells; halo cells outside the box hold 0 and are never updated, so every
// in-box cell sees exactly the operands of the one-pass kernels above in the same
// order (an absent neighbour adds +0), and the results are the same bits.

// LDS window of grid g (s x s) with origin (ox, oy); cells outside the box are 0.
template <int RW, int RH, int NT>
__device__ __forceinline__ void nm_load(float* __restrict__ dst, const float* __restrict__ src, int s, int ox, int oy) {
  constexpr int CNT = RW * RH, K = (CNT + NT - 1) / NT;
  float tmp[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    int q = threadIdx.x + k * NT;
    if (q < CNT) {
      int ly = q / RW, lx = q - ly * RW;
      int gx = ox + lx, gy = oy + ly;
      tmp[k] = (gx >= 0 && gx < s && gy >= 0 && gy < s) ? src[(size_t)gy * s + gx] : 0.f;
    }
  }
#pragma unroll
  for (int k = 0; k < K; k++) {
    int q = threadIdx.x + k * NT;
    if (q < CNT) dst[q] = tmp[k];
  }
}

// nm_load split in two: issue() starts the loads into registers, commit() stores them,
// so a kernel's windows are all in flight at once
template <int RW, int RH, int NT>
struct NmWindow {
  static constexpr int CNT = RW * RH, K = (CNT + NT - 1) / NT;
  float f[K];
  __device__ __forceinline__ void issue(const float* __restrict__ src, int s, int ox, int oy) {
#pragma unroll
    for (int k = 0; k < K; k++) {
      int q = threadIdx.x + k * NT;
      if (q < CNT) {
        int ly = q / RW, lx = q - ly * RW;
        int gx = ox + lx, gy = oy + ly;
        f[k] = (gx >= 0 && gx < s && gy >= 0 && gy < s) ? src[(size_t)gy * s + gx] : 0.f;
      }
    }
  }
  __device__ __forceinline__ void commit(float* __restrict__ dst) const {
#pragma unroll
    for (int k = 0; k < K; k++) {
      int q = threadIdx.x + k * NT;
      if (q < CNT) dst[q] = f[k];
    }
  }
};

// One colour half-sweep on the LDS region [LO, RW-LO) x [LO, RH-LO); cells outside
// the box are skipped (they stay 0).
template <int RW, int RH, int LO, int NT>
__device__ __forceinline__ void nm_half_sweep(float* __restrict__ sx, const float* __restrict__ sr, int colour,
                                              int par0, float h2, int ox, int oy, int s) {
  constexpr int w = RW - 2 * LO, hgt = RH - 2 * LO, hw = w / 2;
  constexpr int CNT = hw * hgt, K = (CNT + NT - 1) / NT;
  float acc[K];
  int kk[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    int q = threadIdx.x + k * NT;
    kk[k] = -1;
    if (q < CNT) {
      int ry = q / hw, rx = q - ry * hw;
      int ly = LO + ry;
      int lx = LO + 2 * rx + ((colour + par0 + LO + ly) & 1);
      int gi = ox + lx, gj = oy + ly;
      if (gi >= 0 && gi < s && gj >= 0 && gj < s) {
        int c = ly * RW + lx;
        kk[k] = c;
        int nk = 4 - (gi == 0) - (gi == s - 1) - (gj == 0) - (gj == s - 1);
        float sum = sx[c - 1] + sx[c + 1] + sx[c - RW] + sx[c + RW];
        acc[k] = nm_div(sum - h2 * sr[c], nk);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < K; k++)
    if (kk[k] >= 0) sx[kk[k]] = acc[k];
}

template <int RW, int RH, int S, int SEND, int LOOFF, int NT>
struct NmSweeps {
  static __device__ __forceinline__ void run(float* sx, const float* sr, int par0, float h2, int ox, int oy, int s) {
    nm_half_sweep<RW, RH, S + LOOFF, NT>(sx, sr, S & 1, par0, h2, ox, oy, s);
    __syncthreads();
    NmSweeps<RW, RH, S + 1, SEND, LOOFF, NT>::run(sx, sr, par0, h2, ox, oy, s);
  }
};
template <int RW, int RH, int SEND, int LOOFF, int NT>
struct NmSweeps<RW, RH, SEND, SEND, LOOFF, NT> {
  static __device__ __forceinline__ void run(float*, const float*, int, float, int, int, int) {}
};

// Down: SW sweeps from zero on a tile of grid g, the pre-smoothed correction stored,
// and the restricted residual of each 2x2 block of the tile to the coarse rhs.
template <int SW, class TS>
__global__ void __launch_bounds__(TS::NT)
k_nmt_down(float* __restrict__ x, const float* __restrict__ rhs, int s, int tiles_x, int tpg, float h2, float ih2,
           float* __restrict__ crhs) {
  constexpr int TX = TS::TX, TY = TS::TY, NT = TS::NT, H = 2 * SW;
  constexpr int RW = TX + 2 * H, RH = TY + 2 * H;
  __shared__ float sx[RH * RW];
  __shared__ float sr[RH * RW];
  const int bid = xcd_tile((int)gridDim.x), g = bid / tpg, t = bid - g * tpg;
  const int tx0 = (t % tiles_x) * TX, ty0 = (t / tiles_x) * TY;
  const int ox = tx0 - H, oy = ty0 - H;
  const size_t gb = (size_t)g * s * s;
  nm_load<RW, RH, NT>(sr, rhs + gb, s, ox, oy);
  __syncthreads();
  const int par0 = (ox + oy) & 1;
  // red from zero on the whole region (black and outside cells 0), then shrinking regions
  for (int q = threadIdx.x; q < RW * RH; q += NT) {
    int ly = q / RW, lx = q - ly * RW;
    int gi = ox + lx, gj = oy + ly;
    bool in = gi >= 0 && gi < s && gj >= 0 && gj < s;
    bool red = ((lx + ly + par0) & 1) == 0;
    int nk = 4 - (gi == 0) - (gi == s - 1) - (gj == 0) - (gj == s - 1);
    sx[q] = (in && red) ? nm_div(0.f - h2 * sr[q], nk) : 0.f;
  }
  __syncthreads();
  NmSweeps<RW, RH, 1, 2 * SW, 0, NT>::run(sx, sr, par0, h2, ox, oy, s);
  for (int q = threadIdx.x; q < TX * TY; q += NT) {
    int iy = q / TX, ix = q - iy * TX;
    x[gb + (size_t)(ty0 + iy) * s + tx0 + ix] = sx[(iy + H) * RW + ix + H];
  }
  constexpr int CX = TX / 2, CY = TY / 2;
  const int cs = s >> 1;
  const size_t cb = (size_t)g * cs * cs;
  for (int q = threadIdx.x; q < CX * CY; q += NT) {
    int cy = q / CX, cx = q - cy * CX;
    int I = (tx0 >> 1) + cx, J = (ty0 >> 1) + cy;
    float acc = 0.f;
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int a = 0; a < 2; a++) {
        int i = 2 * I + a, j = 2 * J + b;
        int k = (2 * cy + b + H) * RW + (2 * cx + a + H);
        float xc = sx[k], sum = 0.f;
        if (i > 0) sum += sx[k - 1] - xc;
        if (i < s - 1) sum += sx[k + 1] - xc;
        if (j > 0) sum += sx[k - RW] - xc;
        if (j < s - 1) sum += sx[k + RW] - xc;
        acc += sr[k] - sum * ih2;
      }
    crhs[cb + (size_t)J * cs + I] = acc * 0.25f;
  }
}

// Up: x = xin + bilinear prolongation of the coarse correction (coarse indices clamped
// into the box), then SW sweeps; the tile interior goes to xout.
template <int SW, class TS>
__global__ void __launch_bounds__(TS::NT)
k_nmt_up(const float* __restrict__ xin, float* __restrict__ xout, const float* __restrict__ rhs,
         const float* __restrict__ xc, int s, int tiles_x, int tpg, float h2) {
  constexpr int TX = TS::TX, TY = TS::TY, NT = TS::NT, H = 2 * SW;
  constexpr int RW = TX + 2 * H, RH = TY + 2 * H;
  constexpr int CW = RW / 2 + 3, CH = RH / 2 + 3;
  __shared__ float sx[RH * RW];
  __shared__ float sr[RH * RW];
  __shared__ float sc[CH * CW];
  const int bid = xcd_tile((int)gridDim.x), g = bid / tpg, t = bid - g * tpg;
  const int tx0 = (t % tiles_x) * TX, ty0 = (t / tiles_x) * TY;
  const int ox = tx0 - H, oy = ty0 - H;
  const int cs = s >> 1;
  const size_t gb = (size_t)g * s * s, cb = (size_t)g * cs * cs;
  const int cox = (ox >> 1) - 1, coy = (oy >> 1) - 1;
  // coarse footprint (indices clamped into the box), x and rhs: all loads in flight at once
  constexpr int KC = (CW * CH + NT - 1) / NT;
  float fc[KC];
#pragma unroll
  for (int k = 0; k < KC; k++) {
    int q = threadIdx.x + k * NT;
    if (q < CW * CH) {
      int ly = q / CW, lx = q - ly * CW;
      int cx = min(max(cox + lx, 0), cs - 1), cy = min(max(coy + ly, 0), cs - 1);
      fc[k] = xc[cb + (size_t)cy * cs + cx];
    }
  }
  NmWindow<RW, RH, NT> wx, wr;
  wx.issue(xin + gb, s, ox, oy);
  wr.issue(rhs + gb, s, ox, oy);
#pragma unroll
  for (int k = 0; k < KC; k++) {
    int q = threadIdx.x + k * NT;
    if (q < CW * CH) sc[q] = fc[k];
  }
  wx.commit(sx);
  wr.commit(sr);
  __syncthreads();
  for (int q = threadIdx.x; q < RW * RH; q += NT) {
    int ly = q / RW, lx = q - ly * RW;
    int i = ox + lx, j = oy + ly;
    if (i < 0 || i >= s || j < 0 || j >= s) continue;
    int i0 = (i & 1) ? (i >> 1) : (i >> 1) - 1, j0 = (j & 1) ? (j >> 1) : (j >> 1) - 1;
    float fx = (i & 1) ? 0.25f : 0.75f, fy = (j & 1) ? 0.25f : 0.75f;
    // clamped coarse indices, as local footprint coordinates
    int x0 = max(i0, 0) - cox, x1 = min(i0 + 1, cs - 1) - cox, y0 = max(j0, 0) - coy, y1 = min(j0 + 1, cs - 1) - coy;
    float a = sc[y0 * CW + x0], b = sc[y0 * CW + x1], c = sc[y1 * CW + x0], d = sc[y1 * CW + x1];
    float top = a * (1.f - fx) + b * fx, bot = c * (1.f - fx) + d * fx;
    sx[q] = sx[q] + (top * (1.f - fy) + bot * fy);
  }
  __syncthreads();
  NmSweeps<RW, RH, 0, 2 * SW, 1, NT>::run(sx, sr, (ox + oy) & 1, h2, ox, oy, s);
  for (int q = threadIdx.x; q < TX * TY; q += NT) {
    int iy = q / TX, ix = q - iy * TX;
    xout[gb + (size_t)(ty0 + iy) * s + tx0 + ix] = sx[(iy + H) * RW + ix + H];
  }
}

// ---- the level kernels with quad accesses and fixed thread maps (SW even). The
// region origin is a multiple of 4 and so is s, so a quad of four cells is wholly inside
// or wholly outside the box; the passes use PairMap over the whole region (cells outside
// the box are never updated and hold 0, as in the kernels above), every in-box cell
// seeing the same operands in the same order: the same bits.

// A window of whole quads of grid g (s x s), zero outside the box.
template <int RW, int RH, int NT>
struct NmQWindow {
  static constexpr int RQ = RW / 4, CNT = RQ * RH, K = (CNT + NT - 1) / NT;
  float f[K][4];
  __device__ __forceinline__ void issue(const float* __restrict__ src, int s, int ox, int oy) {
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int q = threadIdx.x + k * NT;
      f[k][0] = f[k][1] = f[k][2] = f[k][3] = 0.f;
      if (q < CNT) {
        const int ly = q / RQ, gx = ox + 4 * (q - ly * RQ), gy = oy + ly;
        if (gx >= 0 && gx < s && gy >= 0 && gy < s) {
          const float4 w = *(const float4*)(src + gy * s + gx);
          f[k][0] = w.x, f[k][1] = w.y, f[k][2] = w.z, f[k][3] = w.w;
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

__device__ __forceinline__ int nm_count(int gi, int gj, int s) {
  return 4 - (gi == 0) - (gi == s - 1) - (gj == 0) - (gj == s - 1);
}

// one colour pass of nm_half_sweep on every in-box interior cell of the region
template <int RW, int RH, int NT>
__device__ __forceinline__ void nm_fm_half_sweep(float* __restrict__ sx, const float* __restrict__ sr,
                                                 const PairMap<RW, RH, NT>& m, int colour, int par0, float h2, int ox,
                                                 int oy, int s) {
  using M = PairMap<RW, RH, NT>;
  float acc[M::ROWS];
  int cc[M::ROWS];
  bool act[M::ROWS];
#pragma unroll
  for (int k = 0; k < M::ROWS; k++) {
    const int ly = m.r + k * M::RG;
    const int off = (colour + par0 + ly) & 1;
    const int lx = 2 * m.p + off;
    const int gi = ox + lx, gj = oy + ly;
    act[k] = m.row_ok(k) && ly >= 1 && ly <= RH - 2 && lx >= 1 && lx <= RW - 2 && gi >= 0 && gi < s && gj >= 0 &&
             gj < s;
    const int c = m.base + k * M::RG * RW + off;
    cc[k] = c;
    if (act[k]) {
      const float sum = sx[c - 1] + sx[c + 1] + sx[c - RW] + sx[c + RW];
      acc[k] = nm_div(sum - h2 * sr[c], nm_count(gi, gj, s));
    }
  }
#pragma unroll
  for (int k = 0; k < M::ROWS; k++)
    if (act[k]) sx[cc[k]] = acc[k];
}

template <int SW, class TS>
__global__ void __launch_bounds__(TS::NT)
k_nmt_downq(float* __restrict__ x, const float* __restrict__ rhs, int s, int tiles_x, int tpg, float h2, float ih2,
            float* __restrict__ crhs) {
  constexpr int TX = TS::TX, TY = TS::TY, NT = TS::NT, H = 2 * SW;
  constexpr int RW = TX + 2 * H, RH = TY + 2 * H;
  static_assert(H % 4 == 0 && TX % 4 == 0, "quad windows");
  __shared__ __align__(16) float sx[RH * RW];
  __shared__ __align__(16) float sr[RH * RW];
  const int bid = xcd_tile((int)gridDim.x), g = bid / tpg, t = bid - g * tpg;
  const int tx0 = (t % tiles_x) * TX, ty0 = (t / tiles_x) * TY;
  const int ox = tx0 - H, oy = ty0 - H;
  const size_t gb = (size_t)g * s * s;
  {
    NmQWindow<RW, RH, NT> wr;
    wr.issue(rhs + gb, s, ox, oy);
    wr.commit(sr);
  }
  __syncthreads();
  const int par0 = (ox + oy) & 1;
  const PairMap<RW, RH, NT> m;
  // red from zero on the whole region (black and outside cells 0)
  {
    using M = PairMap<RW, RH, NT>;
#pragma unroll
    for (int k = 0; k < M::ROWS; k++) {
      if (!m.row_ok(k)) continue;
      const int ly = m.r + k * M::RG, c = m.base + k * M::RG * RW;
      const int gj = oy + ly, gi = ox + 2 * m.p;
      const bool rowin = gj >= 0 && gj < s;
      const bool red0 = ((ly + par0) & 1) == 0;  // is cell 2p red
      const float2 r2 = *(const float2*)(sr + c);
      float2 x2;
      x2.x = (rowin && gi >= 0 && gi < s && red0) ? nm_div(0.f - h2 * r2.x, nm_count(gi, gj, s)) : 0.f;
      x2.y = (rowin && gi + 1 >= 0 && gi + 1 < s && !red0) ? nm_div(0.f - h2 * r2.y, nm_count(gi + 1, gj, s)) : 0.f;
      *(float2*)(sx + c) = x2;
    }
  }
  __syncthreads();
#pragma unroll
  for (int hs = 1; hs < 2 * SW; hs++) {
    nm_fm_half_sweep<RW, RH, NT>(sx, sr, m, hs & 1, par0, h2, ox, oy, s);
    __syncthreads();
  }
  constexpr int XQ = TX / 4;
  for (int q = threadIdx.x; q < XQ * TY; q += NT) {
    const int iy = q / XQ, m4 = 4 * (q - iy * XQ);
    *(float4*)(x + gb + (size_t)(ty0 + iy) * s + tx0 + m4) = *(const float4*)(sx + (iy + H) * RW + H + m4);
  }
  constexpr int CX = TX / 2, CY = TY / 2;
  const int cs = s >> 1;
  const size_t cb = (size_t)g * cs * cs;
  for (int q = threadIdx.x; q < CX * CY; q += NT) {
    const int cy = q / CX, cx = q - cy * CX;
    const int I = (tx0 >> 1) + cx, J = (ty0 >> 1) + cy;
    float acc = 0.f;
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int a = 0; a < 2; a++) {
        const int i = 2 * I + a, j = 2 * J + b;
        const int k = (2 * cy + b + H) * RW + (2 * cx + a + H);
        const float xc = sx[k];
        float sum = 0.f;
        if (i > 0) sum += sx[k - 1] - xc;
        if (i < s - 1) sum += sx[k + 1] - xc;
        if (j > 0) sum += sx[k - RW] - xc;
        if (j < s - 1) sum += sx[k + RW] - xc;
        acc += sr[k] - sum * ih2;
      }
    crhs[cb + (size_t)J * cs + I] = acc * 0.25f;
  }
}

template <int SW, class TS>
__global__ void __launch_bounds__(TS::NT)
k_nmt_upq(const float* __restrict__ xin, float* __restrict__ xout, const float* __restrict__ rhs,
          const float* __restrict__ xc, int s, int tiles_x, int tpg, float h2) {
  constexpr int TX = TS::TX, TY = TS::TY, NT = TS::NT, H = 2 * SW;
  constexpr int RW = TX + 2 * H, RH = TY + 2 * H;
  constexpr int CW = RW / 2 + 3, CH = RH / 2 + 3;
  static_assert(H % 4 == 0 && TX % 4 == 0, "quad windows");
  __shared__ __align__(16) float sx[RH * RW];
  __shared__ __align__(16) float sr[RH * RW];
  __shared__ float sc[CH * CW];
  const int bid = xcd_tile((int)gridDim.x), g = bid / tpg, t = bid - g * tpg;
  const int tx0 = (t % tiles_x) * TX, ty0 = (t / tiles_x) * TY;
  const int ox = tx0 - H, oy = ty0 - H;
  const int cs = s >> 1;
  const size_t gb = (size_t)g * s * s, cb = (size_t)g * cs * cs;
  const int cox = (ox >> 1) - 1, coy = (oy >> 1) - 1;
  // coarse footprint (indices clamped into the box), x and rhs: all loads in flight at once
  constexpr int KC = (CW * CH + NT - 1) / NT;
  float fc[KC];
#pragma unroll
  for (int k = 0; k < KC; k++) {
    const int q = threadIdx.x + k * NT;
    if (q < CW * CH) {
      const int ly = q / CW, lx = q - ly * CW;
      const int cx = min(max(cox + lx, 0), cs - 1), cy = min(max(coy + ly, 0), cs - 1);
      fc[k] = xc[cb + (size_t)cy * cs + cx];
    }
  }
  NmQWindow<RW, RH, NT> wx, wr;
  wx.issue(xin + gb, s, ox, oy);
  wr.issue(rhs + gb, s, ox, oy);
#pragma unroll
  for (int k = 0; k < KC; k++) {
    const int q = threadIdx.x + k * NT;
    if (q < CW * CH) sc[q] = fc[k];
  }
  wr.commit(sr);
  __syncthreads();
  // x + the prolongation, four cells per thread (a quad is wholly in or out of the box)
  {
    using W = NmQWindow<RW, RH, NT>;
#pragma unroll
    for (int k = 0; k < W::K; k++) {
      const int q = threadIdx.x + k * NT;
      if (q < W::CNT) {
        const int ly = q / W::RQ, lx0 = 4 * (q - ly * W::RQ);
        const int j = oy + ly, i4 = ox + lx0;
        float o[4] = {wx.f[k][0], wx.f[k][1], wx.f[k][2], wx.f[k][3]};
        if (j >= 0 && j < s && i4 >= 0 && i4 < s) {
          const int j0 = (j & 1) ? (j >> 1) : (j >> 1) - 1;
          const float fy = (j & 1) ? 0.25f : 0.75f;
          const int y0 = max(j0, 0) - coy, y1 = min(j0 + 1, cs - 1) - coy;
#pragma unroll
          for (int a = 0; a < 4; a++) {
            const int i = i4 + a;
            const int i0 = (i & 1) ? (i >> 1) : (i >> 1) - 1;
            const float fx = (i & 1) ? 0.25f : 0.75f;
            const int x0 = max(i0, 0) - cox, x1 = min(i0 + 1, cs - 1) - cox;
            const float pa = sc[y0 * CW + x0], pb = sc[y0 * CW + x1], pc = sc[y1 * CW + x0], pd = sc[y1 * CW + x1];
            const float top = pa * (1.f - fx) + pb * fx, bot = pc * (1.f - fx) + pd * fx;
            o[a] = o[a] + (top * (1.f - fy) + bot * fy);
          }
        }
        *(float4*)(sx + ly * RW + lx0) = make_float4(o[0], o[1], o[2], o[3]);
      }
    }
  }
  __syncthreads();
  const PairMap<RW, RH, NT> m;
  const int par0 = (ox + oy) & 1;
#pragma unroll
  for (int hs = 0; hs < 2 * SW; hs++) {
    nm_fm_half_sweep<RW, RH, NT>(sx, sr, m, hs & 1, par0, h2, ox, oy, s);
    __syncthreads();
  }
  constexpr int XQ = TX / 4;
  for (int q = threadIdx.x; q < XQ * TY; q += NT) {
    const int iy = q / XQ, m4 = 4 * (q - iy * XQ);
    *(float4*)(xout + gb + (size_t)(ty0 + iy) * s + tx0 + m4) = *(const float4*)(sx + (iy + H) * RW + H + m4);
  }
}

// The levels from s0 <= 64 down to 4x4 of one simulation per workgroup, in LDS:
// pre-smoothing from zero, restriction, the coarsest 4x4 solved with the
// pseudo-inverse of its Neumann Laplacian (after removing the rhs mean), then
// prolongation and post-smoothing back up. rhs_base/x_base: the simulation's
// level-s0 grid in the level buffers.
constexpr int NMC_CELLS = 4096 + 1024 + 256 + 64 + 16;
__global__ void __launch_bounds__(1024)
k_nm_coarse(float* __restrict__ xall, const float* __restrict__ rall, const long* __restrict__ base, int s0,
            int pre, int post, const float* __restrict__ pinv) {
  __shared__ float bx[NMC_CELLS];
  __shared__ float br[NMC_CELLS];
  __shared__ float sp[256];
  __shared__ double red[16];
  const int k = blockIdx.x, t = threadIdx.x;
  const long b = base[k];
  int lg0 = 0;
  while ((1 << lg0) < s0) lg0++;
  const int L = lg0 - 1;  // levels s0 .. 4
  int off[8];
  {
    int o = 0;
    for (int l = 0; l < L; l++) {
      off[l] = o;
      int s = s0 >> l;
      o += s * s;
    }
  }
  for (int q = t; q < s0 * s0; q += 1024) br[q] = rall[b + q];
  for (int q = t; q < 256; q += 1024) sp[q] = pinv[q];
  __syncthreads();
  auto pass = [&](int l, int colour, bool zero) {
    int s = s0 >> l, lg = lg0 - l, hs = s >> 1;
    float h2 = 1.f / (float)(s * s);
    float* x = bx + off[l];
    const float* r = br + off[l];
    for (int q = t; q < s * hs; q += 1024) {
      int j = q / hs, m = q - j * hs;
      int i = 2 * m + ((j + colour) & 1);
      int c = (j << lg) + i;
      float sum = 0.f;
      int kk = 0;
      if (i > 0) { sum += zero ? 0.f : x[c - 1]; kk++; }
      if (i < s - 1) { sum += zero ? 0.f : x[c + 1]; kk++; }
      if (j > 0) { sum += zero ? 0.f : x[c - s]; kk++; }
      if (j < s - 1) { sum += zero ? 0.f : x[c + s]; kk++; }
      x[c] = nm_div(sum - h2 * r[c], kk);
    }
    __syncthreads();
  };
  for (int l = 0; l < L - 1; l++) {
    int s = s0 >> l, cs = s >> 1;
    for (int sw = 0; sw < pre; sw++) {
      pass(l, 0, sw == 0);
      pass(l, 1, false);
    }
    if (pre == 0) {
      for (int q = t; q < s * s; q += 1024) bx[off[l] + q] = 0.f;
      __syncthreads();
    }
    float ih2 = (float)(s * s);
    const float* x = bx + off[l];
    const float* r = br + off[l];
    for (int q = t; q < cs * cs; q += 1024) {
      int J = q / cs, I = q - J * cs;
      float acc = 0.f;
      for (int bb = 0; bb < 2; bb++)
        for (int a = 0; a < 2; a++) {
          int i = 2 * I + a, j = 2 * J + bb, c = j * s + i;
          float xc = x[c], sum = 0.f;
          if (i > 0) sum += x[c - 1] - xc;
          if (i < s - 1) sum += x[c + 1] - xc;
          if (j > 0) sum += x[c - s] - xc;
          if (j < s - 1) sum += x[c + s] - xc;
          acc += r[c] - sum * ih2;
        }
      br[off[l + 1] + q] = acc * 0.25f;
    }
    __syncthreads();
  }
  {
    // coarsest 4x4: remove the rhs mean, x = pinv r (16 x 16, columns broadcast)
    const int l = L - 1;
    float* x = bx + off[l];
    float* r = br + off[l];
    double sum = t < 16 ? (double)r[t] : 0.0;
    sum = wave_sum(sum);
    if ((t & 63) == 0) red[t >> 6] = sum;
    __syncthreads();
    double tot = 0.0;
    for (int w = 0; w < 16; w++) tot += red[w];
    float mean = (float)(tot / 16.0);
    __syncthreads();
    if (t < 16) r[t] -= mean;
    __syncthreads();
    if (t < 16) {
      float acc = 0.f;
      for (int c = 0; c < 16; c++) acc = fmaf(sp[c * 16 + t], r[c], acc);
      x[t] = acc;
    }
    __syncthreads();
  }
  for (int l = L - 2; l >= 0; l--) {
    int s = s0 >> l, cs = s >> 1, lg = lg0 - l;
    float* x = bx + off[l];
    const float* xc = bx + off[l + 1];
    for (int q = t; q < s * s; q += 1024) {
      int j = q >> lg, i = q & (s - 1);
      int i0 = (i & 1) ? (i >> 1) : (i >> 1) - 1, j0 = (j & 1) ? (j >> 1) : (j >> 1) - 1;
      float fx = (i & 1) ? 0.25f : 0.75f, fy = (j & 1) ? 0.25f : 0.75f;
      int x0 = max(i0, 0), x1 = min(i0 + 1, cs - 1), y0 = max(j0, 0), y1 = min(j0 + 1, cs - 1);
      float a = xc[y0 * cs + x0], bb = xc[y0 * cs + x1], c = xc[y1 * cs + x0], d = xc[y1 * cs + x1];
      float top = a * (1.f - fx) + bb * fx, bot = c * (1.f - fx) + d * fx;
      x[q] += top * (1.f - fy) + bot * fy;
    }
    __syncthreads();
    for (int sw = 0; sw < post; sw++) {
      pass(l, 0, false);
      pass(l, 1, false);
    }
  }
  for (int q = t; q < s0 * s0; q += 1024) xall[b + q] = bx[q];
}

}  // namespace fsk
