// fs_kernels.hpp -- HIP kernels (gfx950) of the stable-fluids step on a periodic MAC grid.
//
// Storage is SoA FP32, row-major with i fastest: u[j*nx+i] on the left face of
// cell (i,j), v[j*nx+i] on its bottom face, pressure at the centre
// (the index map of src/core/grid.js:83-105). Launches use 64-wide rows so one
// wavefront reads 64 consecutive faces: every stencil load is coalesced.
//
// Arithmetic is written with explicit fmaf where a fused multiply-add is meant
// and the library is compiled with -ffp-contract=off, so two kernels that
// evaluate the same formula produce the same bits.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fsk {

// Storage of the multigrid correction and right-hand sides: FP32, or BF16 (the
// upper 16 bits of an FP32, rounded to nearest even). Arithmetic is FP32 either
// way; only loads and stores convert.
struct bf16 {
  uint16_t b;
};
__device__ __forceinline__ float ld(float x) { return x; }
__device__ __forceinline__ float ld(bf16 x) { return __uint_as_float((uint32_t)x.b << 16); }
template <typename S>
__device__ __forceinline__ S st(float f);
template <>
__device__ __forceinline__ float st<float>(float f) {
  return f;
}
template <>
__device__ __forceinline__ bf16 st<bf16>(float f) {
  // v_cvt_pk_bf16_f32: round to nearest, ties to even; NaN stays NaN
  return bf16{__builtin_bit_cast(uint16_t, (__bf16)f)};
}
// value as stored: the FP32 a store of type S followed by a load returns
template <typename S>
__device__ __forceinline__ float rnd(float f) {
  return ld(st<S>(f));
}


constexpr int BX = 64;
constexpr int BY = 4;

// Block coordinates of a 2D grid with the dispatch order remapped so that each XCD
// takes a contiguous run of row-major blocks: workgroups are handed round-robin to
// the 8 XCDs in linear order, which would put x-adjacent blocks on different XCDs,
// each of whose L2s then fetches the lines at both ends of every row segment.
__device__ __forceinline__ int2 xcd_block2() {
  const int gx = (int)gridDim.x, nb = gx * (int)gridDim.y;
  const int b = (int)blockIdx.x + (int)blockIdx.y * gx;
  const int t = (nb & 7) ? b : (b & 7) * (nb >> 3) + (b >> 3);
  return int2{t % gx, t / gx};
}

__device__ __forceinline__ int wrapi(int i, int n) {
  if ((unsigned)i < (unsigned)n) return i;
  int r = i % n;
  return r < 0 ? r + n : r;
}
__device__ __forceinline__ int nxt(int i, int n) { return (i + 1 == n) ? 0 : i + 1; }
__device__ __forceinline__ int prv(int i, int n) { return (i == 0) ? n - 1 : i - 1; }

// Periodic bilinear sample at grid coordinates (gx, gy); the CPU form
// a + (b-a)*f of src/core/grid.js:147-165.
__device__ __forceinline__ float bilerp(const float* __restrict__ a, int nx, int ny, float gx, float gy) {
  float fi = floorf(gx), fj = floorf(gy);
  int i0 = wrapi((int)fi, nx), j0 = wrapi((int)fj, ny);
  float fx = gx - fi, fy = gy - fj;
  int i1 = nxt(i0, nx), j1 = nxt(j0, ny);
  const float* r0 = a + (size_t)j0 * nx;
  const float* r1 = a + (size_t)j1 * nx;
  float A = r0[i0], B = r0[i1], Cc = r1[i0], D = r1[i1];
  float top = fmaf(B - A, fx, A);
  float bot = fmaf(D - Cc, fx, Cc);
  return fmaf(bot - top, fy, top);
}

// min/max of the four nodes a bilinear sample at (gx, gy) reads
// (interpBounds, src/cpu/advect.js:113-121).
__device__ __forceinline__ void bounds4(const float* __restrict__ a, int nx, int ny, float gx, float gy,
                                        float& lo, float& hi) {
  int i0 = wrapi((int)floorf(gx), nx), j0 = wrapi((int)floorf(gy), ny);
  int i1 = nxt(i0, nx), j1 = nxt(j0, ny);
  float A = a[(size_t)j0 * nx + i0], B = a[(size_t)j0 * nx + i1];
  float Cc = a[(size_t)j1 * nx + i0], D = a[(size_t)j1 * nx + i1];
  lo = fminf(fminf(A, B), fminf(Cc, D));
  hi = fmaxf(fmaxf(A, B), fmaxf(Cc, D));
}

// Samplers in cell coordinates X = x/dx, Y = y/dy (src/core/grid.js:167-180):
// u nodes sit at (i, j+1/2), v nodes at (i+1/2, j).
__device__ __forceinline__ float sU(const float* __restrict__ u, int nx, int ny, float X, float Y) {
  return bilerp(u, nx, ny, X, Y - 0.5f);
}
__device__ __forceinline__ float sV(const float* __restrict__ v, int nx, int ny, float X, float Y) {
  return bilerp(v, nx, ny, X - 0.5f, Y);
}

// RK2 midpoint backward trace (src/cpu/advect.js:46-54) in cell coordinates;
// cx = dt/dx, cy = dt/dy.
__device__ __forceinline__ void trace(const float* __restrict__ u, const float* __restrict__ v, int nx, int ny,
                                      float X, float Y, float cx, float cy, float& Xb, float& Yb) {
  float u1 = sU(u, nx, ny, X, Y), v1 = sV(v, nx, ny, X, Y);
  float Xm = fmaf(-0.5f * cx, u1, X), Ym = fmaf(-0.5f * cy, v1, Y);
  float u2 = sU(u, nx, ny, Xm, Ym), v2 = sV(v, nx, ny, Xm, Ym);
  Xb = fmaf(-cx, u2, X);
  Yb = fmaf(-cy, v2, Y);
}

// Per-step coefficients, kept on the device so a CFL time step never has to
// visit the host: cx = dt/dx, cy = dt/dy, da = nu dt/dx^2, db = nu dt/dy^2,
// fdt = dt (forcing amplitude factor), drag = exp(-alpha dt).
struct StepCoef {
  double dt;
  float cx, cy, da, db, fdt, drag;
  int64_t step;  // index of the step these coefficients belong to
  double t;      // simulated time at the start of the step
  double t_end;  // and at its end (the next step's start)
};

// Device-side convergence gate of a tolerance-driven projection. flag[k] != 0 means
// "refinement residual k still has a cell above the tolerance"; every block of the
// residual pass that finds one stores 1 there (the same value from any number of
// blocks, so plain stores suffice), and the flags start the step at zero.
//   skip >= 0: the launch does nothing unless flag[skip] is set (a V-cycle on
//              residual k checks k; the residual pass k checks k - 1, whose V-cycle
//              produced the correction it adds)
//   rec >= 0:  a residual pass records into flag[rec]
// Skipped iterations therefore leave their flags at zero, and the first zero flag
// is the iteration at which the gate was met. skip = rec = -1: ungated.
struct Gate {
  int* flag;
  double tol;
  int skip, rec;
};
__device__ __forceinline__ bool gate_closed(const Gate& g) { return g.skip >= 0 && g.flag[g.skip] == 0; }
__device__ __forceinline__ void gate_record(const Gate& g, double block_max) {
  if (g.rec >= 0 && block_max > g.tol) g.flag[g.rec] = 1;
}
__host__ __device__ inline Gate gate_open() { return Gate{nullptr, 0.0, -1, -1}; }

// Where the gradient pass of a device-gated projection finds the pressure, and the
// per-step record it leaves for the host.
//   k >= 0: the projection ran gated iterations 0 .. k; the pressure after the first
//           iteration j whose flag is zero (j = k when none is) is pbuf[(a0 + j + 1) & 1].
//           The pass subtracts its gradient and, when that buffer is not pbuf[keep],
//           also copies it there, so the host knows where the pressure is without
//           reading anything back.
//   k < 0:  the pressure is pbuf[0]; no record.
struct StepRecord {
  double dt;        // the step's dt
  double residual;  // FP64 max|div u* - L p| after the last iteration run
  int vcycles;      // V-cycles run (j)
  int missed;       // 1: no iteration met the gate
  int64_t step;     // step index (0-based), so the host can check the slot
};
struct FinalP {
  const int* flag;
  const double* rpart;  // per-block residual maxima, one row of `cols` per iteration
  int cols;
  int k, a0, keep;
  double* pbuf[2];
  StepRecord* ring;     // host-visible records, slot = step % ring_len (null: no record)
  int ring_len;
  const StepCoef* coef; // dt and the step index
};

// ------------------------------------------------------------------ advection

// Semi-Lagrangian self-advection of both components, each traced from its own
// face (src/cpu/advect.js:74-97; ADVECT_VELOCITY src/gpu/shaders.js:126-146).
__global__ void __launch_bounds__(BX * BY)
k_advect_sl(const float* __restrict__ u, const float* __restrict__ v, float* __restrict__ ou,
            float* __restrict__ ov, int nx, int ny, const StepCoef* __restrict__ cf, float sign) {
  const int2 bk = xcd_block2();
  int i = bk.x * BX + threadIdx.x, j = bk.y * BY + threadIdx.y;
  if (i >= nx || j >= ny) return;
  float cx = sign * cf->cx, cy = sign * cf->cy;
  float Xb, Yb;
  trace(u, v, nx, ny, (float)i, j + 0.5f, cx, cy, Xb, Yb);
  ou[(size_t)j * nx + i] = sU(u, nx, ny, Xb, Yb);
  trace(u, v, nx, ny, i + 0.5f, (float)j, cx, cy, Xb, Yb);
  ov[(size_t)j * nx + i] = sV(v, nx, ny, Xb, Yb);
}

// MacCormack correction + limiter (src/cpu/advect.js:149-206;
// ADVECT_FIELD_BY + MACCORMACK_COMBINE src/gpu/shaders.js:155-231):
// back = A^-1(fwd) traced through the ORIGINAL flow at -dt, out = fwd + (orig-back)/2
// clamped to the four nodes around the forward trace.
__global__ void __launch_bounds__(BX * BY)
k_advect_mc(const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ fu,
            const float* __restrict__ fv, float* __restrict__ ou, float* __restrict__ ov, int nx, int ny,
            const StepCoef* __restrict__ cf) {
  const int2 bk = xcd_block2();
  int i = bk.x * BX + threadIdx.x, j = bk.y * BY + threadIdx.y;
  if (i >= nx || j >= ny) return;
  float cx = cf->cx, cy = cf->cy;
  size_t k = (size_t)j * nx + i;
  float Xb, Yb, lo, hi;
  {  // u face at (i, j+1/2)
    float X = (float)i, Y = j + 0.5f;
    trace(u, v, nx, ny, X, Y, -cx, -cy, Xb, Yb);
    float back = bilerp(fu, nx, ny, Xb, Yb - 0.5f);
    float c = fmaf(0.5f, u[k] - back, fu[k]);
    trace(u, v, nx, ny, X, Y, cx, cy, Xb, Yb);
    bounds4(u, nx, ny, Xb, Yb - 0.5f, lo, hi);
    ou[k] = c < lo ? lo : (c > hi ? hi : c);
  }
  {  // v face at (i+1/2, j)
    float X = i + 0.5f, Y = (float)j;
    trace(u, v, nx, ny, X, Y, -cx, -cy, Xb, Yb);
    float back = bilerp(fv, nx, ny, Xb - 0.5f, Yb);
    float c = fmaf(0.5f, v[k] - back, fv[k]);
    trace(u, v, nx, ny, X, Y, cx, cy, Xb, Yb);
    bounds4(v, nx, ny, Xb - 0.5f, Yb, lo, hi);
    ov[k] = c < lo ? lo : (c > hi ? hi : c);
  }
}

// ----------------------------------------------------------- dye and splats

// Semi-Lagrangian advection of the cell-centred rgba dye (ADVECT_SCALAR,
// src/gpu/shaders.js:234-266): the RK2 trace from the centre, a bilinear sample of the
// four channels, the fade, clamped to [0, 1].
__global__ void __launch_bounds__(BX * BY)
k_advect_dye(const float* __restrict__ u, const float* __restrict__ v, const float4* __restrict__ q,
             float4* __restrict__ out, int nx, int ny, const StepCoef* __restrict__ cf, float fade) {
  const int2 bk = xcd_block2();
  int i = bk.x * BX + threadIdx.x, j = bk.y * BY + threadIdx.y;
  if (i >= nx || j >= ny) return;
  float Xb, Yb;
  trace(u, v, nx, ny, i + 0.5f, j + 0.5f, cf->cx, cf->cy, Xb, Yb);
  const float gx = Xb - 0.5f, gy = Yb - 0.5f;
  const float fi = floorf(gx), fj = floorf(gy);
  const int i0 = wrapi((int)fi, nx), j0 = wrapi((int)fj, ny), i1 = nxt(i0, nx), j1 = nxt(j0, ny);
  const float fx = gx - fi, fy = gy - fj;
  const float4 A = q[(size_t)j0 * nx + i0], B = q[(size_t)j0 * nx + i1];
  const float4 Cc = q[(size_t)j1 * nx + i0], D = q[(size_t)j1 * nx + i1];
  auto lerp2 = [&](float a, float b, float c, float d) {
    float top = fmaf(b - a, fx, a), bot = fmaf(d - c, fx, c);
    float r = fmaf(bot - top, fy, top) * (1.f - fade);
    return r < 0.f ? 0.f : (r > 1.f ? 1.f : r);
  };
  out[(size_t)j * nx + i] = float4{lerp2(A.x, B.x, Cc.x, D.x), lerp2(A.y, B.y, Cc.y, D.y),
                                   lerp2(A.z, B.z, Cc.z, D.z), lerp2(A.w, B.w, Cc.w, D.w)};
}

// Splats (SPLAT, src/gpu/shaders.js:474-511), applied in order to each cell: the weight
// g = exp(-|d|^2 / max(radius^2, 1e-12)) of the periodic shortest separation d of the
// cell centre from the splat (round half to even, as WGSL's round); a force splat adds
// (dx g, dy g) to the cell's u and v faces, a dye splat (r g, g g, b g, 0) to its dye.
// Every splat carries its own parameters (the reference binds the first one's for all,
// src/gpu/solver.js:546-553 against :745-768).
struct Splat {
  float x, y, dx, dy, r, g, b, radius;
  int is_dye;
};
constexpr int MAX_SPLATS = 16;
struct SplatSet {
  Splat s[MAX_SPLATS];
  int n;
};
__global__ void __launch_bounds__(BX * BY)
k_splat(float* __restrict__ u, float* __restrict__ v, float4* __restrict__ dye, int nx, int ny, float dx, float dy,
        const SplatSet ss) {
  const int2 bk = xcd_block2();
  int i = bk.x * BX + threadIdx.x, j = bk.y * BY + threadIdx.y;
  if (i >= nx || j >= ny) return;
  const float px = ((float)i + 0.5f) * dx, py = ((float)j + 0.5f) * dy;
  const float Lx = (float)nx * dx, Ly = (float)ny * dy;
  const size_t k = (size_t)j * nx + i;
  float uu = u[k], vv = v[k];
  float4 c = dye ? dye[k] : float4{0.f, 0.f, 0.f, 0.f};
  for (int m = 0; m < ss.n; m++) {
    const Splat& sp = ss.s[m];
    float ddx = px - sp.x, ddy = py - sp.y;
    ddx = ddx - Lx * rintf(ddx / Lx);
    ddy = ddy - Ly * rintf(ddy / Ly);
    const float w = expf(-(ddx * ddx + ddy * ddy) / fmaxf(sp.radius * sp.radius, 1e-12f));
    if (sp.is_dye) {
      c.x += sp.r * w;
      c.y += sp.g * w;
      c.z += sp.b * w;
    } else {
      uu += sp.dx * w;
      vv += sp.dy * w;
    }
  }
  u[k] = uu;
  v[k] = vv;
  if (dye) dye[k] = c;
}

// Corner-vorticity kernels (confinement, render) work on tiles of VT_X x VT_Y cells, two
// rows per thread of a BX x BY block: u and v on the tile and a ring are staged in LDS
// once, and the corner vorticity v_x - u_y of every corner the tile reads is computed
// once, with the same operations per corner as a thread computing it alone.
constexpr int VT_X = BX, VT_Y = 2 * BY;
inline dim3 vt_grid(int nx, int ny) {
  return dim3((unsigned)((nx + VT_X - 1) / VT_X), (unsigned)((ny + VT_Y - 1) / VT_Y));
}
// the corner (a, b): (v(a, b) - v(a - 1, b)) / dx - (u(a, b) - u(a, b - 1)) / dy
__device__ __forceinline__ float corner_curl(float vab, float vlb, float uab, float uab1, float dx, float dy) {
  return (vab - vlb) / dx - (uab - uab1) / dy;
}
// u, v of the cells (i0 - R + lx, j0 - R + ly), W x H of each, wrapped (periodic)
template <int W, int H, int R>
__device__ __forceinline__ void stage_uv(const float* __restrict__ u, const float* __restrict__ v, float* su,
                                         float* sv, int nx, int ny, int i0, int j0) {
  for (int q = threadIdx.x + threadIdx.y * BX; q < W * H; q += BX * BY) {
    const int ly = q / W, lx = q - ly * W;
    const size_t g = (size_t)wrapi(j0 - R + ly, ny) * nx + wrapi(i0 - R + lx, nx);
    su[q] = u[g];
    sv[q] = v[g];
  }
}

// Vorticity confinement as the reference's GPU step applies it (VORTICITY_CONFINEMENT,
// src/gpu/shaders.js:437-467; a non-physical force, off unless eps > 0): the corner
// vorticity w at (i, j) (the lower-left corner of cell (i, j)), the unit vector n along
// the central difference of |w| over the four neighbouring corners (|grad| + 1e-20
// guards a uniform flow), and the force eps dx dt (n_y w, -n_x w) added to the cell's
// own u and v faces. Reads (u, v), writes (ou, ov); periodic; launched on vt_grid.
__global__ void __launch_bounds__(BX * BY)
k_confine(const float* __restrict__ u, const float* __restrict__ v, float* __restrict__ ou, float* __restrict__ ov,
          int nx, int ny, float dx, float dy, float eps, const StepCoef* __restrict__ cf) {
  // window: cells i0-2 .. i0+VT_X, j0-2 .. j0+VT_Y; corners i0-1 .. i0+VT_X, j0-1 .. j0+VT_Y
  constexpr int W = VT_X + 3, H = VT_Y + 3, CW = VT_X + 2, CH = VT_Y + 2;
  __shared__ float su[W * H], sv[W * H], sc[CW * CH];
  const int2 bk = xcd_block2();
  const int i0 = bk.x * VT_X, j0 = bk.y * VT_Y;
  stage_uv<W, H, 2>(u, v, su, sv, nx, ny, i0, j0);
  __syncthreads();
  for (int q = threadIdx.x + threadIdx.y * BX; q < CW * CH; q += BX * BY) {
    const int cy = q / CW, cx = q - cy * CW;  // corner (i0 - 1 + cx, j0 - 1 + cy): window (cx + 1, cy + 1)
    const int k = (cy + 1) * W + cx + 1;
    sc[q] = corner_curl(sv[k], sv[k - 1], su[k], su[k - W], dx, dy);
  }
  __syncthreads();
  const float s = eps * dx * cf->fdt;
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const int li = threadIdx.x, lj = threadIdx.y + r * BY;
    const int i = i0 + li, j = j0 + lj;
    if (i >= nx || j >= ny) continue;
    const int c = (lj + 1) * CW + li + 1;  // corner (i, j)
    const float w = sc[c];
    const float wl = fabsf(sc[c - 1]), wr = fabsf(sc[c + 1]);
    const float wb = fabsf(sc[c - CW]), wt = fabsf(sc[c + CW]);
    float gx = (wr - wl) / (2.f * dx), gy = (wt - wb) / (2.f * dy);
    const float m = sqrtf(gx * gx + gy * gy) + 1e-20f;
    gx = gx / m;
    gy = gy / m;
    const int k = (lj + 2) * W + li + 2;
    const size_t g = (size_t)j * nx + i;
    ou[g] = su[k] + s * (gy * w);
    ov[g] = sv[k] + s * (-gx * w);
  }
}

// ----------------------------------------------------------- display: render and bloom
//
// Display only: these read the fields and write pixels that nothing in the step reads.
// Pixels are rgba half floats (IEEE binary16, round to nearest even), as the reference's
// rgba16float targets. The WGSL built-ins are spelled out: smoothstep(e0, e1, x) =
// t^2 (3 - 2t) with t = clamp((x - e0) / (e1 - e0), 0, 1), mix(a, b, t) = a (1 - t) + b t.

struct __align__(8) Half4 {
  _Float16 x, y, z, w;
};
__device__ __forceinline__ float smoothstep_f(float e0, float e1, float x) {
  float t = (x - e0) / (e1 - e0);
  t = fminf(fmaxf(t, 0.f), 1.f);
  return t * t * (3.f - 2.f * t);
}
__device__ __forceinline__ float mix_f(float a, float b, float t) { return a * (1.f - t) + b * t; }
__device__ __forceinline__ float3 mix3(float3 a, float3 b, float t) {
  return float3{mix_f(a.x, b.x, t), mix_f(a.y, b.y, t), mix_f(a.z, b.z, t)};
}
__device__ __forceinline__ float3 scale3(float3 a, float s) { return float3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float3 add3(float3 a, float3 b) { return float3{a.x + b.x, a.y + b.y, a.z + b.z}; }
// Rec. 709 luminance, summed left to right
__device__ __forceinline__ float lum709(float3 c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; }
__device__ __forceinline__ Half4 to_half4(float3 c) {
  return Half4{(_Float16)c.x, (_Float16)c.y, (_Float16)c.z, (_Float16)1.0f};
}
__device__ __forceinline__ float3 rgb_of(Half4 h) { return float3{(float)h.x, (float)h.y, (float)h.z}; }

// RENDER (src/gpu/shaders.js:677-780): the corner vorticity (v_x - u_y at the four
// corners of the cell, averaged to its centre) through a signed neon ramp (mode < 0.5), or
// the face speed |(u, v)| of the cell (mode >= 0.5), both with the auto-exposure `scale`;
// a white-hot term from the speed; a luminance lift and a little colour from the dye
// (dye may be null: no dye yet); non-finite channels painted 0. Periodic neighbours;
// launched on vt_grid.
__global__ void __launch_bounds__(BX * BY)
k_render(const float* __restrict__ u, const float* __restrict__ v, const float4* __restrict__ dye,
         Half4* __restrict__ out, int nx, int ny, float dx, float dy, float scale, float mode) {
  // window: cells i0-1 .. i0+VT_X, j0-1 .. j0+VT_Y; corners i0 .. i0+VT_X, j0 .. j0+VT_Y
  constexpr int W = VT_X + 2, H = VT_Y + 2, CW = VT_X + 1, CH = VT_Y + 1;
  __shared__ float su[W * H], sv[W * H], sc[CW * CH];
  const int2 bk = xcd_block2();
  const int i0 = bk.x * VT_X, j0 = bk.y * VT_Y;
  stage_uv<W, H, 1>(u, v, su, sv, nx, ny, i0, j0);
  __syncthreads();
  for (int q = threadIdx.x + threadIdx.y * BX; q < CW * CH; q += BX * BY) {
    const int cy = q / CW, cx = q - cy * CW;  // corner (i0 + cx, j0 + cy): window (cx + 1, cy + 1)
    const int k = (cy + 1) * W + cx + 1;
    sc[q] = corner_curl(sv[k], sv[k - 1], su[k], su[k - W], dx, dy);
  }
  __syncthreads();
  const float3 MAGENTA{1.00f, 0.10f, 0.62f}, CYAN{0.10f, 0.92f, 1.00f}, VIOLET{0.16f, 0.03f, 0.42f},
      VOID_{0.004f, 0.003f, 0.016f};
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const int li = threadIdx.x, lj = threadIdx.y + r * BY;
    const int i = i0 + li, j = j0 + lj;
    if (i >= nx || j >= ny) continue;
    const int c = lj * CW + li;
    const float w = 0.25f * (sc[c] + sc[c + 1] + sc[c + CW] + sc[c + CW + 1]);
    const int k = (lj + 1) * W + li + 1;
    const float cu = su[k], cv = sv[k];
    const float speed = sqrtf(cu * cu + cv * cv);
    float3 col;
    if (mode < 0.5f) {
      const float sig = tanhf(fminf(fmaxf(w * scale, -12.f), 12.f));
      const float mag = fabsf(sig);
      const float3 hue = sig >= 0.f ? MAGENTA : CYAN;
      const float3 ramp =
          mix3(mix3(VOID_, VIOLET, smoothstep_f(0.f, 0.22f, mag)), hue, smoothstep_f(0.12f, 0.62f, mag));
      col = scale3(ramp, 0.30f + 4.6f * mag * mag);
    } else {
      const float t = tanhf(fminf(fmaxf(speed * scale, -12.f), 12.f));
      const float3 hue = mix3(VIOLET, CYAN, smoothstep_f(0.15f, 0.95f, t));
      col = scale3(mix3(VOID_, hue, smoothstep_f(0.f, 0.6f, t)), 0.35f + 2.9f * t * t);
    }
    const float sp = tanhf(fminf(fmaxf(speed * scale * 0.45f, 0.f), 12.f));
    const float sp2 = sp * sp;
    col = add3(col, scale3(float3{0.55f, 0.78f, 1.0f}, (sp2 * sp2 * sp) * 0.85f));
    const size_t g = (size_t)j * nx + i;
    const float4 d = dye ? dye[g] : float4{0.f, 0.f, 0.f, 0.f};
    const float3 drgb{d.x, d.y, d.z};
    col = add3(add3(col, scale3(scale3(col, lum709(drgb)), 1.5f)), scale3(drgb, 0.20f));
    col = float3{col.x == col.x ? col.x : 0.f, col.y == col.y ? col.y : 0.f, col.z == col.z ? col.z : 0.f};
    out[g] = to_half4(col);
  }
}

// BLOOM_BRIGHT (src/gpu/shaders.js:783-812): the 2x2 box average of the render at half
// resolution (bw x bh pixels, the render's row stride nx), kept by a soft knee
// smoothstep(thr, thr + 0.6, luminance).
__global__ void __launch_bounds__(BX * BY)
k_bloom_bright(const Half4* __restrict__ src, Half4* __restrict__ dst, int bw, int bh, int nx, float thr) {
  const int2 bk = xcd_block2();
  const int i = bk.x * BX + threadIdx.x, j = bk.y * BY + threadIdx.y;
  if (i >= bw || j >= bh) return;
  const size_t r0 = (size_t)(2 * j) * nx + 2 * i, r1 = r0 + nx;
  float3 acc = rgb_of(src[r0]);
  acc = add3(acc, rgb_of(src[r0 + 1]));
  acc = add3(acc, rgb_of(src[r1]));
  acc = add3(acc, rgb_of(src[r1 + 1]));
  acc = scale3(acc, 0.25f);
  const float knee = smoothstep_f(thr, thr + 0.6f, lum709(acc));
  dst[(size_t)j * bw + i] = to_half4(scale3(acc, knee));
}

// BLOOM_BLUR (src/gpu/shaders.js:816-843): one axis (0 = x, 1 = y) of the separable
// 13-tap Gaussian (sigma ~ 3), normalised by its weights, taps clamped to the image (the
// image is not periodic).
__global__ void __launch_bounds__(BX * BY)
k_bloom_blur(const Half4* __restrict__ src, Half4* __restrict__ dst, int bw, int bh, int axis) {
  const int2 bk = xcd_block2();
  const int i = bk.x * BX + threadIdx.x, j = bk.y * BY + threadIdx.y;
  if (i >= bw || j >= bh) return;
  constexpr float W[7] = {0.1964f, 0.1747f, 0.1210f, 0.0656f, 0.0278f, 0.0092f, 0.0024f};
  auto at = [&](int a, int b) {
    a = a < 0 ? 0 : (a > bw - 1 ? bw - 1 : a);
    b = b < 0 ? 0 : (b > bh - 1 ? bh - 1 : b);
    return rgb_of(src[(size_t)b * bw + a]);
  };
  const int ox = axis ? 0 : 1, oy = axis ? 1 : 0;
  float3 acc = scale3(at(i, j), W[0]);
  float norm = W[0];
#pragma unroll
  for (int k = 1; k < 7; k++) {
    acc = add3(acc, scale3(add3(at(i + ox * k, j + oy * k), at(i - ox * k, j - oy * k)), W[k]));
    norm = norm + 2.f * W[k];
  }
  dst[(size_t)j * bw + i] = to_half4(float3{acc.x / norm, acc.y / norm, acc.z / norm});
}

// Explicit 5-point diffusion of both components (src/cpu/solver.js:56-69;
// DIFFUSE src/gpu/shaders.js:401-425); a = nu dt/dx^2, b = nu dt/dy^2.
__global__ void __launch_bounds__(BX * BY)
k_diffuse(const float* __restrict__ u, const float* __restrict__ v, float* __restrict__ ou,
          float* __restrict__ ov, int nx, int ny, const StepCoef* __restrict__ cf) {
  const int2 bk = xcd_block2();
  int i = bk.x * BX + threadIdx.x, j = bk.y * BY + threadIdx.y;
  if (i >= nx || j >= ny) return;
  float a = cf->da, b = cf->db;
  int im = prv(i, nx), ip = nxt(i, nx), jm = prv(j, ny), jp = nxt(j, ny);
  size_t k = (size_t)j * nx + i;
  size_t kl = (size_t)j * nx + im, kr = (size_t)j * nx + ip, kb = (size_t)jm * nx + i, kt = (size_t)jp * nx + i;
  float c = u[k];
  ou[k] = c + a * (u[kl] - 2.f * c + u[kr]) + b * (u[kb] - 2.f * c + u[kt]);
  c = v[k];
  ov[k] = c + a * (v[kl] - 2.f * c + v[kr]) + b * (v[kb] - 2.f * c + v[kt]);
}

// ------------------------------------------- fused MacCormack + explicit diffusion
//
// One launch for k_advect_sl (forward) + k_advect_mc + k_diffuse, with the same bits.
// A workgroup owns an AT x AT tile. It stages u, v on the tile with an AU-cell ring in
// LDS, computes the forward semi-Lagrangian result on an AF-cell ring, the corrected
// and limited result on a 1-cell ring, and the diffusion of that on the tile. Every
// sample uses the same floor, fraction and FMA sequence as the global samplers; a
// sample whose four nodes are not all inside its LDS window reads them from global
// memory, and a forward value needed outside its window is recomputed exactly as the
// forward pass computes it, so the result is independent of dt.
// Ring widths: when every trace moves less than one cell (cx max|u| and cy max|v| of
// the window at most 0.99), a sample's four nodes lie within one cell of the node it
// starts from, so the corrected nodes (tile + 1) back-sample forward nodes on tile + 2
// (AF), whose traces read u, v on tile + 3 (AU).
constexpr int AT = 32, AF = 2, AU = AF + 1;
constexpr int ANT = 512;  // threads per tile
constexpr int AUW = AT + 2 * AU, AFW = AT + 2 * AF, AMW = AT + 2;

struct AdvWin {
  const float* w;  // LDS window of side AUW, origin (ox, oy) in unwrapped cell indices
  int ox, oy;
  int olin;        // oy * AUW + ox: the window position of cell (i, j) is j * AUW + i - olin
  const float* g;  // the field in global memory
};

// Window position of the node (fi, fj) given as floats holding integers (floors of
// sample coordinates, |fi|, |fj| < 2^23): fj * W + fi is exact in FP32, so one FMA
// and one conversion replace two conversions, two subtractions and a multiply.
template <int W>
__device__ __forceinline__ int lin_index(float fi, float fj, int olin) {
  return (int)fmaf(fj, (float)W, fi) - olin;
}

// The four nodes (i0, j0), (i0+1, j0), (i0, j0+1), (i0+1, j0+1) at wrapped indices from
// global memory: the rare path, kept out of line so the common one stays small.
__device__ __noinline__ float4 nodes4_global(const float* __restrict__ g, int nx, int ny, int fi, int fj) {
  int i0 = wrapi(fi, nx), j0 = wrapi(fj, ny);
  int i1 = nxt(i0, nx), j1 = nxt(j0, ny);
  const float* r0 = g + (size_t)j0 * nx;
  const float* r1 = g + (size_t)j1 * nx;
  return float4{r0[i0], r0[i1], r1[i0], r1[i1]};
}

// Position of wrapped-or-not cell index fi in a window staged from the wrapped cells
// origin + l: (fi - origin) mod n, with one correction (enough for |fi| < 2n; other
// indices land outside the window and take the global path).
__device__ __forceinline__ int win_index(int fi, int origin, int n) {
  int l = fi - origin;
  l = l < 0 ? l + n : (l >= n ? l - n : l);
  return l;
}

// The four nodes of a bilinear sample at cell (fi, fj): from the window, or (FAST
// false and not all inside it) from global memory. WRAP false: the tile's windows
// do not cross the periodic seam, so window positions are plain differences.
template <bool FAST, bool WRAP = true>
__device__ __forceinline__ float4 nodes4_w(const AdvWin& a, int nx, int ny, int fi, int fj) {
  int li = WRAP ? win_index(fi, a.ox, nx) : fi - a.ox, lj = WRAP ? win_index(fj, a.oy, ny) : fj - a.oy;
  if (FAST || ((unsigned)li < (unsigned)(AUW - 1) && (unsigned)lj < (unsigned)(AUW - 1))) {
    const float* r0 = a.w + lj * AUW + li;
    return float4{r0[0], r0[1], r0[AUW], r0[AUW + 1]};
  }
  return nodes4_global(a.g, nx, ny, fi, fj);
}

// nodes4_w at floor values fi, fj (floats): on the fast inner path the window
// position comes from lin_index
template <bool FAST, bool WRAP>
__device__ __forceinline__ float4 nodes4_wf(const AdvWin& a, int nx, int ny, float fi, float fj) {
  if (FAST && !WRAP) {
    const float* r0 = a.w + lin_index<AUW>(fi, fj, a.olin);
    return float4{r0[0], r0[1], r0[AUW], r0[AUW + 1]};
  }
  return nodes4_w<FAST, WRAP>(a, nx, ny, (int)fi, (int)fj);
}

// bilerp() with the four nodes from the window when they are all inside it
template <bool FAST, bool WRAP = true>
__device__ __forceinline__ float bilerp_w(const AdvWin& a, int nx, int ny, float gx, float gy) {
  float fi = floorf(gx), fj = floorf(gy);
  float fx = gx - fi, fy = gy - fj;
  float4 q = nodes4_wf<FAST, WRAP>(a, nx, ny, fi, fj);
  float top = fmaf(q.y - q.x, fx, q.x);
  float bot = fmaf(q.w - q.z, fx, q.z);
  return fmaf(bot - top, fy, top);
}

// bilerp_w at a point whose fractions are exactly 1/2 (a face's other velocity
// component, sampled at the face): the nodes (i0, j0) .. (i0 + 1, j0 + 1) indexed
// directly, the same FMAs, so the same bits without the floor and fraction arithmetic
template <bool FAST, bool WRAP = true>
__device__ __forceinline__ float mid_w(const AdvWin& a, int nx, int ny, int i0, int j0) {
  float4 q = nodes4_w<FAST, WRAP>(a, nx, ny, i0, j0);
  float top = fmaf(q.y - q.x, 0.5f, q.x);
  float bot = fmaf(q.w - q.z, 0.5f, q.z);
  return fmaf(bot - top, 0.5f, top);
}

template <bool FAST, bool WRAP = true>
__device__ __forceinline__ void bounds4_w(const AdvWin& a, int nx, int ny, float gx, float gy, float& lo, float& hi) {
  float4 q = nodes4_wf<FAST, WRAP>(a, nx, ny, floorf(gx), floorf(gy));
  lo = fminf(fminf(q.x, q.y), fminf(q.z, q.w));
  hi = fmaxf(fmaxf(q.x, q.y), fmaxf(q.z, q.w));
}

// The value of node (i, j) itself: from the window when inside it, else from global
// memory at the wrapped indices.
template <bool FAST, bool WRAP = true>
__device__ __forceinline__ float node_w(const AdvWin& a, int nx, int ny, int i, int j) {
  int li = WRAP ? win_index(i, a.ox, nx) : i - a.ox, lj = WRAP ? win_index(j, a.oy, ny) : j - a.oy;
  if (FAST || ((unsigned)li < (unsigned)AUW && (unsigned)lj < (unsigned)AUW)) return a.w[lj * AUW + li];
  return a.g[(size_t)wrapi(j, ny) * nx + wrapi(i, nx)];
}

// RK2 departure point from (X, Y) given the velocity (u1, v1) there
template <bool FAST, bool WRAP = true>
__device__ __forceinline__ void trace2_w(const AdvWin& u, const AdvWin& v, int nx, int ny, float X, float Y, float u1,
                                         float v1, float cx, float cy, float& Xb, float& Yb) {
  float Xm = fmaf(-0.5f * cx, u1, X), Ym = fmaf(-0.5f * cy, v1, Y);
  float u2 = bilerp_w<FAST, WRAP>(u, nx, ny, Xm, Ym - 0.5f), v2 = bilerp_w<FAST, WRAP>(v, nx, ny, Xm - 0.5f, Ym);
  Xb = fmaf(-cx, u2, X);
  Yb = fmaf(-cy, v2, Y);
}

// The velocity at the u face (i, j + 1/2) and at the v face (i + 1/2, j): the face's
// own component is its node value (the bilinear weights there are exactly 0; a zero
// of either sign then moves the trace's midpoint identically, as a sum with X or
// Y >= 0 that is exact), the other one is interpolated.
template <bool FAST, bool WRAP = true>
__device__ __forceinline__ void vel_u_face(const AdvWin& u, const AdvWin& v, int nx, int ny, int i, int j, float& u1,
                                           float& v1) {
  u1 = node_w<FAST, WRAP>(u, nx, ny, i, j);
  v1 = mid_w<FAST, WRAP>(v, nx, ny, i - 1, j);  // bilerp at (i - 1/2, j + 1/2)
}
template <bool FAST, bool WRAP = true>
__device__ __forceinline__ void vel_v_face(const AdvWin& u, const AdvWin& v, int nx, int ny, int i, int j, float& u1,
                                           float& v1) {
  u1 = mid_w<FAST, WRAP>(u, nx, ny, i, j - 1);  // bilerp at (i + 1/2, j - 1/2)
  v1 = node_w<FAST, WRAP>(v, nx, ny, i, j);
}

// forward semi-Lagrangian value of component COMP at node (i, j) (wrapped indices),
// as k_advect_sl computes it
template <int COMP, bool FAST, bool WRAP = true>
__device__ __forceinline__ float fwd_node(const AdvWin& u, const AdvWin& v, int nx, int ny, int i, int j, float cx,
                                          float cy) {
  float Xb, Yb, u1, v1;
  if (COMP == 0) {
    vel_u_face<FAST, WRAP>(u, v, nx, ny, i, j, u1, v1);
    trace2_w<FAST, WRAP>(u, v, nx, ny, (float)i, j + 0.5f, u1, v1, cx, cy, Xb, Yb);
    return bilerp_w<FAST, WRAP>(u, nx, ny, Xb, Yb - 0.5f);
  }
  vel_v_face<FAST, WRAP>(u, v, nx, ny, i, j, u1, v1);
  trace2_w<FAST, WRAP>(u, v, nx, ny, i + 0.5f, (float)j, u1, v1, cx, cy, Xb, Yb);
  return bilerp_w<FAST, WRAP>(v, nx, ny, Xb - 0.5f, Yb);
}

// fwd_node, and the range of the four nodes its final sample reads: the bounds the
// MacCormack limiter of this node takes (interpBounds at the same departure point,
// src/cpu/advect.js:113-121, :166-183)
template <int COMP, bool FAST, bool WRAP = true>
__device__ __forceinline__ float fwd_node_bounds(const AdvWin& u, const AdvWin& v, int nx, int ny, int i, int j,
                                                 float cx, float cy, float& lo, float& hi) {
  float Xb, Yb, u1, v1, gx, gy;
  if (COMP == 0) {
    vel_u_face<FAST, WRAP>(u, v, nx, ny, i, j, u1, v1);
    trace2_w<FAST, WRAP>(u, v, nx, ny, (float)i, j + 0.5f, u1, v1, cx, cy, Xb, Yb);
    gx = Xb;
    gy = Yb - 0.5f;
  } else {
    vel_v_face<FAST, WRAP>(u, v, nx, ny, i, j, u1, v1);
    trace2_w<FAST, WRAP>(u, v, nx, ny, i + 0.5f, (float)j, u1, v1, cx, cy, Xb, Yb);
    gx = Xb - 0.5f;
    gy = Yb;
  }
  const AdvWin& a = COMP == 0 ? u : v;
  float fi = floorf(gx), fj = floorf(gy);
  float fx = gx - fi, fy = gy - fj;
  float4 q = nodes4_wf<FAST, WRAP>(a, nx, ny, fi, fj);
  lo = fminf(fminf(q.x, q.y), fminf(q.z, q.w));
  hi = fmaxf(fmaxf(q.x, q.y), fmaxf(q.z, q.w));
  float top = fmaf(q.y - q.x, fx, q.x);
  float bot = fmaf(q.w - q.z, fx, q.z);
  return fmaf(bot - top, fy, top);
}

// the four forward nodes around (i0, j0) at wrapped indices, as the global forward
// field holds them, recomputed (out of line)
template <int COMP>
__device__ __noinline__ float4 fwd_nodes4_far(const AdvWin u, const AdvWin v, int nx, int ny, int i0, int j0,
                                              float cx, float cy) {
  int wi0 = wrapi(i0, nx), wj0 = wrapi(j0, ny), wi1 = nxt(wi0, nx), wj1 = nxt(wj0, ny);
  return float4{fwd_node<COMP, false>(u, v, nx, ny, wi0, wj0, cx, cy),
                fwd_node<COMP, false>(u, v, nx, ny, wi1, wj0, cx, cy),
                fwd_node<COMP, false>(u, v, nx, ny, wi0, wj1, cx, cy),
                fwd_node<COMP, false>(u, v, nx, ny, wi1, wj1, cx, cy)};
}

// MacCormack back-sample of the forward field: bilerp of fwd, nodes from the forward
// window when inside it, else recomputed
template <int COMP, bool FAST, bool WRAP = true>
__device__ __forceinline__ float back_w(const float* __restrict__ sf, int fox, int foy, const AdvWin& u,
                                        const AdvWin& v, int nx, int ny, float gx, float gy, float cx, float cy) {
  float fi = floorf(gx), fj = floorf(gy);
  float fx = gx - fi, fy = gy - fj;
  float4 q;
  if (FAST && !WRAP) {
    const float* r0 = sf + lin_index<AFW>(fi, fj, foy * AFW + fox);
    q = float4{r0[0], r0[1], r0[AFW], r0[AFW + 1]};
  } else {
    int i0 = (int)fi, j0 = (int)fj;
    int li = WRAP ? win_index(i0, fox, nx) : i0 - fox, lj = WRAP ? win_index(j0, foy, ny) : j0 - foy;
    if (FAST || ((unsigned)li < (unsigned)(AFW - 1) && (unsigned)lj < (unsigned)(AFW - 1))) {
      const float* r0 = sf + lj * AFW + li;
      q = float4{r0[0], r0[1], r0[AFW], r0[AFW + 1]};
    } else {
      q = fwd_nodes4_far<COMP>(u, v, nx, ny, i0, j0, cx, cy);
    }
  }
  float top = fmaf(q.y - q.x, fx, q.x);
  float bot = fmaf(q.w - q.z, fx, q.z);
  return fmaf(bot - top, fy, top);
}

// Forward pass and corrected, limited pass of one tile (the diffusion follows in the
// kernel). FAST: every sample is known to lie in its window.
// LDS: the forward values on the AF ring (sfu, sfv) and, on the tile and its 1-cell
// ring, one float4 per corrected node (sbd): first the limiter bounds of the forward
// trace from that node (lo u, hi u, lo v, hi v), which only that node's correction
// reads, then the corrected pair (u, v) written over them by the same thread. The
// corrected node (lx, ly) of the 1-cell ring is the forward node (lx + 1, ly + 1):
// both windows hold the same wrapped nodes in the same order.
template <bool FAST, bool WRAP>
__device__ __forceinline__ void advect_tile(const AdvWin& wu, const AdvWin& wv, float* __restrict__ sfu,
                                            float* __restrict__ sfv, float4* __restrict__ sbd,
                                            const float* __restrict__ u, const float* __restrict__ v, int nx, int ny,
                                            int tx0, int ty0, float cx, float cy) {
  const int uox = wu.ox, uoy = wu.oy;
  // ---- forward pass on the tile and an AF ring; nodes at wrapped indices, as the
  // forward kernel evaluates them (its nodes are the wrapped cells of the grid)
  const int fox = tx0 - AF, foy = ty0 - AF;
  for (int q = threadIdx.x; q < AFW * AFW; q += ANT) {
    int ly = q / AFW, lx = q - ly * AFW;
    int i = WRAP ? wrapi(fox + lx, nx) : fox + lx, j = WRAP ? wrapi(foy + ly, ny) : foy + ly;
    // the forward trace from a node is also the limiter's trace at that node: keep the
    // bounds of its departure point where the node is corrected
    float4 bd;
    sfu[q] = fwd_node_bounds<0, FAST, WRAP>(wu, wv, nx, ny, i, j, cx, cy, bd.x, bd.y);
    sfv[q] = fwd_node_bounds<1, FAST, WRAP>(wu, wv, nx, ny, i, j, cx, cy, bd.z, bd.w);
    const unsigned mx = (unsigned)(lx - (AF - 1)), my = (unsigned)(ly - (AF - 1));
    if (mx < (unsigned)AMW && my < (unsigned)AMW) sbd[my * AMW + mx] = bd;
  }
  __syncthreads();
  // ---- correction + limiter on the tile and a 1-cell ring
  for (int q = threadIdx.x; q < AMW * AMW; q += ANT) {
    int ly = q / AMW, lx = q - ly * AMW;
    int i = WRAP ? wrapi(tx0 - 1 + lx, nx) : tx0 - 1 + lx, j = WRAP ? wrapi(ty0 - 1 + ly, ny) : ty0 - 1 + ly;
    // this node's place in the u, v window and in the forward window
    const int wiu = WRAP ? win_index(i, uox, nx) : i - uox, wju = WRAP ? win_index(j, uoy, ny) : j - uoy;
    const bool inu = FAST || ((unsigned)wiu < (unsigned)AUW && (unsigned)wju < (unsigned)AUW);
    const int qf = (ly + AF - 1) * AFW + lx + AF - 1;
    const float4 bd = sbd[q];
    float Xb, Yb, u1, v1, cu, cv;
    // the backward and the forward trace start from the same velocity at the face,
    // whose own component is the face's value
    {  // u face at (i, j+1/2)
      float X = (float)i, Y = j + 0.5f;
      u1 = inu ? wu.w[wju * AUW + wiu] : u[(size_t)j * nx + i];
      v1 = mid_w<FAST, WRAP>(wv, nx, ny, i - 1, j);  // bilerp at (X - 1/2, Y)
      trace2_w<FAST, WRAP>(wu, wv, nx, ny, X, Y, u1, v1, -cx, -cy, Xb, Yb);
      float back = back_w<0, FAST, WRAP>(sfu, fox, foy, wu, wv, nx, ny, Xb, Yb - 0.5f, cx, cy);
      float c = fmaf(0.5f, u1 - back, sfu[qf]);
      cu = c < bd.x ? bd.x : (c > bd.y ? bd.y : c);
    }
    {  // v face at (i+1/2, j)
      float X = i + 0.5f, Y = (float)j;
      u1 = mid_w<FAST, WRAP>(wu, nx, ny, i, j - 1);  // bilerp at (X, Y - 1/2)
      v1 = inu ? wv.w[wju * AUW + wiu] : v[(size_t)j * nx + i];
      trace2_w<FAST, WRAP>(wu, wv, nx, ny, X, Y, u1, v1, -cx, -cy, Xb, Yb);
      float back = back_w<1, FAST, WRAP>(sfv, fox, foy, wu, wv, nx, ny, Xb - 0.5f, Yb, cx, cy);
      float c = fmaf(0.5f, v1 - back, sfv[qf]);
      cv = c < bd.z ? bd.z : (c > bd.w ? bd.w : c);
    }
    *(float2*)&sbd[q] = float2{cu, cv};
  }
}

// LDS of one tile, 40.5 KB, so that four tiles share a CU
struct AdvLds {
  float su[AUW * AUW], sv[AUW * AUW];
  float sfu[AFW * AFW], sfv[AFW * AFW];
  float4 sbd[AMW * AMW];  // limiter bounds, then the corrected (u, v), per corrected node
  float wmax[2][ANT / 64];
};
static_assert(sizeof(AdvLds) <= 40960, "four advection tiles per CU");

// At most 64 VGPRs (eight wavefronts per SIMD, four tiles per CU with the LDS above):
// the fast paths need about 40; the general path's extra registers go to scratch.
__global__ void __launch_bounds__(ANT) __attribute__((amdgpu_waves_per_eu(8)))
k_advect_mc_diffuse(const float* __restrict__ u, const float* __restrict__ v, float* __restrict__ ou,
                    float* __restrict__ ov, int nx, int ny, int tiles_x, int nblocks,
                    const StepCoef* __restrict__ cf, int diffuse) {
  __shared__ AdvLds L;
  float* su = L.su;
  float* sv = L.sv;
  int b = (int)blockIdx.x;
  if ((nblocks & 7) == 0) b = (b & 7) * (nblocks >> 3) + (b >> 3);
  const int tx0 = (b % tiles_x) * AT, ty0 = (b / tiles_x) * AT;
  const float cx = cf->cx, cy = cf->cy;
  // ---- u, v on the tile and an AU ring (unwrapped origin, periodic data), with
  // the largest |u|, |v| of the window
  const int uox = tx0 - AU, uoy = ty0 - AU;
  float mu = 0.f, mv = 0.f;
  {
    // every load of the window in flight before the first LDS store
    constexpr int NL = (AUW * AUW + ANT - 1) / ANT;
    float ua[NL], va[NL];
#pragma unroll
    for (int k = 0; k < NL; k++) {
      const int q = (int)threadIdx.x + k * ANT;
      if (q < AUW * AUW) {
        int ly = q / AUW, lx = q - ly * AUW;
        size_t g = (size_t)wrapi(uoy + ly, ny) * nx + wrapi(uox + lx, nx);
        ua[k] = u[g];
        va[k] = v[g];
      }
    }
#pragma unroll
    for (int k = 0; k < NL; k++) {
      const int q = (int)threadIdx.x + k * ANT;
      if (q < AUW * AUW) {
        su[q] = ua[k];
        sv[q] = va[k];
        mu = fmaxf(mu, fabsf(ua[k]));
        mv = fmaxf(mv, fabsf(va[k]));
      }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    mu = fmaxf(mu, __shfl_xor(mu, o, 64));
    mv = fmaxf(mv, __shfl_xor(mv, o, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    L.wmax[0][threadIdx.x >> 6] = mu;
    L.wmax[1][threadIdx.x >> 6] = mv;
  }
  __syncthreads();
  mu = L.wmax[0][0];
  mv = L.wmax[1][0];
#pragma unroll
  for (int w = 1; w < ANT / 64; w++) {
    mu = fmaxf(mu, L.wmax[0][w]);
    mv = fmaxf(mv, L.wmax[1][w]);
  }
  const AdvWin wu{su, uox, uoy, uoy * AUW + uox, u}, wv{sv, uox, uoy, uoy * AUW + uox, v};
  // every velocity a trace reads comes from the window, so a trace moves at most
  // cx max|u| by cy max|v| cells; the rings hold every sample for moves up to one
  // cell (NaN fails the test and takes the general path)
  const bool fast = cx * mu <= 0.99f && cy * mv <= 0.99f && nx >= AUW && ny >= AUW;
  // a window that does not cross the periodic seam needs no wrapped indices
  const bool inner = uox >= 0 && uoy >= 0 && uox + AUW <= nx && uoy + AUW <= ny;
  if (fast && inner)
    advect_tile<true, false>(wu, wv, L.sfu, L.sfv, L.sbd, u, v, nx, ny, tx0, ty0, cx, cy);
  else if (fast)
    advect_tile<true, true>(wu, wv, L.sfu, L.sfv, L.sbd, u, v, nx, ny, tx0, ty0, cx, cy);
  else
    advect_tile<false, true>(wu, wv, L.sfu, L.sfv, L.sbd, u, v, nx, ny, tx0, ty0, cx, cy);
  __syncthreads();
  // ---- diffusion (or the corrected field itself) on the tile
  const float a = cf->da, bb = cf->db;
  const float4* sm = L.sbd;
  for (int q = threadIdx.x; q < AT * AT; q += ANT) {
    int iy = q / AT, ix = q - iy * AT;
    int i = tx0 + ix, j = ty0 + iy;
    if (i >= nx || j >= ny) continue;
    int m = (iy + 1) * AMW + ix + 1;
    size_t k = (size_t)j * nx + i;
    const float2 c = *(const float2*)&sm[m];
    if (!diffuse) {
      ou[k] = c.x;
      ov[k] = c.y;
      continue;
    }
    const float2 l = *(const float2*)&sm[m - 1], r = *(const float2*)&sm[m + 1];
    const float2 d = *(const float2*)&sm[m - AMW], t = *(const float2*)&sm[m + AMW];
    ou[k] = c.x + a * (l.x - 2.f * c.x + r.x) + bb * (d.x - 2.f * c.x + t.x);
    ov[k] = c.y + a * (l.y - 2.f * c.y + r.y) + bb * (d.y - 2.f * c.y + t.y);
  }
}

// Forcing and linear drag, in place, for two sources of modes:
//   band forcing (src/cpu/solver.js:200-237): (u + f dt) exp(-alpha dt), 24 seeded modes
//   a mode list (setForcing, src/gpu/solver.js:370-414; FORCING src/gpu/shaders.js:532-566):
//     u exp(-alpha dt) + f dt, the phases drifting by drift * t (after_drag = 1)
// modes[m] = (kx*dx/(2pi), ky*dy/(2pi), phase/(2pi)) in turns per cell; cu/cv carry
// amp/sqrt(M)*ky/|k| and -amp/sqrt(M)*kx/|k| (times dt on the device).
struct Mode { double tx, ty, tph; float cu, cv; };
constexpr int MAX_MODES = 32;
struct ModeSet {
  Mode m[MAX_MODES];
  int n;
  int after_drag;      // 1: the mode list's order
  double drift_turns;  // phase drift per unit time, in turns
};

// Evaluated separably: with A = frac(tx x) and B = frac(ty y + phase) reduced to
// [0, 1) turns in FP64 (so large wavenumbers keep full FP32 accuracy),
// c cos 2pi(A + B) = (c cos 2piA) cos 2piB + (-c sin 2piA) sin 2piB: 4 (nx + ny) sincospi
// per mode and step instead of 2 nx ny cospi, and two fused multiply-adds per mode and
// face. A block of FX x FY cells stages its columns' and rows' entries of the tables in LDS.
constexpr int FX = 64, FY = 32;

// Per-step tables of the separable forcing: col[m][i] = (cu cos 2piA, -cu sin 2piA) at the
// u face (x = i) and (cv cos 2piA, -cv sin 2piA) at the v face (x = i + 1/2);
// row[m][j] = (cos 2piB, sin 2piB) at the u face (y = j + 1/2) and at the v face (y = j).
__global__ void __launch_bounds__(256)
k_forcing_tables(const ModeSet* __restrict__ ms, int nx, int ny, float4* __restrict__ col, float4* __restrict__ row,
                 const StepCoef* __restrict__ cf) {
  const int q = blockIdx.x * 256 + threadIdx.x, per = nx + ny;
  if (q >= ms->n * per) return;
  const int m = q / per, r = q - m * per;
  const Mode& md = ms->m[m];
  double a, b;
  if (r < nx) {
    a = md.tx * r;
    b = md.tx * (r + 0.5);
  } else {
    const int j = r - nx;
    double ph = md.tph + ms->drift_turns * cf->t;
    ph -= floor(ph);
    a = md.ty * (j + 0.5) + ph;
    b = md.ty * j + ph;
  }
  a -= floor(a);
  b -= floor(b);
  float ca, sa, cb, sb;
  sincospif(2.f * (float)a, &sa, &ca);
  sincospif(2.f * (float)b, &sb, &cb);
  if (r < nx) col[(size_t)m * nx + r] = float4{md.cu * ca, -(md.cu * sa), md.cv * cb, -(md.cv * sb)};
  else row[(size_t)m * ny + (r - nx)] = float4{ca, sa, cb, sb};
}

// The forcing and drag of FX x FY cells per block of 256 threads (a column each, FY / 4
// rows apart by 4): the field's and the tables' loads are all issued first, then the
// tables are staged in LDS, then the sum over the modes in mode order, then the drag.
__global__ void __launch_bounds__(256)
k_forcing_sep(float* __restrict__ u, float* __restrict__ v, int nx, int ny, const ModeSet* __restrict__ ms,
              const float4* __restrict__ col, const float4* __restrict__ row, const StepCoef* __restrict__ cf) {
  __shared__ float4 sc[MAX_MODES][FX], sr[MAX_MODES][FY];
  const int i0 = blockIdx.x * FX, j0 = blockIdx.y * FY;
  const int nm = ms->n;
  const int tx = threadIdx.x & (FX - 1), ty = threadIdx.x / FX;  // 4 rows of threads
  const int i = i0 + tx;
  constexpr int R = FY / 4;  // rows per thread: ty, ty + 4, ...
  float ur[R], vr[R];
#pragma unroll
  for (int rr = 0; rr < R; rr++) {
    const int j = j0 + ty + 4 * rr;
    ur[rr] = vr[rr] = 0.f;
    if (i < nx && j < ny) {
      const size_t k = (size_t)j * nx + i;
      ur[rr] = u[k];
      vr[rr] = v[k];
    }
  }
  // every table load in flight before the first LDS store
  constexpr int SPT = (MAX_MODES * (FX + FY) + 255) / 256;
  float4 t[SPT];
#pragma unroll
  for (int k = 0; k < SPT; k++) {
    const int q = (int)threadIdx.x + k * 256;
    t[k] = float4{0.f, 0.f, 0.f, 0.f};
    if (q < nm * (FX + FY)) {
      const int m = q / (FX + FY), r = q - m * (FX + FY);
      if (r < FX) {
        if (i0 + r < nx) t[k] = col[(size_t)m * nx + i0 + r];
      } else if (j0 + r - FX < ny) {
        t[k] = row[(size_t)m * ny + j0 + r - FX];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < SPT; k++) {
    const int q = (int)threadIdx.x + k * 256;
    if (q < nm * (FX + FY)) {
      const int m = q / (FX + FY), r = q - m * (FX + FY);
      if (r < FX) sc[m][r] = t[k];
      else sr[m][r - FX] = t[k];
    }
  }
  __syncthreads();
  if (i >= nx) return;
  const float fdt = cf->fdt, drag = cf->drag;
  const bool after_drag = ms->after_drag != 0;
  float fu[R], fv[R];
#pragma unroll
  for (int rr = 0; rr < R; rr++) fu[rr] = fv[rr] = 0.f;
  for (int m = 0; m < nm; m++) {
    const float4 c = sc[m][tx];
#pragma unroll
    for (int rr = 0; rr < R; rr++) {
      const float4 w = sr[m][ty + 4 * rr];
      fu[rr] = fmaf(c.y, w.y, fmaf(c.x, w.x, fu[rr]));
      fv[rr] = fmaf(c.w, w.w, fmaf(c.z, w.z, fv[rr]));
    }
  }
#pragma unroll
  for (int rr = 0; rr < R; rr++) {
    const int j = j0 + ty + 4 * rr;
    if (j >= ny) break;
    const size_t k = (size_t)j * nx + i;
    if (after_drag) {
      u[k] = ur[rr] * drag + fu[rr] * fdt;
      v[k] = vr[rr] * drag + fv[rr] * fdt;
    } else {
      u[k] = (ur[rr] + fu[rr] * fdt) * drag;
      v[k] = (vr[rr] + fv[rr] * fdt) * drag;
    }
  }
}

// The same forcing on the matrix cores, for grids whose sides are multiples of 64: the
// field is a product F = A B with K = 2 (modes) <= 64, A[j][2m] = cos 2piB_m(j),
// A[j][2m + 1] = sin 2piB_m(j) from the row tables and B[2m][i], B[2m + 1][i] the column
// tables' coefficient-folded cosine and sine, for u and for v. Each operand value is split
// into a BF16 head and tail (x = h + l, h = bf16(x), l = bf16(x - h)) and a b is summed as
// ha hb + ha lb + la hb in FP32 (v_mfma_f32_16x16x32_bf16), ~2^-16 relative per term: the
// force to ~1e-5 relative, far below the FP32 rounding of the velocity it is added to.
// A block of four wavefronts covers 64 x 64 cells, each wavefront 16 rows by four 16 x 16
// output blocks; lane l of the MFMA supplies A[l % 16][8 (l / 16) + t], B[8 (l / 16) + t][l % 16]
// (t = 0..7) and receives D[4 (l / 16) + q][l % 16] (q = 0..3).
typedef __bf16 fk_bf16x8 __attribute__((ext_vector_type(8)));
typedef float fk_f32x4 __attribute__((ext_vector_type(4)));
constexpr int FMX = 64;  // cells per block side
__device__ __forceinline__ void fk_split(const float (&x)[8], fk_bf16x8& h, fk_bf16x8& l) {
#pragma unroll
  for (int t = 0; t < 8; t++) {
    const __bf16 hx = (__bf16)x[t];
    h[t] = hx;
    l[t] = (__bf16)(x[t] - (float)hx);
  }
}
__device__ __forceinline__ fk_f32x4 fk_mma3(fk_f32x4 acc, fk_bf16x8 ah, fk_bf16x8 al, fk_bf16x8 bh, fk_bf16x8 bl) {
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
}

__global__ void __launch_bounds__(256)
k_forcing_mma(float* __restrict__ u, float* __restrict__ v, int nx, int ny, const ModeSet* __restrict__ ms,
              const float4* __restrict__ col, const float4* __restrict__ row, const StepCoef* __restrict__ cf) {
  const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int i0 = (int)blockIdx.x * FMX, jw = (int)blockIdx.y * FMX + 16 * w;  // the wavefront's rows jw .. jw + 15
  const int nm = ms->n;
  // the field this lane updates: rows jw + 4 lg + q, columns i0 + 16 c + li
  float ur[4][4], vr[4][4];
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const size_t k = (size_t)(jw + 4 * lg + q) * nx + i0 + 16 * c + li;
      ur[c][q] = u[k];
      vr[c][q] = v[k];
    }
  // A fragments of row jw + li: modes 16 kk + 4 lg + 0..3, (cos, sin) each, for u and v
  fk_bf16x8 auh[2], aul[2], avh[2], avl[2];
#pragma unroll
  for (int kk = 0; kk < 2; kk++) {
    float xu[8], xv[8];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const int m = 16 * kk + 4 * lg + t;
      const float4 r = m < nm ? row[(size_t)m * ny + jw + li] : float4{0.f, 0.f, 0.f, 0.f};
      xu[2 * t] = r.x, xu[2 * t + 1] = r.y, xv[2 * t] = r.z, xv[2 * t + 1] = r.w;
    }
    fk_split(xu, auh[kk], aul[kk]);
    fk_split(xv, avh[kk], avl[kk]);
  }
  const float fdt = cf->fdt, drag = cf->drag;
  const bool after_drag = ms->after_drag != 0;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int i = i0 + 16 * c + li;
    fk_f32x4 fu = {0.f, 0.f, 0.f, 0.f}, fv = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < 2; kk++) {
      float xu[8], xv[8];
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const int m = 16 * kk + 4 * lg + t;
        const float4 b = m < nm ? col[(size_t)m * nx + i] : float4{0.f, 0.f, 0.f, 0.f};
        xu[2 * t] = b.x, xu[2 * t + 1] = b.y, xv[2 * t] = b.z, xv[2 * t + 1] = b.w;
      }
      fk_bf16x8 buh, bul, bvh, bvl;
      fk_split(xu, buh, bul);
      fk_split(xv, bvh, bvl);
      fu = fk_mma3(fu, auh[kk], aul[kk], buh, bul);
      fv = fk_mma3(fv, avh[kk], avl[kk], bvh, bvl);
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const size_t k = (size_t)(jw + 4 * lg + q) * nx + i;
      if (after_drag) {
        u[k] = ur[c][q] * drag + fu[q] * fdt;
        v[k] = vr[c][q] * drag + fv[q] * fdt;
      } else {
        u[k] = (ur[c][q] + fu[q] * fdt) * drag;
        v[k] = (vr[c][q] + fv[q] * fdt) * drag;
      }
    }
  }
}

// --------------------------------------------------- projection, FP32 paths

// Divergence at centres (src/core/grid.js:191-206; DIVERGENCE src/gpu/shaders.js:269-288).
__global__ void __launch_bounds__(BX * BY)
k_divergence(const float* __restrict__ u, const float* __restrict__ v, float* __restrict__ div, int nx, int ny,
             float idx, float idy) {
  int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
  if (i >= nx || j >= ny) return;
  size_t k = (size_t)j * nx + i;
  float du = u[(size_t)j * nx + nxt(i, nx)] - u[k];
  float dv = v[(size_t)nxt(j, ny) * nx + i] - v[k];
  div[k] = du * idx + dv * idy;
}

__global__ void k_fill_f32(float* __restrict__ a, size_t n, float val) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) a[k] = val;
}
__global__ void k_fill_f64(double* __restrict__ a, size_t n, double val) {
  size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) a[k] = val;
}

// One Jacobi sweep, ping-pong (JACOBI src/gpu/shaders.js:302-325).
__global__ void __launch_bounds__(BX * BY)
k_jacobi(const float* __restrict__ p, const float* __restrict__ rhs, float* __restrict__ q, int nx, int ny,
         float h2, float omega) {
  int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
  if (i >= nx || j >= ny) return;
  size_t k = (size_t)j * nx + i;
  float s = p[(size_t)j * nx + prv(i, nx)] + p[(size_t)j * nx + nxt(i, nx)] + p[(size_t)prv(j, ny) * nx + i] +
            p[(size_t)nxt(j, ny) * nx + i];
  float jac = (s - h2 * rhs[k]) * 0.25f;
  float cur = p[k];
  q[k] = (omega == 1.f) ? jac : fmaf(omega, jac - cur, cur);
}

// One colour of red-black Gauss-Seidel, in place (RED_BLACK_GS src/gpu/shaders.js:340-372).
// Cells of the other colour are untouched, which equals the reference's copy-through.
// Each thread owns the cell of its colour in a horizontal pair, so all lanes work.
// x_zero: the iterate is logically zero (its stored values are stale) -- used for
// the first pass of a correction that starts from zero (src/gpu/multigrid.js:306-307).
template <typename T>
__global__ void __launch_bounds__(BX * BY)
k_rb_colour(T* __restrict__ x, const T* __restrict__ rhs, int nx, int ny, float h2, float omega, int colour,
            int x_zero) {
  int ip = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
  int i = 2 * ip + ((j + colour) & 1);
  if (i >= nx || j >= ny) return;
  size_t k = (size_t)j * nx + i;
  float s = 0.f;
  if (!x_zero)
    s = (float)x[(size_t)j * nx + prv(i, nx)] + (float)x[(size_t)j * nx + nxt(i, nx)] +
        (float)x[(size_t)prv(j, ny) * nx + i] + (float)x[(size_t)nxt(j, ny) * nx + i];
  float gs = (s - h2 * (float)rhs[k]) * 0.25f;
  float cur = x_zero ? 0.f : (float)x[k];
  x[k] = (T)((omega == 1.f) ? gs : fmaf(omega, gs - cur, cur));
}

// u -= grad p with FP32 pressure (SUBTRACT_GRADIENT src/gpu/shaders.js:375-398).
__global__ void __launch_bounds__(BX * BY)
k_gradient32(const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ p,
             float* __restrict__ ou, float* __restrict__ ov, int nx, int ny, float idx, float idy) {
  int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
  if (i >= nx || j >= ny) return;
  size_t k = (size_t)j * nx + i;
  float pc = p[k];
  ou[k] = u[k] - (pc - p[(size_t)j * nx + prv(i, nx)]) * idx;
  ov[k] = v[k] - (pc - p[(size_t)prv(j, ny) * nx + i]) * idy;
}

// ------------------------------------------------------------ reductions

__device__ __forceinline__ double wave_max(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
  return x;
}
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

// Block-wide max of a non-negative double, stored to dst[block] (one plain store per
// block: a single-address atomic from thousands of blocks serialises at memory).
// NaN is turned into +inf by the callers because fmax drops NaN operands.
__device__ __forceinline__ double block_max(double x) {
  __shared__ double sm[16];
  int lane = threadIdx.x + threadIdx.y * blockDim.x;
  int nw = (blockDim.x * blockDim.y + 63) / 64;
  __syncthreads();  // a previous call in this kernel may still be reading sm
  x = wave_max(x);
  if ((lane & 63) == 0) sm[lane >> 6] = x;
  __syncthreads();
  double m = sm[0];
  for (int w = 1; w < nw; w++) m = fmax(m, sm[w]);
  return m;
}

__device__ __forceinline__ int flat_block() { return blockIdx.x + blockIdx.y * gridDim.x; }
__device__ __forceinline__ bool thread0() { return threadIdx.x == 0 && threadIdx.y == 0; }
__device__ __forceinline__ double nan_to_inf(double a) {
  return (a != a) ? __longlong_as_double(0x7ff0000000000000LL) : a;
}

// ----------------------------------------------- FP64 refinement residual

// r = div(u*) - L(p + e) in FP64, p' = p + e stored FP64, r stored FP32 as the
// multigrid right-hand side, the block's max|r| stored to rpart[block].
// e == nullptr means a zero correction, p == nullptr a zero initial pressure. The divergence and Laplacian follow
// src/core/grid.js:191-206 and src/cpu/projection.js:48-62.
template <typename S>
__global__ void __launch_bounds__(BX * BY)
k_residual64(const float* __restrict__ u, const float* __restrict__ v, const double* __restrict__ p,
             const float* __restrict__ e, double* __restrict__ pout, S* __restrict__ r32, int nx, int ny,
             double idx, double idy, double idx2, double idy2, double* __restrict__ rpart, Gate gate) {
  if (gate_closed(gate)) return;
  int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
  double amax = 0.0;
  if (i < nx && j < ny) {
    int im = prv(i, nx), ip = nxt(i, nx), jm = prv(j, ny), jp = nxt(j, ny);
    size_t k = (size_t)j * nx + i;
    size_t kl = (size_t)j * nx + im, kr = (size_t)j * nx + ip, kb = (size_t)jm * nx + i, kt = (size_t)jp * nx + i;
    double pc = 0, pl = 0, pr = 0, pb = 0, pt = 0;
    if (p) { pc = p[k]; pl = p[kl]; pr = p[kr]; pb = p[kb]; pt = p[kt]; }
    if (e) { pc += e[k]; pl += e[kl]; pr += e[kr]; pb += e[kb]; pt += e[kt]; }
    double lap = (pl - 2 * pc + pr) * idx2 + (pb - 2 * pc + pt) * idy2;
    double dv = ((double)u[kr] - (double)u[k]) * idx + ((double)v[kt] - (double)v[k]) * idy;
    double r = dv - lap;
    pout[k] = pc;
    r32[k] = st<S>((float)r);
    amax = nan_to_inf(fabs(r));
  }
  double m = block_max(amax);
  if (thread0()) {
    rpart[flat_block()] = m;
    gate_record(gate, m);
  }
}

// ------------------------------------------------------------- multigrid
// Cell-centred V-cycle of src/gpu/multigrid.js:277-330: red-black smoothing,
// full-weighting 2x2 restriction of the residual, bilinear prolongation added
// to the finer correction. The residual and the restriction are fused: each
// coarse cell evaluates the residual of its four children (RESIDUAL :48-66,
// RESTRICT :72-93) and writes their mean.

template <typename T>
__global__ void __launch_bounds__(BX * BY)
k_mg_rb(T* __restrict__ x, const T* __restrict__ rhs, int n, float h2, int colour, int x_zero) {
  int ip = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
  int i = 2 * ip + ((j + colour) & 1);
  if (i >= n || j >= n) return;
  size_t k = (size_t)j * n + i;
  float s = 0.f;
  if (!x_zero)
    s = (float)x[(size_t)j * n + prv(i, n)] + (float)x[(size_t)j * n + nxt(i, n)] +
        (float)x[(size_t)prv(j, n) * n + i] + (float)x[(size_t)nxt(j, n) * n + i];
  x[k] = (T)((s - h2 * (float)rhs[k]) * 0.25f);
}

template <typename T>
__global__ void __launch_bounds__(BX * BY)
k_mg_restrict(const T* __restrict__ x, const T* __restrict__ rhs, int n, float ih2, T* __restrict__ crhs,
              int x_zero) {
  int I = blockIdx.x * BX + threadIdx.x, J = blockIdx.y * BY + threadIdx.y;
  int cn = n >> 1;
  if (I >= cn || J >= cn) return;
  float s = 0.f;
#pragma unroll
  for (int b = 0; b < 2; b++)
#pragma unroll
    for (int a = 0; a < 2; a++) {
      int i = 2 * I + a, j = 2 * J + b;
      size_t k = (size_t)j * n + i;
      float lap = 0.f;
      if (!x_zero) {
        float xc = (float)x[k];
        lap = ((float)x[(size_t)j * n + prv(i, n)] + (float)x[(size_t)j * n + nxt(i, n)] +
               (float)x[(size_t)prv(j, n) * n + i] + (float)x[(size_t)nxt(j, n) * n + i] - 4.f * xc) * ih2;
      }
      s += (float)rhs[k] - lap;
    }
  crhs[(size_t)J * cn + I] = (T)(s * 0.25f);
}

template <typename T>
__global__ void __launch_bounds__(BX * BY)
k_mg_prolong(const T* __restrict__ xc, T* __restrict__ x, int n, int x_zero) {
  int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
  if (i >= n || j >= n) return;
  int cn = n >> 1;
  // fine centre (i+1/2) sits at coarse coordinate (i+1/2)/2 - 1/2 (PROLONG :99-131)
  int i0 = (i & 1) ? (i >> 1) : (i >> 1) - 1;
  int j0 = (j & 1) ? (j >> 1) : (j >> 1) - 1;
  float fx = (i & 1) ? 0.25f : 0.75f, fy = (j & 1) ? 0.25f : 0.75f;
  int x0 = wrapi(i0, cn), x1 = nxt(x0, cn), y0 = wrapi(j0, cn), y1 = nxt(y0, cn);
  float a = (float)xc[(size_t)y0 * cn + x0], b = (float)xc[(size_t)y0 * cn + x1];
  float c = (float)xc[(size_t)y1 * cn + x0], d = (float)xc[(size_t)y1 * cn + x1];
  float top = a * (1.f - fx) + b * fx, bot = c * (1.f - fx) + d * fx;
  float corr = top * (1.f - fy) + bot * fy;
  size_t k = (size_t)j * n + i;
  x[k] = (T)((x_zero ? 0.f : (float)x[k]) + corr);
}

// u = u* - grad p with FP64 pressure, stored FP32; the block maxima of |u|, |v|
// of the result go to part[2*block + {0,1}] for the CFL time step of the next step.
// One wavefront per BX x BY block: lane l takes 4 consecutive cells of row l / 16
// (16-byte loads and stores when nx % 4 == 0), so the maxima are wavefront
// reductions with no barrier. Launched with 64 threads on grid2(nx, ny).
// The pressure is chosen as FinalP says; block 0 also writes the step record.
__global__ void __launch_bounds__(64)
k_gradient64(const float* __restrict__ u, const float* __restrict__ v, const FinalP fp, float* __restrict__ ou,
             float* __restrict__ ov, int nx, int ny, double idx, double idy, double* __restrict__ part) {
  static_assert(BX == 64 && BY == 4, "16 lanes x 4 cells per row, 4 rows");
  int jc = 0, fin = 0;
  if (fp.k >= 0) {
    jc = fp.k;
    for (int q = 0; q < fp.k; q++)
      if (fp.flag[q] == 0) {
        jc = q;
        break;
      }
    fin = (fp.a0 + jc + 1) & 1;
  }
  const double* __restrict__ p = fp.pbuf[fin];
  double* __restrict__ pk = (fp.k >= 0 && fin != fp.keep) ? fp.pbuf[fp.keep] : nullptr;
  const int2 bk = xcd_block2();
  const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
  const int i0 = bk.x * BX + 4 * lx, j = bk.y * BY + ly;
  double mu = 0.0, mv = 0.0;
  if (j < ny && i0 < nx) {
    const size_t row = (size_t)j * nx, rowb = (size_t)prv(j, ny) * nx;
    if ((nx & 3) == 0) {
      const float4 uu = *(const float4*)(u + row + i0), vv = *(const float4*)(v + row + i0);
      const double2 c01 = *(const double2*)(p + row + i0), c23 = *(const double2*)(p + row + i0 + 2);
      const double2 b01 = *(const double2*)(p + rowb + i0), b23 = *(const double2*)(p + rowb + i0 + 2);
      const double pl = p[row + prv(i0, nx)];
      const float4 nu4{(float)((double)uu.x - (c01.x - pl) * idx), (float)((double)uu.y - (c01.y - c01.x) * idx),
                       (float)((double)uu.z - (c23.x - c01.y) * idx), (float)((double)uu.w - (c23.y - c23.x) * idx)};
      const float4 nv4{(float)((double)vv.x - (c01.x - b01.x) * idy), (float)((double)vv.y - (c01.y - b01.y) * idy),
                       (float)((double)vv.z - (c23.x - b23.x) * idy), (float)((double)vv.w - (c23.y - b23.y) * idy)};
      *(float4*)(ou + row + i0) = nu4;
      *(float4*)(ov + row + i0) = nv4;
      if (pk) {
        *(double2*)(pk + row + i0) = c01;
        *(double2*)(pk + row + i0 + 2) = c23;
      }
      mu = fmax(fmax(nan_to_inf(fabs((double)nu4.x)), nan_to_inf(fabs((double)nu4.y))),
                fmax(nan_to_inf(fabs((double)nu4.z)), nan_to_inf(fabs((double)nu4.w))));
      mv = fmax(fmax(nan_to_inf(fabs((double)nv4.x)), nan_to_inf(fabs((double)nv4.y))),
                fmax(nan_to_inf(fabs((double)nv4.z)), nan_to_inf(fabs((double)nv4.w))));
    } else {
      for (int c = 0; c < 4 && i0 + c < nx; c++) {
        const size_t k = row + i0 + c;
        const double pc = p[k];
        const float nu_ = (float)((double)u[k] - (pc - p[row + prv(i0 + c, nx)]) * idx);
        const float nv_ = (float)((double)v[k] - (pc - p[rowb + i0 + c]) * idy);
        ou[k] = nu_;
        ov[k] = nv_;
        if (pk) pk[k] = pc;
        mu = fmax(mu, nan_to_inf(fabs((double)nu_)));
        mv = fmax(mv, nan_to_inf(fabs((double)nv_)));
      }
    }
  }
  mu = wave_max(mu);
  mv = wave_max(mv);
  if (threadIdx.x == 0) {
    part[2 * flat_block()] = mu;
    part[2 * flat_block() + 1] = mv;
  }
  if (fp.ring && fp.k >= 0 && flat_block() == 0) {
    // the residual the last iteration run left: the row of iteration jc
    const double* rowp = fp.rpart + (size_t)jc * fp.cols;
    double m = 0.0;
    for (int q = threadIdx.x; q < fp.cols; q += 64) m = fmax(m, rowp[q]);
    m = wave_max(m);
    if (threadIdx.x == 0) {
      const int64_t st = fp.coef->step;
      StepRecord* r = fp.ring + (size_t)(st % fp.ring_len);
      r->dt = fp.coef->dt;
      r->residual = m;
      r->vcycles = jc;
      r->missed = (jc == fp.k && fp.flag[fp.k] != 0) ? 1 : 0;
      r->step = st;
    }
  }
}

// Pressure extrapolated in time, p_guess = 2 p_n - p_{n-1}, into `out`, and the
// history advanced (hist <- p_n); with no history yet (have_hist = 0) p_guess = p_n.
// Two cells per thread with 16-byte accesses (n even).
__global__ void __launch_bounds__(256)
k_p_extrapolate(const double* __restrict__ pn, double* __restrict__ hist, double* __restrict__ out, size_t n,
                int have_hist) {
  const size_t q = ((size_t)blockIdx.x * 256 + threadIdx.x) * 2;
  if (q >= n) return;
  const double2 a = *(const double2*)(pn + q);
  double2 g = a;
  if (have_hist) {
    const double2 b = *(const double2*)(hist + q);
    g = double2{2.0 * a.x - b.x, 2.0 * a.y - b.y};
  }
  *(double2*)(out + q) = g;
  *(double2*)(hist + q) = a;
}

// block maxima of |u|, |v| of a field (for a CFL step before any projection ran).
__global__ void __launch_bounds__(BX * BY)
k_maxabs(const float* __restrict__ u, const float* __restrict__ v, int nx, int ny, double* __restrict__ part) {
  int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
  double mu = 0.0, mv = 0.0;
  if (i < nx && j < ny) {
    size_t k = (size_t)j * nx + i;
    mu = nan_to_inf(fabs((double)u[k]));
    mv = nan_to_inf(fabs((double)v[k]));
  }
  double a = block_max(mu), b = block_max(mv);
  if (thread0()) {
    part[2 * flat_block()] = a;
    part[2 * flat_block() + 1] = b;
  }
}

__device__ __forceinline__ void fill_coef(StepCoef* cf, double dt, double dx, double dy, double nu, double drag) {
  cf->dt = dt;
  cf->cx = (float)(dt / dx);
  cf->cy = (float)(dt / dy);
  cf->da = (float)(nu * dt / (dx * dx));
  cf->db = (float)(nu * dt / (dy * dy));
  cf->fdt = (float)dt;
  cf->drag = drag > 0 ? (float)exp(-drag * dt) : 1.0f;
}

// What the first launch of a step sets up on the device (one workgroup of 1024):
//   dt: `value` itself (use_cfl = 0), or value / (max|u|/dx + max|v|/dy)
//       (src/core/grid.js:322-327) from the block maxima part[2*b + {0,1}] of the
//       last gradient (or maxabs) pass; then the step's coefficients and index
//   the step's start time t (and, for a step, its end time t + dt)
//   the step's forcing modes, copied from the argument into ms_dev (copy_modes)
//   the convergence flags 0 .. nflags-1 of the gated projection, cleared
__global__ void __launch_bounds__(1024)
k_step_head(const double* __restrict__ part, int nparts, int use_cfl, double value, double dx, double dy,
            double nu, double drag, double dt_max, int64_t step, int advance_t, StepCoef* __restrict__ cf,
            int* __restrict__ flags, int nflags, const ModeSet ms, ModeSet* __restrict__ ms_dev, int copy_modes) {
  __shared__ double su[16], sv[16];
  for (int q = threadIdx.x; q < nflags; q += 1024) flags[q] = 0;
  if (copy_modes) {
    const int* src = reinterpret_cast<const int*>(&ms);
    int* dst = reinterpret_cast<int*>(ms_dev);
    constexpr int W = (int)(sizeof(ModeSet) / sizeof(int));
    for (int q = threadIdx.x; q < W; q += 1024) dst[q] = src[q];
  }
  double dt = value;
  if (use_cfl) {
    // the block maxima as (|u|, |v|) pairs, four loads in flight per thread (a maximum
    // does not depend on the order it is taken in)
    double mu = 0.0, mv = 0.0;
    const double2* p2 = reinterpret_cast<const double2*>(part);
#pragma unroll 4
    for (int b = threadIdx.x; b < nparts; b += 1024) {
      const double2 m = p2[b];
      mu = fmax(mu, m.x);
      mv = fmax(mv, m.y);
    }
    mu = wave_max(mu);
    mv = wave_max(mv);
    if ((threadIdx.x & 63) == 0) {
      su[threadIdx.x >> 6] = mu;
      sv[threadIdx.x >> 6] = mv;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < 16; w++) {
        mu = fmax(mu, su[w]);
        mv = fmax(mv, sv[w]);
      }
      mu = fmax(mu, su[0]);
      mv = fmax(mv, sv[0]);
      const double rate = mu / dx + mv / dy;
      if (!(rate <= 1.7976931348623157e308)) dt = __longlong_as_double(0x7ff8000000000000LL);  // non-finite field
      else dt = rate > 0 ? fmin(value / rate, dt_max) : dt_max;
    }
  }
  if (threadIdx.x == 0) {
    const double t0 = cf->t_end;
    fill_coef(cf, dt, dx, dy, nu, drag);
    cf->step = step;
    cf->t = t0;
    if (advance_t) cf->t_end = t0 + dt;
  }
}

// Per-block FP64 diagnostics of the stored FP32 field (src/core/grid.js:290-319):
// part[b*DIAG_W + q] = {sum KE_c, max|div|, sum div^2, sum u, sum v, max|u|, max|v|,
// nonfinite, sum omega_c^2}, omega_c the centre average of the four corner vorticities
// (vorticityCorner / vorticityCenter, src/core/grid.js:228-257).
constexpr int DIAG_W = 9;
__device__ __forceinline__ bool diag_is_max(int q) { return q == 1 || q == 5 || q == 6; }

// The diagnostics' terms of cell (i, j).
__device__ __forceinline__ void diag_cell(const float* __restrict__ u, const float* __restrict__ v, int nx, int ny,
                                          int i, int j, double idx, double idy, double vals[DIAG_W]) {
  size_t k = (size_t)j * nx + i;
  const int ip = nxt(i, nx), im = prv(i, nx), jp = nxt(j, ny), jm = prv(j, ny);
  double uc = u[k], vc = v[k];
  double ur = u[(size_t)j * nx + ip], vt = v[(size_t)jp * nx + i];
  double cu = 0.5 * (uc + ur), cv = 0.5 * (vc + vt);
  double dv = (ur - uc) * idx + (vt - vc) * idy;
  double md = fabs(dv);
  if (md != md) md = __longlong_as_double(0x7ff0000000000000LL);
  // corner vorticity w(a, b) = (v(a,b) - v(a-1,b)) / dx - (u(a,b) - u(a,b-1)) / dy at the
  // corners (i,j), (i+1,j), (i,j+1), (i+1,j+1)
  auto U = [&](int a, int b) { return (double)u[(size_t)b * nx + a]; };
  auto V = [&](int a, int b) { return (double)v[(size_t)b * nx + a]; };
  const double w00 = (V(i, j) - V(im, j)) * idx - (U(i, j) - U(i, jm)) * idy;
  const double w10 = (V(ip, j) - V(i, j)) * idx - (U(ip, j) - U(ip, jm)) * idy;
  const double w01 = (V(i, jp) - V(im, jp)) * idx - (U(i, jp) - U(i, j)) * idy;
  const double w11 = (V(ip, jp) - V(i, jp)) * idx - (U(ip, jp) - U(ip, j)) * idy;
  const double wc = 0.25 * (w00 + w10 + w01 + w11);
  vals[0] = cu * cu + cv * cv;
  vals[1] = md;
  vals[2] = dv * dv;
  vals[3] = uc;
  vals[4] = vc;
  vals[5] = nan_to_inf(fabs(uc));
  vals[6] = nan_to_inf(fabs(vc));
  vals[7] = (isfinite(uc) ? 0 : 1) + (isfinite(vc) ? 0 : 1);
  vals[8] = wc * wc;
}

// The block's fold of its threads' terms (BX x BY threads) to part[blockslot * DIAG_W + q].
__device__ __forceinline__ void diag_block_store(double vals[DIAG_W], double* __restrict__ part, size_t blockslot) {
  __shared__ double sm[DIAG_W][4];
  int lane = threadIdx.x + threadIdx.y * BX;
#pragma unroll
  for (int q = 0; q < DIAG_W; q++) {
    double x = diag_is_max(q) ? wave_max(vals[q]) : wave_sum(vals[q]);
    if ((lane & 63) == 0) sm[q][lane >> 6] = x;
  }
  __syncthreads();
  if (lane < DIAG_W) {
    int q = lane;
    double x = sm[q][0];
    for (int w = 1; w < 4; w++) x = diag_is_max(q) ? fmax(x, sm[q][w]) : x + sm[q][w];
    part[blockslot * DIAG_W + q] = x;
  }
}

__global__ void __launch_bounds__(BX * BY)
k_diag(const float* __restrict__ u, const float* __restrict__ v, int nx, int ny, double idx, double idy,
       double* __restrict__ part) {
  int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
  double vals[DIAG_W] = {};
  if (i < nx && j < ny) diag_cell(u, v, nx, ny, i, j, idx, idy, vals);
  diag_block_store(vals, part, (size_t)blockIdx.y * gridDim.x + blockIdx.x);
}

// The same terms with a fixed number of workgroups (gridDim.x, 1D, BX x BY threads)
// striding over the BX x BY blocks of the grid, so the partials have a fixed size
// whatever the grid: every workgroup writes its slot, also when it found no block.
__global__ void __launch_bounds__(BX * BY)
k_diag_strided(const float* __restrict__ u, const float* __restrict__ v, int nx, int ny, double idx, double idy,
               double* __restrict__ part) {
  const int bx = (nx + BX - 1) / BX, nb = bx * ((ny + BY - 1) / BY);
  double acc[DIAG_W] = {};
  for (int b = blockIdx.x; b < nb; b += gridDim.x) {
    int i = (b % bx) * BX + threadIdx.x, j = (b / bx) * BY + threadIdx.y;
    if (i < nx && j < ny) {
      double vals[DIAG_W];
      diag_cell(u, v, nx, ny, i, j, idx, idy, vals);
#pragma unroll
      for (int q = 0; q < DIAG_W; q++) acc[q] = diag_is_max(q) ? fmax(acc[q], vals[q]) : acc[q] + vals[q];
    }
  }
  diag_block_store(acc, part, blockIdx.x);
}

// The folded sums and maxima as the diagnostics report them (one thread):
// {KE, max|div|, l2 div, momentum x, momentum y, max|u|, max|v|, non-finite faces, enstrophy}.
__global__ void k_diag_finish(const double* __restrict__ raw, double ncells, double dA, double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  out[0] = 0.5 * raw[0] / ncells;
  out[1] = raw[1];
  out[2] = sqrt(raw[2] / ncells);
  out[3] = raw[3] * dA;
  out[4] = raw[4] * dA;
  out[5] = raw[5];
  out[6] = raw[6];
  out[7] = raw[7];
  out[8] = 0.5 * raw[8] / ncells;
}

// Deterministic fold of the per-block partials (one workgroup, fixed order).
__global__ void __launch_bounds__(256)
k_diag_final(const double* __restrict__ part, int nblocks, double* __restrict__ out) {
  __shared__ double sm[DIAG_W][256];
  int t = threadIdx.x;
  double acc[DIAG_W] = {};
  for (int b = t; b < nblocks; b += 256)
    for (int q = 0; q < DIAG_W; q++) {
      double x = part[(size_t)b * DIAG_W + q];
      acc[q] = diag_is_max(q) ? fmax(acc[q], x) : acc[q] + x;
    }
  for (int q = 0; q < DIAG_W; q++) sm[q][t] = acc[q];
  __syncthreads();
  if (t < DIAG_W) {
    int q = t;
    double x = sm[q][0];
    for (int k = 1; k < 256; k++) x = diag_is_max(q) ? fmax(x, sm[q][k]) : x + sm[q][k];
    out[q] = x;
  }
}

// ------------------------------------------------------------ immersed disk

// Disk state for one step: centre, radius, rigid velocity, mask width in world
// units, cell area, domain size and the number of forcing passes.
struct Disk {
  double x, y, r, vx, vy, omega, width, dA, lx, ly, hx, hy;
  int passes;
};

__device__ __forceinline__ double wrap_delta(double d, double l) { return d - l * floor(d / l + 0.5); }

// chi = 0.5 (1 - tanh((|p - c| - r) / width)) (src/cpu/solid.js:101-107); 0 when
// it would fall below the 1e-6 cut of applyCoupling (src/cpu/solid.js:189-190).
// Beyond t = 8, chi < 1.2e-7, so tanh is not evaluated there.
__device__ __forceinline__ double disk_chi(const Disk& d, double ddx, double ddy) {
  double t = (hypot(ddx, ddy) - d.r) / d.width;
  if (t > 8.0) return 0.0;
  double chi = 0.5 * (1.0 - tanh(t));
  return chi < 1e-6 ? 0.0 : chi;
}

// Direct forcing u += chi (U_s - u) on this cell's u-face and v-face, `passes`
// times in FP64 registers, then rounded once to FP32 (src/cpu/solid.js:165-233;
// SOLID_COUPLE src/gpu/shaders.js:587-653). The fluid impulse (x, y) and the
// torque impulse about the centre, dA included, are reduced per block in FP64.
__global__ void __launch_bounds__(BX * BY)
k_solid(float* __restrict__ u, float* __restrict__ v, int nx, int ny, const Disk d, double* __restrict__ part) {
  int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
  double fx = 0, fy = 0, tq = 0;
  if (i < nx && j < ny) {
    size_t k = (size_t)j * nx + i;
    // u-face at (i hx, (j + 1/2) hy)
    double ux = wrap_delta(i * d.hx - d.x, d.lx), uy = wrap_delta((j + 0.5) * d.hy - d.y, d.ly);
    double cu = disk_chi(d, ux, uy);
    if (cu > 0) {
      double su = d.vx - d.omega * uy;
      double q = u[k], q0 = q;
      for (int p = 0; p < d.passes; p++) q += cu * (su - q);
      u[k] = (float)q;
      fx = (q - q0) * d.dA;
      tq = -uy * (q - q0) * d.dA;
    }
    // v-face at ((i + 1/2) hx, j hy)
    double vx = wrap_delta((i + 0.5) * d.hx - d.x, d.lx), vy = wrap_delta(j * d.hy - d.y, d.ly);
    double cv = disk_chi(d, vx, vy);
    if (cv > 0) {
      double sv = d.vy + d.omega * vx;
      double q = v[k], q0 = q;
      for (int p = 0; p < d.passes; p++) q += cv * (sv - q);
      v[k] = (float)q;
      fy = (q - q0) * d.dA;
      tq += vx * (q - q0) * d.dA;
    }
  }
  __shared__ double sm[3][4];
  int lane = threadIdx.x + threadIdx.y * BX;
  double a = wave_sum(fx), b = wave_sum(fy), c = wave_sum(tq);
  if ((lane & 63) == 0) {
    sm[0][lane >> 6] = a;
    sm[1][lane >> 6] = b;
    sm[2][lane >> 6] = c;
  }
  __syncthreads();
  if (lane < 3) {
    double x = sm[lane][0] + sm[lane][1] + sm[lane][2] + sm[lane][3];
    part[(size_t)flat_block() * 3 + lane] = x;
  }
}

// Fixed-order fold of k_solid's per-block partials into out[3] (one workgroup).
__global__ void __launch_bounds__(256)
k_solid_final(const double* __restrict__ part, int nblocks, double* __restrict__ out) {
  __shared__ double sm[3][256];
  int t = threadIdx.x;
  double acc[3] = {0, 0, 0};
  for (int b = t; b < nblocks; b += 256)
    for (int q = 0; q < 3; q++) acc[q] += part[(size_t)b * 3 + q];
  for (int q = 0; q < 3; q++) sm[q][t] = acc[q];
  __syncthreads();
  if (t < 3) {
    double x = 0;
    for (int k = 0; k < 256; k++) x += sm[t][k];
    out[t] = x;
  }
}

}  // namespace fsk
