// Semi-Lagrangian / MacCormack advection (VelocityGrid.warp semantics,
// data_structures.py:1690-1737) and its exact adjoint on gfx950.
//
// The reference composes a warp from 2-4 SampleGridLut launches (each with a
// device sync), a 4-channel LUT materialisation and ~6 TF elementwise ops.
// Here a MacCormack step is two kernels:
//   A  sl(x)  = S(rho, x + ts*v(x))                        (gather, 8 taps)
//   B  back   = S(sl, x - ts*v(x)); corr = sl + (rho - back)/2
//      MC_SMOOTH: min/max of rho's 8 corners at x + ts*v revert corr to sl
//      NEGATIVE : max(., 0)
// Velocity is read centered or MAC-staggered; the face average is folded
// into the loads (VelocityGrid.centered, data_structures.py:1634-1653), so
// no centered copy is materialised.
//
// Exact backward (B1, B2, optional face-gather B3):
//   B1  g_sl(x) += g;  if not reverted: g_rho(x) += g/2, scatter -g/2 into
//       g_sl at x - ts*v, g_v(x) -= ts * <grad S(sl), -g/2>
//   B2  scatter g_sl into g_rho at x + ts*v, g_v(x) += ts * <grad S(rho), g_sl>
//   B3  staggered: g_face = (g_v(left cell) + g_v(right cell)) / 2
// For 1-2 channels the scatters of B1 / B2 accumulate in LDS tiles per brick
// of cells and are summed by the consumer (no global atomics unless a
// displacement leaves the tile halo); the reference-normalised gradient and
// 3-4 channels scatter with hardware fp32 atomics.
#include <algorithm>

#include "ngt_host.hpp"

namespace ngt {
namespace {

constexpr int kBX = 64, kBY = 4;

struct AdvArgs {
  Dims d;
  int n_data, n_vel, separate;
  float ts;
  const float* vx;
  const float* vy;
  const float* vz;
  int staggered;
};

NGT_DEVICE void pair_indices(const AdvArgs& a, int pair, int& n, int& v) {
  if (a.separate) { n = pair / a.n_vel; v = pair % a.n_vel; }
  else { n = pair; v = pair; }
}

NGT_DEVICE void load_velocity(const AdvArgs& a, int v, int x, int y, int z, float& ux, float& uy, float& uz) {
  const Dims& d = a.d;
  if (!a.staggered) {
    const float* p = a.vx + (static_cast<int64_t>(v) * d.x * d.y * d.z + cell_offset(d, x, y, z)) * 3;
    ux = p[0]; uy = p[1]; uz = p[2];
  } else {
    const int64_t sx = static_cast<int64_t>(v) * d.z * d.y * (d.x + 1);
    const int64_t rowx = sx + (static_cast<int64_t>(z) * d.y + y) * (d.x + 1) + x;
    ux = (a.vx[rowx + 1] + a.vx[rowx]) * 0.5f;
    const int64_t sy = static_cast<int64_t>(v) * d.z * (d.y + 1) * d.x;
    const int64_t rowy = sy + (static_cast<int64_t>(z) * (d.y + 1) + y) * d.x + x;
    uy = (a.vy[rowy + d.x] + a.vy[rowy]) * 0.5f;
    const int64_t sz = static_cast<int64_t>(v) * (d.z + 1) * d.y * d.x;
    const int64_t rowz = sz + (static_cast<int64_t>(z) * d.y + y) * d.x + x;
    uz = (a.vz[rowz + static_cast<int64_t>(d.y) * d.x] + a.vz[rowz]) * 0.5f;
  }
}

struct ThreadCell {
  int x, y, z, pair;
  bool ok;
};

NGT_DEVICE ThreadCell thread_cell(const AdvArgs& a, const VBlock& vb) {
  ThreadCell t;
  t.x = vb.x * kBX + threadIdx.x;
  t.y = vb.y * kBY + threadIdx.y;
  t.z = vb.z % a.d.z;
  t.pair = vb.z / a.d.z;
  t.ok = t.x < a.d.x && t.y < a.d.y;
  return t;
}

// A: semi-Lagrangian step (also the whole order-1 warp).
template <int C, int BM, bool NEG>
NGT_DEVICE void advect_sl_kernel_body(const VBlock& vb, const AdvArgs& a, const float* __restrict__ rho,
                                                        float* __restrict__ out) {
  const ThreadCell t = thread_cell(a, vb);
  if (!t.ok) return;
  int n, v;
  pair_indices(a, t.pair, n, v);
  float ux, uy, uz;
  load_velocity(a, v, t.x, t.y, t.z, ux, uy, uz);
  const int64_t cells = static_cast<int64_t>(a.d.x) * a.d.y * a.d.z;
  const float px = static_cast<float>(t.x) + a.ts * ux;
  const float py = static_cast<float>(t.y) + a.ts * uy;
  const float pz = static_cast<float>(t.z) + a.ts * uz;
  Vec<C> r = read_linear<C, BM>(rho + n * cells * C, a.d, px, py, pz);
  if constexpr (NEG) {
#pragma unroll
    for (int i = 0; i < C; ++i) r.v[i] = fmaxf(r.v[i], 0.f);
  }
  vstore<C>(out + (t.pair * cells + cell_offset(a.d, t.x, t.y, t.z)) * C, r);
}

template <int C, int BM, bool NEG>
__global__ __launch_bounds__(256) void advect_sl_kernel(AdvArgs a, const float* __restrict__ rho,
                                                        float* __restrict__ out) {
  advect_sl_kernel_body<C, BM, NEG>(VBlock{static_cast<int>(blockIdx.x), static_cast<int>(blockIdx.y), static_cast<int>(blockIdx.z)}, a, rho, out);
}

// B: MacCormack correction + clamp.
template <int C, int BM, int CLAMPM>
NGT_DEVICE void advect_mc_kernel_body(const VBlock& vb, const AdvArgs& a, const float* __restrict__ rho,
                                                        const float* __restrict__ sl, float* __restrict__ out) {
  const ThreadCell t = thread_cell(a, vb);
  if (!t.ok) return;
  int n, v;
  pair_indices(a, t.pair, n, v);
  float ux, uy, uz;
  load_velocity(a, v, t.x, t.y, t.z, ux, uy, uz);
  const int64_t cells = static_cast<int64_t>(a.d.x) * a.d.y * a.d.z;
  const int64_t cell = cell_offset(a.d, t.x, t.y, t.z);
  const float* src = rho + n * cells * C;
  const float* slp = sl + t.pair * cells * C;
  const float fx = static_cast<float>(t.x), fy = static_cast<float>(t.y), fz = static_cast<float>(t.z);
  const float tx = a.ts * ux, ty = a.ts * uy, tz = a.ts * uz;
  const Vec<C> fwd = vload<C>(slp + cell * C);
  const Vec<C> back = read_linear<C, BM>(slp, a.d, fx + (-tx), fy + (-ty), fz + (-tz));
  const Vec<C> cur = vload<C>(src + cell * C);
  Vec<C> r;
#pragma unroll
  for (int i = 0; i < C; ++i) r.v[i] = fwd.v[i] + 0.5f * (cur.v[i] - back.v[i]);
  if constexpr (CLAMPM == NGT_CLAMP_MC_SMOOTH) {
    Vec<C> mn, mx;
    read_minmax2<C, BM>(src, a.d, fx + tx, fy + ty, fz + tz, mn, mx);
#pragma unroll
    for (int i = 0; i < C; ++i)
      if (r.v[i] < mn.v[i] || r.v[i] > mx.v[i]) r.v[i] = fwd.v[i];
  }
  if constexpr (CLAMPM == NGT_CLAMP_NEGATIVE) {
#pragma unroll
    for (int i = 0; i < C; ++i) r.v[i] = fmaxf(r.v[i], 0.f);
  }
  vstore<C>(out + (t.pair * cells + cell) * C, r);
}

template <int C, int BM, int CLAMPM>
__global__ __launch_bounds__(256) void advect_mc_kernel(AdvArgs a, const float* __restrict__ rho,
                                                        const float* __restrict__ sl, float* __restrict__ out) {
  advect_mc_kernel_body<C, BM, CLAMPM>(VBlock{static_cast<int>(blockIdx.x), static_cast<int>(blockIdx.y), static_cast<int>(blockIdx.z)}, a, rho, sl, out);
}

// B1 (order 2): route g to the SL field and the direct rho term; scatter the
// back-step adjoint into g_sl.  REF: scatter into tmp with hit counts.
template <int C, int BM, int CLAMPM, bool REF>
__global__ __launch_bounds__(256) void advect_mc_bwd_kernel(AdvArgs a, const float* __restrict__ rho,
                                                            const float* __restrict__ sl,
                                                            const float* __restrict__ g_out,
                                                            float* __restrict__ g_sl, float* __restrict__ g_rho,
                                                            float* __restrict__ g_vc, float* __restrict__ tmp,
                                                            uint32_t* __restrict__ hits) {
  const ThreadCell t = thread_cell(a, VBlock{static_cast<int>(blockIdx.x), static_cast<int>(blockIdx.y), static_cast<int>(blockIdx.z)});
  if (!t.ok) return;
  int n, v;
  pair_indices(a, t.pair, n, v);
  float ux, uy, uz;
  load_velocity(a, v, t.x, t.y, t.z, ux, uy, uz);
  const int64_t cells = static_cast<int64_t>(a.d.x) * a.d.y * a.d.z;
  const int64_t cell = cell_offset(a.d, t.x, t.y, t.z);
  const float* src = rho + n * cells * C;
  const float* slp = sl + t.pair * cells * C;
  const float fx = static_cast<float>(t.x), fy = static_cast<float>(t.y), fz = static_cast<float>(t.z);
  const float tx = a.ts * ux, ty = a.ts * uy, tz = a.ts * uz;
  const float bx = fx + (-tx), by = fy + (-ty), bz = fz + (-tz);
  Vec<C> g = vload<C>(g_out + (t.pair * cells + cell) * C);
  const Vec<C> fwd = vload<C>(slp + cell * C);
  Vec<C> dx, dy, dz;
  const Vec<C> back = read_linear_gradient<C, BM>(slp, a.d, bx, by, bz, dx, dy, dz);
  const Vec<C> cur = vload<C>(src + cell * C);
  Vec<C> corr;
#pragma unroll
  for (int i = 0; i < C; ++i) corr.v[i] = fwd.v[i] + 0.5f * (cur.v[i] - back.v[i]);
  Vec<C> keep = vfill<C>(1.f);  // 1 where the corrected value is kept
  if constexpr (CLAMPM == NGT_CLAMP_MC_SMOOTH) {
    Vec<C> mn, mx;
    read_minmax2<C, BM>(src, a.d, fx + tx, fy + ty, fz + tz, mn, mx);
#pragma unroll
    for (int i = 0; i < C; ++i)
      if (corr.v[i] < mn.v[i] || corr.v[i] > mx.v[i]) keep.v[i] = 0.f;
  }
  if constexpr (CLAMPM == NGT_CLAMP_NEGATIVE) {
#pragma unroll
    for (int i = 0; i < C; ++i)
      if (!(corr.v[i] >= 0.f)) g.v[i] = 0.f;  // tf.maximum passes the gradient where x >= 0
  }
  float* gsl = g_sl + (t.pair * cells) * C;
  Vec<C> gback;
#pragma unroll
  for (int i = 0; i < C; ++i) {
    atomic_add(gsl + cell * C + i, g.v[i]);
    atomic_add(g_rho + (n * cells + cell) * C + i, 0.5f * keep.v[i] * g.v[i]);
    gback.v[i] = -0.5f * keep.v[i] * g.v[i];
  }
  if constexpr (REF) {
    scatter_linear<C, BM, true>(tmp + (t.pair * cells) * C, a.d, bx, by, bz, gback, hits + t.pair * cells);
  } else {
    scatter_linear<C, BM, false>(gsl, a.d, bx, by, bz, gback);
  }
  float* gv = g_vc + (static_cast<int64_t>(v) * cells + cell) * 3;
  atomic_add(gv + 0, -a.ts * vdot(gback, dx));
  atomic_add(gv + 1, -a.ts * vdot(gback, dy));
  atomic_add(gv + 2, -a.ts * vdot(gback, dz));
}

// B2: adjoint of the semi-Lagrangian read.  For order 1, g_sl is g_out with
// the NEGATIVE mask applied on the fly (NEG).
template <int C, int BM, bool NEG, bool REF>
__global__ __launch_bounds__(256) void advect_sl_bwd_kernel(AdvArgs a, const float* __restrict__ rho,
                                                            const float* __restrict__ g_sl,
                                                            const float* __restrict__ sl_out,
                                                            float* __restrict__ g_rho, float* __restrict__ g_vc,
                                                            float* __restrict__ tmp, uint32_t* __restrict__ hits) {
  const ThreadCell t = thread_cell(a, VBlock{static_cast<int>(blockIdx.x), static_cast<int>(blockIdx.y), static_cast<int>(blockIdx.z)});
  if (!t.ok) return;
  int n, v;
  pair_indices(a, t.pair, n, v);
  float ux, uy, uz;
  load_velocity(a, v, t.x, t.y, t.z, ux, uy, uz);
  const int64_t cells = static_cast<int64_t>(a.d.x) * a.d.y * a.d.z;
  const int64_t cell = cell_offset(a.d, t.x, t.y, t.z);
  const float* src = rho + n * cells * C;
  const float px = static_cast<float>(t.x) + a.ts * ux;
  const float py = static_cast<float>(t.y) + a.ts * uy;
  const float pz = static_cast<float>(t.z) + a.ts * uz;
  Vec<C> g = vload<C>(g_sl + (t.pair * cells + cell) * C);
  if constexpr (NEG) {
    const Vec<C> r = read_linear<C, BM>(src, a.d, px, py, pz);
#pragma unroll
    for (int i = 0; i < C; ++i)
      if (!(r.v[i] >= 0.f)) g.v[i] = 0.f;
  }
  (void)sl_out;
  if constexpr (REF) {
    scatter_linear<C, BM, true>(tmp + (t.pair * cells) * C, a.d, px, py, pz, g, hits + t.pair * cells);
  } else {
    scatter_linear<C, BM, false>(g_rho + n * cells * C, a.d, px, py, pz, g);
  }
  Vec<C> dx, dy, dz;
  read_gradient<C, BM>(src, a.d, px, py, pz, dx, dy, dz);
  float* gv = g_vc + (static_cast<int64_t>(v) * cells + cell) * 3;
  atomic_add(gv + 0, a.ts * vdot(g, dx));
  atomic_add(gv + 1, a.ts * vdot(g, dy));
  atomic_add(gv + 2, a.ts * vdot(g, dz));
}

// dst(pair or n) += tmp / hits (reference count normalisation), one cell per
// work-item; dst_index maps the pair slot to the destination slot.
template <int C>
__global__ __launch_bounds__(256) void add_normalized_kernel(AdvArgs a, const float* __restrict__ tmp,
                                                             const uint32_t* __restrict__ hits,
                                                             float* __restrict__ dst, int to_data) {
  const ThreadCell t = thread_cell(a, VBlock{static_cast<int>(blockIdx.x), static_cast<int>(blockIdx.y), static_cast<int>(blockIdx.z)});
  if (!t.ok) return;
  int n, v;
  pair_indices(a, t.pair, n, v);
  const int64_t cells = static_cast<int64_t>(a.d.x) * a.d.y * a.d.z;
  const int64_t cell = cell_offset(a.d, t.x, t.y, t.z);
  const uint32_t h = hits[t.pair * cells + cell];
  if (h == 0u) return;
  const float w = 1.0f / static_cast<float>(h);
  const int slot = to_data ? n : t.pair;
#pragma unroll
  for (int i = 0; i < C; ++i)
    atomic_add(dst + (slot * cells + cell) * C + i, tmp[(t.pair * cells + cell) * C + i] * w);
}

// B3: centered-velocity gradient -> MAC face gradients (adjoint of the face
// average).  A work-group covers 64 x 4 (x, y) columns of one velocity over
// kFaceZ planes; each work-item walks its column in z, so g_vc is read once:
// the z - 1 term is the previous plane's value kept in a register, the x - 1
// and y - 1 terms hit the lines the neighbouring work-items just read.  The
// last cell of an axis also writes the closing face (x = d.x, y = d.y,
// z = d.z), which has only the left cell.
constexpr int kFaceZ = 8;

inline int64_t face_tiles(const Dims& d) {
  return static_cast<int64_t>((d.x + 63) / 64) * ((d.y + 3) / 4) * ((d.z + kFaceZ - 1) / kFaceZ);
}

NGT_DEVICE void faces_from_centered_grad_kernel_body(const VBlock& vb, Dims d, int n_vel, const float* __restrict__ g_vc,
                                                     float* __restrict__ gx, float* __restrict__ gy,
                                                     float* __restrict__ gz) {
  const int ntx = (d.x + 63) / 64, nty = (d.y + 3) / 4, nzc = (d.z + kFaceZ - 1) / kFaceZ;
  const int per_v = ntx * nty * nzc;
  const int v = vb.x / per_v;
  const int t = vb.x - v * per_v;
  const int x = (t % ntx) * 64 + static_cast<int>(threadIdx.x & 63u);
  const int y = ((t / ntx) % nty) * 4 + static_cast<int>(threadIdx.x >> 6);
  const int z0 = (t / (ntx * nty)) * kFaceZ, z1 = min(z0 + kFaceZ, d.z);
  if (v >= n_vel || x >= d.x || y >= d.y) return;
  const int64_t nx = static_cast<int64_t>(d.z) * d.y * (d.x + 1);
  const int64_t ny = static_cast<int64_t>(d.z) * (d.y + 1) * d.x;
  const int64_t nz = static_cast<int64_t>(d.z + 1) * d.y * d.x;
  const int64_t cells = static_cast<int64_t>(d.x) * d.y * d.z;
  const float* g = g_vc + static_cast<int64_t>(v) * cells * 3;
  float* ox = gx + static_cast<int64_t>(v) * nx;
  float* oy = gy + static_cast<int64_t>(v) * ny;
  float* oz = gz + static_cast<int64_t>(v) * nz;
  float prev = z0 > 0 ? g[cell_offset(d, x, y, z0 - 1) * 3 + 2] : 0.f;
  for (int z = z0; z < z1; ++z) {
    const int64_t c = cell_offset(d, x, y, z);
    const float g0 = g[c * 3 + 0], g1 = g[c * 3 + 1], g2 = g[c * 3 + 2];
    const float l0 = x > 0 ? g[(c - 1) * 3 + 0] : 0.f;
    const float l1 = y > 0 ? g[(c - d.x) * 3 + 1] : 0.f;
    const int64_t fx = (static_cast<int64_t>(z) * d.y + y) * (d.x + 1) + x;
    ox[fx] = 0.5f * (l0 + g0);
    if (x == d.x - 1) ox[fx + 1] = 0.5f * g0;
    const int64_t fy = (static_cast<int64_t>(z) * (d.y + 1) + y) * d.x + x;
    oy[fy] = 0.5f * (l1 + g1);
    if (y == d.y - 1) oy[fy + d.x] = 0.5f * g1;
    oz[c] = 0.5f * (prev + g2);
    prev = g2;
  }
  if (z1 == d.z) oz[cells + static_cast<int64_t>(y) * d.x + x] = 0.5f * prev;
}

__global__ __launch_bounds__(256) void faces_from_centered_grad_kernel(Dims d, int n_vel, const float* __restrict__ g_vc,
                                                                      float* __restrict__ gx, float* __restrict__ gy,
                                                                      float* __restrict__ gz) {
  faces_from_centered_grad_kernel_body(VBlock{static_cast<int>(blockIdx.x), static_cast<int>(blockIdx.y), static_cast<int>(blockIdx.z)}, d, n_vel, g_vc, gx, gy, gz);
}

// ---------------------------------------------------------------------------
// Brick-tiled exact backward (C <= 2).  A work-group owns a brick of
// kBrickX x kBrickY x kBrickZ cells and an LDS tile of the brick grown by
// kHalo cells on every side.  The trilinear scatter of the brick's samples
// lands in the tile whenever the displacement s = ts*v per axis satisfies
// -kHalo < s <= kHalo (the corners floor(x - s), floor(x - s) + 1 then lie in
// [x - kHalo, x + kHalo]), i.e. for displacements of up to 2 cells per step;
// rarer corners further out (faster flows, WRAP across the domain edge) go
// to an overflow field with global fp32 atomics.  Each tile is stored whole
// to a per-brick partial buffer with plain coalesced stores, and the
// consumer gathers, for every cell, the (at most 2 per axis) tiles that
// cover it in a fixed order: no global atomics in the common case.  (Within
// a tile the LDS float atomics add in no fixed order.)  The halo sets the
// partial-buffer traffic: tile cells per brick cell 2.11 with kHalo = 2,
// 2.86 with 3.  16 cells in z (rather than 8) shorten that ratio and give
// the column walks of the scatter more shared corner planes; 32 was slower.
constexpr int kBrickX = 32, kBrickY = 8, kBrickZ = 16, kHalo = 2;
// a work-group is kBrickX x kBrickY x kBrickZT work-items; each walks
// kBrickZ / kBrickZT cells of its (x, y) column
constexpr int kBrickZT = 2, kBrickZPer = kBrickZ / kBrickZT;
constexpr int kBrickThreads = kBrickX * kBrickY * kBrickZT;
constexpr int kTileX = kBrickX + 2 * kHalo, kTileY = kBrickY + 2 * kHalo, kTileZ = kBrickZ + 2 * kHalo;
constexpr int kTileCells = kTileX * kTileY * kTileZ;

struct BrickGrid {
  int nx, ny, nz;
  __host__ __device__ int count() const { return nx * ny * nz; }
};

inline BrickGrid brick_grid(const Dims& d) {
  return BrickGrid{static_cast<int>(blocks_for(d.x, kBrickX)), static_cast<int>(blocks_for(d.y, kBrickY)),
                   static_cast<int>(blocks_for(d.z, kBrickZ))};
}

// The upper-z corner plane of the last in-tile scatter of a z-column walk,
// held in registers: the next cell of the column usually has these four
// corners as its lower plane, and the two contributions then reach the LDS
// tile with one atomic add per corner instead of two.
template <int C>
struct HeldPlane {
  int base = -1;  // tile slot of corner (x0, y0), -1 = nothing held
  int dx = 0;     // slot step from x0 to x1: 0 or 1, or across the tile where WRAP wraps x1
  int dy = 0;     // slot step from y0 to y1: 0 or kTileX, or across the tile under WRAP
  Vec<C> v[4];    // corners (x0,y0), (x1,y0), (x0,y1), (x1,y1)
};

template <int C>
NGT_DEVICE void plane_add(float* tile, int base, int dx, int dy, const Vec<C> (&v)[4]) {
  const int off[4] = {0, dx, dy, dx + dy};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    float* t = tile + (base + off[k]) * C;
#pragma unroll
    for (int i = 0; i < C; ++i) lds_add(t + i, v[k].v[i]);
  }
}

// Value of the previous / next lane of the 16-lane DPP row (row_shr:1 /
// row_shl:1); 0 where that lane is outside the row or disabled.
NGT_DEVICE int from_prev_lane(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false); }
NGT_DEVICE int from_next_lane(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x101, 0xF, 0xF, false); }
NGT_DEVICE float from_prev_lane(float v) {
  return __builtin_bit_cast(float, from_prev_lane(__builtin_bit_cast(int, v)));
}
// Value of lane (l + 32) % 64 (v_permlane32_swap of the wave's two halves).
NGT_DEVICE int from_other_half(int v) {
  const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
  return __lane_id() < 32 ? static_cast<int>(r[1]) : static_cast<int>(r[0]);
}
NGT_DEVICE float from_other_half(float v) {
  return __builtin_bit_cast(float, from_other_half(__builtin_bit_cast(int, v)));
}

// plane_add for the lanes with `on` set, with the corners that neighbouring
// lanes share added once.  A wave holds two brick rows, y (lanes 0-31) and
// y + 1 (lanes 32-63), of consecutive x cells, so usually
//  - a lane's x1 column is the x0 column of the next lane of its DPP row: it
//    hands that column over;
//  - then a lane of row y has as its y1 row the y0 row of lane + 32, which
//    takes it where it still adds those slots itself.
// Every lane of the calling loop executes it (the exchanges read the other
// lanes' registers); a lane whose partner is outside the row or disabled
// sees 0 and keeps its own adds.
template <int C>
NGT_DEVICE void plane_add_row_merged(float* tile, bool on, int base, int dx, int dy, Vec<C> (&v)[4]) {
  static_assert(kBrickX == 32 && kBrickY % 2 == 0, "the two halves of a wave are consecutive brick rows");
  // keys carry the exact slot steps (WRAP can make them anything across the
  // tile): x column = (first slot, y row step), y row = (first slot, x step);
  // 0 = nothing offered
  static_assert(kTileX <= 64 && kTileY <= 16 && kTileCells < (1 << 18), "merge key ranges");
  const int ry = dy / kTileX + 16;  // 5 bits
  const int rx = dx + 64;           // 7 bits
  // x: my x1 column (slots base + dx, base + dx + dy) to the next lane
  const int give_key = on ? ((base + dx + 1) << 5) | ry : 0;
  const bool take_x = on && from_prev_lane(give_key) == (((base + 1) << 5) | ry);
#pragma unroll
  for (int i = 0; i < C; ++i) {
    const float r0 = from_prev_lane(v[1].v[i]), r1 = from_prev_lane(v[3].v[i]);
    if (take_x) {
      v[0].v[i] += r0;
      v[2].v[i] += r1;
    }
  }
  const bool own1 = !(from_next_lane(take_x ? 1 : 0) != 0);  // still adds its x1 column
  // y: the y1 row (slots base + dy, base + dy + dx) of a lane of row y to lane + 32
  const bool upper = __lane_id() >= 32;
  const int ykey = (on && !upper) ? ((((base + dy + 1) << 7) | rx) << 1) | (own1 ? 1 : 0) : 0;
  const int got = from_other_half(ykey);
  const bool take_y = on && upper && (got >> 1) == (((base + 1) << 7) | rx);
  const bool take_y3 = take_y && (got & 1) && own1;  // the x1 slot of the row, where both still own it
#pragma unroll
  for (int i = 0; i < C; ++i) {
    const float r2 = from_other_half(v[2].v[i]), r3 = from_other_half(v[3].v[i]);
    if (take_y) v[0].v[i] += r2;
    if (take_y3) v[1].v[i] += r3;
  }
  const int back = from_other_half((take_y ? 1 : 0) | (take_y3 ? 2 : 0));
  const bool own2 = upper || !(back & 1);
  const bool own3 = own1 && (upper || !(back & 2));
  if (on) {
    const int off[4] = {0, dx, dy, dx + dy};
    const bool own[4] = {true, own1, own2, own3};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (!own[k]) continue;
      float* t = tile + (base + off[k]) * C;
#pragma unroll
      for (int i = 0; i < C; ++i) lds_add(t + i, v[k].v[i]);
    }
  }
}

// Adds what `held` holds to the tile and empties it.  Executed by every lane
// of the walk at its end.
template <int C>
NGT_DEVICE void flush_plane(float* tile, HeldPlane<C>& held) {
  plane_add_row_merged<C>(tile, held.base >= 0, held.base, held.dx, held.dy, held.v);
  held.base = -1;
}

// Scatter val * w to the 8 corners of q: into the LDS tile (origin o*) when
// the corner lies in it, else atomically into `ovf` (a full field).  Called
// by every lane of a z-column walk for its consecutive cells (`on` false for
// a lane that scatters nothing), with one `held` plane per walk that
// flush_plane empties at the end of the walk.
template <int C, int BM>
NGT_DEVICE void scatter_linear_tile(float* tile, HeldPlane<C>& held, bool on, int ox, int oy, int oz,
                                    float* __restrict__ ovf, const Dims& d, float qx, float qy, float qz,
                                    const Vec<C>& val) {
  const Corners c = make_corners<BM>(d, qx, qy, qz);
  bool inside = on;
  if constexpr (BM == BORDER) inside = on && c.inside;
  const float fw[2][3] = {{1.f - c.wx, 1.f - c.wy, 1.f - c.wz}, {c.wx, c.wy, c.wz}};
  const int lx0 = c.x0 - ox, lx1 = c.x1 - ox, ly0 = c.y0 - oy, ly1 = c.y1 - oy, lz0 = c.z0 - oz, lz1 = c.z1 - oz;
  const bool in_tile = inside && static_cast<unsigned>(lx0) < static_cast<unsigned>(kTileX) &&
                       static_cast<unsigned>(lx1) < static_cast<unsigned>(kTileX) &&
                       static_cast<unsigned>(ly0) < static_cast<unsigned>(kTileY) &&
                       static_cast<unsigned>(ly1) < static_cast<unsigned>(kTileY) &&
                       static_cast<unsigned>(lz0) < static_cast<unsigned>(kTileZ) &&
                       static_cast<unsigned>(lz1) < static_cast<unsigned>(kTileZ);
  // the common case, all 8 corners in the tile (a BORDER corner outside the
  // grid adds 0 to its halo slot instead of being skipped, which leaves every
  // sum unchanged)
  float wx[2] = {fw[0][0], fw[1][0]}, wy[2] = {fw[0][1], fw[1][1]}, wz[2] = {fw[0][2], fw[1][2]};
  if constexpr (BM == BORDER) {
    wx[0] = c.vx0 ? wx[0] : 0.f;
    wx[1] = c.vx1 ? wx[1] : 0.f;
    wy[0] = c.vy0 ? wy[0] : 0.f;
    wy[1] = c.vy1 ? wy[1] : 0.f;
    wz[0] = c.vz0 ? wz[0] : 0.f;
    wz[1] = c.vz1 ? wz[1] : 0.f;
  }
  const int dx = lx1 - lx0, dy = (ly1 - ly0) * kTileX;
  const int base0 = (lz0 * kTileY + ly0) * kTileX + lx0;
  const int base1 = base0 + (lz1 - lz0) * (kTileX * kTileY);
  Vec<C> lo[4], hi[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int bx = k & 1, by = k >> 1;
    const float w0 = wx[bx] * wy[by] * wz[0], w1 = wx[bx] * wy[by] * wz[1];
#pragma unroll
    for (int i = 0; i < C; ++i) {
      lo[k].v[i] = val.v[i] * w0;
      hi[k].v[i] = val.v[i] * w1;
    }
  }
  if (in_tile) {
    if (held.base == base0 && held.dx == dx && held.dy == dy) {
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int i = 0; i < C; ++i) lo[k].v[i] += held.v[k].v[i];
    } else if (held.base >= 0) {
      plane_add<C>(tile, held.base, held.dx, held.dy, held.v);
    }
  }
  plane_add_row_merged<C>(tile, in_tile, base0, dx, dy, lo);
  if (in_tile) {
    held.base = base1;
    held.dx = dx;
    held.dy = dy;
#pragma unroll
    for (int k = 0; k < 4; ++k) held.v[k] = hi[k];
    return;
  }
  if (!inside) return;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int bx = k & 1, by = (k >> 1) & 1, bz = (k >> 2) & 1;
    if constexpr (BM == BORDER) {
      const bool ok = (bx ? c.vx1 : c.vx0) && (by ? c.vy1 : c.vy0) && (bz ? c.vz1 : c.vz0);
      if (!ok) continue;
    }
    const int ix = bx ? c.x1 : c.x0, iy = by ? c.y1 : c.y0, iz = bz ? c.z1 : c.z0;
    const float w = fw[bx][0] * fw[by][1] * fw[bz][2];
    const int lx = ix - ox, ly = iy - oy, lz = iz - oz;
    if (static_cast<unsigned>(lx) < static_cast<unsigned>(kTileX) &&
        static_cast<unsigned>(ly) < static_cast<unsigned>(kTileY) &&
        static_cast<unsigned>(lz) < static_cast<unsigned>(kTileZ)) {
      float* t = tile + ((lz * kTileY + ly) * kTileX + lx) * C;
#pragma unroll
      for (int i = 0; i < C; ++i) lds_add(t + i, val.v[i] * w);
    } else {
      const int64_t o = cell_offset(d, ix, iy, iz);
#pragma unroll
      for (int i = 0; i < C; ++i) atomic_add(ovf + o * C + i, val.v[i] * w);
    }
  }
}

// Sum of the tile partials of `slot` covering cell (x, y, z), bricks in
// ascending z, y, x order.
template <int C>
NGT_DEVICE Vec<C> gather_tiles(const float* __restrict__ tiles, int64_t slot, const BrickGrid& bg, int x, int y,
                               int z) {
  // the (up to) 8 bricks whose tiles hold the cell: every load is issued
  // before any is added, and the sum runs in the order z, y, x of the bricks
  const int bx1 = min((x + kHalo) / kBrickX, bg.nx - 1), by1 = min((y + kHalo) / kBrickY, bg.ny - 1);
  const int bz1 = min((z + kHalo) / kBrickZ, bg.nz - 1);
  Vec<C> t[8];
  bool ok[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int bz = bz1 - 1 + (q >> 2), by = by1 - 1 + ((q >> 1) & 1), bx = bx1 - 1 + (q & 1);
    const int lz = z - (bz * kBrickZ - kHalo), ly = y - (by * kBrickY - kHalo), lx = x - (bx * kBrickX - kHalo);
    ok[q] = bz >= 0 && by >= 0 && bx >= 0 && lz >= 0 && lz < kTileZ && ly >= 0 && ly < kTileY && lx >= 0 &&
            lx < kTileX;
    t[q] = vzero<C>();
    if (ok[q]) {
      const int64_t brick = (slot * bg.nz + bz) * bg.ny * bg.nx + static_cast<int64_t>(by) * bg.nx + bx;
      t[q] = vload<C>(tiles + (brick * kTileCells + (lz * kTileY + ly) * kTileX + lx) * C);
    }
  }
  Vec<C> s = vzero<C>();
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    if (!ok[q]) continue;
#pragma unroll
    for (int i = 0; i < C; ++i) s.v[i] += t[q].v[i];
  }
  return s;
}

struct BrickCell {
  int pair, brick, ox, oy, oz, x, y, z0;
};

NGT_DEVICE BrickCell brick_cell(const BrickGrid& bg, const VBlock& vb) {
  BrickCell b;
  b.pair = vb.x / bg.count();
  b.brick = vb.x % bg.count();
  const int bx = b.brick % bg.nx, by = (b.brick / bg.nx) % bg.ny, bz = b.brick / (bg.nx * bg.ny);
  b.x = bx * kBrickX + static_cast<int>(threadIdx.x);
  b.y = by * kBrickY + static_cast<int>(threadIdx.y);
  b.z0 = bz * kBrickZ;
  b.ox = bx * kBrickX - kHalo;
  b.oy = by * kBrickY - kHalo;
  b.oz = b.z0 - kHalo;
  return b;
}

template <int C>
NGT_DEVICE void tile_clear(float* tile) {
  const int tid = (threadIdx.z * kBrickY + threadIdx.y) * kBrickX + threadIdx.x;
  for (int i = tid; i < kTileCells * C; i += kBrickThreads) tile[i] = 0.f;
  __syncthreads();
}

template <int C>
NGT_DEVICE void tile_store(const float* tile, float* __restrict__ tiles, const BrickGrid& bg, const BrickCell& b) {
  __syncthreads();
  const int tid = (threadIdx.z * kBrickY + threadIdx.y) * kBrickX + threadIdx.x;
  float* dst = tiles + (static_cast<int64_t>(b.pair) * bg.count() + b.brick) * kTileCells * C;
  for (int i = tid; i < kTileCells * C; i += kBrickThreads) dst[i] = tile[i];
}

// Velocity-gradient contribution of one cell: plain store / read-add-store
// when each velocity has one (data, velocity) pair, atomics when several
// pairs share it.
NGT_DEVICE void add_vel_grad(float* gv, float gx, float gy, float gz, int mode) {
  if (mode == 0) {
    gv[0] = gx; gv[1] = gy; gv[2] = gz;
  } else if (mode == 1) {
    gv[0] += gx; gv[1] += gy; gv[2] += gz;
  } else {
    atomic_add(gv + 0, gx); atomic_add(gv + 1, gy); atomic_add(gv + 2, gz);
  }
}

// B1 brick form: g_dir = g (direct path to the SL field), r_dir = g/2 where the
// corrected value is kept (direct path to rho), the back-step adjoint into
// the tile (overflow into ovf), and the back-step velocity gradient.
template <int C, int BM, int CLAMPM>
NGT_DEVICE void advect_mc_bwd_brick_kernel_body(const VBlock& vb, const AdvArgs& a, BrickGrid bg, const float* __restrict__ rho,
                                                                  const float* __restrict__ sl,
                                                                  const float* __restrict__ g_out,
                                                                  float* __restrict__ g_dir, float* __restrict__ r_dir,
                                                                  float* __restrict__ ovf, float* __restrict__ tiles,
                                                                  float* __restrict__ g_vc, int vel_mode) {
  __shared__ float tile[kTileCells * C];
  tile_clear<C>(tile);
  const BrickCell b = brick_cell(bg, vb);
  int n, v;
  pair_indices(a, b.pair, n, v);
  const int64_t cells = static_cast<int64_t>(a.d.x) * a.d.y * a.d.z;
  const float* src = rho + n * cells * C;
  const float* slp = sl + b.pair * cells * C;
  float* ovp = ovf + b.pair * cells * C;
  // every lane of the wave runs the walk (the tile scatter exchanges values
  // between neighbouring lanes): a lane outside the grid reads its clamped
  // edge cell, stores nothing and scatters nothing
  const bool lane_ok = b.x < a.d.x && b.y < a.d.y;
  const int cx = min(b.x, a.d.x - 1), cy = min(b.y, a.d.y - 1);
  {
    const int zb = b.z0 + static_cast<int>(threadIdx.z) * kBrickZPer;
    const int z1 = min(zb + kBrickZPer, a.d.z);
    HeldPlane<C> held;
#pragma unroll 2
    for (int z = zb; z < z1; ++z) {
      float ux, uy, uz;
      load_velocity(a, v, cx, cy, z, ux, uy, uz);
      const int64_t cell = cell_offset(a.d, cx, cy, z);
      const float fx = static_cast<float>(cx), fy = static_cast<float>(cy), fz = static_cast<float>(z);
      const float tx = a.ts * ux, ty = a.ts * uy, tz = a.ts * uz;
      const float bx = fx + (-tx), by = fy + (-ty), bz = fz + (-tz);
      Vec<C> g = vload<C>(g_out + (b.pair * cells + cell) * C);
      const Vec<C> fwd = vload<C>(slp + cell * C);
      // the back-step read and its position gradient share one corner fetch
      Vec<C> dx, dy, dz;
      const Vec<C> back = read_linear_gradient<C, BM>(slp, a.d, bx, by, bz, dx, dy, dz);
      const Vec<C> cur = vload<C>(src + cell * C);
      Vec<C> corr;
#pragma unroll
      for (int i = 0; i < C; ++i) corr.v[i] = fwd.v[i] + 0.5f * (cur.v[i] - back.v[i]);
      Vec<C> keep = vfill<C>(1.f);
      if constexpr (CLAMPM == NGT_CLAMP_MC_SMOOTH) {
        Vec<C> mn, mx;
        read_minmax2<C, BM>(src, a.d, fx + tx, fy + ty, fz + tz, mn, mx);
#pragma unroll
        for (int i = 0; i < C; ++i)
          if (corr.v[i] < mn.v[i] || corr.v[i] > mx.v[i]) keep.v[i] = 0.f;
      }
      if constexpr (CLAMPM == NGT_CLAMP_NEGATIVE) {
#pragma unroll
        for (int i = 0; i < C; ++i)
          if (!(corr.v[i] >= 0.f)) g.v[i] = 0.f;
      }
      Vec<C> rd, gback;
#pragma unroll
      for (int i = 0; i < C; ++i) {
        rd.v[i] = 0.5f * keep.v[i] * g.v[i];
        gback.v[i] = -rd.v[i];
      }
      if (lane_ok) {
        if (g_dir) vstore<C>(g_dir + (b.pair * cells + cell) * C, g);
        vstore<C>(r_dir + (b.pair * cells + cell) * C, rd);
        add_vel_grad(g_vc + (static_cast<int64_t>(v) * cells + cell) * 3, -a.ts * vdot(gback, dx),
                     -a.ts * vdot(gback, dy), -a.ts * vdot(gback, dz), vel_mode == 1 ? 0 : vel_mode);
      }
      scatter_linear_tile<C, BM>(tile, held, lane_ok, b.ox, b.oy, b.oz, ovp, a.d, bx, by, bz, gback);
    }
    flush_plane<C>(tile, held);
  }
  tile_store<C>(tile, tiles, bg, b);
}

template <int C, int BM, int CLAMPM>
__global__ __launch_bounds__(kBrickThreads) void advect_mc_bwd_brick_kernel(AdvArgs a, BrickGrid bg, const float* __restrict__ rho,
                                                                  const float* __restrict__ sl,
                                                                  const float* __restrict__ g_out,
                                                                  float* __restrict__ g_dir, float* __restrict__ r_dir,
                                                                  float* __restrict__ ovf, float* __restrict__ tiles,
                                                                  float* __restrict__ g_vc, int vel_mode) {
  advect_mc_bwd_brick_kernel_body<C, BM, CLAMPM>(VBlock{static_cast<int>(blockIdx.x), static_cast<int>(blockIdx.y), static_cast<int>(blockIdx.z)}, a, bg, rho, sl, g_out, g_dir, r_dir, ovf, tiles, g_vc, vel_mode);
}

// B2 brick form: g_sl = g_dir + ovf + tiles1 (order 2) or g_out with the
// NEGATIVE mask (order 1); scatter into the tile (overflow straight into
// g_rho, which starts at zero), SL velocity gradient.
template <int C, int BM, bool NEG, bool ORDER2>
NGT_DEVICE void advect_sl_bwd_brick_kernel_body(const VBlock& vb, const AdvArgs& a, BrickGrid bg, const float* __restrict__ rho,
                                                                  const float* __restrict__ g_in,
                                                                  const float* __restrict__ ovf1,
                                                                  const float* __restrict__ tiles1,
                                                                  float* __restrict__ g_rho, float* __restrict__ tiles2,
                                                                  float* __restrict__ g_vc, int vel_mode) {
  __shared__ float tile[kTileCells * C];
  tile_clear<C>(tile);
  const BrickCell b = brick_cell(bg, vb);
  int n, v;
  pair_indices(a, b.pair, n, v);
  const int64_t cells = static_cast<int64_t>(a.d.x) * a.d.y * a.d.z;
  const float* src = rho + n * cells * C;
  float* ovp = g_rho + n * cells * C;
  // every lane of the wave runs the walk (the tile scatter exchanges values
  // between neighbouring lanes): a lane outside the grid reads its clamped
  // edge cell, stores nothing and scatters nothing
  const bool lane_ok = b.x < a.d.x && b.y < a.d.y;
  const int cx = min(b.x, a.d.x - 1), cy = min(b.y, a.d.y - 1);
  {
    const int zb = b.z0 + static_cast<int>(threadIdx.z) * kBrickZPer;
    const int z1 = min(zb + kBrickZPer, a.d.z);
    HeldPlane<C> held;
#pragma unroll 2
    for (int z = zb; z < z1; ++z) {
      float ux, uy, uz;
      load_velocity(a, v, cx, cy, z, ux, uy, uz);
      const int64_t cell = cell_offset(a.d, cx, cy, z);
      const float px = static_cast<float>(cx) + a.ts * ux;
      const float py = static_cast<float>(cy) + a.ts * uy;
      const float pz = static_cast<float>(z) + a.ts * uz;
      Vec<C> g = vload<C>(g_in + (b.pair * cells + cell) * C);
      if constexpr (ORDER2) {
        const Vec<C> o = vload<C>(ovf1 + (b.pair * cells + cell) * C);
        const Vec<C> t = gather_tiles<C>(tiles1, b.pair, bg, cx, cy, z);
#pragma unroll
        for (int i = 0; i < C; ++i) g.v[i] += o.v[i] + t.v[i];
      }
      if constexpr (NEG) {
        const Vec<C> r = read_linear<C, BM>(src, a.d, px, py, pz);
#pragma unroll
        for (int i = 0; i < C; ++i)
          if (!(r.v[i] >= 0.f)) g.v[i] = 0.f;
      }
      scatter_linear_tile<C, BM>(tile, held, lane_ok, b.ox, b.oy, b.oz, ovp, a.d, px, py, pz, g);
      Vec<C> dx, dy, dz;
      read_gradient<C, BM>(src, a.d, px, py, pz, dx, dy, dz);
      if (lane_ok)
        add_vel_grad(g_vc + (static_cast<int64_t>(v) * cells + cell) * 3, a.ts * vdot(g, dx), a.ts * vdot(g, dy),
                     a.ts * vdot(g, dz), vel_mode);
    }
    flush_plane<C>(tile, held);
  }
  tile_store<C>(tile, tiles2, bg, b);
}

template <int C, int BM, bool NEG, bool ORDER2>
__global__ __launch_bounds__(kBrickThreads) void advect_sl_bwd_brick_kernel(AdvArgs a, BrickGrid bg, const float* __restrict__ rho,
                                                                  const float* __restrict__ g_in,
                                                                  const float* __restrict__ ovf1,
                                                                  const float* __restrict__ tiles1,
                                                                  float* __restrict__ g_rho, float* __restrict__ tiles2,
                                                                  float* __restrict__ g_vc, int vel_mode) {
  advect_sl_bwd_brick_kernel_body<C, BM, NEG, ORDER2>(VBlock{static_cast<int>(blockIdx.x), static_cast<int>(blockIdx.y), static_cast<int>(blockIdx.z)}, a, bg, rho, g_in, ovf1, tiles1, g_rho, tiles2, g_vc, vel_mode);
}

// g_rho(n) += sum over the pairs of n of r_dir (order 2) + the B2 tiles.
template <int C>
NGT_DEVICE void advect_rho_gather_kernel_body(const VBlock& vb, const AdvArgs& a, BrickGrid bg, const float* __restrict__ r_dir,
                                                                const float* __restrict__ tiles2,
                                                                float* __restrict__ g_rho) {
  const int x = vb.x * kBX + threadIdx.x;
  const int y = vb.y * kBY + threadIdx.y;
  const int z = vb.z % a.d.z;
  const int n = vb.z / a.d.z;
  if (x >= a.d.x || y >= a.d.y) return;
  const int64_t cells = static_cast<int64_t>(a.d.x) * a.d.y * a.d.z;
  const int64_t cell = cell_offset(a.d, x, y, z);
  Vec<C> s = vload<C>(g_rho + (n * cells + cell) * C);
  const int np = a.separate ? a.n_vel : 1;
  for (int j = 0; j < np; ++j) {
    const int pair = a.separate ? n * a.n_vel + j : n;
    if (r_dir) {
      const Vec<C> r = vload<C>(r_dir + (pair * cells + cell) * C);
#pragma unroll
      for (int i = 0; i < C; ++i) s.v[i] += r.v[i];
    }
    const Vec<C> t = gather_tiles<C>(tiles2, pair, bg, x, y, z);
#pragma unroll
    for (int i = 0; i < C; ++i) s.v[i] += t.v[i];
  }
  vstore<C>(g_rho + (n * cells + cell) * C, s);
}

template <int C>
__global__ __launch_bounds__(256) void advect_rho_gather_kernel(AdvArgs a, BrickGrid bg, const float* __restrict__ r_dir,
                                                                const float* __restrict__ tiles2,
                                                                float* __restrict__ g_rho) {
  advect_rho_gather_kernel_body<C>(VBlock{static_cast<int>(blockIdx.x), static_cast<int>(blockIdx.y), static_cast<int>(blockIdx.z)}, a, bg, r_dir, tiles2, g_rho);
}

// The MAC face gradients that cell (x, y, z) owns, from the centered velocity
// gradient g (3 per cell): its lower face on each axis, and the upper one at
// the grid's far side.  The same sums as faces_from_centered_grad_kernel.
NGT_DEVICE void faces_of_cell(const Dims& d, const float* __restrict__ g, float* __restrict__ ox,
                              float* __restrict__ oy, float* __restrict__ oz, int x, int y, int z) {
  const int64_t plane = static_cast<int64_t>(d.x) * d.y;
  const int64_t c = cell_offset(d, x, y, z);
  const float g0 = g[c * 3 + 0], g1 = g[c * 3 + 1], g2 = g[c * 3 + 2];
  const float l0 = x > 0 ? g[(c - 1) * 3 + 0] : 0.f;
  const float l1 = y > 0 ? g[(c - d.x) * 3 + 1] : 0.f;
  const float l2 = z > 0 ? g[(c - plane) * 3 + 2] : 0.f;
  const int64_t fx = (static_cast<int64_t>(z) * d.y + y) * (d.x + 1) + x;
  ox[fx] = 0.5f * (l0 + g0);
  if (x == d.x - 1) ox[fx + 1] = 0.5f * g0;
  const int64_t fy = (static_cast<int64_t>(z) * (d.y + 1) + y) * d.x + x;
  oy[fy] = 0.5f * (l1 + g1);
  if (y == d.y - 1) oy[fy + d.x] = 0.5f * g1;
  oz[c] = 0.5f * (l2 + g2);
  if (z == d.z - 1) oz[c + plane] = 0.5f * g2;
}

// advect_rho_gather_kernel and the face gradients in one pass, for one data
// grid with one staggered velocity (grid: one work-group row per z plane).
template <int C>
__global__ __launch_bounds__(256) void advect_rho_faces_kernel(AdvArgs a, BrickGrid bg, const float* __restrict__ r_dir,
                                                               const float* __restrict__ tiles2,
                                                               float* __restrict__ g_rho,
                                                               const float* __restrict__ g_vc, float* __restrict__ gx,
                                                               float* __restrict__ gy, float* __restrict__ gz) {
  const VBlock vb{static_cast<int>(blockIdx.x), static_cast<int>(blockIdx.y), static_cast<int>(blockIdx.z)};
  advect_rho_gather_kernel_body<C>(vb, a, bg, r_dir, tiles2, g_rho);
  const int x = vb.x * kBX + static_cast<int>(threadIdx.x), y = vb.y * kBY + static_cast<int>(threadIdx.y);
  if (x >= a.d.x || y >= a.d.y) return;
  faces_of_cell(a.d, g_vc, gx, gy, gz, x, y, vb.z);
}

#define NGT_SWITCH_C(C_, ...)                              \
  switch (C_) {                                            \
    case 1: { constexpr int CC = 1; __VA_ARGS__; } break;  \
    case 2: { constexpr int CC = 2; __VA_ARGS__; } break;  \
    case 3: { constexpr int CC = 3; __VA_ARGS__; } break;  \
    default: { constexpr int CC = 4; __VA_ARGS__; } break; \
  }
#define NGT_SWITCH_C12(C_, ...)                           \
  switch (C_) {                                           \
    case 1: { constexpr int CC = 1; __VA_ARGS__; } break; \
    default: { constexpr int CC = 2; __VA_ARGS__; } break; \
  }
#define NGT_SWITCH_BM(B_, ...)                                      \
  switch (B_) {                                                     \
    case BORDER: { constexpr int BB = BORDER; __VA_ARGS__; } break; \
    case CLAMP: { constexpr int BB = CLAMP; __VA_ARGS__; } break;   \
    default: { constexpr int BB = WRAP; __VA_ARGS__; } break;       \
  }
#define NGT_SWITCH_CLAMP(K_, ...)                                                          \
  switch (K_) {                                                                            \
    case NGT_CLAMP_MC_SMOOTH: { constexpr int KK = NGT_CLAMP_MC_SMOOTH; __VA_ARGS__; } break; \
    case NGT_CLAMP_NEGATIVE: { constexpr int KK = NGT_CLAMP_NEGATIVE; __VA_ARGS__; } break;   \
    default: { constexpr int KK = NGT_CLAMP_NONE; __VA_ARGS__; } break;                      \
  }

struct AdvPlan {
  AdvArgs a;
  int pairs;
  int64_t cells;
  int C;
};

int plan_advect(const int64_t* input_shape, const float* vx, const float* vy, const float* vz, int num_velocities,
                float timestep, int order, int clamp, int boundary, int separate, AdvPlan& p) {
  int rc = check_grid_shape(input_shape, "input", true);
  if (rc) return rc;
  NGT_REQUIRE(vx != nullptr, NGT_ERR_INVALID_ARGUMENT, "velocity must be given");
  NGT_REQUIRE((vy == nullptr) == (vz == nullptr), NGT_ERR_INVALID_ARGUMENT,
              "staggered velocity needs all three components");
  NGT_REQUIRE(num_velocities > 0, NGT_ERR_INVALID_ARGUMENT, "velocity batch must not be empty");
  NGT_REQUIRE(order == 1 || order == 2, NGT_ERR_INVALID_ARGUMENT, "Unsupported warp order");
  NGT_REQUIRE(clamp >= 0 && clamp <= 2, NGT_ERR_INVALID_ARGUMENT, "invalid clamp mode");
  NGT_REQUIRE(boundary >= 0 && boundary <= 2, NGT_ERR_INVALID_ARGUMENT, "invalid boundary");
  if (!separate)
    NGT_REQUIRE(num_velocities == input_shape[0], NGT_ERR_INVALID_ARGUMENT,
                "velocity batch must match data batch when not using separate velocity batches.");
  const Shape5 s5 = shape5(input_shape);
  p.a.d = s5.dims();
  p.a.n_data = static_cast<int>(s5.n);
  p.a.n_vel = num_velocities;
  p.a.separate = separate ? 1 : 0;
  p.a.ts = timestep;
  p.a.vx = vx; p.a.vy = vy; p.a.vz = vz;
  p.a.staggered = vy != nullptr ? 1 : 0;
  p.pairs = separate ? p.a.n_data * num_velocities : p.a.n_data;
  p.cells = s5.cells();
  p.C = static_cast<int>(s5.c);
  (void)timestep;
  return NGT_OK;
}

dim3 adv_grid(const AdvPlan& p) {
  return dim3(blocks_for(p.a.d.x, kBX), blocks_for(p.a.d.y, kBY), static_cast<unsigned>(p.a.d.z * p.pairs));
}

void launch_faces(const AdvPlan& p, const float* g_vc, float* gx, float* gy, float* gz, hipStream_t s) {
  const Dims d = p.a.d;
  hipLaunchKernelGGL(faces_from_centered_grad_kernel, dim3(static_cast<unsigned>(face_tiles(d) * p.a.n_vel)), dim3(256),
                     0, s, d, p.a.n_vel, g_vc, gx, gy, gz);
}

// ---------------------------------------------------------------------------
// Uneven batches: instances of different sizes (one density, one velocity
// each) in one persistent launch per kernel.  The plan blob holds a header,
// one AdvBatchItem per instance and the tile-range tables; its device copy is
// what the kernels read.
struct AdvBatchHeader {
  int n, order, clamp, boundary, staggered;
  int tiles_cell, tiles_brick, tiles_fill;
  size_t off_items, off_cell, off_brick, off_fill;
  int64_t ws_bytes;
};

struct AdvBatchItem {
  AdvArgs a;  // its velocity pointers are read from vx, vy, vz below
  BrickGrid bg;
  int nbx, nby;  // 64 x 4 cell tiles in x / y
  int64_t cells;
  gptr<const float> vx, vy, vz;
  gptr<const float> input;
  gptr<float> output;
  gptr<float> sl;
  gptr<const float> g_out;
  gptr<float> g_in;
  gptr<float> gvx;
  gptr<float> gvy;
  gptr<float> gvz;
  int64_t ws_gsl, ws_gvc, ws_t2, ws_ovf, ws_t1;  // workspace offsets (floats)
};

constexpr int kFillChunk = 1024;

struct AdvBatchView {
  const AdvBatchHeader* h;
  const AdvBatchItem* items;
  const int* cell;
  const int* brick;
  const int* fill;
};

NGT_DEVICE AdvBatchView adv_batch_view(const void* plan) {
  AdvBatchView v;
  const char* p = as_constant<char>(plan);
  v.h = reinterpret_cast<const AdvBatchHeader*>(p);
  v.items = reinterpret_cast<const AdvBatchItem*>(p + v.h->off_items);
  v.cell = reinterpret_cast<const int*>(p + v.h->off_cell);
  v.brick = reinterpret_cast<const int*>(p + v.h->off_brick);
  v.fill = reinterpret_cast<const int*>(p + v.h->off_fill);
  return v;
}

// An instance's arguments and buffers as generic pointers converted from
// their global-typed plan fields (gptr), so that their accesses stay global.
struct AdvBatchBufs {
  AdvArgs a;
  const float* input;
  float* output;
  float* sl;
  const float* g_out;
  float* g_in;
  float* gvx;
  float* gvy;
  float* gvz;
};

NGT_DEVICE AdvBatchBufs batch_bufs(const AdvBatchItem& it) {
  AdvBatchBufs b;
  b.a = it.a;
  b.a.vx = (const float*)it.vx;
  b.a.vy = (const float*)it.vy;
  b.a.vz = (const float*)it.vz;
  b.input = (const float*)it.input;
  b.output = (float*)it.output;
  b.sl = (float*)it.sl;
  b.g_out = (const float*)it.g_out;
  b.g_in = (float*)it.g_in;
  b.gvx = (float*)it.gvx;
  b.gvy = (float*)it.gvy;
  b.gvz = (float*)it.gvz;
  return b;
}

// A cell tile of the batched per-cell kernels is 64 x 4 cells in x / y over
// kCellZ consecutive z planes (the work-group visits the planes in turn, so
// that finding the tile's instance is paid once per kCellZ planes).
constexpr int kCellZ = 4;

NGT_DEVICE VBlock cell_tile(const AdvBatchItem& it, int local) {
  return VBlock{local % it.nbx, (local / it.nbx) % it.nby, local / (it.nbx * it.nby) * kCellZ};
}

// Calls f(VBlock) for the planes z0 .. z0 + kCellZ - 1 of the tile inside the grid.
template <typename F>
NGT_DEVICE void for_tile_planes(const AdvBatchItem& it, const VBlock& tile, F&& f) {
  const int z1 = min(tile.z + kCellZ, it.a.d.z);
  for (int z = tile.z; z < z1; ++z) f(VBlock{tile.x, tile.y, z});
}

template <int BM, bool NEG, bool TO_SL>
__global__ __launch_bounds__(256) void advect_sl_batch_kernel(const void* __restrict__ plan) {
  const AdvBatchView v = adv_batch_view(plan);
  for (int t = blockIdx.x; t < v.h->tiles_cell; t += gridDim.x) {
    const int i = __builtin_amdgcn_readfirstlane(find_item(v.cell, v.h->n, t));
    const AdvBatchItem& it = v.items[i];
    const AdvBatchBufs b = batch_bufs(it);
    for_tile_planes(it, cell_tile(it, t - v.cell[i]), [&](const VBlock& vb) {
      advect_sl_kernel_body<1, BM, NEG>(vb, b.a, b.input, TO_SL ? b.sl : b.output);
    });
  }
}

template <int BM, int CLAMPM>
__global__ __launch_bounds__(256) void advect_mc_batch_kernel(const void* __restrict__ plan) {
  const AdvBatchView v = adv_batch_view(plan);
  for (int t = blockIdx.x; t < v.h->tiles_cell; t += gridDim.x) {
    const int i = __builtin_amdgcn_readfirstlane(find_item(v.cell, v.h->n, t));
    const AdvBatchItem& it = v.items[i];
    const AdvBatchBufs b = batch_bufs(it);
    for_tile_planes(it, cell_tile(it, t - v.cell[i]), [&](const VBlock& vb) {
      advect_mc_kernel_body<1, BM, CLAMPM>(vb, b.a, b.input, b.sl, b.output);
    });
  }
}

// zero what the adjoint's atomics add into: the overflow field of the first
// brick kernel (order 2; the first kernel then stores the direct rho term
// into input_grad, every cell once, before the second kernel's overflow adds
// to it), or input_grad itself (order 1)
__global__ __launch_bounds__(256) void advect_fill_batch_kernel(const void* __restrict__ plan, float* __restrict__ ws) {
  const AdvBatchView v = adv_batch_view(plan);
  for (int t = blockIdx.x; t < v.h->tiles_fill; t += gridDim.x) {
    const int i = __builtin_amdgcn_readfirstlane(find_item(v.fill, v.h->n, t));
    const AdvBatchItem& it = v.items[i];
    float* dst = v.h->order == 2 ? ws + it.ws_ovf : (float*)it.g_in;
    const int64_t b = static_cast<int64_t>(t - v.fill[i]) * kFillChunk;
    const int64_t e1 = min(b + kFillChunk, it.cells);
    if ((reinterpret_cast<uintptr_t>(dst) & 15) == 0 && e1 - b == kFillChunk) {
      static_assert(kFillChunk == 4 * 256, "one 16-byte store per work-item of a 256-wide work-group");
      reinterpret_cast<float4*>(dst + b)[threadIdx.x] = make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
      for (int64_t e = b + threadIdx.x; e < e1; e += blockDim.x) dst[e] = 0.f;
    }
  }
}

template <int BM, int CLAMPM>
__global__ __launch_bounds__(kBrickThreads) void advect_mc_bwd_batch_kernel(const void* __restrict__ plan, float* __restrict__ ws) {
  const AdvBatchView v = adv_batch_view(plan);
  for (int t = blockIdx.x; t < v.h->tiles_brick; t += gridDim.x) {
    const int i = __builtin_amdgcn_readfirstlane(find_item(v.brick, v.h->n, t));
    const AdvBatchItem& it = v.items[i];
    const AdvBatchBufs b = batch_bufs(it);
    float* g_vc = v.h->staggered ? ws + it.ws_gvc : b.gvx;
    // the direct rho term goes straight to input_grad; the masked output
    // gradient is kept only where NEGATIVE changes it
    advect_mc_bwd_brick_kernel_body<1, BM, CLAMPM>(VBlock{t - v.brick[i], 0, 0}, b.a, it.bg, b.input, b.sl,
                                                   b.g_out,
                                                   CLAMPM == NGT_CLAMP_NEGATIVE ? ws + it.ws_gsl : nullptr,
                                                   b.g_in, ws + it.ws_ovf, ws + it.ws_t1, g_vc, 1);
    __syncthreads();  // the LDS tile is cleared again by the next brick
  }
}

template <int BM, bool NEG, bool ORDER2>
__global__ __launch_bounds__(kBrickThreads) void advect_sl_bwd_batch_kernel(const void* __restrict__ plan, float* __restrict__ ws) {
  const AdvBatchView v = adv_batch_view(plan);
  for (int t = blockIdx.x; t < v.h->tiles_brick; t += gridDim.x) {
    const int i = __builtin_amdgcn_readfirstlane(find_item(v.brick, v.h->n, t));
    const AdvBatchItem& it = v.items[i];
    const AdvBatchBufs b = batch_bufs(it);
    float* g_vc = v.h->staggered ? ws + it.ws_gvc : b.gvx;
    const float* g = ORDER2 && v.h->clamp == NGT_CLAMP_NEGATIVE ? ws + it.ws_gsl : b.g_out;
    advect_sl_bwd_brick_kernel_body<1, BM, NEG, ORDER2>(VBlock{t - v.brick[i], 0, 0}, b.a, it.bg, b.input,
                                                        g, ws + it.ws_ovf,
                                                        ws + it.ws_t1, b.g_in, ws + it.ws_t2, g_vc, ORDER2 ? 1 : 0);
    __syncthreads();
  }
}

// The density gather and, for staggered velocities, the face gradients of
// the same cells.
__global__ __launch_bounds__(256) void advect_rho_gather_batch_kernel(const void* __restrict__ plan, float* __restrict__ ws) {
  const AdvBatchView v = adv_batch_view(plan);
  const bool staggered = v.h->staggered != 0;
  for (int t = blockIdx.x; t < v.h->tiles_cell; t += gridDim.x) {
    const int i = __builtin_amdgcn_readfirstlane(find_item(v.cell, v.h->n, t));
    const AdvBatchItem& it = v.items[i];
    float* g_in = (float*)it.g_in;
    for_tile_planes(it, cell_tile(it, t - v.cell[i]), [&](const VBlock& vb) {
      advect_rho_gather_kernel_body<1>(vb, it.a, it.bg, nullptr, ws + it.ws_t2, g_in);
      const int x = vb.x * kBX + static_cast<int>(threadIdx.x), y = vb.y * kBY + static_cast<int>(threadIdx.y);
      if (staggered && x < it.a.d.x && y < it.a.d.y)
        faces_of_cell(it.a.d, ws + it.ws_gvc, (float*)it.gvx, (float*)it.gvy, (float*)it.gvz, x, y, vb.z);
    });
  }
}

}  // namespace
}  // namespace ngt

using namespace ngt;

extern "C" {

size_t ngt_advect_workspace_size(const int64_t* input_shape, int num_velocities, int separate_velocity_batch,
                                 int order, int backward) {
  if (input_shape == nullptr || num_velocities <= 0) return 0;
  const int64_t cells = input_shape[1] * input_shape[2] * input_shape[3];
  const int64_t C = input_shape[4];
  const int64_t pairs = separate_velocity_batch ? input_shape[0] * num_velocities : input_shape[0];
  int64_t floats = 0;
  if (order == 2) floats += pairs * cells * C;          // SL field (when the caller keeps none)
  if (backward) {
    floats += pairs * cells * C;                         // g_sl
    floats += static_cast<int64_t>(num_velocities) * cells * 3;  // centered velocity gradient
    const int64_t ref_extra = pairs * cells * C          // reference: scatter staging
                              + pairs * cells;           // reference: hit counts (uint32)
    int64_t brick_extra = 0;
    if (C <= 2) {
      const Dims d{static_cast<int>(input_shape[3]), static_cast<int>(input_shape[2]),
                   static_cast<int>(input_shape[1])};
      const int64_t tiles = pairs * brick_grid(d).count() * static_cast<int64_t>(kTileCells) * C;
      brick_extra = tiles + (order == 2 ? 2 * pairs * cells * C + tiles : 0);  // tiles2 (+ r_dir, ovf1, tiles1)
    }
    floats += std::max(ref_extra, brick_extra);
  }
  return static_cast<size_t>(floats) * sizeof(float) + 256;
}

int ngt_advect_fwd(const float* input, const int64_t* input_shape, const float* vel_x, const float* vel_y,
                   const float* vel_z, int num_velocities, float timestep, int order, int clamp, int boundary,
                   int separate_velocity_batch, float* output, float* sl_out, void* workspace, size_t workspace_bytes,
                   ngt_stream_t stream) {
  AdvPlan p;
  int rc = plan_advect(input_shape, vel_x, vel_y, vel_z, num_velocities, timestep, order, clamp, boundary,
                       separate_velocity_batch, p);
  if (rc) return rc;
  hipStream_t s = as_stream(stream);
  const dim3 grid = adv_grid(p), block(kBX, kBY);
  if (order == 1) {
    const bool neg = clamp == NGT_CLAMP_NEGATIVE;
    NGT_SWITCH_C(p.C, NGT_SWITCH_BM(boundary,
        if (neg) hipLaunchKernelGGL((advect_sl_kernel<CC, BB, true>), grid, block, 0, s, p.a, input, output);
        else hipLaunchKernelGGL((advect_sl_kernel<CC, BB, false>), grid, block, 0, s, p.a, input, output)));
    return check_launch("ngt_advect_fwd");
  }
  float* sl = sl_out;
  if (sl == nullptr) {
    const size_t need = static_cast<size_t>(p.pairs * p.cells * p.C) * sizeof(float);
    NGT_REQUIRE(workspace && workspace_bytes >= need, NGT_ERR_WORKSPACE, "workspace too small for ngt_advect_fwd");
    sl = static_cast<float*>(workspace);
  }
  NGT_SWITCH_C(p.C, NGT_SWITCH_BM(boundary,
      hipLaunchKernelGGL((advect_sl_kernel<CC, BB, false>), grid, block, 0, s, p.a, input, sl);
      NGT_SWITCH_CLAMP(clamp,
          hipLaunchKernelGGL((advect_mc_kernel<CC, BB, KK>), grid, block, 0, s, p.a, input, sl, output))));
  return check_launch("ngt_advect_fwd");
}

int ngt_advect_bwd(const float* input, const int64_t* input_shape, const float* vel_x, const float* vel_y,
                   const float* vel_z, int num_velocities, float timestep, int order, int clamp, int boundary,
                   int separate_velocity_batch, const float* output_grad, const float* sl, int gradient_semantics,
                   float* input_grad, float* vel_x_grad, float* vel_y_grad, float* vel_z_grad, void* workspace,
                   size_t workspace_bytes, ngt_stream_t stream) {
  AdvPlan p;
  int rc = plan_advect(input_shape, vel_x, vel_y, vel_z, num_velocities, timestep, order, clamp, boundary,
                       separate_velocity_batch, p);
  if (rc) return rc;
  NGT_REQUIRE(vel_x_grad != nullptr && (!p.a.staggered || (vel_y_grad && vel_z_grad)), NGT_ERR_INVALID_ARGUMENT,
              "velocity gradient outputs must be given");
  const size_t need = ngt_advect_workspace_size(input_shape, num_velocities, separate_velocity_batch, order, 1);
  NGT_REQUIRE(workspace && workspace_bytes >= need, NGT_ERR_WORKSPACE, "workspace too small for ngt_advect_bwd");
  const bool ref = gradient_semantics == NGT_GRAD_REFERENCE;
  hipStream_t s = as_stream(stream);
  const dim3 grid = adv_grid(p), block(kBX, kBY);
  const int64_t field = p.pairs * p.cells * p.C;
  float* ws = static_cast<float*>(workspace);
  float* sl_buf = ws;
  float* g_sl = sl_buf + (order == 2 ? field : 0);
  float* g_vc = g_sl + field;
  float* tmp = g_vc + static_cast<int64_t>(p.a.n_vel) * p.cells * 3;
  uint32_t* hits = reinterpret_cast<uint32_t*>(tmp + field);
  if (!p.a.staggered) g_vc = vel_x_grad;  // centered: accumulate in place

  const float* slp = sl;
  if (order == 2 && slp == nullptr) {
    NGT_SWITCH_C(p.C, NGT_SWITCH_BM(boundary,
        hipLaunchKernelGGL((advect_sl_kernel<CC, BB, false>), grid, block, 0, s, p.a, input, sl_buf)));
    slp = sl_buf;
  }
  const bool brick_path = !ref && p.C <= 2;
  // the brick path clears (or has B1 store) input_grad in its order-2 branch
  if (!(brick_path && order == 2)) fill_zero(input_grad, p.a.n_data * p.cells * p.C, s);
  if (brick_path) {
    // brick-tiled exact adjoint: workspace after g_vc = [tiles2 | r_dir | ovf1 | tiles1]
    const BrickGrid bg = brick_grid(p.a.d);
    const int64_t tiles_n = static_cast<int64_t>(p.pairs) * bg.count() * kTileCells * p.C;
    float* tiles2 = tmp;
    float* r_dir = tiles2 + tiles_n;
    float* ovf1 = r_dir + field;
    float* tiles1 = ovf1 + field;
    const bool shared_v = p.a.separate && p.a.n_data > 1;
    const int vel_mode = shared_v ? 2 : (order == 2 ? 1 : 0);
    if (shared_v) fill_zero(g_vc, static_cast<int64_t>(p.a.n_vel) * p.cells * 3, s);
    const dim3 bgrid(static_cast<unsigned>(bg.count() * p.pairs)), bblock(kBrickX, kBrickY, kBrickZT);
    const bool neg = order == 1 && clamp == NGT_CLAMP_NEGATIVE;
    // one (data, velocity) pair per data entry: B1 stores the direct rho term
    // straight into input_grad (each cell once, before B2's overflow atomics
    // add to it); without the NEGATIVE mask B1's output gradient is g_out
    // itself and B2 reads that
    const bool direct = !p.a.separate || p.a.n_vel == 1;
    const bool pass_g = clamp != NGT_CLAMP_NEGATIVE;
    if (order == 2) {
      if (direct) fill_zero(ovf1, field, s);
      else fill_zero2(input_grad, p.a.n_data * p.cells * p.C, ovf1, field, s);
      float* g_dir = pass_g ? nullptr : g_sl;
      float* rd = direct ? input_grad : r_dir;
      const float* g_b2 = pass_g ? output_grad : g_sl;
      NGT_SWITCH_C12(p.C, NGT_SWITCH_BM(boundary, NGT_SWITCH_CLAMP(clamp,
          hipLaunchKernelGGL((advect_mc_bwd_brick_kernel<CC, BB, KK>), bgrid, bblock, 0, s, p.a, bg, input, slp,
                             output_grad, g_dir, rd, ovf1, tiles1, g_vc, vel_mode);
          hipLaunchKernelGGL((advect_sl_bwd_brick_kernel<CC, BB, false, true>), bgrid, bblock, 0, s, p.a, bg, input,
                             g_b2, ovf1, tiles1, input_grad, tiles2, g_vc, vel_mode))));
    } else {
      NGT_SWITCH_C12(p.C, NGT_SWITCH_BM(boundary,
          if (neg) hipLaunchKernelGGL((advect_sl_bwd_brick_kernel<CC, BB, true, false>), bgrid, bblock, 0, s, p.a, bg,
                                      input, output_grad, nullptr, nullptr, input_grad, tiles2, g_vc, vel_mode);
          else hipLaunchKernelGGL((advect_sl_bwd_brick_kernel<CC, BB, false, false>), bgrid, bblock, 0, s, p.a, bg,
                                  input, output_grad, nullptr, nullptr, input_grad, tiles2, g_vc, vel_mode)));
    }
    const dim3 dgrid(blocks_for(p.a.d.x, kBX), blocks_for(p.a.d.y, kBY), static_cast<unsigned>(p.a.d.z * p.a.n_data));
    const float* rd_in = order == 2 && !direct ? r_dir : nullptr;
    if (p.a.staggered && p.a.n_data == 1 && p.a.n_vel == 1) {
      NGT_SWITCH_C12(p.C, hipLaunchKernelGGL((advect_rho_faces_kernel<CC>), dgrid, block, 0, s, p.a, bg, rd_in, tiles2,
                                             input_grad, g_vc, vel_x_grad, vel_y_grad, vel_z_grad));
    } else {
      NGT_SWITCH_C12(p.C, hipLaunchKernelGGL((advect_rho_gather_kernel<CC>), dgrid, block, 0, s, p.a, bg, rd_in, tiles2,
                                             input_grad));
      if (p.a.staggered) launch_faces(p, g_vc, vel_x_grad, vel_y_grad, vel_z_grad, s);
    }
    return check_launch("ngt_advect_bwd");
  }
  fill_zero(g_vc, static_cast<int64_t>(p.a.n_vel) * p.cells * 3, s);
  const float* gsl_in = output_grad;
  if (order == 2) {
    fill_zero(g_sl, field, s);
    if (ref) {
      fill_zero(tmp, field, s);
      fill_zero_u32(hits, p.pairs * p.cells, s);
    }
    NGT_SWITCH_C(p.C, NGT_SWITCH_BM(boundary, NGT_SWITCH_CLAMP(clamp,
        if (ref) hipLaunchKernelGGL((advect_mc_bwd_kernel<CC, BB, KK, true>), grid, block, 0, s, p.a, input, slp,
                                    output_grad, g_sl, input_grad, g_vc, tmp, hits);
        else hipLaunchKernelGGL((advect_mc_bwd_kernel<CC, BB, KK, false>), grid, block, 0, s, p.a, input, slp,
                                output_grad, g_sl, input_grad, g_vc, tmp, hits))));
    if (ref) {
      NGT_SWITCH_C(p.C, hipLaunchKernelGGL((add_normalized_kernel<CC>), grid, block, 0, s, p.a, tmp, hits, g_sl, 0));
    }
    gsl_in = g_sl;
  }
  if (ref) {
    fill_zero(tmp, field, s);
    fill_zero_u32(hits, p.pairs * p.cells, s);
  }
  const bool neg = order == 1 && clamp == NGT_CLAMP_NEGATIVE;
  NGT_SWITCH_C(p.C, NGT_SWITCH_BM(boundary,
      if (neg) {
        if (ref) hipLaunchKernelGGL((advect_sl_bwd_kernel<CC, BB, true, true>), grid, block, 0, s, p.a, input, gsl_in,
                                    nullptr, input_grad, g_vc, tmp, hits);
        else hipLaunchKernelGGL((advect_sl_bwd_kernel<CC, BB, true, false>), grid, block, 0, s, p.a, input, gsl_in,
                                nullptr, input_grad, g_vc, tmp, hits);
      } else {
        if (ref) hipLaunchKernelGGL((advect_sl_bwd_kernel<CC, BB, false, true>), grid, block, 0, s, p.a, input,
                                    gsl_in, nullptr, input_grad, g_vc, tmp, hits);
        else hipLaunchKernelGGL((advect_sl_bwd_kernel<CC, BB, false, false>), grid, block, 0, s, p.a, input, gsl_in,
                                nullptr, input_grad, g_vc, tmp, hits);
      }));
  if (ref) {
    NGT_SWITCH_C(p.C, hipLaunchKernelGGL((add_normalized_kernel<CC>), grid, block, 0, s, p.a, tmp, hits, input_grad, 1));
  }
  if (p.a.staggered) launch_faces(p, g_vc, vel_x_grad, vel_y_grad, vel_z_grad, s);
  return check_launch("ngt_advect_bwd");
}


size_t ngt_advect_batch_plan_bytes(int num_instances) {
  if (num_instances <= 0) return 0;
  BlobLayout L;
  L.add<AdvBatchHeader>(1);
  L.add<AdvBatchItem>(num_instances);
  for (int k = 0; k < 3; ++k) L.add<int>(num_instances + 1);
  return L.bytes;
}

int ngt_advect_batch_plan(const ngt_advect_instance* inst, int n, int order, int clamp, int boundary, void* plan,
                          size_t plan_bytes, size_t* workspace_bytes) {
  NGT_REQUIRE(inst != nullptr && n > 0, NGT_ERR_INVALID_ARGUMENT, "batch must not be empty");
  NGT_REQUIRE(order == 1 || order == 2, NGT_ERR_INVALID_ARGUMENT, "Unsupported warp order");
  NGT_REQUIRE(clamp >= 0 && clamp <= 2, NGT_ERR_INVALID_ARGUMENT, "invalid clamp mode");
  NGT_REQUIRE(boundary >= 0 && boundary <= 2, NGT_ERR_INVALID_ARGUMENT, "invalid boundary");
  NGT_REQUIRE(plan != nullptr && plan_bytes >= ngt_advect_batch_plan_bytes(n), NGT_ERR_WORKSPACE,
              "plan buffer too small for ngt_advect_batch_plan");
  BlobLayout L;
  const size_t off_h = L.add<AdvBatchHeader>(1);
  const size_t off_items = L.add<AdvBatchItem>(n);
  const size_t off_cell = L.add<int>(n + 1), off_brick = L.add<int>(n + 1), off_fill = L.add<int>(n + 1);
  AdvBatchHeader* h = blob_at<AdvBatchHeader>(plan, off_h);
  AdvBatchItem* items = blob_at<AdvBatchItem>(plan, off_items);
  int* cell = blob_at<int>(plan, off_cell);
  int* brick = blob_at<int>(plan, off_brick);
  int* fill = blob_at<int>(plan, off_fill);
  const int staggered = inst[0].vel_y != nullptr ? 1 : 0;
  int64_t tc = 0, tb = 0, tl = 0, ws = 0;
  cell[0] = brick[0] = fill[0] = 0;
  for (int i = 0; i < n; ++i) {
    const ngt_advect_instance& q = inst[i];
    NGT_REQUIRE(q.dhw[0] > 0 && q.dhw[1] > 0 && q.dhw[2] > 0, NGT_ERR_INVALID_ARGUMENT, "empty batch instance");
    NGT_REQUIRE(q.input && q.vel_x, NGT_ERR_INVALID_ARGUMENT, "batch instance without input or velocity");
    NGT_REQUIRE((q.vel_y != nullptr) == (staggered != 0) && (q.vel_z != nullptr) == (staggered != 0),
                NGT_ERR_INVALID_ARGUMENT, "all batch instances must use the same velocity layout");
    const int64_t sh[5] = {1, q.dhw[0], q.dhw[1], q.dhw[2], 1};
    int rc = check_grid_shape(sh, "batch instance");
    if (rc) return rc;
    AdvBatchItem& it = items[i];
    it.a.d = Dims{q.dhw[2], q.dhw[1], q.dhw[0]};
    it.a.n_data = 1; it.a.n_vel = 1; it.a.separate = 0;
    it.a.ts = q.timestep;
    it.a.vx = q.vel_x; it.a.vy = q.vel_y; it.a.vz = q.vel_z;
    it.a.staggered = staggered;
    it.bg = brick_grid(it.a.d);
    it.nbx = static_cast<int>(blocks_for(it.a.d.x, kBX));
    it.nby = static_cast<int>(blocks_for(it.a.d.y, kBY));
    it.cells = static_cast<int64_t>(q.dhw[0]) * q.dhw[1] * q.dhw[2];
    it.vx = (gptr<const float>)q.vel_x; it.vy = (gptr<const float>)q.vel_y; it.vz = (gptr<const float>)q.vel_z;
    it.input = (gptr<const float>)q.input; it.output = (gptr<float>)q.output; it.sl = (gptr<float>)q.sl;
    it.g_out = (gptr<const float>)q.output_grad; it.g_in = (gptr<float>)q.input_grad;
    it.gvx = (gptr<float>)q.vel_x_grad; it.gvy = (gptr<float>)q.vel_y_grad; it.gvz = (gptr<float>)q.vel_z_grad;
    const int64_t tiles = static_cast<int64_t>(it.bg.count()) * kTileCells;
    auto take = [&](int64_t floats) {
      const int64_t off = ws;
      ws += (floats + 63) / 64 * 64;  // 256-byte aligned regions
      return off;
    };
    it.ws_gsl = order == 2 && clamp == NGT_CLAMP_NEGATIVE ? take(it.cells) : 0;
    it.ws_gvc = staggered ? take(it.cells * 3) : 0;
    it.ws_t2 = take(tiles);
    it.ws_ovf = order == 2 ? take(it.cells) : 0;
    it.ws_t1 = order == 2 ? take(tiles) : 0;
    tc += static_cast<int64_t>(it.nbx) * it.nby * ((q.dhw[0] + kCellZ - 1) / kCellZ);
    tb += it.bg.count();
    tl += (it.cells + kFillChunk - 1) / kFillChunk;
    NGT_REQUIRE(tc < (int64_t(1) << 31) && tl < (int64_t(1) << 31), NGT_ERR_INVALID_ARGUMENT, "batch too large");
    cell[i + 1] = static_cast<int>(tc); brick[i + 1] = static_cast<int>(tb); fill[i + 1] = static_cast<int>(tl);
  }
  h->n = n; h->order = order; h->clamp = clamp; h->boundary = boundary; h->staggered = staggered;
  h->tiles_cell = static_cast<int>(tc); h->tiles_brick = static_cast<int>(tb);
  h->tiles_fill = static_cast<int>(tl);
  h->off_items = off_items; h->off_cell = off_cell; h->off_brick = off_brick; h->off_fill = off_fill;
  h->ws_bytes = ws * static_cast<int64_t>(sizeof(float));
  if (workspace_bytes) *workspace_bytes = static_cast<size_t>(h->ws_bytes);
  return NGT_OK;
}

int ngt_advect_batch_fwd(const void* plan_host, const void* plan_device, int blocks_per_cu, ngt_stream_t stream) {
  NGT_REQUIRE(plan_host && plan_device, NGT_ERR_INVALID_ARGUMENT, "plan must be given");
  const AdvBatchHeader& h = *static_cast<const AdvBatchHeader*>(plan_host);
  const AdvBatchItem* items = blob_at<AdvBatchItem>(plan_host, h.off_items);
  for (int i = 0; i < h.n; ++i)
    NGT_REQUIRE(items[i].output && (h.order == 1 || items[i].sl), NGT_ERR_INVALID_ARGUMENT,
                "batch instance without output (or sl buffer for order 2)");
  hipStream_t s = as_stream(stream);
  const dim3 g(persistent_grid(h.tiles_cell, blocks_per_cu)), b(kBX, kBY);
  if (h.order == 1) {
    const bool neg = h.clamp == NGT_CLAMP_NEGATIVE;
    NGT_SWITCH_BM(h.boundary,
        if (neg) hipLaunchKernelGGL((advect_sl_batch_kernel<BB, true, false>), g, b, 0, s, plan_device);
        else hipLaunchKernelGGL((advect_sl_batch_kernel<BB, false, false>), g, b, 0, s, plan_device));
  } else {
    NGT_SWITCH_BM(h.boundary,
        hipLaunchKernelGGL((advect_sl_batch_kernel<BB, false, true>), g, b, 0, s, plan_device);
        NGT_SWITCH_CLAMP(h.clamp, hipLaunchKernelGGL((advect_mc_batch_kernel<BB, KK>), g, b, 0, s, plan_device)));
  }
  return check_launch("ngt_advect_batch_fwd");
}

int ngt_advect_batch_bwd(const void* plan_host, const void* plan_device, void* workspace, size_t workspace_bytes,
                         int blocks_per_cu, ngt_stream_t stream) {
  NGT_REQUIRE(plan_host && plan_device, NGT_ERR_INVALID_ARGUMENT, "plan must be given");
  const AdvBatchHeader& h = *static_cast<const AdvBatchHeader*>(plan_host);
  NGT_REQUIRE(workspace != nullptr && workspace_bytes >= static_cast<size_t>(h.ws_bytes), NGT_ERR_WORKSPACE,
              "workspace too small for ngt_advect_batch_bwd");
  const AdvBatchItem* items = blob_at<AdvBatchItem>(plan_host, h.off_items);
  for (int i = 0; i < h.n; ++i)
    NGT_REQUIRE(items[i].g_out && items[i].g_in && items[i].gvx && (!h.staggered || (items[i].gvy && items[i].gvz)) &&
                    (h.order == 1 || items[i].sl),
                NGT_ERR_INVALID_ARGUMENT, "batch instance without gradient buffers (or sl for order 2)");
  hipStream_t s = as_stream(stream);
  float* ws = static_cast<float*>(workspace);
  const dim3 b(kBX, kBY);
  hipLaunchKernelGGL(advect_fill_batch_kernel, dim3(persistent_grid(h.tiles_fill, blocks_per_cu)), dim3(256), 0, s,
                     plan_device, ws);
  const dim3 gb(persistent_grid(h.tiles_brick, blocks_per_cu)), bb(kBrickX, kBrickY, kBrickZT);
  if (h.order == 2) {
    NGT_SWITCH_BM(h.boundary, NGT_SWITCH_CLAMP(h.clamp,
        hipLaunchKernelGGL((advect_mc_bwd_batch_kernel<BB, KK>), gb, bb, 0, s, plan_device, ws);
        hipLaunchKernelGGL((advect_sl_bwd_batch_kernel<BB, false, true>), gb, bb, 0, s, plan_device, ws)));
  } else {
    const bool neg = h.clamp == NGT_CLAMP_NEGATIVE;
    NGT_SWITCH_BM(h.boundary,
        if (neg) hipLaunchKernelGGL((advect_sl_bwd_batch_kernel<BB, true, false>), gb, bb, 0, s, plan_device, ws);
        else hipLaunchKernelGGL((advect_sl_bwd_batch_kernel<BB, false, false>), gb, bb, 0, s, plan_device, ws));
  }
  hipLaunchKernelGGL(advect_rho_gather_batch_kernel, dim3(persistent_grid(h.tiles_cell, blocks_per_cu)), b, 0, s,
                     plan_device, ws);
  return check_launch("ngt_advect_batch_bwd");
}

}  // extern "C"
