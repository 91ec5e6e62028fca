// RaymarchGridTransform (+Grad) and the fused density renderer on gfx950.
//
// One work-item per pixel; each wavefront owns an 8x8 pixel tile (one
// 64-wide work-group), which keeps the 64 rays of a wavefront spatially
// coherent and their 8-tap gathers inside a few cache lines per step.  All
// (batch, camera) pairs of a call run in one launch (grid.z), with camera
// matrices passed by value as kernel arguments instead of __constant__
// uploads per camera (raymarch_grid.cu.cc:399-430).
//
// Ray lattice (raymarch_grid.cu.cc:173-200): start = IDXtoOS(x+.5, y+.5, .5),
// end = IDXtoOS(x+.5, y+.5, zMax+.5), step = (end-start)/zMax, D samples;
// BORDER clips to the cell-centre AABB widened by one cell.  Samples are
// evaluated as start + k*step (the reference accumulates samplePos += step).
//
// Exact backward: a forward walk over the same lattice recomputing the
// samples; the per-sample cell gradient uses
//   d sigma_k = g_d - (<L_D, g_L> - sum_{j<k} e_j <l_j, g_L>)
// (BEER_LAMBERT), so no accumulator is rebuilt by subtraction; ALPHA walks
// backward.  The volume gradient is scattered with hardware fp32 atomics.
// Reference semantics (gradient_semantics = REFERENCE) reproduce the
// reference: reverse walk from `end`, subtraction reconstruction, corner
// weights accumulated and the gradient divided by them per camera.
#include "ngt_host.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

namespace ngt {
namespace {

struct RayArgs {
  const float* input;
  Dims din;
  int out_w, out_h, out_d;
  float inv_w, inv_h, inv_d;
  float light_r, light_g, light_b, depth_step;  // fused density renderer only
  const float* bricked;                         // bricked copy of input (kernels with BRICKED = true)
  BrickDims bd;
  int n_data;
  // fused L2 image loss (render backward): when l2_target is set the image
  // gradient is 2 * l2_scale * l2_loss_grad[0] * (image - target) on RGB and
  // 0 on the optical depth, formed in the gradient pass instead of read
  const float* l2_target;     // [N*V, H, W, 3]
  const float* l2_loss_grad;  // one float on the device (nullptr: 1)
  float l2_scale;
  // fused L2 image loss (render forward): when set, each work-group of the
  // forward writes the sum over its pixels and RGB of (image - target)^2 to
  // l2_partial[linear block index of the launch]
  float* l2_partial;
};

NGT_DEVICE int kMaxRaySteps(const RayArgs& a) { return 16 * (a.din.x + a.din.y + a.din.z + a.out_d); }

// Ray samples whose gathers are issued together, measured on the 128^3 /
// 5-view render with one wavefront per 8x8 pixel tile: the image walk is
// fastest with 2 (43.7 us; 1: 48.9, 4: 47.2), the recording walk (more work
// per sample) with 4, one quad of its sample record per round (54.7 us; 1:
// 54.9, 2: 58.1, 8: 59.6).  With few rays per SIMD a walk is bound by its
// per-sample recurrence as much as by the latency of its corner loads, and
// wider rounds add registers and past-the-end samples.
constexpr int kUnroll = 2;
constexpr int kRecordUnroll = 4;

struct Ray {
  float sx, sy, sz;  // start
  float ex, ey, ez;  // end
  float dx, dy, dz;  // step
  int steps;
  bool hit;
};

// AABB entry / exit distances (raymarch_grid.cu.cc:109-156).  fmaxf / fminf
// drop NaNs produced by axis-parallel rays (0 * inf).
NGT_DEVICE void ray_box(const Dims& din, float ox, float oy, float oz, float rx, float ry, float rz, float border,
                        float& tmin, float& tmax) {
  const float bx = static_cast<float>(din.x), by = static_cast<float>(din.y), bz = static_cast<float>(din.z);
  const float hx = bx * 0.5f + border, hy = by * 0.5f + border, hz = bz * 0.5f + border;
  const float cx = ox - bx * 0.5f, cy = oy - by * 0.5f, cz = oz - bz * 0.5f;
  const float ix = 1.0f / rx, iy = 1.0f / ry, iz = 1.0f / rz;
  const float nx = ix * cx, ny = iy * cy, nz = iz * cz;
  const float kx = fabsf(ix) * hx, ky = fabsf(iy) * hy, kz = fabsf(iz) * hz;
  tmin = fmaxf(fmaxf(0.0f - nx - kx, 0.0f - ny - ky), 0.0f - nz - kz);
  tmax = fminf(fminf(kx - nx, ky - ny), kz - nz);
}

template <int BM>
NGT_DEVICE Ray make_ray(const RayArgs& a, const CamParams& cp, int x, int y) {
  Ray r;
  const float zmax = static_cast<float>(max(a.out_d - 1, 1));
  const float px = static_cast<float>(x) + 0.5f, py = static_cast<float>(y) + 0.5f;
  idx_to_os(cp, px, py, 0.5f, a.inv_w, a.inv_h, a.inv_d, r.sx, r.sy, r.sz);
  idx_to_os(cp, px, py, zmax + 0.5f, a.inv_w, a.inv_h, a.inv_d, r.ex, r.ey, r.ez);
  r.dx = (r.ex - r.sx) / zmax; r.dy = (r.ey - r.sy) / zmax; r.dz = (r.ez - r.sz) / zmax;
  r.steps = a.out_d;
  r.hit = true;
  if constexpr (BM == BORDER) {
    float ux = r.ex - r.sx, uy = r.ey - r.sy, uz = r.ez - r.sz;
    const float inv_len = rsqrtf(ux * ux + uy * uy + uz * uz);
    ux *= inv_len; uy *= inv_len; uz *= inv_len;
    float tmin, tmax;
    ray_box(a.din, r.sx, r.sy, r.sz, ux, uy, uz, -0.5f, tmin, tmax);
    if (tmin > tmax || tmax < 0.f) {
      r.hit = false;
      r.steps = 0;
      return r;
    }
    r.ex = r.sx + ux * (tmax + 1.f); r.ey = r.sy + uy * (tmax + 1.f); r.ez = r.sz + uz * (tmax + 1.f);
    const float t0 = fmaxf(tmin - 1.f, 0.f);
    r.sx = r.sx + ux * t0; r.sy = r.sy + uy * t0; r.sz = r.sz + uz * t0;
    const float step_len = sqrtf(r.dx * r.dx + r.dy * r.dy + r.dz * r.dz);
    const float lx = r.sx - r.ex, ly = r.sy - r.ey, lz = r.sz - r.ez;
    const float ray_len = sqrtf(lx * lx + ly * ly + lz * lz);
    r.steps = static_cast<int>(ray_len / step_len) + 2;
  }
  // Bound the walk for degenerate cameras (step length ~ 0): no sane frustum
  // needs more samples than a few times the grid diagonal plus the depth.
  r.steps = min(max(r.steps, 0), kMaxRaySteps(a));
  return r;
}

// The per-pixel ray kernels run one wavefront per 8x8 pixel tile: rays of
// very different lengths (the image's edges miss the grid) balance over the
// CUs in finer units than 16x16 work-groups (the 128^3 / 5-view render
// forward 52 -> 43 us, the recording forward 60 -> 58 us).
constexpr int kRayBlock = 64, kRayTile = 8;

// Pixel of a work-item: a 64-wide work-group covers one 8x8 tile; a
// 256-wide one covers a 16x16 tile as four 8x8 wavefront tiles.
NGT_DEVICE void pixel_of_thread(const VBlock& vb, int& x, int& y) {
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  if (blockDim.x == 64) {
    x = vb.x * 8 + (lane & 7);
    y = vb.y * 8 + (lane >> 3);
  } else {
    x = vb.x * 16 + (wave & 1) * 8 + (lane & 7);
    y = vb.y * 16 + (wave >> 1) * 8 + (lane >> 3);
  }
}

inline dim3 ray_grid_dims(int out_w, int out_h, int count) {
  return dim3(blocks_for(out_w, kRayTile), blocks_for(out_h, kRayTile), count);
}


// Cell value read by a sample: the C-channel input (LIGHT = false) or the
// fused renderer's (rho*rgb*ds, rho*ds) from the 1-channel density.
// BRICKED (LINEAR only): src is a slice of the bricked copy.
template <int CIN, int FM, int BM, bool BRICKED>
NGT_DEVICE Vec<CIN> read_cell(const RayArgs& a, const float* __restrict__ src, float qx, float qy, float qz) {
  if constexpr (BRICKED) {
    static_assert(FM == LINEAR, "the bricked layout serves LINEAR sampling only");
    const float x[1] = {qx}, y[1] = {qy}, z[1] = {qz};
    Vec<CIN> r[1];
    read_linear_batch<CIN, BM, 1, true>(src, a.din, a.bd, x, y, z, r);
    return r[0];
  } else {
    return read_filtered<CIN, FM, BM>(src, a.din, qx, qy, qz);
  }
}

template <int C, int FM, int BM, bool LIGHT, bool BRICKED = false>
NGT_DEVICE Vec<C> sample_cell(const RayArgs& a, const float* __restrict__ src, float qx, float qy, float qz) {
  if constexpr (LIGHT) {
    const Vec<1> rho = read_cell<1, FM, BM, BRICKED>(a, src, qx, qy, qz);
    const float s = rho.v[0] * a.depth_step;
    Vec<C> c;
    c.v[0] = a.light_r * s; c.v[1] = a.light_g * s; c.v[2] = a.light_b * s; c.v[3] = s;
    return c;
  } else {
    return read_cell<C, FM, BM, BRICKED>(a, src, qx, qy, qz);
  }
}

// Slice of data entry n read by a kernel (the bricked copy when BRICKED).
template <int CIN, bool BRICKED>
NGT_DEVICE const float* ray_source(const RayArgs& a, int n) {
  if constexpr (BRICKED) return a.bricked + static_cast<int64_t>(n) * a.bd.slice_cells * CIN;
  else return a.input + static_cast<int64_t>(n) * a.din.x * a.din.y * a.din.z * CIN;
}

// The light cell (rho rgb ds, rho ds) of a density sample.  Its products are
// rounded on their own (never fused into a later add), so that every walk,
// however its samples are unrolled, forms the same optical depths and
// images.
NGT_DEVICE Vec<4> light_cell(const RayArgs& a, float rho) {
#pragma clang fp contract(off)
  const float s = rho * a.depth_step;
  Vec<4> c;
  c.v[0] = a.light_r * s; c.v[1] = a.light_g * s; c.v[2] = a.light_b * s;
  c.v[3] = s;
  return c;
}

// One sample of a light-cell walk with BEER_LAMBERT or ADDITIVE blending, in
// scalar form.  The light is the same at every sample, so the colour is
// light * Q with Q = sum_k e_k s_k (BEER_LAMBERT; e_k = exp(-dd_k), dd_k the
// optical depth through sample k) or Q = sum_k s_k (ADDITIVE), where s_k =
// rho_k ds: one accumulator in place of three.  Returns e_k (1 for ADDITIVE).
template <int MODE>
NGT_DEVICE float light_blend_step(float& dd, float& Q, float s) {
  if constexpr (MODE == BEER_LAMBERT) {
    dd = fmaxf(dd + s, 0.f);
    const float e = exp_neg(dd);
    Q = fmaf(e, s, Q);
    return e;
  } else {
    Q += s;
    return 1.f;
  }
}

// The pixel (rgb, optical depth or summed depth) of a light-cell walk.
template <int MODE>
NGT_DEVICE Vec<4> light_pixel(const RayArgs& a, float dd, float Q) {
  Vec<4> p;
  p.v[0] = a.light_r * Q; p.v[1] = a.light_g * Q; p.v[2] = a.light_b * Q;
  p.v[3] = MODE == BEER_LAMBERT ? dd : Q;
  return p;
}

// U cell values at once; for LINEAR all their corner loads are issued
// before any is used.
template <int C, int FM, int BM, bool LIGHT, bool BRICKED, int U>
NGT_DEVICE void sample_cells(const RayArgs& a, const float* __restrict__ src, const float (&qx)[U],
                             const float (&qy)[U], const float (&qz)[U], Vec<C> (&cells)[U]) {
  if constexpr (FM == LINEAR) {
    constexpr int CIN = LIGHT ? 1 : C;
    Vec<CIN> raw[U];
    read_linear_batch<CIN, BM, U, BRICKED>(src, a.din, a.bd, qx, qy, qz, raw);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if constexpr (LIGHT) {
        cells[u] = light_cell(a, raw[u].v[0]);
      } else {
#pragma unroll
        for (int i = 0; i < C; ++i) cells[u].v[i] = raw[u].v[i];
      }
    }
  } else {
#pragma unroll
    for (int u = 0; u < U; ++u) cells[u] = sample_cell<C, FM, BM, LIGHT>(a, src, qx[u], qy[u], qz[u]);
  }
}

// Scatter a cell gradient at a sample position (exact adjoint of sample_cell).
template <int C, int FM, int BM, bool LIGHT>
NGT_DEVICE void scatter_cell(const RayArgs& a, const float* __restrict__ src, float* __restrict__ g, float qx, float qy,
                             float qz, const Vec<C>& gc) {
  if constexpr (LIGHT) {
    Vec<1> gr;
    gr.v[0] = a.depth_step * (a.light_r * gc.v[0] + a.light_g * gc.v[1] + a.light_b * gc.v[2] + gc.v[3]);
    scatter_filtered<1, FM, BM>(src, g, a.din, qx, qy, qz, gr);
  } else {
    scatter_filtered<C, FM, BM>(src, g, a.din, qx, qy, qz, gc);
  }
}

template <int C, int FM, int BM, int MODE, bool LIGHT, bool BRICKED>
NGT_DEVICE void raymarch_fwd_pixel(const RayArgs& a, const CamBatch& cb, const VBlock& vb, int x, int y,
                                   float* __restrict__ out, float& l2) {
  const CamParams& cp = cb.cam[vb.z];
  constexpr int CIN = LIGHT ? 1 : C;
  const float* src = ray_source<CIN, BRICKED>(a, cp.data_index);
  const Ray r = make_ray<BM>(a, cp, x, y);
  Vec<C> acc = vzero<C>();
  // Samples are fetched kUnroll at a time so their gathers are in flight
  // together; the blend recurrence then consumes them in order.  The last
  // round's positions past the ray end are fetched (in-grid addresses) and
  // blended as empty cells, which leaves the accumulator unchanged in every
  // blend mode (no branch per sample, so no load is sunk into one).
  if constexpr (LIGHT && (MODE == BEER_LAMBERT || MODE == ADDITIVE)) {
    float dd = 0.f, Q = 0.f;
    for (int k = 0; k < r.steps; k += kUnroll) {
      float qx[kUnroll], qy[kUnroll], qz[kUnroll];
      Vec<C> cells[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const float kf = static_cast<float>(k + u);
        qx[u] = fmaf(kf, r.dx, r.sx) - 0.5f; qy[u] = fmaf(kf, r.dy, r.sy) - 0.5f; qz[u] = fmaf(kf, r.dz, r.sz) - 0.5f;
      }
      sample_cells<C, FM, BM, LIGHT, BRICKED>(a, src, qx, qy, qz, cells);
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) light_blend_step<MODE>(dd, Q, k + u < r.steps ? cells[u].v[3] : 0.f);
    }
    acc = light_pixel<MODE>(a, dd, Q);
  } else {
  for (int k = 0; k < r.steps; k += kUnroll) {
    float qx[kUnroll], qy[kUnroll], qz[kUnroll];
    Vec<C> cells[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const float kf = static_cast<float>(k + u);
      qx[u] = fmaf(kf, r.dx, r.sx) - 0.5f; qy[u] = fmaf(kf, r.dy, r.sy) - 0.5f; qz[u] = fmaf(kf, r.dz, r.sz) - 0.5f;
    }
    sample_cells<C, FM, BM, LIGHT, BRICKED>(a, src, qx, qy, qz, cells);
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const bool live = k + u < r.steps;
#pragma unroll
      for (int i = 0; i < C; ++i) cells[u].v[i] = live ? cells[u].v[i] : 0.f;
      blend_step<C, MODE>(acc, cells[u]);
    }
  }
  }
  const int64_t pix = (static_cast<int64_t>(cp.out_index) * a.out_h + y) * a.out_w + x;
  vstore<C>(out + pix * C, acc);
  if constexpr (C >= 3) {
    if (a.l2_partial) {
      const float* t = a.l2_target + pix * 3;
      const float d0 = acc.v[0] - t[0], d1 = acc.v[1] - t[1], d2 = acc.v[2] - t[2];
      l2 = d0 * d0 + d1 * d1 + d2 * d2;
    }
  }
}

// Slot of a work-group's L2 loss partial in a launch whose grid is the tile
// grid (the batched kernels pass their flat tile index instead).
NGT_DEVICE int64_t grid_slot(const VBlock& vb) {
  return (static_cast<int64_t>(vb.z) * gridDim.y + vb.y) * gridDim.x + vb.x;
}

// The work-group's partial of the fused L2 loss (a shuffle tree per
// wavefront, then the 4 wavefronts in order) into a.l2_partial[slot].
NGT_DEVICE void store_l2_partial(const RayArgs& a, float l2, int64_t slot) {
  __shared__ float s_w[4];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) l2 += __shfl_down(l2, o, 64);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = l2;
  __syncthreads();
  if (threadIdx.x == 0) a.l2_partial[slot] = blockDim.x == 64 ? s_w[0] : (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
}

template <int C, int FM, int BM, int MODE, bool LIGHT, bool BRICKED>
NGT_DEVICE void raymarch_fwd_kernel_body(const VBlock& vb, const RayArgs& a, const CamBatch& cb, float* __restrict__ out,
                                         int64_t slot) {
  int x, y;
  pixel_of_thread(vb, x, y);
  float l2 = 0.f;
  if (x < a.out_w && y < a.out_h) raymarch_fwd_pixel<C, FM, BM, MODE, LIGHT, BRICKED>(a, cb, vb, x, y, out, l2);
  if (a.l2_partial) store_l2_partial(a, l2, slot);  // uniform
}

template <int C, int FM, int BM, int MODE, bool LIGHT, bool BRICKED>
__global__ __launch_bounds__(256) void raymarch_fwd_kernel(RayArgs a, CamBatch cb, float* __restrict__ out) {
  const VBlock vb{static_cast<int>(blockIdx.x), static_cast<int>(blockIdx.y), static_cast<int>(blockIdx.z)};
  raymarch_fwd_kernel_body<C, FM, BM, MODE, LIGHT, BRICKED>(vb, a, cb, out, grid_slot(vb));
}

// Per-sample cell gradients of one ray (exact adjoint of the forward walk),
// handed to ``sink.one(k, qx, qy, qz, gc)``.
template <int C, int FM, int BM, int MODE, bool LIGHT, bool BRICKED, typename Sink>
NGT_DEVICE void walk_ray_gradient(const RayArgs& a, const float* __restrict__ src, const Ray& r, Vec<C> g,
                                  const float* __restrict__ out_pix, Sink& sink) {
  if constexpr (C > 1 && MODE == ALPHA) {
    Vec<C> acc = vload<C>(out_pix);
    for (int k = r.steps - 1; k >= 0; --k) {
      const float kf = static_cast<float>(k);
      const float qx = fmaf(kf, r.dx, r.sx) - 0.5f, qy = fmaf(kf, r.dy, r.sy) - 0.5f, qz = fmaf(kf, r.dz, r.sz) - 0.5f;
      const Vec<C> cell = sample_cell<C, FM, BM, LIGHT, BRICKED>(a, src, qx, qy, qz);
      const float aprev = 1.f - (1.f - acc.v[C - 1]) / (1.f - cell.v[C - 1]);
      const float t = 1.f - aprev;
      float lg = 0.f;
#pragma unroll
      for (int i = 0; i < C - 1; ++i) lg += cell.v[i] * g.v[i];
      Vec<C> gc;
#pragma unroll
      for (int i = 0; i < C - 1; ++i) gc.v[i] = t * g.v[i];
      gc.v[C - 1] = t * g.v[C - 1];
      g.v[C - 1] = (1.f - cell.v[C - 1]) * g.v[C - 1] - lg;
      acc.v[C - 1] = aprev;
      sink.one(k, qx, qy, qz, gc);
    }
  } else {
    float total = 0.f;
    if constexpr (C > 1 && MODE == BEER_LAMBERT) {
      const Vec<C> L = vload<C>(out_pix);
#pragma unroll
      for (int i = 0; i < C - 1; ++i) total += L.v[i] * g.v[i];
    }
    float prefix = 0.f, dacc = 0.f;
    auto step = [&](const Vec<C>& cell) {
      Vec<C> gc;
      if constexpr (C == 1 || MODE == ADDITIVE) {
        gc = g;
        (void)cell;
      } else {
        float lg = 0.f;
#pragma unroll
        for (int i = 0; i < C - 1; ++i) lg += cell.v[i] * g.v[i];
        if constexpr (MODE == BEER_LAMBERT) {
          dacc = fmaxf(dacc + cell.v[C - 1], 0.f);
          const float e = exp_neg(dacc);
#pragma unroll
          for (int i = 0; i < C - 1; ++i) gc.v[i] = e * g.v[i];
          gc.v[C - 1] = g.v[C - 1] - (total - prefix);
          prefix += e * lg;
        } else {  // ALPHA_ADDITIVE
          const float t = 1.f - cell.v[C - 1];
#pragma unroll
          for (int i = 0; i < C - 1; ++i) gc.v[i] = t * g.v[i];
          gc.v[C - 1] = g.v[C - 1] - lg;
        }
      }
      return gc;
    };
    constexpr bool kNeedCell = !(C == 1 || MODE == ADDITIVE);
    for (int k = 0; k < r.steps; k += kUnroll) {
      float qx[kUnroll], qy[kUnroll], qz[kUnroll];
      Vec<C> cells[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const float kf = static_cast<float>(k + u);
        qx[u] = fmaf(kf, r.dx, r.sx) - 0.5f; qy[u] = fmaf(kf, r.dy, r.sy) - 0.5f; qz[u] = fmaf(kf, r.dz, r.sz) - 0.5f;
        cells[u] = vzero<C>();
      }
      if constexpr (kNeedCell) sample_cells<C, FM, BM, LIGHT, BRICKED>(a, src, qx, qy, qz, cells);
      // samples past the ray end step as empty cells (state unchanged); only
      // the live ones reach the sink
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const bool live = k + u < r.steps;
#pragma unroll
        for (int i = 0; i < C; ++i) cells[u].v[i] = live ? cells[u].v[i] : 0.f;
        const Vec<C> gc = step(cells[u]);
        if (live) sink.one(k + u, qx[u], qy[u], qz[u], gc);
      }
    }
  }
}

// Cell gradient -> gradient of the stored channels (1 for the fused density
// renderer: d cell / d rho = ds * (rgb, 1)).
template <int C, bool LIGHT>
struct GradChannels {
  static constexpr int value = LIGHT ? 1 : C;
};

template <int C, bool LIGHT>
NGT_DEVICE Vec<GradChannels<C, LIGHT>::value> fold_cell_grad(const RayArgs& a, const Vec<C>& gc) {
  if constexpr (LIGHT) {
    Vec<1> r;
    r.v[0] = a.depth_step * (a.light_r * gc.v[0] + a.light_g * gc.v[1] + a.light_b * gc.v[2] + gc.v[3]);
    return r;
  } else {
    return gc;
  }
}

// Atomic path: scatter each sample's gradient into the volume.
template <int C, int FM, int BM, int MODE, bool LIGHT>
__global__ __launch_bounds__(256) void raymarch_bwd_exact_kernel(RayArgs a, CamBatch cb, const float* __restrict__ out,
                                                                 const float* __restrict__ out_grad,
                                                                 float* __restrict__ in_grad) {
  int x, y;
  pixel_of_thread(VBlock{static_cast<int>(blockIdx.x), static_cast<int>(blockIdx.y), static_cast<int>(blockIdx.z)}, x, y);
  if (x >= a.out_w || y >= a.out_h) return;
  const CamParams& cp = cb.cam[blockIdx.z];
  constexpr int CIN = LIGHT ? 1 : C;
  const int64_t slice = static_cast<int64_t>(a.din.x) * a.din.y * a.din.z * CIN;
  const float* src = a.input + cp.data_index * slice;
  float* gdst = in_grad + cp.data_index * slice;
  const Ray r = make_ray<BM>(a, cp, x, y);
  if (r.steps <= 0) return;
  const int64_t pix = ((static_cast<int64_t>(cp.out_index) * a.out_h + y) * a.out_w + x) * C;
  struct {
    const RayArgs& a;
    const float* src;
    float* gdst;
    NGT_DEVICE void one(int, float qx, float qy, float qz, const Vec<C>& gc) {
      scatter_filtered<CIN, FM, BM>(src, gdst, a.din, qx, qy, qz, fold_cell_grad<C, LIGHT>(a, gc));
    }
  } sink{a, src, gdst};
  walk_ray_gradient<C, FM, BM, MODE, LIGHT, false>(a, src, r, vload<C>(out_grad + pix), out + pix, sink);
}

// ---------------------------------------------------------------------------
// Brick-gather adjoint (BORDER, LINEAR).  Pass 1 stores every sample's
// gradient G (layout at g_offset) and the ray parameters; pass 2 runs one
// work-group per brick of voxels that collects, from the rays whose pixel
// centres fall in the projection of the brick's trilinear support, exactly the
// samples with a corner in the brick - the same terms the scatter would add -
// accumulates them in LDS and writes each voxel once, with no global atomics.

// Recorded per-sample gradients G of one (data, camera) pair: the image's
// pixels (row-major) in groups of kGPix, each group holding its pixels'
// samples in quads of kGQuad consecutive k, (k, pixel) at
//   (pixel / kGPix) * kcap * kGPix + (k / kGQuad) * kGPix * kGQuad
//   + (pixel % kGPix) * kGQuad + k % kGQuad
// (kcap is a multiple of kGQuad).  The recording walk stores a pixel's quad
// with one 16-byte store, so the 8 pixels of a wavefront's tile row write one
// 128-byte line; the gather reads a ray's consecutive samples 4 to a quad.
constexpr int kGPix = 8, kGQuad = 4;
inline __host__ __device__ int64_t g_pixels(int64_t npix) { return (npix + kGPix - 1) / kGPix * kGPix; }
// offset of sample k from sample 0 of the same pixel
NGT_DEVICE int64_t g_step(int k) { return static_cast<int64_t>(k / kGQuad) * (kGPix * kGQuad) + k % kGQuad; }
NGT_DEVICE int64_t g_offset(int k, int64_t pixel, int kcap) {
  return (pixel / kGPix) * kcap * kGPix + (pixel % kGPix) * kGQuad + g_step(k);
}

// A ray of the gather adjoint, and how its stored per-sample values become
// gradients: G = g_scale * (stored - g_off) (1 and 0 when the stored values
// are the gradients themselves).
// The 8 floats a gather reads per sample come first and the record is
// padded to 48 bytes, so they are two aligned 16-byte loads.
struct alignas(16) RayRec {
  float sx, sy, sz, dx, dy, dz;
  float g_scale, g_off;
  int steps;    // samples recorded (at most the plan's kcap)
  int pad[3];   // pad[0]: the ray's own sample count (steps < pad[0] would mean a cut ray)
};
// Beside the records, one float per 8x8 pixel tile of each pair: a bound on
// |g_scale * (G - g_off)| over the samples of the tile's rays, from which the
// gather takes its fixed-point unit (a brick reads the bounds of the tiles
// its pixel rectangle touches, 1/64 of the loads of a per-ray bound).

inline __host__ __device__ int bound_tiles(int out_w, int out_h) {
  return ((out_w + kRayTile - 1) / kRayTile) * ((out_h + kRayTile - 1) / kRayTile);
}

// The wavefront's 8x8 tile bound: the largest of its lanes' bounds (NaN
// propagates), stored by lane 0.  Every lane of the wavefront calls it.
NGT_DEVICE void store_tile_bound(const RayArgs& a, float* __restrict__ bound, int pair, int x, int y, float b) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const float t = __shfl_xor(b, o, 64);
    b = (t > b || t != t) ? t : b;
  }
  if ((threadIdx.x & 63) == 0) {
    const int ntx = (a.out_w + kRayTile - 1) / kRayTile;
    bound[static_cast<int64_t>(pair) * bound_tiles(a.out_w, a.out_h) + (y / kRayTile) * ntx + x / kRayTile] = b;
  }
}

// Image gradient of the fused L2 loss at one pixel (img: the pixel's C
// channels, tpix: its first target value).
template <int C>
NGT_DEVICE Vec<C> l2_image_grad(const RayArgs& a, const float* __restrict__ img, int64_t tpix) {
  const float s = 2.f * a.l2_scale * (a.l2_loss_grad ? a.l2_loss_grad[0] : 1.f);
  const Vec<C> v = vload<C>(img);
  Vec<C> g = vzero<C>();
#pragma unroll
  for (int c = 0; c < (C < 3 ? C : 3); ++c) g.v[c] = s * (v.v[c] - a.l2_target[tpix + c]);
  return g;
}

template <int C, int BM, int MODE, bool LIGHT, bool BRICKED>
NGT_DEVICE float raymarch_bwd_prepass_pixel(const VBlock& vb, const RayArgs& a, const CamBatch& cb, int pair_base,
                                            int x, int y, const float* __restrict__ out,
                                            const float* __restrict__ out_grad, float* __restrict__ G,
                                            RayRec* __restrict__ rays, int kcap) {
  const CamParams& cp = cb.cam[vb.z];
  constexpr int CIN = LIGHT ? 1 : C;
  constexpr int CG = GradChannels<C, LIGHT>::value;
  const float* src = ray_source<CIN, BRICKED>(a, cp.data_index);
  Ray r = make_ray<BM>(a, cp, x, y);
  const int all_steps = r.steps;
  r.steps = min(r.steps, kcap);
  const int pair = pair_base + static_cast<int>(vb.z);
  const int64_t npix = static_cast<int64_t>(a.out_w) * a.out_h;
  const int64_t pixel = static_cast<int64_t>(y) * a.out_w + x;
  RayRec rec{};
  rec.sx = r.sx; rec.sy = r.sy; rec.sz = r.sz; rec.dx = r.dx; rec.dy = r.dy; rec.dz = r.dz;
  rec.steps = r.steps;
  rec.pad[0] = all_steps;
  rec.g_scale = 1.f;
  rec.g_off = 0.f;
  rays[pair * npix + pixel] = rec;
  if (r.steps <= 0) return 0.f;
  const int64_t pix = ((static_cast<int64_t>(cp.out_index) * a.out_h + y) * a.out_w + x) * C;
  struct {
    const RayArgs& a;
    float* gp;
    int64_t pixel;
    int kcap;
    float gmax;
    NGT_DEVICE void one(int k, float, float, float, const Vec<C>& gc) {
      const Vec<CG> g = fold_cell_grad<C, LIGHT>(a, gc);
      vstore<CG>(gp + g_offset(k, pixel, kcap) * CG, g);
#pragma unroll
      for (int i = 0; i < CG; ++i) gmax = fmaxf(gmax, fabsf(g.v[i])) + (g.v[i] - g.v[i]);  // NaN stays NaN
    }
  } sink{a, G + static_cast<int64_t>(pair) * kcap * g_pixels(npix) * CG, pixel, kcap, 0.f};
  const Vec<C> g_img = a.l2_target ? l2_image_grad<C>(a, out + pix, pix / C * 3) : vload<C>(out_grad + pix);
  walk_ray_gradient<C, LINEAR, BM, MODE, LIGHT, BRICKED>(a, src, r, g_img, out + pix, sink);
  return sink.gmax;  // the gather's fixed-point unit comes from this bound
}

template <int C, int BM, int MODE, bool LIGHT, bool BRICKED>
NGT_DEVICE void raymarch_bwd_prepass_kernel_body(const VBlock& vb, const RayArgs& a, const CamBatch& cb, int pair_base,
                                                 const float* __restrict__ out, const float* __restrict__ out_grad,
                                                 float* __restrict__ G, RayRec* __restrict__ rays,
                                                 float* __restrict__ bound, int kcap) {
  int x, y;
  pixel_of_thread(vb, x, y);
  float b = 0.f;
  if (x < a.out_w && y < a.out_h)
    b = raymarch_bwd_prepass_pixel<C, BM, MODE, LIGHT, BRICKED>(vb, a, cb, pair_base, x, y, out, out_grad, G, rays,
                                                                 kcap);
  store_tile_bound(a, bound, pair_base + static_cast<int>(vb.z), x, y, b);
}

template <int C, int BM, int MODE, bool LIGHT, bool BRICKED>
__global__ __launch_bounds__(256) void raymarch_bwd_prepass_kernel(RayArgs a, CamBatch cb, int pair_base,
                                                                   const float* __restrict__ out,
                                                                   const float* __restrict__ out_grad,
                                                                   float* __restrict__ G, RayRec* __restrict__ rays,
                                                                   float* __restrict__ bound, int kcap) {
  raymarch_bwd_prepass_kernel_body<C, BM, MODE, LIGHT, BRICKED>(VBlock{static_cast<int>(blockIdx.x), static_cast<int>(blockIdx.y), static_cast<int>(blockIdx.z)}, a, cb, pair_base, out, out_grad, G, rays, bound, kcap);
}

// The fused render + L2 forward that also records what the gather adjoint
// needs: each ray's record and its per-sample gradients of the loss for a
// unit upstream gradient (the backward scales the volume gradient by the
// actual one).  A ray's second walk runs right after its first, while its
// samples are still in the caches.
// One-walk recording (fused density render, BEER_LAMBERT or ADDITIVE, L2
// loss on RGB only, so the optical-depth channel's image gradient is 0).  With
// s_j = rho_j ds the sample's cell is (light s_j, s_j), and the exact blend
// adjoint gives, for the image gradient g,
//   d loss / d rho_k = ds <light, g> (e_k - sum_{j>=k} e_j s_j)     (BEER_LAMBERT)
//                    = ds <light, g> (h_k - Q_N),  h_k = e_k + Q_{k-1},
// with e_j = exp(-d_j) and Q_k = sum_{j<=k} e_j s_j; for ADDITIVE it is
// ds <light, g> (h_k = 1, Q_N = 0).  The walk stores h_k per sample and the
// ray record carries g_scale = ds <light, g> (g for a unit loss gradient) and
// g_off = Q_N, known once the walk ends.  The image and its loss partial are
// exactly those of raymarch_fwd_kernel.
template <int MODE>
NGT_DEVICE void raymarch_fwd_record_pixel(const RayArgs& a, const CamBatch& cb, const VBlock& vb, int x, int y,
                                          float* __restrict__ out, float& l2, int pair_base, float* __restrict__ G,
                                          RayRec* __restrict__ rays, float& bnd, int kcap) {
  static_assert(MODE == BEER_LAMBERT || MODE == ADDITIVE, "one-walk recording: BEER_LAMBERT or ADDITIVE");
  const CamParams& cp = cb.cam[vb.z];
  const float* src = ray_source<1, true>(a, cp.data_index);
  const Ray r = make_ray<BORDER>(a, cp, x, y);
  const int64_t npix = static_cast<int64_t>(a.out_w) * a.out_h;
  const int64_t pixel = static_cast<int64_t>(y) * a.out_w + x;
  const int pair = pair_base + vb.z;
  float* gp = G + static_cast<int64_t>(pair) * kcap * g_pixels(npix) + g_offset(0, pixel, kcap);
  const int rec_steps = min(r.steps, kcap);
  float dd = 0.f, q = 0.f;  // optical depth, Q_{k-1}
  float hmax = 0.f;
  constexpr int U = kRecordUnroll;
  static_assert(U == kGQuad, "a round of the recording walk stores one quad of samples");
  for (int k = 0; k < r.steps; k += U) {
    float qx[U], qy[U], qz[U];
    bool live[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const float kf = static_cast<float>(k + u);
      qx[u] = fmaf(kf, r.dx, r.sx) - 0.5f; qy[u] = fmaf(kf, r.dy, r.sy) - 0.5f; qz[u] = fmaf(kf, r.dz, r.sz) - 0.5f;
      live[u] = k + u < r.steps;
    }
    BrickedTaps<1, U> taps;
    bricked_taps_issue<1, BORDER, U>(src, a.din, a.bd, qx, qy, qz, taps);
    Vec<1> raw[U];
    bricked_taps_finish<1, BORDER, U>(taps, raw, live);  // samples past the end are empty
    Vec<4> hq;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const float q_prev = q;
      const float e = light_blend_step<MODE>(dd, q, light_cell(a, raw[u].v[0]).v[3]);
      const float h = MODE == BEER_LAMBERT ? e + q_prev : 1.f;
      hq.v[u] = h;
      // a NaN h needs no propagation here: h = e + Q_{k-1} with e in [0, 1],
      // so it is NaN only when Q is, and the bound includes |Q_N|
      hmax = fmaxf(hmax, fabsf(h));
    }
    // the whole quad: k is a multiple of 4 below rec_steps <= kcap, and kcap
    // is a multiple of 4, so the quad lies inside the pixel's kcap samples;
    // its samples past rec_steps are never read
    if (k < rec_steps) vstore<4>(gp + g_step(k), hq);
  }
  const Vec<4> acc = light_pixel<MODE>(a, dd, q);
  const int64_t pix = (static_cast<int64_t>(cp.out_index) * a.out_h + y) * a.out_w + x;
  vstore<4>(out + pix * 4, acc);
  const float* t = a.l2_target + pix * 3;
  const float d0 = acc.v[0] - t[0], d1 = acc.v[1] - t[1], d2 = acc.v[2] - t[2];
  l2 = d0 * d0 + d1 * d1 + d2 * d2;
  RayRec rec{};
  rec.sx = r.sx; rec.sy = r.sy; rec.sz = r.sz; rec.dx = r.dx; rec.dy = r.dy; rec.dz = r.dz;
  rec.steps = rec_steps;
  rec.pad[0] = r.steps;
  rec.g_scale = a.depth_step * 2.f * a.l2_scale * (a.light_r * d0 + a.light_g * d1 + a.light_b * d2);
  rec.g_off = MODE == BEER_LAMBERT ? q : 0.f;
  rays[static_cast<int64_t>(pair) * npix + pixel] = rec;
  // bound on |g_scale (h_k - g_off)|: the gather's fixed-point unit comes from it
  bnd = fabsf(rec.g_scale) * (hmax + fabsf(rec.g_off));
}

template <int MODE>
NGT_DEVICE void raymarch_fwd_record_kernel_body(const VBlock& vb, const RayArgs& a, const CamBatch& cb,
                                                float* __restrict__ out, int pair_base, float* __restrict__ G,
                                                RayRec* __restrict__ rays, float* __restrict__ bound, int kcap,
                                                int64_t slot) {
  if constexpr (MODE == BEER_LAMBERT || MODE == ADDITIVE) {
    int x, y;
    pixel_of_thread(vb, x, y);
    float l2 = 0.f;
    float b = 0.f;
    if (x < a.out_w && y < a.out_h) raymarch_fwd_record_pixel<MODE>(a, cb, vb, x, y, out, l2, pair_base, G, rays, b, kcap);
    store_tile_bound(a, bound, pair_base + vb.z, x, y, b);
    store_l2_partial(a, l2, slot);
  } else {  // ALPHA blends: the image first, then the adjoint walk of each ray
    raymarch_fwd_kernel_body<4, LINEAR, BORDER, MODE, true, true>(vb, a, cb, out, slot);
    raymarch_bwd_prepass_kernel_body<4, BORDER, MODE, true, true>(vb, a, cb, pair_base, out, nullptr, G, rays, bound,
                                                                   kcap);
  }
}

template <int MODE>
__global__ __launch_bounds__(256) void raymarch_fwd_record_kernel(RayArgs a, CamBatch cb, float* __restrict__ out,
                                                                  int pair_base, float* __restrict__ G,
                                                                  RayRec* __restrict__ rays, float* __restrict__ bound,
                                                                  int kcap) {
  const VBlock vb{static_cast<int>(blockIdx.x), static_cast<int>(blockIdx.y), static_cast<int>(blockIdx.z)};
  raymarch_fwd_record_kernel_body<MODE>(vb, a, cb, out, pair_base, G, rays, bound, kcap, grid_slot(vb));
}

// Inclusive scan of one int per work-item over a 256-wide work-group: a
// shuffle scan inside each wavefront, then the wavefront totals through LDS
// (one barrier).  The caller synchronises before s_wtot is written again.
NGT_DEVICE int block_inclusive_scan(int v, int* s_wtot) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  if (lane == 63) s_wtot[wave] = v;
  __syncthreads();
  int pre = 0;
#pragma unroll
  for (int w = 0; w < 3; ++w)
    if (w < wave) pre += s_wtot[w];
  return v + pre;
}

// Pass 2: one work-group per brick of BX x BY x BZ voxels.  For every view it
// projects the brick's sample-support box (positions whose trilinear corners
// can fall in the brick: p in [b0 - 0.5, b1 + 1.5)) to a pixel rectangle,
// walks the rays of those pixels over the k interval inside the box, and adds
// each sample's stored gradient times its corner weights to the brick's
// voxels.  The brick owns its voxels, so the result is written with plain
// stores; a sample on a brick boundary is visited by each brick it touches
// and contributes only that brick's corners.
//
// The sums are kept in LDS as 64-bit fixed point and added with integer LDS
// atomics (ds_add_u64, at the LDS's full rate; float LDS atomics run some
// 30x slower here).  The brick's unit is 2^-S with S chosen from the largest
// gradient bound of the 8x8 pixel tiles its rays belong to (the bound array,
// written by the walk that recorded the gradients), so that every corner
// contribution rounds to
// an integer below 2^30 in magnitude: the rounding error of a contribution is
// at most 2^-30 of the brick's largest sample gradient, and integer sums are
// exact and independent of the order of the atomics (the gradient is
// reproducible bit for bit).  A non-finite bound makes the brick's gradient
// NaN.
template <int CG, int BX, int BY, int BZ>
NGT_DEVICE void raymarch_bwd_brick_kernel_body(const VBlock& vb, const RayArgs& a, const CamBatch& cb, int pair_base,
                                                                 const RayRec* __restrict__ rays,
                                                                 const float* __restrict__ bound,
                                                                 const float* __restrict__ G, int kcap,
                                                                 float* __restrict__ in_grad, int accumulate,
                                                                 const float* __restrict__ grad_scale) {
  constexpr int NV = BX * BY * BZ;
  // accumulator: the brick grown by one cell on every side, so that the 8
  // corners of every sample whose floor cell lies in b0-1 .. b1 land in it
  constexpr int PX = BX + 2, PY = BY + 2, PZ = BZ + 2, NP = PX * PY * PZ;
  __shared__ unsigned long long acc[CG][NP];
  __shared__ int rect[kMaxCamsPerLaunch][4];  // x0, y0, nx, ny (nx = 0: view skipped)
  __shared__ int ray_begin[kMaxCamsPerLaunch + 1];
  __shared__ float s_wmax[4];
  const int tid = threadIdx.x;
  const int bricks_x = (a.din.x + BX - 1) / BX, bricks_y = (a.din.y + BY - 1) / BY;
  const int bricks_z = (a.din.z + BZ - 1) / BZ;
  const int bid = vb.x;
  const int n = bid / (bricks_x * bricks_y * bricks_z);
  const int brem = bid % (bricks_x * bricks_y * bricks_z);
  const int b0x = (brem % bricks_x) * BX;
  const int b0y = ((brem / bricks_x) % bricks_y) * BY;
  const int b0z = (brem / (bricks_x * bricks_y)) * BZ;
  const int b1x = min(b0x + BX, a.din.x) - 1, b1y = min(b0y + BY, a.din.y) - 1, b1z = min(b0z + BZ, a.din.z) - 1;
  for (int i = tid; i < CG * NP; i += 256) (&acc[0][0])[i] = 0ull;

  // pixel rectangles of the views: wavefront 0 projects the 8 corners of the
  // support box for up to 8 views at once (lane = view * 8 + corner), then
  // reduces each view's 8 lanes with shuffles
  if (tid < 64) {
    const int ci = tid >> 3, c = tid & 7;
    const bool view = ci < cb.count && cb.cam[ci < cb.count ? ci : 0].data_index == n;
    float umin = 3.0e38f, umax = -3.0e38f, vmin = 3.0e38f, vmax = -3.0e38f;
    int behind = 0;
    if (view) {
      const CamParams& cp = cb.cam[ci];
      const float px = (c & 1) ? static_cast<float>(b1x) + 1.5f : static_cast<float>(b0x) - 0.5f;
      const float py = (c & 2) ? static_cast<float>(b1y) + 1.5f : static_cast<float>(b0y) - 0.5f;
      const float pz = (c & 4) ? static_cast<float>(b1z) + 1.5f : static_cast<float>(b0z) - 0.5f;
      float vx, vy, vz, vw;
      mat4_apply(cp.mv, px, py, pz, 1.f, vx, vy, vz, vw);
      float cx, cy, cz, cw;
      mat4_apply(cp.p, vx, vy, vz, vw, cx, cy, cz, cw);
      if (!(cw > 0.f)) {
        behind = 1;
      } else {
        const float iw = 1.f / cw;
        umin = umax = (cx * iw * 0.5f + 0.5f) * static_cast<float>(a.out_w) - 0.5f;
        vmin = vmax = (cy * iw * 0.5f + 0.5f) * static_cast<float>(a.out_h) - 0.5f;
      }
    }
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) {
      umin = fminf(umin, __shfl_xor(umin, o, 64)); umax = fmaxf(umax, __shfl_xor(umax, o, 64));
      vmin = fminf(vmin, __shfl_xor(vmin, o, 64)); vmax = fmaxf(vmax, __shfl_xor(vmax, o, 64));
      behind |= __shfl_xor(behind, o, 64);
    }
    // every lane of a view's group holds its rectangle; lane 0 of the group
    // stores it, and the groups' ray counts are scanned with shuffles
    int x0 = 0, y0 = 0, nx = 0, ny = 0;
    if (view) {
      int x1 = a.out_w - 1, y1 = a.out_h - 1;
      if (!behind) {
        // clamp before the float -> int conversions (far-off corners project anywhere)
        const float fw = static_cast<float>(a.out_w), fh = static_cast<float>(a.out_h);
        umin = fminf(fmaxf(umin, -2.f), fw + 1.f); umax = fminf(fmaxf(umax, -2.f), fw + 1.f);
        vmin = fminf(fmaxf(vmin, -2.f), fh + 1.f); vmax = fminf(fmaxf(vmax, -2.f), fh + 1.f);
        x0 = max(0, static_cast<int>(ceilf(umin - 1e-3f))); x1 = min(x1, static_cast<int>(floorf(umax + 1e-3f)));
        y0 = max(0, static_cast<int>(ceilf(vmin - 1e-3f))); y1 = min(y1, static_cast<int>(floorf(vmax + 1e-3f)));
      }
      nx = max(0, x1 - x0 + 1);
      ny = max(0, y1 - y0 + 1);
    }
    const int cnt = nx * ny;  // 0 for the groups past cb.count
    int begin = 0, all = 0;
#pragma unroll
    for (int g = 0; g < kMaxCamsPerLaunch; ++g) {
      const int cg = __shfl(cnt, g * 8, 64);
      begin += g < ci ? cg : 0;
      all += cg;
    }
    if (c == 0 && ci < cb.count) {
      rect[ci][0] = x0; rect[ci][1] = y0; rect[ci][2] = nx; rect[ci][3] = ny;
      ray_begin[ci] = begin;
    }
    if (tid == 0) ray_begin[cb.count] = all;
  }
  __syncthreads();

  const int64_t npix = static_cast<int64_t>(a.out_w) * a.out_h;
  const int total = ray_begin[cb.count];
  const int fb0m1x = b0x - 1, fb0m1y = b0y - 1, fb0m1z = b0z - 1;
  // ray r of the brick's list (r < total): its record, and the G offset of its sample 0
  auto fetch_ray = [&](int r, RayRec& rr, int64_t& gofs) {
    int ci = 0;
    while (ci + 1 < cb.count && r >= ray_begin[ci + 1]) ++ci;
    const int local = r - ray_begin[ci];
    const int px = rect[ci][0] + local % rect[ci][2];
    const int py = rect[ci][1] + local / rect[ci][2];
    gofs = static_cast<int64_t>(pair_base + ci) * kcap * g_pixels(npix) +
           g_offset(0, static_cast<int64_t>(py) * a.out_w + px, kcap);
    rr = rays[(pair_base + ci) * static_cast<int>(npix) + py * a.out_w + px];
  };
  // each work-item's ray of the next chunk is loaded one chunk ahead, so that
  // its latency overlaps the current chunk's samples (the first one overlaps
  // the gradient bounds' loads)
  RayRec next_rr{};
  int64_t next_gofs = 0;
  if (tid < total) fetch_ray(tid, next_rr, next_gofs);
  // the brick's fixed-point unit from the largest gradient bound of its rays' tiles
  {
    float m = 0.f;
    const int ntx = (a.out_w + kRayTile - 1) / kRayTile, per_pair = bound_tiles(a.out_w, a.out_h);
    for (int ci = 0; ci < cb.count; ++ci) {
      if (rect[ci][2] == 0) continue;
      const int tx0 = rect[ci][0] / kRayTile, tx1 = (rect[ci][0] + rect[ci][2] - 1) / kRayTile;
      const int ty0 = rect[ci][1] / kRayTile, ty1 = (rect[ci][1] + rect[ci][3] - 1) / kRayTile;
      const int ntiles = (tx1 - tx0 + 1) * (ty1 - ty0 + 1);
      for (int t = tid; t < ntiles; t += 256) {
        const int tx = tx0 + t % (tx1 - tx0 + 1), ty = ty0 + t / (tx1 - tx0 + 1);
        const float bnd = bound[static_cast<int64_t>(pair_base + ci) * per_pair + ty * ntx + tx];
        m = (bnd > m || bnd != bnd) ? bnd : m;  // NaN propagates
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const float t = __shfl_xor(m, o, 64);
      m = (t > m || t != t) ? t : m;
    }
    if ((tid & 63) == 0) s_wmax[tid >> 6] = m;
  }
  __syncthreads();
  float wmax = s_wmax[0];
#pragma unroll
  for (int w = 1; w < 4; ++w) wmax = (s_wmax[w] > wmax || s_wmax[w] != s_wmax[w]) ? s_wmax[w] : wmax;
  const bool finite = wmax < 3.0e38f;  // false for +inf and NaN
  // |contribution| <= wmax < 2^e  ->  |contribution * 2^S| < 2^30 with S = 30 - e
  const int S = wmax > 0.f && finite ? 30 - ilogbf(wmax) - 1 : 0;

  const float fdx = static_cast<float>(a.din.x), fdy = static_cast<float>(a.din.y), fdz = static_cast<float>(a.din.z);
  // the support box widened by 1e-3 cells: far above the rounding of the
  // slab test, so no sample in the box is lost; the per-sample test below
  // decides exactly
  const float lo3[3] = {static_cast<float>(b0x) - 0.501f, static_cast<float>(b0y) - 0.501f,
                        static_cast<float>(b0z) - 0.501f};
  const float hi3[3] = {static_cast<float>(b1x) + 1.501f, static_cast<float>(b1y) + 1.501f,
                        static_cast<float>(b1z) + 1.501f};
  // Rays are taken 256 at a time: each work-item clips one ray to the support
  // box (slab test) and counts its samples there;
  // after a block-wide scan the chunk's samples are dealt out in sub-chunks of
  // kItems, kPer per work-item, and each sample adds its 8 corner
  // contributions to the accumulator.  (A corner at floor+1 where the
  // reference uses ceil = floor at integer positions carries weight exactly 0.)
  // The walk is bound by the latency of its loads at 4 waves / SIMD: 8 samples'
  // G loads in flight per work-item measured fastest at 128^3 x 5 views (52 us;
  // 1: 57.7, 4: 54.4; 16 needs more than 128 VGPRs, which halves the waves).
  constexpr int kItems = 2048, kPer = kItems / 256;
  __shared__ int s_off[257];
  __shared__ int s_kb[256];        // each ray's first k in the box minus its first item
  __shared__ int64_t s_gofs[256];  // G offset of each ray's sample 0
  __shared__ float4 s_ps[256];     // each ray's start and step x
  __shared__ float4 s_pd[256];     // each ray's step y, z and loss-gradient scale, offset
  __shared__ int s_wtot[4];
  __shared__ unsigned char s_ray[kItems];
  for (int base_r = 0; base_r < total; base_r += 256) {
    const int r = base_r + tid;
    const RayRec rr = next_rr;
    const int64_t gofs = r < total ? next_gofs : 0;
    if (r + 256 < total) fetch_ray(r + 256, next_rr, next_gofs);
    int cnt = 0, k0 = 0;
    if (r < total) {
      s_ps[tid] = make_float4(rr.sx, rr.sy, rr.sz, rr.dx);
      s_pd[tid] = make_float4(rr.dy, rr.dz, rr.g_scale, rr.g_off);
      if (rr.steps > 0) {
        const float sv[3] = {rr.sx, rr.sy, rr.sz};
        const float dv[3] = {rr.dx, rr.dy, rr.dz};
        float lo = -1.0e30f, hi = 1.0e30f;
        bool miss = false;
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
          if (fabsf(dv[ax]) > 1e-30f) {
            const float inv = 1.f / dv[ax];
            const float t0 = (lo3[ax] - sv[ax]) * inv, t1 = (hi3[ax] - sv[ax]) * inv;
            lo = fmaxf(lo, fminf(t0, t1));
            hi = fminf(hi, fmaxf(t0, t1));
          } else if (!(sv[ax] > lo3[ax] - 1e-3f && sv[ax] < hi3[ax] + 1e-3f)) {
            miss = true;
          }
        }
        if (!miss && lo <= hi) {
          const float fsteps = static_cast<float>(rr.steps);
          k0 = max(0, static_cast<int>(ceilf(fminf(fmaxf(lo, -1.f), fsteps))));
          const int k1 = min(rr.steps - 1, static_cast<int>(floorf(fminf(fmaxf(hi, -2.f), fsteps))));
          cnt = max(0, k1 - k0 + 1);
        }
      }
    }
    // inclusive block scan of the counts
    const int incl = block_inclusive_scan(cnt, s_wtot);
    s_off[tid + 1] = incl;
    if (tid == 0) s_off[0] = 0;
    s_kb[tid] = k0 - (incl - cnt);
    s_gofs[tid] = gofs;
    __syncthreads();
    const int items = s_off[256];
    for (int i0 = 0; i0 < items; i0 += kItems) {
      const int i_end = min(i0 + kItems, items);
      {  // the ray of every sample of this sub-chunk: each ray marks its own range
        const int r0 = max(s_off[tid], i0), r1 = min(s_off[tid + 1], i_end);
        for (int i = r0; i < r1; ++i) s_ray[i - i0] = static_cast<unsigned char>(tid);
      }
      __syncthreads();
      // a work-item's kPer samples of the sub-chunk: all their G loads are
      // issued before any of their atomics
      float wxs[kPer], wys[kPer], wzs[kPer], gss[kPer], gos[kPer];
      int cell[kPer];
      bool oks[kPer];
      Vec<CG> gvs[kPer];
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int i = i0 + tid + u * 256;
        oks[u] = false;
        if (i < i_end) {
          const int lo_j = s_ray[i - i0];
          const int k = s_kb[lo_j] + i;
          const float4 ps = s_ps[lo_j];
          const float4 pd = s_pd[lo_j];
          const float kf = static_cast<float>(k);
          const float qx = fmaf(kf, ps.w, ps.x) - 0.5f;
          const float qy = fmaf(kf, pd.x, ps.y) - 0.5f;
          const float qz = fmaf(kf, pd.y, ps.z) - 0.5f;
          const float fx = floorf(qx), fy = floorf(qy), fz = floorf(qz);
          wxs[u] = qx - fx; wys[u] = qy - fy; wzs[u] = qz - fz;
          const int lx = static_cast<int>(fmaxf(fminf(fx, 1e8f), -1e8f)) - fb0m1x;
          const int ly = static_cast<int>(fmaxf(fminf(fy, 1e8f), -1e8f)) - fb0m1y;
          const int lz = static_cast<int>(fmaxf(fminf(fz, 1e8f), -1e8f)) - fb0m1z;
          // forward (BORDER) mask, floor cell in b0-1 .. b1 on every axis
          oks[u] = (qx + 0.5f >= 0.f) & (qx + 0.5f < fdx) & (qy + 0.5f >= 0.f) & (qy + 0.5f < fdy) &
                   (qz + 0.5f >= 0.f) & (qz + 0.5f < fdz) & (lx >= 0) & (lx <= b1x - fb0m1x) & (ly >= 0) &
                   (ly <= b1y - fb0m1y) & (lz >= 0) & (lz <= b1z - fb0m1z);
          cell[u] = (lz * PY + ly) * PX + lx;
          gss[u] = pd.z; gos[u] = pd.w;
          if (oks[u]) gvs[u] = vload<CG>(G + (s_gofs[lo_j] + g_step(k)) * CG);
        }
      }
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        if (!oks[u]) continue;
        // corner weights: z, then y, then the x pair
        const float wz0 = 1.f - wzs[u], wy0 = 1.f - wys[u], wx0 = 1.f - wxs[u];
        const float w00 = wz0 * wy0, w01 = wz0 * wys[u], w10 = wzs[u] * wy0, w11 = wzs[u] * wys[u];
        const float wyz[4] = {w00, w01, w10, w11};
        unsigned long long* const p0 = &acc[0][cell[u]];
#pragma unroll
        for (int ch = 0; ch < CG; ++ch) {
          const float g = ldexpf(gss[u] * (gvs[u].v[ch] - gos[u]), S);
#pragma unroll
          for (int yz = 0; yz < 4; ++yz) {
            const float t = g * wyz[yz];
            const int off = ch * NP + (yz >> 1) * PX * PY + (yz & 1) * PX;
            // truncation: at most one unit (2^-S) off, far below the sum's rounding
            const long long c0 = static_cast<long long>(static_cast<int>(t * wx0));
            const long long c1 = static_cast<long long>(static_cast<int>(t * wxs[u]));
            atomicAdd(p0 + off, static_cast<unsigned long long>(c0));
            atomicAdd(p0 + off + 1, static_cast<unsigned long long>(c1));
          }
        }
      }
      __syncthreads();  // s_ray is rewritten by the next sub-chunk
    }
  }
  __syncthreads();
  const int64_t base = static_cast<int64_t>(n) * a.din.x * a.din.y * a.din.z;
  const float scale = grad_scale ? grad_scale[0] : 1.f;
  // the brick's own voxels, kOwn per work-item: every value is formed first,
  // then all are stored together (a store loop that read-added per voxel
  // behind a select waited for each voxel's load in turn: 11 us at 128^3)
  constexpr int kOwn = (NV + 255) / 256;
  float vals[kOwn][CG];
  int64_t offs[kOwn];
  bool own[kOwn];
#pragma unroll
  for (int m = 0; m < kOwn; ++m) {
    const int i = tid + 256 * m;
    const int lx = i % BX, ly = (i / BX) % BY, lz = i / (BX * BY);
    const int x = b0x + lx, y = b0y + ly, z = b0z + lz;
    own[m] = i < NV && x <= b1x && y <= b1y && z <= b1z;
    offs[m] = (base + cell_offset(a.din, min(x, b1x), min(y, b1y), min(z, b1z))) * CG;
    const int pi = ((min(lz, BZ - 1) + 1) * PY + min(ly, BY - 1) + 1) * PX + min(lx, BX - 1) + 1;
#pragma unroll
    for (int c = 0; c < CG; ++c) {
      const long long q = static_cast<long long>(acc[c][pi]);
      vals[m][c] = finite ? ldexpf(static_cast<float>(q), -S) * scale : __int_as_float(0x7fc00000);
    }
  }
  if (accumulate) {
#pragma unroll
    for (int m = 0; m < kOwn; ++m) {
      if (!own[m]) continue;
#pragma unroll
      for (int c = 0; c < CG; ++c) in_grad[offs[m] + c] += vals[m][c];
    }
  } else {
#pragma unroll
    for (int m = 0; m < kOwn; ++m) {
      if (!own[m]) continue;
#pragma unroll
      for (int c = 0; c < CG; ++c) in_grad[offs[m] + c] = vals[m][c];
    }
  }
}

template <int CG, int BX, int BY, int BZ>
__global__ __launch_bounds__(256) void raymarch_bwd_brick_kernel(RayArgs a, CamBatch cb, int pair_base,
                                                                 const RayRec* __restrict__ rays,
                                                                 const float* __restrict__ bound,
                                                                 const float* __restrict__ G, int kcap,
                                                                 float* __restrict__ in_grad, int accumulate,
                                                                 const float* __restrict__ grad_scale) {
  raymarch_bwd_brick_kernel_body<CG, BX, BY, BZ>(VBlock{static_cast<int>(blockIdx.x), static_cast<int>(blockIdx.y), static_cast<int>(blockIdx.z)}, a, cb, pair_base, rays, bound, G, kcap, in_grad, accumulate, grad_scale);
}

// Reference gradient of one camera: reverse walk from `end` with the
// reference's reconstruction, scatter with corner weights into a per-camera
// buffer (raymarch_grid.cu.cc:482-521).
template <int C, int BM, int MODE>
__global__ __launch_bounds__(256) void raymarch_bwd_reference_kernel(RayArgs a, CamBatch cb,
                                                                     const float* __restrict__ out,
                                                                     const float* __restrict__ out_grad,
                                                                     float* __restrict__ gbuf,
                                                                     float* __restrict__ wsum) {
  int x, y;
  pixel_of_thread(VBlock{static_cast<int>(blockIdx.x), static_cast<int>(blockIdx.y), static_cast<int>(blockIdx.z)}, x, y);
  if (x >= a.out_w || y >= a.out_h) return;
  const CamParams& cp = cb.cam[blockIdx.z];
  const float* src = a.input + static_cast<int64_t>(cp.data_index) * a.din.x * a.din.y * a.din.z * C;
  const Ray r = make_ray<BM>(a, cp, x, y);
  if (!r.hit) return;
  const int64_t pix = ((static_cast<int64_t>(cp.out_index) * a.out_h + y) * a.out_w + x) * C;
  Vec<C> acc = vload<C>(out + pix);
  Vec<C> g = vload<C>(out_grad + pix);
  for (int j = 0; j < r.steps; ++j) {
    const float jf = static_cast<float>(j);
    const float px = r.ex - jf * r.dx, py = r.ey - jf * r.dy, pz = r.ez - jf * r.dz;
    const Vec<C> cell = read_linear<C, BM>(src, a.din, px - 0.5f, py - 0.5f, pz - 0.5f);
    const Vec<C> gc = blend_grad_reference<C, MODE>(g, cell, acc);
    scatter_linear<C, BM, true>(gbuf, a.din, px - 0.5f, py - 0.5f, pz - 0.5f, gc, nullptr, wsum);
  }
}

// in_grad += gbuf / wsum where wsum > 0; clears both buffers for the next
// camera (raymarch_grid.cu.cc:458-480, 700-718).
template <int C>
__global__ __launch_bounds__(256) void normalize_by_weight_kernel(float* __restrict__ gbuf, float* __restrict__ wsum,
                                                                  float* __restrict__ in_grad, int64_t cells) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= cells) return;
  const float w = wsum[i];
  if (w > 0.f) {
    const float inv = 1.0f / w;
    const Vec<C> gb = vload<C>(gbuf + i * C);
    Vec<C> gi = vload<C>(in_grad + i * C);
#pragma unroll
    for (int c = 0; c < C; ++c) gi.v[c] += gb.v[c] * inv;
    vstore<C>(in_grad + i * C, gi);
    vstore<C>(gbuf + i * C, vzero<C>());
    wsum[i] = 0.f;
  }
}

#define NGT_SWITCH_C(C_, ...)                              \
  switch (C_) {                                            \
    case 1: { constexpr int CC = 1; __VA_ARGS__; } break;  \
    case 2: { constexpr int CC = 2; __VA_ARGS__; } break;  \
    default: { constexpr int CC = 4; __VA_ARGS__; } break; \
  }
#define NGT_SWITCH_BM(B_, ...)                                      \
  switch (B_) {                                                     \
    case BORDER: { constexpr int BB = BORDER; __VA_ARGS__; } break; \
    case CLAMP: { constexpr int BB = CLAMP; __VA_ARGS__; } break;   \
    default: { constexpr int BB = WRAP; __VA_ARGS__; } break;       \
  }
#define NGT_SWITCH_FM2(F_, ...)                                        \
  switch (F_) {                                                        \
    case LINEAR: { constexpr int FF = LINEAR; __VA_ARGS__; } break;    \
    default: { constexpr int FF = NEAREST; __VA_ARGS__; } break;       \
  }
#define NGT_SWITCH_MODE(M_, ...)                                                   \
  switch (M_) {                                                                    \
    case BEER_LAMBERT: { constexpr int MM = BEER_LAMBERT; __VA_ARGS__; } break;     \
    case ALPHA: { constexpr int MM = ALPHA; __VA_ARGS__; } break;                   \
    case ALPHA_ADDITIVE: { constexpr int MM = ALPHA_ADDITIVE; __VA_ARGS__; } break; \
    default: { constexpr int MM = ADDITIVE; __VA_ARGS__; } break;                   \
  }

int check_ray_args(const int64_t* shape, const float* M, const float* V, const float* P, const float* F,
                   int num_cameras, const int32_t* out_dhw, int interpolation, int boundary, int mode,
                   int separate, std::vector<Pair>& pairs, bool light) {
  int rc = check_grid_shape(shape, light ? "density" : "input");
  if (rc) return rc;
  if (light) NGT_REQUIRE(shape[4] == 1, NGT_ERR_INVALID_ARGUMENT, "density must have 1 channel");
  NGT_REQUIRE(M && V && P && F, NGT_ERR_INVALID_ARGUMENT, "matrices and frustum must be given");
  NGT_REQUIRE(num_cameras > 0, NGT_ERR_INVALID_ARGUMENT, "camera batch size mismatch");
  NGT_REQUIRE(out_dhw && out_dhw[0] > 0 && out_dhw[1] > 0 && out_dhw[2] > 0, NGT_ERR_INVALID_ARGUMENT,
              "Invalid output_shape");
  NGT_REQUIRE(boundary >= 0 && boundary <= 2, NGT_ERR_INVALID_ARGUMENT, "invalid boundary mode");
  NGT_REQUIRE(mode == BEER_LAMBERT || mode == ALPHA || mode == ALPHA_ADDITIVE || mode == ADDITIVE,
              NGT_ERR_INVALID_ARGUMENT, "invalid blend mode");
  NGT_REQUIRE(interpolation >= 0 && interpolation <= 3, NGT_ERR_INVALID_ARGUMENT, "invalid filter mode");
  // MIN / MAX raymarching reads through the reference's lerp-based extrema
  // tracker (sampling_v2.hpp:149-161), whose first update computes
  // FLOAT_MAX + (v - FLOAT_MAX) = 0; not reproduced.
  NGT_REQUIRE(interpolation == LINEAR || interpolation == NEAREST, NGT_ERR_UNIMPLEMENTED,
              "raymarching supports LINEAR and NEAREST interpolation");
  return make_pairs(static_cast<int>(shape[0]), num_cameras, separate != 0, pairs);
}

RayArgs make_ray_args(const float* input, const int64_t* shape, const int32_t* out_dhw) {
  RayArgs a;
  a.input = input;
  a.din = shape5(shape).dims();
  a.bricked = nullptr;
  a.bd = brick_dims(a.din);
  a.n_data = static_cast<int>(shape[0]);
  a.out_d = out_dhw[0]; a.out_h = out_dhw[1]; a.out_w = out_dhw[2];
  a.inv_w = 1.f / a.out_w; a.inv_h = 1.f / a.out_h; a.inv_d = 1.f / a.out_d;
  a.light_r = a.light_g = a.light_b = 0.f;
  a.depth_step = 1.f;
  a.l2_target = nullptr;
  a.l2_loss_grad = nullptr;
  a.l2_scale = 0.f;
  a.l2_partial = nullptr;
  return a;
}

template <typename F>
void for_cam_batches(const std::vector<Pair>& pairs, const float* M, const float* V, const float* P, const float* Fr,
                     F&& fn) {
  for (size_t b = 0; b < pairs.size(); b += kMaxCamsPerLaunch) {
    CamBatch cb;
    cb.count = static_cast<int>(std::min<size_t>(kMaxCamsPerLaunch, pairs.size() - b));
    for (int i = 0; i < cb.count; ++i) {
      const Pair& p = pairs[b + i];
      cb.cam[i] = make_cam(M + 16 * p.n, V + 16 * p.v, P + 16 * p.v, Fr + 6 * p.v, p.n, p.out);
    }
    fn(cb);
  }
}

// Copy of an NDHWC grid in the haloed brick layout.  A work-group writes one
// row of bricks (all bx for one (by, bz)), 256 stored slots per pass, so its
// writes are one contiguous run and the source rows it reads (5 y x 5 z rows
// of the grid) are fetched once by one work-group instead of by the work-groups
// of neighbouring bricks on other XCDs.  Halo slots read the boundary mode's
// value (0 for BORDER, the clamped or wrapped cell otherwise); the 3 padding
// slots of a brick are written as 0.
template <int C, int BM>
NGT_DEVICE void to_bricked_kernel_body(const VBlock& vb, const float* __restrict__ in, Dims d, BrickDims b,
                                       float* __restrict__ out) {
  const int n = vb.y;
  const uint32_t row = static_cast<uint32_t>(vb.x);  // by + nby * bz
  const uint32_t nby = b.nbxy / b.nbx;
  const uint32_t by = row % nby, bz = row / nby;
  const uint32_t first = row * b.nbx * kBrickFloats, count = b.nbx * kBrickFloats;
  const int64_t cells = static_cast<int64_t>(d.x) * d.y * d.z;
  for (uint32_t t = threadIdx.x; t < count; t += 256u) {
    const uint32_t l = t % kBrickFloats, bx = t / kBrickFloats;
    Vec<C> v = vzero<C>();
    if (l < 125u) {
      const uint32_t k = l / 25u, j = (l / 5u) % 5u, i = l % 5u;
      int x = static_cast<int>(4u * bx + i) - 1, y = static_cast<int>(4u * by + j) - 1,
          z = static_cast<int>(4u * bz + k) - 1;
      bool ok = true;
      if constexpr (BM == BORDER) {
        ok = x >= 0 && x < d.x && y >= 0 && y < d.y && z >= 0 && z < d.z;
      } else if constexpr (BM == CLAMP) {
        x = min(max(x, 0), d.x - 1); y = min(max(y, 0), d.y - 1); z = min(max(z, 0), d.z - 1);
      } else {
        x = wrap_index(x, d.x); y = wrap_index(y, d.y); z = wrap_index(z, d.z);
      }
      if (ok) v = vload<C>(in + (n * cells + cell_offset(d, x, y, z)) * C);
    }
    vstore<C>(out + (static_cast<int64_t>(n) * b.slice_cells + first + t) * C, v);
  }
}

template <int C, int BM>
__global__ __launch_bounds__(256) void to_bricked_kernel(const float* __restrict__ in, Dims d, BrickDims b,
                                                         float* __restrict__ out) {
  to_bricked_kernel_body<C, BM>(VBlock{static_cast<int>(blockIdx.x), static_cast<int>(blockIdx.y), 0}, in, d, b, out);
}

// work-groups of a brick copy: one per row of bricks
unsigned brick_rows(const BrickDims& b) { return b.slice_cells / kBrickFloats / b.nbx; }

void launch_to_bricked(const float* in, const Dims& d, const BrickDims& b, int n_data, int C, int boundary, float* out,
                       hipStream_t s) {
  const dim3 grid(brick_rows(b), static_cast<unsigned>(n_data));
  NGT_SWITCH_C(C, NGT_SWITCH_BM(boundary,
      hipLaunchKernelGGL((to_bricked_kernel<CC, BB>), grid, dim3(256), 0, s, in, d, b, out)));
}

// Bytes of the haloed brick copy of a grid; 0 when its slices would not be
// addressable with 32-bit offsets (callers then use the plain kernels).
size_t bricked_bytes(const int64_t* shape, int channels) {
  const Dims d{static_cast<int>(shape[3]), static_cast<int>(shape[2]), static_cast<int>(shape[1])};
  if (!bricked_fits(d, channels)) return 0;
  return static_cast<size_t>(shape[0]) * brick_dims(d).slice_cells * channels * sizeof(float);
}

// ------------------------------------------------------ gather planning ----

// Upper bound of the samples per BORDER ray: steps = int(len / |step|) + 2
// with len <= (cell-centre AABB diagonal + 2).  |step| is the exact minimum
// over the image: when the inverse projection's last row has no x / y terms
// (every OpenGL perspective or orthographic camera, centred or not), the
// view-space depth of an NDC depth does not depend on the pixel, so the
// object-space step s(ix, iy) = (IDXtoOS(ix, iy, zmax + .5) - IDXtoOS(ix, iy,
// .5)) / zmax is affine in the pixel coordinates and |s|^2 is a convex
// quadratic, minimised in closed form over the rectangle of pixel centres
// (interior stationary point or the minimum of each edge); any other
// projection evaluates every pixel.  1 % slack absorbs the kernels' fp32
// rounding of the ray setup.
double min_ray_step(const CamParams& cp, const RayArgs& a) {
  const double zmax = static_cast<double>(std::max(a.out_d - 1, 1));
  auto to_os = [&](double ix, double iy, double iz, double o[3]) {
    const double n[4] = {ix / a.out_w * 2.0 - 1.0, iy / a.out_h * 2.0 - 1.0, iz / a.out_d * 2.0 - 1.0, 1.0};
    double vs[4];
    for (int r = 0; r < 4; ++r) {
      vs[r] = 0.0;
      for (int c = 0; c < 4; ++c) vs[r] += static_cast<double>(cp.p_inv[r * 4 + c]) * n[c];
    }
    for (int r = 0; r < 4; ++r) vs[r] /= vs[3];
    for (int r = 0; r < 3; ++r) {
      o[r] = 0.0;
      for (int c = 0; c < 4; ++c) o[r] += static_cast<double>(cp.mv_inv[r * 4 + c]) * vs[c];
    }
  };
  auto step_at = [&](double ix, double iy, double s[3]) {
    double s0[3], s1[3];
    to_os(ix, iy, 0.5, s0);
    to_os(ix, iy, zmax + 0.5, s1);
    for (int r = 0; r < 3; ++r) s[r] = (s1[r] - s0[r]) / zmax;
  };
  auto dot = [](const double* u, const double* v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
  const double u0 = 0.5, u1 = a.out_w - 0.5, v0 = 0.5, v1 = a.out_h - 0.5;
  double best = 1e300;
  if (cp.p_inv[12] == 0.f && cp.p_inv[13] == 0.f) {
    // s(u, v) = c + u A + v B
    double c[3], A[3], B[3], sa[3], sb[3];
    step_at(0.0, 0.0, c);
    step_at(1.0, 0.0, sa);
    step_at(0.0, 1.0, sb);
    for (int r = 0; r < 3; ++r) { A[r] = sa[r] - c[r]; B[r] = sb[r] - c[r]; }
    const double aa = dot(A, A), ab = dot(A, B), bb = dot(B, B), ac = dot(A, c), bc = dot(B, c);
    auto f = [&](double u, double v) {
      double w[3];
      for (int r = 0; r < 3; ++r) w[r] = c[r] + u * A[r] + v * B[r];
      return dot(w, w);
    };
    auto clampd = [](double x, double lo, double hi) { return std::min(std::max(x, lo), hi); };
    const double det = aa * bb - ab * ab;
    if (det > 1e-12 * aa * bb) {
      const double u = (-ac * bb + bc * ab) / det, v = (-bc * aa + ac * ab) / det;
      if (u >= u0 && u <= u1 && v >= v0 && v <= v1) best = f(u, v);
    }
    for (double u : {u0, u1}) {  // edges u = const: minimise over v
      const double v = bb > 0.0 ? clampd(-(bc + u * ab) / bb, v0, v1) : v0;
      best = std::min(best, f(u, v));
    }
    for (double v : {v0, v1}) {  // edges v = const: minimise over u
      const double u = aa > 0.0 ? clampd(-(ac + v * ab) / aa, u0, u1) : u0;
      best = std::min(best, f(u, v));
    }
  } else {
    for (int y = 0; y < a.out_h; ++y)
      for (int x = 0; x < a.out_w; ++x) {
        double s[3];
        step_at(x + 0.5, y + 0.5, s);
        best = std::min(best, dot(s, s));
      }
  }
  return std::sqrt(std::max(best, 0.0));
}

int ray_step_cap(const std::vector<Pair>& pairs, const float* M, const float* V, const float* P, const float* Fr,
                 const RayArgs& a) {
  double min_step = 1e300;
  for (const Pair& p : pairs) {
    const CamParams cp = make_cam(M + 16 * p.n, V + 16 * p.v, P + 16 * p.v, Fr + 6 * p.v, p.n, p.out);
    const double l = min_ray_step(cp, a);
    if (l > 0.0 && l < min_step) min_step = l;
  }
  const double diag = std::sqrt(static_cast<double>(a.din.x - 1) * (a.din.x - 1) +
                                static_cast<double>(a.din.y - 1) * (a.din.y - 1) +
                                static_cast<double>(a.din.z - 1) * (a.din.z - 1));
  // beyond this the per-sample record would outgrow the grid by far: -1, and
  // the callers take the scatter adjoint (no gather plan) instead
  const int hard_cap = 16 * (a.din.x + a.din.y + a.din.z + a.out_d);
  if (!(min_step < 1e300)) return -1;
  const double k = std::ceil((diag + 2.0) / (0.99 * min_step)) + 3.0;
  if (!(k <= hard_cap)) return -1;
  // a multiple of the record's sample quad (and of the walks' rounds)
  constexpr int kRound = kGQuad > kUnroll ? kGQuad : kUnroll;
  static_assert(kRound % kUnroll == 0 && kRound % kGQuad == 0, "cap granularity");
  const int cap = static_cast<int>(k);
  return (cap + kRound - 1) / kRound * kRound;
}

struct GatherPlan {
  int kcap;
  RayRec* rays;
  float* bound;  // one per ray, beside the records
  float* G;
  float* bricked;  // bricked copy of the sampled grid (LINEAR: always, it follows G)
};

constexpr size_t kAlign = 256;

size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

size_t gather_workspace_bytes(const int64_t* shape, const float* M, const float* V, const float* P, const float* F,
                              int num_cameras, const int32_t* out_dhw, int interpolation, int boundary, int separate,
                              int grad_channels) {
  if (!shape || !M || !V || !P || !F || !out_dhw || num_cameras <= 0) return 0;
  if (interpolation != LINEAR || boundary != BORDER || bricked_bytes(shape, grad_channels) == 0) return 0;
  std::vector<Pair> pairs;
  if (make_pairs(static_cast<int>(shape[0]), num_cameras, separate != 0, pairs) != NGT_OK) return 0;
  const RayArgs a = make_ray_args(nullptr, shape, out_dhw);
  const int kcap = ray_step_cap(pairs, M, V, P, F, a);
  if (kcap <= 0) return 0;
  const size_t npix = static_cast<size_t>(a.out_w) * a.out_h;
  return align_up(pairs.size() * npix * sizeof(RayRec)) +
         align_up(pairs.size() * bound_tiles(a.out_w, a.out_h) * sizeof(float)) +
         align_up(pairs.size() * static_cast<size_t>(kcap) * g_pixels(npix) * grad_channels * sizeof(float)) +
         align_up(bricked_bytes(shape, grad_channels));
}

// with_bricked: the workspace also holds a haloed brick copy for the prepass
// walk (not when the per-sample gradients are recorded by the forward, whose
// gather reads no grid values).
bool plan_gather(const std::vector<Pair>& pairs, const float* M, const float* V, const float* P, const float* F,
                 const RayArgs& a, int interpolation, int boundary, int grad_channels, void* workspace,
                 size_t workspace_bytes, GatherPlan& gp, bool with_bricked = true) {
  if (interpolation != LINEAR || boundary != BORDER || workspace == nullptr || !bricked_fits(a.din, grad_channels))
    return false;
  gp.kcap = ray_step_cap(pairs, M, V, P, F, a);
  if (gp.kcap <= 0) return false;
  const size_t npix = static_cast<size_t>(a.out_w) * a.out_h;
  const size_t rays_bytes = align_up(pairs.size() * npix * sizeof(RayRec));
  const size_t bound_bytes = align_up(pairs.size() * bound_tiles(a.out_w, a.out_h) * sizeof(float));
  const size_t g_bytes =
      align_up(pairs.size() * static_cast<size_t>(gp.kcap) * g_pixels(npix) * grad_channels * sizeof(float));
  const size_t b_bytes =
      with_bricked ? align_up(static_cast<size_t>(a.n_data) * a.bd.slice_cells * grad_channels * sizeof(float)) : 0;
  if (workspace_bytes < rays_bytes + bound_bytes + g_bytes + b_bytes) return false;
  char* w = static_cast<char*>(workspace);
  gp.rays = reinterpret_cast<RayRec*>(w);
  gp.bound = reinterpret_cast<float*>(w + rays_bytes);
  gp.G = reinterpret_cast<float*>(w + rays_bytes + bound_bytes);
  gp.bricked = with_bricked ? reinterpret_cast<float*>(w + rays_bytes + bound_bytes + g_bytes) : nullptr;
  return true;
}

// bricked_in: a haloed brick copy of the input made by the caller (BORDER
// halo), or nullptr to make one in the plan's workspace.
// recorded: the plan's rays and per-sample gradients were written by the
// forward (raymarch_fwd_record_kernel) for a unit loss gradient; only the
// gather runs, scaled by *grad_scale.
template <bool LIGHT>
void run_gather(const std::vector<Pair>& pairs, const float* M, const float* V, const float* P, const float* F,
                const RayArgs& a_in, int blending_mode, int C, const float* output, const float* output_grad,
                const GatherPlan& gp, float* in_grad, int n_data, hipStream_t s, const float* bricked_in = nullptr,
                bool recorded = false, const float* grad_scale = nullptr) {
  RayArgs a = a_in;
  if (recorded) {
    a.bricked = nullptr;  // the gather reads no grid values
  } else if (bricked_in) {
    a.bricked = bricked_in;
  } else {
    launch_to_bricked(a.input, a.din, a.bd, n_data, LIGHT ? 1 : C, BORDER, gp.bricked, s);
    a.bricked = gp.bricked;
  }
  for (size_t b = 0; b < pairs.size(); b += kMaxCamsPerLaunch) {
    CamBatch cb;
    cb.count = static_cast<int>(std::min<size_t>(kMaxCamsPerLaunch, pairs.size() - b));
    for (int i = 0; i < cb.count; ++i) {
      const Pair& p = pairs[b + i];
      cb.cam[i] = make_cam(M + 16 * p.n, V + 16 * p.v, P + 16 * p.v, F + 6 * p.v, p.n, p.out);
    }
    const dim3 rgrid = ray_grid_dims(a.out_w, a.out_h, cb.count);
    const int base = static_cast<int>(b);
    if (recorded) {
      // per-sample gradients recorded by the forward
    } else if constexpr (LIGHT) {
      NGT_SWITCH_MODE(blending_mode,
          hipLaunchKernelGGL((raymarch_bwd_prepass_kernel<4, BORDER, MM, true, true>), rgrid, dim3(kRayBlock), 0, s, a, cb,
                             base, output, output_grad, gp.G, gp.rays, gp.bound, gp.kcap));
    } else {
      NGT_SWITCH_C(C, NGT_SWITCH_MODE(blending_mode,
          hipLaunchKernelGGL((raymarch_bwd_prepass_kernel<CC, BORDER, MM, false, true>), rgrid, dim3(kRayBlock), 0, s, a, cb,
                             base, output, output_grad, gp.G, gp.rays, gp.bound, gp.kcap)));
    }
    // bricks: larger ones visit fewer samples twice (the support box is one
    // cell wider than the brick), but the 64-bit accumulators of 2 or 4
    // channels must fit the LDS
    constexpr int BX = 16, BY = 8, BZ = 8;
    constexpr int BX1 = 16, BY1 = 16, BZ1 = 8;
    const int accumulate = b > 0 ? 1 : 0;
    const int CG = LIGHT ? 1 : C;
    if (CG == 1) {
      const unsigned bricks = blocks_for(a.din.x, BX1) * blocks_for(a.din.y, BY1) * blocks_for(a.din.z, BZ1);
      hipLaunchKernelGGL((raymarch_bwd_brick_kernel<1, BX1, BY1, BZ1>), dim3(bricks * n_data), dim3(256), 0, s, a,
                         cb, base, gp.rays, gp.bound, gp.G, gp.kcap, in_grad, accumulate,
                         recorded ? grad_scale : nullptr);
    } else {
      const unsigned bricks = blocks_for(a.din.x, BX) * blocks_for(a.din.y, BY) * blocks_for(a.din.z, BZ);
      NGT_SWITCH_C(CG,
          hipLaunchKernelGGL((raymarch_bwd_brick_kernel<CC, BX, BY, BZ>), dim3(bricks * n_data), dim3(256), 0, s, a,
                             cb, base, gp.rays, gp.bound, gp.G, gp.kcap, in_grad, accumulate,
                             recorded ? grad_scale : nullptr));
    }
  }
}

// ---------------------------------------------------------------------------
// Fused L2 image loss: sum over pixels and RGB of (image - target)^2, as
// per-work-group partial sums written by the forward kernel (a shuffle tree
// per wavefront, then the 4 wavefronts in order) and one fixed-order final
// sum in double (deterministic).

__global__ __launch_bounds__(256) void l2_finish_kernel(const float* __restrict__ partial, int n, float weight,
                                                        float* __restrict__ loss) {
  __shared__ double s_w[4];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) v += static_cast<double>(partial[i]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) *loss = static_cast<float>(weight * ((s_w[0] + s_w[1]) + (s_w[2] + s_w[3])));
}

// Image gradient of the fused L2 loss for the backward paths that read an
// image gradient buffer.
__global__ __launch_bounds__(256) void l2_image_grad_kernel(RayArgs a, const float* __restrict__ img, int64_t npix,
                                                            float* __restrict__ dimg) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (p >= npix) return;
  vstore<4>(dimg + p * 4, l2_image_grad<4>(a, img + p * 4, p * 3));
}

// ---------------------------------------------------------------------------
// Uneven batches of fused density renders: every instance (one density grid,
// 1-8 views of one image size) in one persistent launch per kernel.  The
// per-sample records of the gather adjoint (RayRec per pixel, one gradient
// factor per sample) are the only large buffer: instances are grouped into
// consecutive chunks whose records fit a byte budget, and the recording and
// gather kernels run chunk by chunk over one shared record region.

// Gather brick shapes (voxels per work-group): 8x8x8, 16x8x8, 16x16x8.
constexpr int kGatherShapes = 3;
constexpr size_t kDefaultRecordBudget = size_t(4) << 30;

struct RenBatchHeader {
  int n, mode, nchunks;
  int tiles_ray, tiles_brickify;
  int tiles_brick[kGatherShapes];
  size_t off_items, off_ray, off_brickify, off_chunk;
  size_t off_brick[kGatherShapes];
  int64_t ws_partial, ws_record, record_bytes, ws_bytes;  // workspace byte offsets and sizes
};

struct RenBatchItem {
  RayArgs a;  // its density pointer is read from `density` below
  CamBatch cb;
  int ntx, nty;        // 8 x 8 pixel tiles per view (one wavefront each)
  int kcap;
  int brick_tiles;     // brick rows of the haloed copy (one work-group each)
  gptr<const float> density;
  gptr<float> images;
  gptr<const float> images_grad;
  gptr<float> density_grad;
  gptr<const float> l2_target;
  gptr<float> loss;
  int64_t ws_bricked;        // byte offset of the haloed copy in the workspace
  int64_t rec_rays, rec_bound, rec_G;  // byte offsets of the records in the record region
};

struct RenBatchView {
  const RenBatchHeader* h;
  const RenBatchItem* items;
  const int* ray;
  const int* brickify;
};

NGT_DEVICE RenBatchView ren_batch_view(const void* plan) {
  RenBatchView v;
  const char* p = as_constant<char>(plan);
  v.h = reinterpret_cast<const RenBatchHeader*>(p);
  v.items = reinterpret_cast<const RenBatchItem*>(p + v.h->off_items);
  v.ray = reinterpret_cast<const int*>(p + v.h->off_ray);
  v.brickify = reinterpret_cast<const int*>(p + v.h->off_brickify);
  return v;
}

NGT_DEVICE VBlock ray_tile(const RenBatchItem& it, int local) {
  const int per_view = it.ntx * it.nty;
  const int view = local / per_view, r = local % per_view;
  return VBlock{r % it.ntx, r / it.ntx, view};
}

// The instance's ray arguments with its haloed copy (and, for the L2 step,
// its target, weight 1 and the loss partials) filled in.
NGT_DEVICE RayArgs batch_ray_args(const RenBatchView& v, const RenBatchItem& it, const char* ws, bool l2) {
  RayArgs a = it.a;
  a.input = (const float*)it.density;
  a.bricked = reinterpret_cast<const float*>(ws + it.ws_bricked);
  a.l2_target = nullptr;
  a.l2_loss_grad = nullptr;
  a.l2_partial = nullptr;
  if (l2) {
    a.l2_target = (const float*)it.l2_target;
    a.l2_scale = 1.f;
    a.l2_partial = reinterpret_cast<float*>(const_cast<char*>(ws) + v.h->ws_partial);
  }
  return a;
}

__global__ __launch_bounds__(256) void render_brickify_batch_kernel(const void* __restrict__ plan, char* __restrict__ ws) {
  const RenBatchView v = ren_batch_view(plan);
  for (int t = blockIdx.x; t < v.h->tiles_brickify; t += gridDim.x) {
    const int i = __builtin_amdgcn_readfirstlane(find_item(v.brickify, v.h->n, t));
    const RenBatchItem& it = v.items[i];
    const int local = t - v.brickify[i];
    to_bricked_kernel_body<1, BORDER>(VBlock{local, 0, 0}, (const float*)it.density, it.a.din, it.a.bd,
                                      reinterpret_cast<float*>(ws + it.ws_bricked));
  }
}

template <int MODE>
__global__ __launch_bounds__(256) void render_fwd_batch_kernel(const void* __restrict__ plan, const char* __restrict__ ws) {
  const RenBatchView v = ren_batch_view(plan);
  for (int t = blockIdx.x; t < v.h->tiles_ray; t += gridDim.x) {
    const int i = __builtin_amdgcn_readfirstlane(find_item(v.ray, v.h->n, t));
    const RenBatchItem& it = v.items[i];
    raymarch_fwd_kernel_body<4, LINEAR, BORDER, MODE, true, true>(ray_tile(it, t - v.ray[i]),
                                                                  batch_ray_args(v, it, ws, false), it.cb,
                                                                  (float*)it.images, t);
  }
}

// Ray tiles t0 .. t1 (one chunk): the backward's walk that records each
// sample's gradient for the given image gradients.
template <int MODE>
__global__ __launch_bounds__(256) void render_prepass_batch_kernel(const void* __restrict__ plan, char* __restrict__ ws,
                                                                   int t0, int t1) {
  const RenBatchView v = ren_batch_view(plan);
  char* rec = ws + v.h->ws_record;
  for (int t = t0 + static_cast<int>(blockIdx.x); t < t1; t += gridDim.x) {
    const int i = __builtin_amdgcn_readfirstlane(find_item(v.ray, v.h->n, t));
    const RenBatchItem& it = v.items[i];
    raymarch_bwd_prepass_kernel_body<4, BORDER, MODE, true, true>(
        ray_tile(it, t - v.ray[i]), batch_ray_args(v, it, ws, false), it.cb, 0, (float*)it.images,
        (const float*)it.images_grad,
        reinterpret_cast<float*>(rec + it.rec_G), reinterpret_cast<RayRec*>(rec + it.rec_rays),
        reinterpret_cast<float*>(rec + it.rec_bound), it.kcap);
  }
}

// Ray tiles t0 .. t1 (one chunk) of the L2 step: the image, its loss partial
// (slot t) and the recorded per-sample factors in one walk.
template <int MODE>
__global__ __launch_bounds__(256) void render_record_batch_kernel(const void* __restrict__ plan, char* __restrict__ ws,
                                                                  int t0, int t1) {
  const RenBatchView v = ren_batch_view(plan);
  char* rec = ws + v.h->ws_record;
  for (int t = t0 + static_cast<int>(blockIdx.x); t < t1; t += gridDim.x) {
    const int i = __builtin_amdgcn_readfirstlane(find_item(v.ray, v.h->n, t));
    const RenBatchItem& it = v.items[i];
    raymarch_fwd_record_kernel_body<MODE>(ray_tile(it, t - v.ray[i]), batch_ray_args(v, it, ws, true), it.cb,
                                          (float*)it.images, 0, reinterpret_cast<float*>(rec + it.rec_G),
                                          reinterpret_cast<RayRec*>(rec + it.rec_rays),
                                          reinterpret_cast<float*>(rec + it.rec_bound), it.kcap, t);
    __syncthreads();  // the loss partial's LDS is written again by the next tile
  }
}

// Brick tiles t0 .. t1 (one chunk) of gather shape SHAPE.
template <int SHAPE>
__global__ __launch_bounds__(256) void render_brick_batch_kernel(const void* __restrict__ plan, char* __restrict__ ws,
                                                                 int t0, int t1) {
  constexpr int BX = SHAPE == 0 ? 8 : 16, BY = SHAPE == 2 ? 16 : 8, BZ = 8;
  const RenBatchView v = ren_batch_view(plan);
  const int* table = reinterpret_cast<const int*>(as_constant<char>(plan) + v.h->off_brick[SHAPE]);
  const char* rec = ws + v.h->ws_record;
  for (int t = t0 + static_cast<int>(blockIdx.x); t < t1; t += gridDim.x) {
    const int i = __builtin_amdgcn_readfirstlane(find_item(table, v.h->n, t));
    const RenBatchItem& it = v.items[i];
    raymarch_bwd_brick_kernel_body<1, BX, BY, BZ>(VBlock{t - table[i], 0, 0}, it.a, it.cb, 0,
                                                  reinterpret_cast<const RayRec*>(rec + it.rec_rays),
                                                  reinterpret_cast<const float*>(rec + it.rec_bound),
                                                  reinterpret_cast<const float*>(rec + it.rec_G), it.kcap,
                                                  (float*)it.density_grad, 0, nullptr);
    __syncthreads();  // the brick's LDS arrays are rewritten by the next brick
  }
}

// One work-group per instance: its loss, the fixed-order double sum of its
// ray tiles' partials.
__global__ __launch_bounds__(256) void render_l2_finish_batch_kernel(const void* __restrict__ plan,
                                                                     const char* __restrict__ ws) {
  const RenBatchView v = ren_batch_view(plan);
  const int i = blockIdx.x;
  const float* partial = reinterpret_cast<const float*>(ws + v.h->ws_partial);
  __shared__ double s_w[4];
  double sum = 0.0;
  for (int t = v.ray[i] + static_cast<int>(threadIdx.x); t < v.ray[i + 1]; t += 256) sum += static_cast<double>(partial[t]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) sum += __shfl_down(sum, o, 64);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = sum;
  __syncthreads();
  if (threadIdx.x == 0) *(float*)v.items[i].loss = static_cast<float>((s_w[0] + s_w[1]) + (s_w[2] + s_w[3]));
}

// Pixels of the N x V images of a fused render.
int64_t render_pixels(const int64_t* density_shape, int num_cameras, const int32_t* output_dhw) {
  return density_shape[0] * num_cameras * static_cast<int64_t>(output_dhw[1]) * output_dhw[2];
}

// The fused density render forward.  bricked_in: a haloed brick copy of
// density made by the caller with this boundary (nullptr: made here in the
// workspace when it is large enough).  l2_target / l2_partial: the fused L2
// loss's target and one partial per forward work-group (both or neither).
int render_density_fwd_impl(const float* density, const int64_t* density_shape, const float* light_rgb,
                            float depth_step, const float* matrix_m, const float* matrix_v, const float* matrix_p,
                            const float* frustum, int num_cameras, const int32_t* output_dhw, int interpolation,
                            int boundary, int blending_mode, float* output, void* workspace, size_t workspace_bytes,
                            const float* bricked_in, const float* l2_target, float* l2_partial, ngt_stream_t stream,
                            void* record = nullptr, size_t record_bytes = 0, float l2_scale = 0.f) {
  std::vector<Pair> pairs;
  int rc = check_ray_args(density_shape, matrix_m, matrix_v, matrix_p, frustum, num_cameras, output_dhw, interpolation,
                          boundary, blending_mode, 1, pairs, true);
  if (rc) return rc;
  NGT_REQUIRE(light_rgb != nullptr, NGT_ERR_INVALID_ARGUMENT, "light_rgb must be given");
  RayArgs a = make_ray_args(density, density_shape, output_dhw);
  a.light_r = light_rgb[0]; a.light_g = light_rgb[1]; a.light_b = light_rgb[2];
  a.depth_step = depth_step;
  hipStream_t s = as_stream(stream);
  const size_t need = ngt_render_density_fwd_workspace_size(density_shape, interpolation);
  const bool bricked = bricked_in != nullptr || (need > 0 && workspace != nullptr && workspace_bytes >= need);
  if (bricked_in) {
    a.bricked = bricked_in;
  } else if (bricked) {
    launch_to_bricked(density, a.din, a.bd, a.n_data, 1, boundary, static_cast<float*>(workspace), s);
    a.bricked = static_cast<const float*>(workspace);
  }
  a.l2_target = l2_target;
  a.l2_scale = l2_scale;  // the recorded image gradient 2 * l2_scale * (image - target)
  // record: the backward's gather plan filled by the forward (fused L2 only)
  GatherPlan gp;
  if (record) {
    NGT_REQUIRE(l2_target != nullptr && bricked && interpolation == LINEAR && boundary == BORDER,
                NGT_ERR_INVALID_ARGUMENT, "recording the per-sample gradients needs LINEAR, BORDER and the L2 loss");
    NGT_REQUIRE(plan_gather(pairs, matrix_m, matrix_v, matrix_p, frustum, a, interpolation, boundary, 1, record,
                            record_bytes, gp, false),
                NGT_ERR_WORKSPACE, "record buffer too small (ngt_render_density_l2_record_size)");
  }
  int64_t partial_base = 0;
  int pair_base = 0;
  for_cam_batches(pairs, matrix_m, matrix_v, matrix_p, frustum, [&](const CamBatch& cb) {
    const dim3 grid = ray_grid_dims(a.out_w, a.out_h, cb.count);
    a.l2_partial = l2_partial ? l2_partial + partial_base : nullptr;
    partial_base += static_cast<int64_t>(grid.x) * grid.y * grid.z;
    const int base = pair_base;
    pair_base += cb.count;
    if (record) {
      NGT_SWITCH_MODE(blending_mode,
          hipLaunchKernelGGL((raymarch_fwd_record_kernel<MM>), grid, dim3(kRayBlock), 0, s, a, cb, output, base, gp.G,
                             gp.rays, gp.bound, gp.kcap));
    } else if (bricked) {
      NGT_SWITCH_BM(boundary, NGT_SWITCH_MODE(blending_mode,
          hipLaunchKernelGGL((raymarch_fwd_kernel<4, LINEAR, BB, MM, true, true>), grid, dim3(kRayBlock), 0, s, a, cb,
                             output)));
    } else {
      NGT_SWITCH_FM2(interpolation, NGT_SWITCH_BM(boundary, NGT_SWITCH_MODE(blending_mode,
          hipLaunchKernelGGL((raymarch_fwd_kernel<4, FF, BB, MM, true, false>), grid, dim3(kRayBlock), 0, s, a, cb,
                             output))));
    }
  });
  return check_launch("ngt_render_density_fwd");
}

}  // namespace
}  // namespace ngt

// ---- allocation-free fused render + L2 forward ----------------------------
// The scratch of ngt_render_density_l2_fwd (the haloed brick copy of the
// density and the per-work-group loss partials) lives in a file-scope device
// array instead of a caller workspace, for callers that cannot allocate
// device memory.  Each GPU the library runs on holds its own copy of the
// array until the process ends; host code addresses it only through the
// pointer hipGetSymbolAddress returns for the current device (the array's
// name is a host-side stand-in).  Every call on a GPU shares that GPU's
// array: calls that overlap on one GPU race.  Nothing in it needs clearing:
// the brick copy and every partial are written before they are read.
constexpr size_t kStaticScratchBytes = size_t(48) << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace ngt {
namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// The current device's address of `scratch` (looked up once per device).
int scratch_for_current_device(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  NGT_REQUIRE(e == hipSuccess, NGT_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  NGT_REQUIRE(dev >= 0 && dev < kMaxScratchDevices, NGT_ERR_HIP, "device index out of range for the static scratch");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    NGT_REQUIRE(e == hipSuccess && p != nullptr, NGT_ERR_HIP,
                std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return NGT_OK;
}
}  // namespace
}  // namespace ngt


using namespace ngt;

extern "C" {

size_t ngt_raymarch_grid_transform_fwd_workspace_size(const int64_t* input_shape, int interpolation) {
  if (input_shape == nullptr || interpolation != LINEAR) return 0;
  return bricked_bytes(input_shape, static_cast<int>(input_shape[4]));
}

int ngt_raymarch_grid_transform_fwd(const float* input, const int64_t* input_shape, const float* matrix_m,
                                    const float* matrix_v, const float* matrix_p, const float* frustum,
                                    int num_cameras, const int32_t* output_dhw, int interpolation, int boundary,
                                    int blending_mode, int separate_camera_batch, float* output, void* workspace,
                                    size_t workspace_bytes, ngt_stream_t stream) {
  std::vector<Pair> pairs;
  int rc = check_ray_args(input_shape, matrix_m, matrix_v, matrix_p, frustum, num_cameras, output_dhw, interpolation,
                          boundary, blending_mode, separate_camera_batch, pairs, false);
  if (rc) return rc;
  RayArgs a = make_ray_args(input, input_shape, output_dhw);
  hipStream_t s = as_stream(stream);
  const int C = static_cast<int>(input_shape[4]);
  const size_t need = ngt_raymarch_grid_transform_fwd_workspace_size(input_shape, interpolation);
  const bool bricked = need > 0 && workspace != nullptr && workspace_bytes >= need;
  if (bricked) {
    launch_to_bricked(input, a.din, a.bd, a.n_data, C, boundary, static_cast<float*>(workspace), s);
    a.bricked = static_cast<const float*>(workspace);
  }
  for_cam_batches(pairs, matrix_m, matrix_v, matrix_p, frustum, [&](const CamBatch& cb) {
    const dim3 grid = ray_grid_dims(a.out_w, a.out_h, cb.count);
    if (bricked) {
      NGT_SWITCH_C(C, NGT_SWITCH_BM(boundary, NGT_SWITCH_MODE(blending_mode,
          hipLaunchKernelGGL((raymarch_fwd_kernel<CC, LINEAR, BB, MM, false, true>), grid, dim3(kRayBlock), 0, s, a, cb,
                             output))));
    } else {
      NGT_SWITCH_C(C, NGT_SWITCH_FM2(interpolation, NGT_SWITCH_BM(boundary, NGT_SWITCH_MODE(blending_mode,
          hipLaunchKernelGGL((raymarch_fwd_kernel<CC, FF, BB, MM, false, false>), grid, dim3(kRayBlock), 0, s, a, cb,
                             output)))));
    }
  });
  return check_launch("ngt_raymarch_grid_transform_fwd");
}

size_t ngt_raymarch_grid_transform_bwd_workspace_size(const int64_t* input_shape, const float* matrix_m,
                                                      const float* matrix_v, const float* matrix_p,
                                                      const float* frustum, int num_cameras,
                                                      const int32_t* output_dhw, int interpolation, int boundary,
                                                      int separate_camera_batch, int gradient_semantics) {
  if (input_shape == nullptr) return 0;
  if (gradient_semantics == NGT_GRAD_REFERENCE) {
    const int64_t cells = input_shape[1] * input_shape[2] * input_shape[3];
    return static_cast<size_t>(cells * (input_shape[4] + 1)) * sizeof(float);
  }
  return gather_workspace_bytes(input_shape, matrix_m, matrix_v, matrix_p, frustum, num_cameras, output_dhw,
                                interpolation, boundary, separate_camera_batch, static_cast<int>(input_shape[4]));
}

int ngt_raymarch_step_cap(const int64_t* input_shape, const float* matrix_m, const float* matrix_v,
                          const float* matrix_p, const float* frustum, int num_cameras, const int32_t* output_dhw,
                          int separate_camera_batch) {
  if (!input_shape || !matrix_m || !matrix_v || !matrix_p || !frustum || !output_dhw || num_cameras <= 0) return -1;
  std::vector<Pair> pairs;
  if (make_pairs(static_cast<int>(input_shape[0]), num_cameras, separate_camera_batch != 0, pairs) != NGT_OK)
    return -1;
  return ray_step_cap(pairs, matrix_m, matrix_v, matrix_p, frustum, make_ray_args(nullptr, input_shape, output_dhw));
}

int ngt_raymarch_grid_transform_bwd(const float* input, const int64_t* input_shape, const float* output,
                                    const float* output_grad, const float* matrix_m, const float* matrix_v,
                                    const float* matrix_p, const float* frustum, int num_cameras,
                                    const int32_t* output_dhw, int interpolation, int boundary, int blending_mode,
                                    int separate_camera_batch, int gradient_semantics, float* input_grad,
                                    void* workspace, size_t workspace_bytes, ngt_stream_t stream) {
  std::vector<Pair> pairs;
  int rc = check_ray_args(input_shape, matrix_m, matrix_v, matrix_p, frustum, num_cameras, output_dhw, interpolation,
                          boundary, blending_mode, separate_camera_batch, pairs, false);
  if (rc) return rc;
  const RayArgs a = make_ray_args(input, input_shape, output_dhw);
  hipStream_t s = as_stream(stream);
  const Shape5 si = shape5(input_shape);
  const int C = static_cast<int>(si.c);
  if (gradient_semantics == NGT_GRAD_REFERENCE) {
    fill_zero(input_grad, si.elems(), s);
    const int64_t cells = si.cells();
    const size_t need = static_cast<size_t>(cells * (C + 1)) * sizeof(float);
    NGT_REQUIRE(workspace && workspace_bytes >= need, NGT_ERR_WORKSPACE,
                "workspace too small for ngt_raymarch_grid_transform_bwd");
    if (interpolation != LINEAR) return check_launch("ngt_raymarch_grid_transform_bwd");  // reference scatters LINEAR only
    float* gbuf = static_cast<float*>(workspace);
    float* wsum = gbuf + si.cells() * C;
    fill_zero(gbuf, si.cells() * (C + 1), s);
    for (const Pair& p : pairs) {
      CamBatch cb;
      cb.count = 1;
      cb.cam[0] = make_cam(matrix_m + 16 * p.n, matrix_v + 16 * p.v, matrix_p + 16 * p.v, frustum + 6 * p.v, p.n, p.out);
      const dim3 grid = ray_grid_dims(a.out_w, a.out_h, 1);
      NGT_SWITCH_C(C, NGT_SWITCH_BM(boundary, NGT_SWITCH_MODE(blending_mode,
          hipLaunchKernelGGL((raymarch_bwd_reference_kernel<CC, BB, MM>), grid, dim3(kRayBlock), 0, s, a, cb, output,
                             output_grad, gbuf, wsum))));
      NGT_SWITCH_C(C,
          hipLaunchKernelGGL((normalize_by_weight_kernel<CC>), dim3(blocks_for(si.cells(), 256)), dim3(256), 0, s, gbuf,
                             wsum, input_grad + p.n * si.cells() * C, si.cells()));
    }
    return check_launch("ngt_raymarch_grid_transform_bwd");
  }
  GatherPlan gp;
  if (plan_gather(pairs, matrix_m, matrix_v, matrix_p, frustum, a, interpolation, boundary, C, workspace,
                  workspace_bytes, gp)) {
    run_gather<false>(pairs, matrix_m, matrix_v, matrix_p, frustum, a, blending_mode, C, output, output_grad, gp,
                      input_grad, static_cast<int>(si.n), s);
  } else {
    fill_zero(input_grad, si.elems(), s);
    for_cam_batches(pairs, matrix_m, matrix_v, matrix_p, frustum, [&](const CamBatch& cb) {
      const dim3 grid = ray_grid_dims(a.out_w, a.out_h, cb.count);
      NGT_SWITCH_C(C, NGT_SWITCH_FM2(interpolation, NGT_SWITCH_BM(boundary, NGT_SWITCH_MODE(blending_mode,
          hipLaunchKernelGGL((raymarch_bwd_exact_kernel<CC, FF, BB, MM, false>), grid, dim3(kRayBlock), 0, s, a, cb, output,
                             output_grad, input_grad)))));
    });
  }
  return check_launch("ngt_raymarch_grid_transform_bwd");
}

size_t ngt_render_density_fwd_workspace_size(const int64_t* density_shape, int interpolation) {
  if (density_shape == nullptr || interpolation != LINEAR) return 0;
  return bricked_bytes(density_shape, 1);
}

int ngt_render_density_fwd(const float* density, const int64_t* density_shape, const float* light_rgb,
                           float depth_step, const float* matrix_m, const float* matrix_v, const float* matrix_p,
                           const float* frustum, int num_cameras, const int32_t* output_dhw, int interpolation,
                           int boundary, int blending_mode, float* output, void* workspace, size_t workspace_bytes,
                           ngt_stream_t stream) {
  return render_density_fwd_impl(density, density_shape, light_rgb, depth_step, matrix_m, matrix_v, matrix_p, frustum,
                                 num_cameras, output_dhw, interpolation, boundary, blending_mode, output, workspace,
                                 workspace_bytes, nullptr, nullptr, nullptr, stream);
}

size_t ngt_render_density_bwd_workspace_size(const int64_t* density_shape, const float* matrix_m,
                                             const float* matrix_v, const float* matrix_p, const float* frustum,
                                             int num_cameras, const int32_t* output_dhw, int interpolation,
                                             int boundary) {
  return gather_workspace_bytes(density_shape, matrix_m, matrix_v, matrix_p, frustum, num_cameras, output_dhw,
                                interpolation, boundary, 1, 1);
}

int ngt_render_density_bwd(const float* density, const int64_t* density_shape, const float* light_rgb,
                           float depth_step, const float* output, const float* output_grad, const float* matrix_m,
                           const float* matrix_v, const float* matrix_p, const float* frustum, int num_cameras,
                           const int32_t* output_dhw, int interpolation, int boundary, int blending_mode,
                           float* density_grad, void* workspace, size_t workspace_bytes, ngt_stream_t stream) {
  std::vector<Pair> pairs;
  int rc = check_ray_args(density_shape, matrix_m, matrix_v, matrix_p, frustum, num_cameras, output_dhw, interpolation,
                          boundary, blending_mode, 1, pairs, true);
  if (rc) return rc;
  NGT_REQUIRE(light_rgb != nullptr, NGT_ERR_INVALID_ARGUMENT, "light_rgb must be given");
  RayArgs a = make_ray_args(density, density_shape, output_dhw);
  a.light_r = light_rgb[0]; a.light_g = light_rgb[1]; a.light_b = light_rgb[2];
  a.depth_step = depth_step;
  hipStream_t s = as_stream(stream);
  GatherPlan gp;
  if (plan_gather(pairs, matrix_m, matrix_v, matrix_p, frustum, a, interpolation, boundary, 1, workspace,
                  workspace_bytes, gp)) {
    run_gather<true>(pairs, matrix_m, matrix_v, matrix_p, frustum, a, blending_mode, 4, output, output_grad, gp,
                     density_grad, static_cast<int>(density_shape[0]), s);
  } else {
    fill_zero(density_grad, shape5(density_shape).elems(), s);
    for_cam_batches(pairs, matrix_m, matrix_v, matrix_p, frustum, [&](const CamBatch& cb) {
      const dim3 grid = ray_grid_dims(a.out_w, a.out_h, cb.count);
      NGT_SWITCH_FM2(interpolation, NGT_SWITCH_BM(boundary, NGT_SWITCH_MODE(blending_mode,
          hipLaunchKernelGGL((raymarch_bwd_exact_kernel<4, FF, BB, MM, true>), grid, dim3(kRayBlock), 0, s, a, cb, output,
                             output_grad, density_grad))));
    });
  }
  return check_launch("ngt_render_density_bwd");
}

size_t ngt_bricked_grid_bytes(const int64_t* shape) {
  if (shape == nullptr || check_grid_shape(shape, "grid") != NGT_OK) return 0;
  return bricked_bytes(shape, static_cast<int>(shape[4]));
}

int ngt_bricked_grid(const float* grid, const int64_t* shape, int boundary, float* bricked, ngt_stream_t stream) {
  int rc = check_grid_shape(shape, "grid");
  if (rc) return rc;
  NGT_REQUIRE(grid != nullptr && bricked != nullptr, NGT_ERR_INVALID_ARGUMENT, "grid and bricked must be given");
  NGT_REQUIRE(boundary >= 0 && boundary <= 2, NGT_ERR_INVALID_ARGUMENT, "invalid boundary mode");
  const int C = static_cast<int>(shape[4]);
  NGT_REQUIRE(C == 1 || C == 2 || C == 4, NGT_ERR_UNIMPLEMENTED, "Only 1,2 and 4 Channel supported.");
  NGT_REQUIRE(bricked_bytes(shape, C) > 0, NGT_ERR_INVALID_ARGUMENT,
              "grid too large for the bricked layout (a slice must hold fewer than 2^32 values)");
  const Dims d = shape5(shape).dims();
  launch_to_bricked(grid, d, brick_dims(d), static_cast<int>(shape[0]), C, boundary, bricked, as_stream(stream));
  return check_launch("ngt_bricked_grid");
}

// partial sums of the fused L2 loss: one per forward work-group (an 8x8 pixel tile)
static int64_t l2_partials(const int64_t* density_shape, int num_cameras, const int32_t* output_dhw) {
  return density_shape[0] * num_cameras * static_cast<int64_t>(blocks_for(output_dhw[2], kRayTile)) *
         blocks_for(output_dhw[1], kRayTile);
}

size_t ngt_render_density_l2_fwd_workspace_size(const int64_t* density_shape, int interpolation, int num_cameras,
                                                const int32_t* output_dhw) {
  if (density_shape == nullptr || output_dhw == nullptr || num_cameras <= 0) return 0;
  return align_up(ngt_render_density_fwd_workspace_size(density_shape, interpolation)) +
         align_up(static_cast<size_t>(l2_partials(density_shape, num_cameras, output_dhw)) * sizeof(float));
}

int ngt_render_density_l2_fwd(const float* density, const int64_t* density_shape, const float* light_rgb,
                              float depth_step, const float* matrix_m, const float* matrix_v, const float* matrix_p,
                              const float* frustum, int num_cameras, const int32_t* output_dhw, int interpolation,
                              int boundary, int blending_mode, const float* target, float loss_weight,
                              const float* bricked, void* record, size_t record_bytes, float* output, float* loss,
                              void* workspace, size_t workspace_bytes, ngt_stream_t stream) {
  NGT_REQUIRE(target != nullptr && loss != nullptr, NGT_ERR_INVALID_ARGUMENT, "target and loss must be given");
  NGT_REQUIRE(output_dhw != nullptr && num_cameras > 0 && density_shape != nullptr, NGT_ERR_INVALID_ARGUMENT,
              "Invalid output_shape");
  const size_t need = ngt_render_density_l2_fwd_workspace_size(density_shape, interpolation, num_cameras, output_dhw);
  NGT_REQUIRE(workspace != nullptr && workspace_bytes >= need, NGT_ERR_WORKSPACE,
              "workspace too small for ngt_render_density_l2_fwd");
  NGT_REQUIRE(bricked == nullptr || interpolation == LINEAR, NGT_ERR_INVALID_ARGUMENT,
              "a bricked copy serves LINEAR interpolation only");
  const size_t brick_ws = align_up(ngt_render_density_fwd_workspace_size(density_shape, interpolation));
  float* partial = reinterpret_cast<float*>(static_cast<char*>(workspace) + brick_ws);
  int rc = render_density_fwd_impl(density, density_shape, light_rgb, depth_step, matrix_m, matrix_v, matrix_p,
                                   frustum, num_cameras, output_dhw, interpolation, boundary, blending_mode, output,
                                   brick_ws ? workspace : nullptr, brick_ws, bricked, target, partial, stream,
                                   record, record_bytes, loss_weight);
  if (rc) return rc;
  hipLaunchKernelGGL(l2_finish_kernel, dim3(1), dim3(256), 0, as_stream(stream), partial,
                     static_cast<int>(l2_partials(density_shape, num_cameras, output_dhw)), loss_weight, loss);
  return check_launch("ngt_render_density_l2_fwd");
}

size_t ngt_static_scratch_bytes(void) { return kStaticScratchBytes; }

int ngt_render_density_l2_fwd_noalloc(const float* density, const int64_t* density_shape, const float* light_rgb,
                                      float depth_step, const float* matrix_m, const float* matrix_v,
                                      const float* matrix_p, const float* frustum, int num_cameras,
                                      const int32_t* output_dhw, int interpolation, int boundary, int blending_mode,
                                      const float* target, float loss_weight, const float* bricked, void* record,
                                      size_t record_bytes, float* output, float* loss, ngt_stream_t stream) {
  NGT_REQUIRE(output_dhw != nullptr && num_cameras > 0 && density_shape != nullptr, NGT_ERR_INVALID_ARGUMENT,
              "Invalid output_shape");
  const size_t need = ngt_render_density_l2_fwd_workspace_size(density_shape, interpolation, num_cameras, output_dhw);
  NGT_REQUIRE(need <= kStaticScratchBytes, NGT_ERR_WORKSPACE,
              "ngt_render_density_l2_fwd_noalloc: the call needs " + std::to_string(need) +
                  " bytes of scratch, more than the static " + std::to_string(kStaticScratchBytes) + " bytes");
  void* ws = nullptr;
  const int rc = scratch_for_current_device(&ws);
  if (rc) return rc;
  return ngt_render_density_l2_fwd(density, density_shape, light_rgb, depth_step, matrix_m, matrix_v, matrix_p,
                                   frustum, num_cameras, output_dhw, interpolation, boundary, blending_mode, target,
                                   loss_weight, bricked, record, record_bytes, output, loss, ws, kStaticScratchBytes,
                                   stream);
}

size_t ngt_render_density_l2_record_size(const int64_t* density_shape, const float* matrix_m, const float* matrix_v,
                                          const float* matrix_p, const float* frustum, int num_cameras,
                                          const int32_t* output_dhw) {
  if (!density_shape || !matrix_m || !matrix_v || !matrix_p || !frustum || !output_dhw || num_cameras <= 0) return 0;
  if (density_shape[4] != 1 || bricked_bytes(density_shape, 1) == 0) return 0;
  std::vector<Pair> pairs;
  if (make_pairs(static_cast<int>(density_shape[0]), num_cameras, true, pairs) != NGT_OK) return 0;
  const RayArgs a = make_ray_args(nullptr, density_shape, output_dhw);
  const int kcap = ray_step_cap(pairs, matrix_m, matrix_v, matrix_p, frustum, a);
  if (kcap <= 0) return 0;
  const size_t npix = static_cast<size_t>(a.out_w) * a.out_h;
  return align_up(pairs.size() * npix * sizeof(RayRec)) +
         align_up(pairs.size() * bound_tiles(a.out_w, a.out_h) * sizeof(float)) +
         align_up(pairs.size() * static_cast<size_t>(kcap) * g_pixels(npix) * sizeof(float));
}

size_t ngt_render_density_l2_bwd_workspace_size(const int64_t* density_shape, const float* matrix_m,
                                                const float* matrix_v, const float* matrix_p, const float* frustum,
                                                int num_cameras, const int32_t* output_dhw, int interpolation,
                                                int boundary) {
  const size_t g = ngt_render_density_bwd_workspace_size(density_shape, matrix_m, matrix_v, matrix_p, frustum,
                                                         num_cameras, output_dhw, interpolation, boundary);
  if (g > 0 || density_shape == nullptr || output_dhw == nullptr || num_cameras <= 0) return g;
  return static_cast<size_t>(render_pixels(density_shape, num_cameras, output_dhw)) * 4 * sizeof(float);
}

int ngt_render_density_l2_bwd(const float* density, const int64_t* density_shape, const float* light_rgb,
                              float depth_step, const float* output, const float* target, float loss_weight,
                              const float* loss_grad, const float* matrix_m, const float* matrix_v,
                              const float* matrix_p, const float* frustum, int num_cameras, const int32_t* output_dhw,
                              int interpolation, int boundary, int blending_mode, const float* bricked,
                              int recorded, float* density_grad, void* workspace, size_t workspace_bytes,
                              ngt_stream_t stream) {
  std::vector<Pair> pairs;
  int rc = check_ray_args(density_shape, matrix_m, matrix_v, matrix_p, frustum, num_cameras, output_dhw, interpolation,
                          boundary, blending_mode, 1, pairs, true);
  if (rc) return rc;
  NGT_REQUIRE(light_rgb != nullptr, NGT_ERR_INVALID_ARGUMENT, "light_rgb must be given");
  NGT_REQUIRE(output != nullptr && target != nullptr && loss_grad != nullptr, NGT_ERR_INVALID_ARGUMENT,
              "output, target and loss_grad must be given");
  // the gather adjoint when the workspace holds its buffers, else the image
  // gradient buffer and the scatter adjoint
  const int64_t npix = render_pixels(density_shape, num_cameras, output_dhw);
  const size_t img_need = static_cast<size_t>(npix) * 4 * sizeof(float);
  // a record made by the forward holds no brick copy when the caller passes one
  const bool own_bricks = !(recorded && bricked != nullptr);
  const size_t g_need =
      own_bricks ? ngt_render_density_bwd_workspace_size(density_shape, matrix_m, matrix_v, matrix_p, frustum,
                                                         num_cameras, output_dhw, interpolation, boundary)
                 : ngt_render_density_l2_record_size(density_shape, matrix_m, matrix_v, matrix_p, frustum,
                                                     num_cameras, output_dhw);
  const bool gather = g_need > 0 && workspace != nullptr && workspace_bytes >= g_need;
  NGT_REQUIRE(gather || (workspace != nullptr && workspace_bytes >= img_need), NGT_ERR_WORKSPACE,
              "workspace too small for ngt_render_density_l2_bwd");
  RayArgs a = make_ray_args(density, density_shape, output_dhw);
  a.light_r = light_rgb[0]; a.light_g = light_rgb[1]; a.light_b = light_rgb[2];
  a.depth_step = depth_step;
  a.l2_target = target;
  a.l2_loss_grad = loss_grad;
  a.l2_scale = loss_weight;
  hipStream_t s = as_stream(stream);
  GatherPlan gp;
  if (gather && plan_gather(pairs, matrix_m, matrix_v, matrix_p, frustum, a, interpolation, boundary, 1, workspace,
                            workspace_bytes, gp, own_bricks)) {
    run_gather<true>(pairs, matrix_m, matrix_v, matrix_p, frustum, a, blending_mode, 4, output, nullptr, gp,
                     density_grad, static_cast<int>(density_shape[0]), s, bricked, recorded != 0, loss_grad);
    return check_launch("ngt_render_density_l2_bwd");
  }
  NGT_REQUIRE(!recorded, NGT_ERR_WORKSPACE, "the recorded per-sample gradients need the gather workspace");
  NGT_REQUIRE(workspace_bytes >= img_need, NGT_ERR_WORKSPACE, "workspace too small for ngt_render_density_l2_bwd");
  float* dimg = static_cast<float*>(workspace);
  hipLaunchKernelGGL(l2_image_grad_kernel, dim3(blocks_for(npix, 256)), dim3(256), 0, s, a, output, npix, dimg);
  a.l2_target = nullptr;
  fill_zero(density_grad, shape5(density_shape).elems(), s);
  for_cam_batches(pairs, matrix_m, matrix_v, matrix_p, frustum, [&](const CamBatch& cb) {
    const dim3 grid = ray_grid_dims(a.out_w, a.out_h, cb.count);
    NGT_SWITCH_FM2(interpolation, NGT_SWITCH_BM(boundary, NGT_SWITCH_MODE(blending_mode,
        hipLaunchKernelGGL((raymarch_bwd_exact_kernel<4, FF, BB, MM, true>), grid, dim3(kRayBlock), 0, s, a, cb, output,
                           dimg, density_grad))));
  });
  return check_launch("ngt_render_density_l2_bwd");
}


size_t ngt_render_batch_plan_bytes(int num_instances) {
  if (num_instances <= 0) return 0;
  BlobLayout L;
  L.add<RenBatchHeader>(1);
  L.add<RenBatchItem>(num_instances);
  for (int k = 0; k < 3 + kGatherShapes; ++k) L.add<int>(num_instances + 1);
  return L.bytes;
}

int ngt_render_batch_plan(const ngt_render_instance* inst, int n, int blending_mode, size_t record_budget, void* plan,
                          size_t plan_bytes, size_t* workspace_bytes) {
  NGT_REQUIRE(inst != nullptr && n > 0, NGT_ERR_INVALID_ARGUMENT, "batch must not be empty");
  NGT_REQUIRE(blending_mode == BEER_LAMBERT || blending_mode == ALPHA || blending_mode == ALPHA_ADDITIVE ||
                  blending_mode == ADDITIVE,
              NGT_ERR_INVALID_ARGUMENT, "invalid blending mode");
  NGT_REQUIRE(plan != nullptr && plan_bytes >= ngt_render_batch_plan_bytes(n), NGT_ERR_WORKSPACE,
              "plan buffer too small for ngt_render_batch_plan");
  if (record_budget == 0) record_budget = kDefaultRecordBudget;
  BlobLayout L;
  const size_t off_h = L.add<RenBatchHeader>(1);
  const size_t off_items = L.add<RenBatchItem>(n);
  const size_t off_ray = L.add<int>(n + 1), off_brickify = L.add<int>(n + 1), off_chunk = L.add<int>(n + 1);
  size_t off_brick[kGatherShapes];
  for (int k = 0; k < kGatherShapes; ++k) off_brick[k] = L.add<int>(n + 1);
  RenBatchHeader* h = blob_at<RenBatchHeader>(plan, off_h);
  RenBatchItem* items = blob_at<RenBatchItem>(plan, off_items);
  int* ray = blob_at<int>(plan, off_ray);
  int* brickify = blob_at<int>(plan, off_brickify);
  int* chunk = blob_at<int>(plan, off_chunk);
  int* brick[kGatherShapes];
  for (int k = 0; k < kGatherShapes; ++k) brick[k] = blob_at<int>(plan, off_brick[k]);
  static const int kShapeX[kGatherShapes] = {8, 16, 16}, kShapeY[kGatherShapes] = {8, 8, 16};
  ray[0] = brickify[0] = chunk[0] = 0;
  for (int k = 0; k < kGatherShapes; ++k) brick[k][0] = 0;
  int64_t tr = 0, tz = 0, tb[kGatherShapes] = {0, 0, 0};
  size_t ws = 0, chunk_bytes = 0, region = 0;
  int nchunks = 0;
  for (int i = 0; i < n; ++i) {
    const ngt_render_instance& q = inst[i];
    NGT_REQUIRE(q.density && q.matrix_m && q.matrix_v && q.matrix_p && q.frustum, NGT_ERR_INVALID_ARGUMENT,
                "batch instance without density or camera parameters");
    NGT_REQUIRE(q.num_views >= 1 && q.num_views <= kMaxCamsPerLaunch, NGT_ERR_INVALID_ARGUMENT,
                "a batch instance has 1 to 8 views");
    NGT_REQUIRE(q.image_dhw[0] > 0 && q.image_dhw[1] > 0 && q.image_dhw[2] > 0, NGT_ERR_INVALID_ARGUMENT,
                "empty image");
    const int64_t sh[5] = {1, q.dhw[0], q.dhw[1], q.dhw[2], 1};
    int rc = check_grid_shape(sh, "batch density");
    if (rc) return rc;
    RenBatchItem& it = items[i];
    it = RenBatchItem{};
    it.a = make_ray_args(q.density, sh, q.image_dhw);
    it.a.light_r = q.light_rgb[0]; it.a.light_g = q.light_rgb[1]; it.a.light_b = q.light_rgb[2];
    it.a.depth_step = q.depth_step;
    std::vector<Pair> pairs;
    make_pairs(1, q.num_views, true, pairs);
    it.cb.count = q.num_views;
    for (int v = 0; v < q.num_views; ++v)
      it.cb.cam[v] = make_cam(q.matrix_m, q.matrix_v + 16 * v, q.matrix_p + 16 * v, q.frustum + 6 * v, 0, v);
    it.kcap = ray_step_cap(pairs, q.matrix_m, q.matrix_v, q.matrix_p, q.frustum, it.a);
    NGT_REQUIRE(it.kcap > 0, NGT_ERR_INVALID_ARGUMENT,
                "batch render: rays too long for the gather adjoint's per-sample record");
    it.ntx = static_cast<int>(blocks_for(it.a.out_w, kRayTile));
    it.nty = static_cast<int>(blocks_for(it.a.out_h, kRayTile));
    NGT_REQUIRE(bricked_fits(it.a.din, 1), NGT_ERR_INVALID_ARGUMENT, "batch density grid too large");
    it.brick_tiles = static_cast<int>(brick_rows(it.a.bd));
    it.density = (gptr<const float>)q.density;
    it.images = (gptr<float>)q.images; it.images_grad = (gptr<const float>)q.images_grad;
    it.density_grad = (gptr<float>)q.density_grad;
    it.l2_target = (gptr<const float>)q.l2_target; it.loss = (gptr<float>)q.loss;
    it.ws_bricked = static_cast<int64_t>(ws);
    ws += align_up(static_cast<size_t>(it.a.bd.slice_cells) * sizeof(float));
    // records: chunk-relative offsets; a new chunk starts where the budget would be exceeded
    const size_t npix = static_cast<size_t>(it.a.out_w) * it.a.out_h;
    const size_t rays_bytes = align_up(q.num_views * npix * sizeof(RayRec));
    const size_t bound_bytes = align_up(q.num_views * static_cast<size_t>(bound_tiles(it.a.out_w, it.a.out_h)) *
                                        sizeof(float));
    const size_t g_bytes = align_up(q.num_views * static_cast<size_t>(it.kcap) * g_pixels(npix) * sizeof(float));
    const size_t inst_bytes = rays_bytes + bound_bytes + g_bytes;
    if (chunk_bytes > 0 && chunk_bytes + inst_bytes > record_budget) {
      chunk[++nchunks] = i;
      chunk_bytes = 0;
    }
    it.rec_rays = static_cast<int64_t>(chunk_bytes);
    it.rec_bound = static_cast<int64_t>(chunk_bytes + rays_bytes);
    it.rec_G = static_cast<int64_t>(chunk_bytes + rays_bytes + bound_bytes);
    chunk_bytes += inst_bytes;
    region = std::max(region, chunk_bytes);
    tr += static_cast<int64_t>(q.num_views) * it.ntx * it.nty;
    tz += it.brick_tiles;
    for (int k = 0; k < kGatherShapes; ++k)
      tb[k] += static_cast<int64_t>(blocks_for(it.a.din.x, kShapeX[k])) * blocks_for(it.a.din.y, kShapeY[k]) *
               blocks_for(it.a.din.z, 8);
    NGT_REQUIRE(tr < (int64_t(1) << 31) && tz < (int64_t(1) << 31) && tb[0] < (int64_t(1) << 31),
                NGT_ERR_INVALID_ARGUMENT, "batch too large");
    ray[i + 1] = static_cast<int>(tr); brickify[i + 1] = static_cast<int>(tz);
    for (int k = 0; k < kGatherShapes; ++k) brick[k][i + 1] = static_cast<int>(tb[k]);
  }
  chunk[++nchunks] = n;
  h->n = n; h->mode = blending_mode; h->nchunks = nchunks;
  h->tiles_ray = static_cast<int>(tr); h->tiles_brickify = static_cast<int>(tz);
  for (int k = 0; k < kGatherShapes; ++k) {
    h->tiles_brick[k] = static_cast<int>(tb[k]);
    h->off_brick[k] = off_brick[k];
  }
  h->off_items = off_items; h->off_ray = off_ray; h->off_brickify = off_brickify; h->off_chunk = off_chunk;
  h->ws_partial = static_cast<int64_t>(ws);
  ws += align_up(static_cast<size_t>(tr) * sizeof(float));
  h->ws_record = static_cast<int64_t>(ws);
  h->record_bytes = static_cast<int64_t>(region);
  ws += region;
  h->ws_bytes = static_cast<int64_t>(ws);
  if (workspace_bytes) *workspace_bytes = ws;
  return NGT_OK;
}

int ngt_render_batch_plan_info(const void* plan_host, int* num_chunks, size_t* record_bytes) {
  NGT_REQUIRE(plan_host != nullptr, NGT_ERR_INVALID_ARGUMENT, "plan must be given");
  const RenBatchHeader& h = *static_cast<const RenBatchHeader*>(plan_host);
  if (num_chunks) *num_chunks = h.nchunks;
  if (record_bytes) *record_bytes = static_cast<size_t>(h.record_bytes);
  return NGT_OK;
}

}  // extern "C"

namespace {

int check_batch_call(const void* plan_host, const void* plan_device, size_t workspace_bytes, int gather_brick,
                     const char* who) {
  NGT_REQUIRE(plan_host && plan_device, NGT_ERR_INVALID_ARGUMENT, "plan must be given");
  const RenBatchHeader& h = *static_cast<const RenBatchHeader*>(plan_host);
  NGT_REQUIRE(workspace_bytes >= static_cast<size_t>(h.ws_bytes), NGT_ERR_WORKSPACE,
              (std::string("workspace too small for ") + who).c_str());
  NGT_REQUIRE(gather_brick >= 0 && gather_brick < kGatherShapes, NGT_ERR_INVALID_ARGUMENT,
              "gather_brick must be 0 (8x8x8), 1 (16x8x8) or 2 (16x16x8)");
  return NGT_OK;
}

// The gather over the brick tiles of chunk c.
void launch_batch_gather(const RenBatchHeader& h, const void* plan_host, const void* plan_device, char* ws, int c,
                         int blocks_per_cu, int shape, hipStream_t s) {
  const int* chunk = blob_at<int>(plan_host, h.off_chunk);
  const int* table = blob_at<int>(plan_host, h.off_brick[shape]);
  const int t0 = table[chunk[c]], t1 = table[chunk[c + 1]];
  const dim3 g(persistent_grid(t1 - t0, blocks_per_cu));
  if (shape == 0) hipLaunchKernelGGL((render_brick_batch_kernel<0>), g, dim3(256), 0, s, plan_device, ws, t0, t1);
  else if (shape == 1) hipLaunchKernelGGL((render_brick_batch_kernel<1>), g, dim3(256), 0, s, plan_device, ws, t0, t1);
  else hipLaunchKernelGGL((render_brick_batch_kernel<2>), g, dim3(256), 0, s, plan_device, ws, t0, t1);
}

}  // namespace

extern "C" {

int ngt_render_batch_fwd(const void* plan_host, const void* plan_device, void* workspace, size_t workspace_bytes,
                         int blocks_per_cu, ngt_stream_t stream) {
  int rc = check_batch_call(plan_host, plan_device, workspace_bytes, 0, "ngt_render_batch_fwd");
  if (rc) return rc;
  const RenBatchHeader& h = *static_cast<const RenBatchHeader*>(plan_host);
  NGT_REQUIRE(workspace != nullptr, NGT_ERR_WORKSPACE, "workspace must be given");
  const RenBatchItem* items = blob_at<RenBatchItem>(plan_host, h.off_items);
  for (int i = 0; i < h.n; ++i) NGT_REQUIRE(items[i].images, NGT_ERR_INVALID_ARGUMENT, "batch instance without images");
  hipStream_t s = as_stream(stream);
  char* ws = static_cast<char*>(workspace);
  hipLaunchKernelGGL(render_brickify_batch_kernel, dim3(persistent_grid(h.tiles_brickify, blocks_per_cu)),
                     dim3(256), 0, s, plan_device, ws);
  const dim3 g(persistent_grid(h.tiles_ray, blocks_per_cu));
  NGT_SWITCH_MODE(h.mode,
      hipLaunchKernelGGL((render_fwd_batch_kernel<MM>), g, dim3(kRayBlock), 0, s, plan_device, ws));
  return check_launch("ngt_render_batch_fwd");
}

int ngt_render_batch_bwd(const void* plan_host, const void* plan_device, void* workspace, size_t workspace_bytes,
                         int blocks_per_cu, int gather_brick, ngt_stream_t stream) {
  int rc = check_batch_call(plan_host, plan_device, workspace_bytes, gather_brick, "ngt_render_batch_bwd");
  if (rc) return rc;
  const RenBatchHeader& h = *static_cast<const RenBatchHeader*>(plan_host);
  NGT_REQUIRE(workspace != nullptr, NGT_ERR_WORKSPACE, "workspace must be given");
  const RenBatchItem* items = blob_at<RenBatchItem>(plan_host, h.off_items);
  for (int i = 0; i < h.n; ++i)
    NGT_REQUIRE(items[i].images && items[i].images_grad && items[i].density_grad, NGT_ERR_INVALID_ARGUMENT,
                "batch instance without images, image gradients or density gradient");
  hipStream_t s = as_stream(stream);
  char* ws = static_cast<char*>(workspace);
  const int* chunk = blob_at<int>(plan_host, h.off_chunk);
  const int* ray = blob_at<int>(plan_host, h.off_ray);
  for (int c = 0; c < h.nchunks; ++c) {
    const int t0 = ray[chunk[c]], t1 = ray[chunk[c + 1]];
    const dim3 g(persistent_grid(t1 - t0, blocks_per_cu));
    NGT_SWITCH_MODE(h.mode,
        hipLaunchKernelGGL((render_prepass_batch_kernel<MM>), g, dim3(kRayBlock), 0, s, plan_device, ws, t0, t1));
    launch_batch_gather(h, plan_host, plan_device, ws, c, blocks_per_cu, gather_brick, s);
  }
  return check_launch("ngt_render_batch_bwd");
}

int ngt_render_batch_l2(const void* plan_host, const void* plan_device, void* workspace, size_t workspace_bytes,
                        int blocks_per_cu, int gather_brick, ngt_stream_t stream) {
  int rc = check_batch_call(plan_host, plan_device, workspace_bytes, gather_brick, "ngt_render_batch_l2");
  if (rc) return rc;
  const RenBatchHeader& h = *static_cast<const RenBatchHeader*>(plan_host);
  NGT_REQUIRE(workspace != nullptr, NGT_ERR_WORKSPACE, "workspace must be given");
  const RenBatchItem* items = blob_at<RenBatchItem>(plan_host, h.off_items);
  for (int i = 0; i < h.n; ++i)
    NGT_REQUIRE(items[i].images && items[i].l2_target && items[i].loss && items[i].density_grad,
                NGT_ERR_INVALID_ARGUMENT, "batch instance without images, target, loss or density gradient");
  hipStream_t s = as_stream(stream);
  char* ws = static_cast<char*>(workspace);
  hipLaunchKernelGGL(render_brickify_batch_kernel, dim3(persistent_grid(h.tiles_brickify, blocks_per_cu)),
                     dim3(256), 0, s, plan_device, ws);
  const int* chunk = blob_at<int>(plan_host, h.off_chunk);
  const int* ray = blob_at<int>(plan_host, h.off_ray);
  for (int c = 0; c < h.nchunks; ++c) {
    const int t0 = ray[chunk[c]], t1 = ray[chunk[c + 1]];
    const dim3 g(persistent_grid(t1 - t0, blocks_per_cu));
    NGT_SWITCH_MODE(h.mode,
        hipLaunchKernelGGL((render_record_batch_kernel<MM>), g, dim3(kRayBlock), 0, s, plan_device, ws, t0, t1));
    launch_batch_gather(h, plan_host, plan_device, ws, c, blocks_per_cu, gather_brick, s);
  }
  hipLaunchKernelGGL(render_l2_finish_batch_kernel, dim3(h.n), dim3(256), 0, s, plan_device, ws);
  return check_launch("ngt_render_batch_l2");
}

}  // extern "C"
