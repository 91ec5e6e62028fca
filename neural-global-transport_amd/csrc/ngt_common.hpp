// Device-side building blocks shared by every NGT kernel (gfx950 / CDNA4).
//
//  * Grids are NDHWC fp32, x fastest (flat = ((z*H + y)*W + x)*C + c), the
//    layout of the reference ops (vector_io.inl:21-37, resample_grid.cc:25).
//  * Sampling positions are in integer-centred cell units (cell i has its
//    centre at q = i) after the caller removed the cell-centre offset
//    (sampling.hpp:802-826).  Corners are floor(q) / ceil(q) and weights
//    fract(q) = q - floor(q); interpolation runs x, then y, then z with
//    lerp(a,b,t) = a + t*(b-a) (sampling.hpp:371-465).
//  * Boundary modes BORDER (0 outside [-0.5, dims-0.5), out-of-grid corners 0),
//    CLAMP, WRAP (sampling.hpp:381-440).
//
// Everything here is header-only and force-inlined into the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define NGT_DEVICE __device__ __forceinline__

namespace ngt {

enum Boundary : int { BORDER = 0, CLAMP = 1, WRAP = 2 };
enum Filter : int { NEAREST = 0, LINEAR = 1, FMIN = 2, FMAX = 3, QUADRATIC = 4 };
enum BlendMode : int { BEER_LAMBERT = 0, ALPHA = 2, ALPHA_ADDITIVE = 3, ADDITIVE = 4 };

constexpr float kFloatMax = 3.402823466e+38f;

// ---------------------------------------------------------------- vectors ---

template <int C>
struct Vec {
  float v[C];
  NGT_DEVICE float& operator[](int i) { return v[i]; }
  NGT_DEVICE float operator[](int i) const { return v[i]; }
};

template <int C>
NGT_DEVICE Vec<C> vzero() {
  Vec<C> r;
#pragma unroll
  for (int i = 0; i < C; ++i) r.v[i] = 0.f;
  return r;
}
template <int C>
NGT_DEVICE Vec<C> vfill(float s) {
  Vec<C> r;
#pragma unroll
  for (int i = 0; i < C; ++i) r.v[i] = s;
  return r;
}

// Vector loads/stores: one dwordx{1,2,4} access per cell for C in {1,2,4}.
template <int C>
NGT_DEVICE Vec<C> vload(const float* __restrict__ p) {
  Vec<C> r;
  if constexpr (C == 4) {
    const float4 t = *reinterpret_cast<const float4*>(p);
    r.v[0] = t.x; r.v[1] = t.y; r.v[2] = t.z; r.v[3] = t.w;
  } else if constexpr (C == 2) {
    const float2 t = *reinterpret_cast<const float2*>(p);
    r.v[0] = t.x; r.v[1] = t.y;
  } else {
#pragma unroll
    for (int i = 0; i < C; ++i) r.v[i] = p[i];
  }
  return r;
}
template <int C>
NGT_DEVICE void vstore(float* __restrict__ p, const Vec<C>& r) {
  if constexpr (C == 4) {
    *reinterpret_cast<float4*>(p) = make_float4(r.v[0], r.v[1], r.v[2], r.v[3]);
  } else if constexpr (C == 2) {
    *reinterpret_cast<float2*>(p) = make_float2(r.v[0], r.v[1]);
  } else {
#pragma unroll
    for (int i = 0; i < C; ++i) p[i] = r.v[i];
  }
}

template <int C>
NGT_DEVICE Vec<C> vlerp(const Vec<C>& a, const Vec<C>& b, float t) {
  // a + t (b - a) with one rounding of the product and sum (an explicit FMA,
  // so that every kernel that interpolates forms the same bits, whatever the
  // compiler's contraction choices in its context)
  Vec<C> r;
#pragma unroll
  for (int i = 0; i < C; ++i) r.v[i] = fmaf(t, b.v[i] - a.v[i], a.v[i]);
  return r;
}
template <int C>
NGT_DEVICE Vec<C> vsub(const Vec<C>& a, const Vec<C>& b) {
  Vec<C> r;
#pragma unroll
  for (int i = 0; i < C; ++i) r.v[i] = a.v[i] - b.v[i];
  return r;
}
template <int C>
NGT_DEVICE Vec<C> vscale(const Vec<C>& a, float s) {
  Vec<C> r;
#pragma unroll
  for (int i = 0; i < C; ++i) r.v[i] = a.v[i] * s;
  return r;
}
template <int C>
NGT_DEVICE float vdot(const Vec<C>& a, const Vec<C>& b) {
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < C; ++i) s += a.v[i] * b.v[i];
  return s;
}

// Hardware fp32 atomic add (global_atomic_add_f32, no CAS loop).
NGT_DEVICE void atomic_add(float* p, float v) { unsafeAtomicAdd(p, v); }

// fp32 atomic add into LDS (ds_add_f32).  p must point into __shared__
// memory: a plain float* reaching an atomic is otherwise compiled as a flat
// atomic, which goes through the vector-memory path.
typedef __attribute__((address_space(3))) float lds_float;
NGT_DEVICE void lds_add(float* p, float v) {
  __hip_atomic_fetch_add((lds_float*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Block coordinates of a work-group's tile.  Kernel bodies take them as an
// argument so the same body runs under a plain grid launch (blockIdx) and
// under the persistent work-list launches of uneven batches.
struct VBlock {
  int x, y, z;
};

// ------------------------------------------------------------------- grid ---

struct Dims {
  int x, y, z;  // W, H, D
};

NGT_DEVICE int wrap_index(int i, int m) {
  // floor-mod; the reference's (i%m)+m*(i<0) differs only at i = -k*m where it
  // yields m (an out-of-range read), reachable only for shifts >= a grid extent
  int r = i % m;
  return r < 0 ? r + m : r;
}

NGT_DEVICE int64_t cell_offset(const Dims& d, int x, int y, int z) {
  return (static_cast<int64_t>(z) * d.y + y) * d.x + x;
}

struct Corners {
  int x0, y0, z0, x1, y1, z1;   // floor / ceil indices (wrapped or clamped)
  float wx, wy, wz;             // ceil weights = fract(q)
  bool vx0, vy0, vz0, vx1, vy1, vz1;  // per-axis in-grid flags (BORDER)
  bool inside;                  // BORDER forward validity of the position
};

template <int BM>
NGT_DEVICE Corners make_corners(const Dims& d, float qx, float qy, float qz) {
  Corners c;
  const float fx = floorf(qx), fy = floorf(qy), fz = floorf(qz);
  c.wx = qx - fx; c.wy = qy - fy; c.wz = qz - fz;
  c.x0 = static_cast<int>(fx); c.y0 = static_cast<int>(fy); c.z0 = static_cast<int>(fz);
  c.x1 = static_cast<int>(ceilf(qx)); c.y1 = static_cast<int>(ceilf(qy)); c.z1 = static_cast<int>(ceilf(qz));
  if constexpr (BM == BORDER) {
    // bitwise & keeps the masks branch-free
    c.inside = (qx + 0.5f >= 0.f) & (qx + 0.5f < static_cast<float>(d.x)) &
               (qy + 0.5f >= 0.f) & (qy + 0.5f < static_cast<float>(d.y)) &
               (qz + 0.5f >= 0.f) & (qz + 0.5f < static_cast<float>(d.z));
    c.vx0 = (c.x0 > -1) & (c.x0 < d.x); c.vx1 = (c.x1 > -1) & (c.x1 < d.x);
    c.vy0 = (c.y0 > -1) & (c.y0 < d.y); c.vy1 = (c.y1 > -1) & (c.y1 < d.y);
    c.vz0 = (c.z0 > -1) & (c.z0 < d.z); c.vz1 = (c.z1 > -1) & (c.z1 < d.z);
  } else {
    c.inside = true;
    c.vx0 = c.vx1 = c.vy0 = c.vy1 = c.vz0 = c.vz1 = true;
    if constexpr (BM == CLAMP) {
      c.x0 = min(max(c.x0, 0), d.x - 1); c.x1 = min(max(c.x1, 0), d.x - 1);
      c.y0 = min(max(c.y0, 0), d.y - 1); c.y1 = min(max(c.y1, 0), d.y - 1);
      c.z0 = min(max(c.z0, 0), d.z - 1); c.z1 = min(max(c.z1, 0), d.z - 1);
    } else {
      c.x0 = wrap_index(c.x0, d.x); c.x1 = wrap_index(c.x1, d.x);
      c.y0 = wrap_index(c.y0, d.y); c.y1 = wrap_index(c.y1, d.y);
      c.z0 = wrap_index(c.z0, d.z); c.z1 = wrap_index(c.z1, d.z);
    }
  }
  return c;
}

// The 8 corner values, index bit0 = x-ceil, bit1 = y-ceil, bit2 = z-ceil.
template <int C>
struct Corner8 {
  Vec<C> v[8];
  bool ok[8];
};

// The 8 corner loads, branch-free: issued unconditionally from in-grid
// addresses (BORDER corners clamped into the grid for the address only; the
// out-of-grid ones are zeroed by mask8 afterwards), so the loads of several
// samples can be in flight together.  Offsets are 32-bit (a grid slice holds
// fewer than 2^32 floats, checked on the host).
template <int C, int BM>
NGT_DEVICE void fetch8_raw(const float* __restrict__ g, const Dims& d, const Corners& c, Vec<C> (&v)[8]) {
  int x0 = c.x0, x1 = c.x1, y0 = c.y0, y1 = c.y1, z0 = c.z0, z1 = c.z1;
  if constexpr (BM == BORDER) {
    x0 = min(max(x0, 0), d.x - 1); x1 = min(max(x1, 0), d.x - 1);
    y0 = min(max(y0, 0), d.y - 1); y1 = min(max(y1, 0), d.y - 1);
    z0 = min(max(z0, 0), d.z - 1); z1 = min(max(z1, 0), d.z - 1);
  }
  const uint32_t W = static_cast<uint32_t>(d.x), H = static_cast<uint32_t>(d.y);
  const uint32_t rows[4] = {(static_cast<uint32_t>(z0) * H + static_cast<uint32_t>(y0)) * W,
                            (static_cast<uint32_t>(z0) * H + static_cast<uint32_t>(y1)) * W,
                            (static_cast<uint32_t>(z1) * H + static_cast<uint32_t>(y0)) * W,
                            (static_cast<uint32_t>(z1) * H + static_cast<uint32_t>(y1)) * W};
  const uint32_t xs[2] = {static_cast<uint32_t>(x0), static_cast<uint32_t>(x1)};
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = vload<C>(g + (rows[k >> 1] + xs[k & 1]) * static_cast<uint32_t>(C));
}

// Haloed brick layout of a grid slice: the grid is cut into 4x4x4-cell
// bricks, each stored with a one-cell halo as a 5x5x5 block (x fastest)
// padded to 128 floats.  Brick (bx, by, bz) holds cells 4*b - 1 .. 4*b + 3 of
// each axis, so every trilinear footprint whose floor corner x0 lies in
// [-1, d.x - 1] sits inside the brick with bx = (x0 + 1) >> 2, its 8 corners
// at the fixed offsets {0, 1, 5, 6, 25, 26, 30, 31} from the floor corner: one
// address per sample (the rest are immediate offsets of the loads), and the
// footprint spans 31 consecutive floats.  Halo cells outside the grid hold
// what the boundary mode reads there: 0 (BORDER, so out-of-grid corners need
// no mask), the clamped edge cell (CLAMP) or the wrapped cell (WRAP).
struct BrickDims {
  uint32_t nbx, nbxy;    // bricks per row / per plane
  uint32_t slice_cells;  // floats per channel of one slice (bricks * 128)
};

constexpr uint32_t kBrickFloats = 128;  // 5^3 = 125 halo cells, padded

inline BrickDims brick_dims(const Dims& d) {
  const uint32_t nbx = d.x / 4 + 1, nby = d.y / 4 + 1, nbz = d.z / 4 + 1;
  const uint64_t cells = static_cast<uint64_t>(nbx) * nby * nbz * kBrickFloats;
  return BrickDims{nbx, nbx * nby, cells < (uint64_t(1) << 32) ? static_cast<uint32_t>(cells) : 0u};
}

// Whether a C-channel slice of the haloed layout is addressable with 32-bit
// offsets (the bricked kernels are used only then).
// (The brick index is formed with 24-bit multiplies: bricks per plane below
// 2^24.)
inline bool bricked_fits(const Dims& d, int C) {
  const BrickDims b = brick_dims(d);
  return b.slice_cells != 0 && static_cast<uint64_t>(b.slice_cells) * C < (uint64_t(1) << 32) &&
         b.nbxy < (1u << 24);
}

// Offset (cells) of the floor corner of a footprint in the haloed layout.
// fx etc. are floor(q) as floats; BORDER and CLAMP clamp them into
// [-1, d - 1] (for BORDER that changes only positions outside the forward
// validity, whose value is zeroed anyway; for CLAMP both corners of a clamped
// axis then read the same edge value, as the clamped reads do), WRAP wraps
// them.  The clamp is one med3 on floor(q) + 1 (exact for |q| < 2^24, and
// any larger or NaN q lands in range as well).
template <int BM>
NGT_DEVICE uint32_t brick_corner(const Dims& d, const BrickDims& b, float fx, float fy, float fz) {
  uint32_t xp, yp, zp;  // corner + 1, in [0, d]
  if constexpr (BM == WRAP) {
    xp = static_cast<uint32_t>(wrap_index(static_cast<int>(fx), d.x) + 1);
    yp = static_cast<uint32_t>(wrap_index(static_cast<int>(fy), d.y) + 1);
    zp = static_cast<uint32_t>(wrap_index(static_cast<int>(fz), d.z) + 1);
  } else {
    xp = static_cast<uint32_t>(__builtin_amdgcn_fmed3f(fx + 1.f, 0.f, static_cast<float>(d.x)));
    yp = static_cast<uint32_t>(__builtin_amdgcn_fmed3f(fy + 1.f, 0.f, static_cast<float>(d.y)));
    zp = static_cast<uint32_t>(__builtin_amdgcn_fmed3f(fz + 1.f, 0.f, static_cast<float>(d.z)));
  }
  const uint32_t brick = __umul24(zp >> 2, b.nbxy) + __umul24(yp >> 2, b.nbx) + (xp >> 2);
  return brick * kBrickFloats + (xp & 3u) + (yp & 3u) * 5u + (zp & 3u) * 25u;
}

// Corner k (bit0 = x-ceil, bit1 = y-ceil, bit2 = z-ceil) relative to the floor corner.
NGT_DEVICE constexpr uint32_t brick_corner_step(int k) { return (k & 1) + ((k >> 1) & 1) * 5 + (k >> 2) * 25; }

NGT_DEVICE bool corner_ok(const Corners& c, int k) {
  return ((k & 1) ? c.vx1 : c.vx0) & ((k & 2) ? c.vy1 : c.vy0) & ((k & 4) ? c.vz1 : c.vz0);
}

template <int C, int BM>
NGT_DEVICE void mask8(const Corners& c, Vec<C> (&v)[8]) {
  if constexpr (BM == BORDER) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const bool ok = corner_ok(c, k);
#pragma unroll
      for (int i = 0; i < C; ++i) v[k].v[i] = ok ? v[k].v[i] : 0.f;
    }
  }
}

template <int C, int BM>
NGT_DEVICE Corner8<C> fetch8(const float* __restrict__ g, const Dims& d, const Corners& c) {
  Corner8<C> r;
  fetch8_raw<C, BM>(g, d, c, r.v);
  mask8<C, BM>(c, r.v);
#pragma unroll
  for (int k = 0; k < 8; ++k) r.ok[k] = BM == BORDER ? corner_ok(c, k) : true;
  return r;
}

// Trilinear value from masked corners; 0 outside the BORDER domain.
template <int C, int BM>
NGT_DEVICE Vec<C> lerp8(const Vec<C> (&v)[8], const Corners& c) {
  const Vec<C> v0 = vlerp(vlerp(v[0], v[1], c.wx), vlerp(v[2], v[3], c.wx), c.wy);
  const Vec<C> v1 = vlerp(vlerp(v[4], v[5], c.wx), vlerp(v[6], v[7], c.wx), c.wy);
  Vec<C> r = vlerp(v0, v1, c.wz);
  if constexpr (BM == BORDER) {
#pragma unroll
    for (int i = 0; i < C; ++i) r.v[i] = c.inside ? r.v[i] : 0.f;
  }
  return r;
}

template <int C, int BM>
NGT_DEVICE Vec<C> read_linear(const float* __restrict__ g, const Dims& d, float qx, float qy, float qz) {
  const Corners c = make_corners<BM>(d, qx, qy, qz);
  Vec<C> v[8];
  fetch8_raw<C, BM>(g, d, c, v);
  mask8<C, BM>(c, v);
  return lerp8<C, BM>(v, c);
}

// The corner loads of U trilinear reads in the haloed brick layout:
// bricked_taps_issue starts them, bricked_taps_finish interpolates.
template <int C, int U>
struct BrickedTaps {
  Vec<C> v[U][8];
  float wx[U], wy[U], wz[U];
  bool inside[U];
};

// The smallest float t with t + 0.5f >= d in float arithmetic (d >= 1 below
// 2^22), so that q + 0.5f < d exactly when q < t: d - 0.5, except for d = 1,
// where the largest float below 0.5 plus 0.5 rounds (to even) up to 1.
NGT_DEVICE float validity_upper(int d) {
  const float t = static_cast<float>(d) - 0.5f;
  const float below = __int_as_float(__float_as_int(t) - 1);
  return below + 0.5f >= static_cast<float>(d) ? below : t;
}

template <int C, int BM, int U>
NGT_DEVICE void bricked_taps_issue(const float* __restrict__ g, const Dims& d, const BrickDims& b,
                                   const float (&qx)[U], const float (&qy)[U], const float (&qz)[U],
                                   BrickedTaps<C, U>& t) {
  // q + 0.5 in [0, d) per axis, compared without the three adds as
  // q >= -0.5 (the sum is exact near -0.5) and q < validity_upper(d)
  float hx = 0.f, hy = 0.f, hz = 0.f;
  if constexpr (BM == BORDER) {
    hx = validity_upper(d.x); hy = validity_upper(d.y); hz = validity_upper(d.z);
  }
  // one address per sample, no per-corner masks
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const float fx = floorf(qx[u]), fy = floorf(qy[u]), fz = floorf(qz[u]);
    t.wx[u] = qx[u] - fx; t.wy[u] = qy[u] - fy; t.wz[u] = qz[u] - fz;
    t.inside[u] = true;
    if constexpr (BM == BORDER)
      t.inside[u] = (qx[u] >= -0.5f) & (qx[u] < hx) & (qy[u] >= -0.5f) & (qy[u] < hy) & (qz[u] >= -0.5f) &
                    (qz[u] < hz);
    const uint32_t base = brick_corner<BM>(d, b, fx, fy, fz);
    const float* p = g + base * static_cast<uint32_t>(C);
#pragma unroll
    for (int k = 0; k < 8; ++k) t.v[u][k] = vload<C>(p + brick_corner_step(k) * C);
  }
}

// keep[u] false zeroes sample u as well (one select with the BORDER mask).
template <int C, int BM, int U>
NGT_DEVICE void bricked_taps_finish(const BrickedTaps<C, U>& t, Vec<C> (&out)[U], const bool (&keep)[U]) {
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const Vec<C> v0 = vlerp(vlerp(t.v[u][0], t.v[u][1], t.wx[u]), vlerp(t.v[u][2], t.v[u][3], t.wx[u]), t.wy[u]);
    const Vec<C> v1 = vlerp(vlerp(t.v[u][4], t.v[u][5], t.wx[u]), vlerp(t.v[u][6], t.v[u][7], t.wx[u]), t.wy[u]);
    out[u] = vlerp(v0, v1, t.wz[u]);
    const bool ok = (BM == BORDER ? t.inside[u] : true) && keep[u];
#pragma unroll
    for (int i = 0; i < C; ++i) out[u].v[i] = ok ? out[u].v[i] : 0.f;
  }
}

// U trilinear reads whose 8*U loads are all issued before any is consumed;
// BRICKED selects the haloed brick layout (BrickDims b).
template <int C, int BM, int U, bool BRICKED>
NGT_DEVICE void read_linear_batch(const float* __restrict__ g, const Dims& d, const BrickDims& b, const float (&qx)[U],
                                  const float (&qy)[U], const float (&qz)[U], Vec<C> (&out)[U]) {
  if constexpr (BRICKED) {
    BrickedTaps<C, U> t;
    bricked_taps_issue<C, BM, U>(g, d, b, qx, qy, qz, t);
    bool keep[U];
#pragma unroll
    for (int u = 0; u < U; ++u) keep[u] = true;
    bricked_taps_finish<C, BM, U>(t, out, keep);
  } else {
    Vec<C> v[U][8];
    Corners c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) c[u] = make_corners<BM>(d, qx[u], qy[u], qz[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) fetch8_raw<C, BM>(g, d, c[u], v[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      mask8<C, BM>(c[u], v[u]);
      out[u] = lerp8<C, BM>(v[u], c[u]);
    }
  }
}

// Nearest cell: int(q + 0.5) truncates toward zero (sampling.hpp:570-584).
template <int BM>
NGT_DEVICE bool nearest_cell(const Dims& d, float qx, float qy, float qz, int& ix, int& iy, int& iz) {
  ix = static_cast<int>(qx + 0.5f); iy = static_cast<int>(qy + 0.5f); iz = static_cast<int>(qz + 0.5f);
  if constexpr (BM == BORDER) {
    return ix >= 0 && ix < d.x && iy >= 0 && iy < d.y && iz >= 0 && iz < d.z;
  } else if constexpr (BM == CLAMP) {
    ix = min(max(ix, 0), d.x - 1); iy = min(max(iy, 0), d.y - 1); iz = min(max(iz, 0), d.z - 1);
  } else {
    ix = wrap_index(ix, d.x); iy = wrap_index(iy, d.y); iz = wrap_index(iz, d.z);
  }
  return true;
}

template <int C, int BM>
NGT_DEVICE Vec<C> read_nearest(const float* __restrict__ g, const Dims& d, float qx, float qy, float qz) {
  int ix, iy, iz;
  if (!nearest_cell<BM>(d, qx, qy, qz, ix, iy, iz)) return vzero<C>();
  return vload<C>(g + cell_offset(d, ix, iy, iz) * C);
}

// Min / max over the 8 corners (sampling.hpp:585-704).  ``sel`` receives, per
// channel, the corner index (0..7) whose value was selected, or -1 when the
// BORDER start value 0 won.
template <int C, int BM, bool IS_MAX>
NGT_DEVICE Vec<C> read_minmax(const float* __restrict__ g, const Dims& d, float qx, float qy, float qz,
                              int* sel = nullptr, Corners* cout = nullptr) {
  const Corners c = make_corners<BM>(d, qx, qy, qz);
  if (cout) *cout = c;
  Vec<C> acc = vfill<C>(IS_MAX ? -kFloatMax : kFloatMax);
  if (sel) {
#pragma unroll
    for (int i = 0; i < C; ++i) sel[i] = -1;
  }
  if constexpr (BM == BORDER) {
    if (!(c.vx0 && c.vx1 && c.vy0 && c.vy1 && c.vz0 && c.vz1)) acc = vzero<C>();
  }
  const Corner8<C> k = fetch8<C, BM>(g, d, c);
#pragma unroll
  for (int j = 0; j < 8; ++j) {
#pragma unroll
    for (int i = 0; i < C; ++i) {
      const bool better = k.ok[j] && (IS_MAX ? (k.v[j].v[i] > acc.v[i]) : (k.v[j].v[i] < acc.v[i]));
      acc.v[i] = better ? k.v[j].v[i] : acc.v[i];
      if (sel) sel[i] = better ? j : sel[i];
    }
  }
  if constexpr (BM == BORDER) {
    if (!c.inside) {
#pragma unroll
      for (int i = 0; i < C; ++i) {
        acc.v[i] = 0.f;
        if (sel) sel[i] = -1;
      }
    }
  }
  return acc;
}

// Min and max over the 8 corners from one fetch (read_minmax semantics).
template <int C, int BM>
NGT_DEVICE void read_minmax2(const float* __restrict__ g, const Dims& d, float qx, float qy, float qz, Vec<C>& mn,
                             Vec<C>& mx) {
  const Corners c = make_corners<BM>(d, qx, qy, qz);
  mn = vfill<C>(kFloatMax);
  mx = vfill<C>(-kFloatMax);
  if constexpr (BM == BORDER) {
    if (!(c.vx0 && c.vx1 && c.vy0 && c.vy1 && c.vz0 && c.vz1)) {
      mn = vzero<C>();
      mx = vzero<C>();
    }
  }
  const Corner8<C> k = fetch8<C, BM>(g, d, c);
#pragma unroll
  for (int j = 0; j < 8; ++j) {
#pragma unroll
    for (int i = 0; i < C; ++i) {
      mn.v[i] = (k.ok[j] && k.v[j].v[i] < mn.v[i]) ? k.v[j].v[i] : mn.v[i];
      mx.v[i] = (k.ok[j] && k.v[j].v[i] > mx.v[i]) ? k.v[j].v[i] : mx.v[i];
    }
  }
  if constexpr (BM == BORDER) {
    if (!c.inside) {
      mn = vzero<C>();
      mx = vzero<C>();
    }
  }
}

// ---------------------------------------------------- quadratic B-spline ---
// 3x3x3 stencil (sampling.hpp:247-365): per axis the cells centre-1, centre,
// centre+1 with centre = floor(q + 0.5), weighted by the quadratic B-spline
// of their distance |q - i|: 3/4 - d^2 below 1/2, (3/2 - d)^2 / 2 below 3/2.
NGT_DEVICE float bspline2(float d) {
  return d < 0.5f ? 0.75f - d * d : (d < 1.5f ? 0.5f * (1.5f - d) * (1.5f - d) : 0.f);
}

struct QuadAxis {
  int i[3];    // unfixed cell indices
  float w[3];  // weights
};

NGT_DEVICE QuadAxis quad_axis(float q) {
  QuadAxis a;
  const int c = static_cast<int>(floorf(q + 0.5f));
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    a.i[t] = c - 1 + t;
    a.w[t] = bspline2(fabsf(q - static_cast<float>(a.i[t])));
  }
  return a;
}

template <int BM>
NGT_DEVICE int quad_fix(int i, int m) {
  if constexpr (BM == CLAMP) return min(max(i, 0), m - 1);
  else return wrap_index(i, m);
}

// read3DQuadratic: BORDER returns 0 everywhere (the reference returns before
// reading, sampling.hpp:284-287); CLAMP / WRAP fix the indices and sum along
// x, then y, then z.
template <int C, int BM>
NGT_DEVICE Vec<C> read_quadratic(const float* __restrict__ g, const Dims& d, float qx, float qy, float qz) {
  if constexpr (BM == BORDER) {
    return vzero<C>();
  } else {
    const QuadAxis ax = quad_axis(qx), ay = quad_axis(qy), az = quad_axis(qz);
    int ix[3], iy[3], iz[3];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      ix[t] = quad_fix<BM>(ax.i[t], d.x); iy[t] = quad_fix<BM>(ay.i[t], d.y); iz[t] = quad_fix<BM>(az.i[t], d.z);
    }
    Vec<C> r = vzero<C>();
#pragma unroll
    for (int tz = 0; tz < 3; ++tz) {
      Vec<C> vz = vzero<C>();
#pragma unroll
      for (int ty = 0; ty < 3; ++ty) {
        Vec<C> vy = vzero<C>();
#pragma unroll
        for (int tx = 0; tx < 3; ++tx) {
          const Vec<C> v = vload<C>(g + cell_offset(d, ix[tx], iy[ty], iz[tz]) * C);
#pragma unroll
          for (int i = 0; i < C; ++i) vy.v[i] += ax.w[tx] * v.v[i];
        }
#pragma unroll
        for (int i = 0; i < C; ++i) vz.v[i] += ay.w[ty] * vy.v[i];
      }
#pragma unroll
      for (int i = 0; i < C; ++i) r.v[i] += az.w[tz] * vz.v[i];
    }
    return r;
  }
}

// Exact adjoint of read_quadratic (nothing for BORDER, whose read is 0).
template <int C, int BM>
NGT_DEVICE void scatter_quadratic(float* __restrict__ g, const Dims& d, float qx, float qy, float qz,
                                  const Vec<C>& val) {
  if constexpr (BM != BORDER) {
    const QuadAxis ax = quad_axis(qx), ay = quad_axis(qy), az = quad_axis(qz);
#pragma unroll
    for (int tz = 0; tz < 3; ++tz)
#pragma unroll
      for (int ty = 0; ty < 3; ++ty)
#pragma unroll
        for (int tx = 0; tx < 3; ++tx) {
          const float w = ax.w[tx] * ay.w[ty] * az.w[tz];
          const int64_t o = cell_offset(d, quad_fix<BM>(ax.i[tx], d.x), quad_fix<BM>(ay.i[ty], d.y),
                                        quad_fix<BM>(az.i[tz], d.z));
#pragma unroll
          for (int i = 0; i < C; ++i) atomic_add(g + o * C + i, val.v[i] * w);
        }
  }
}

// scatterGradQuadratic (resample_grid.cu.cc:904-1080) with hit counts, for
// the reference gradient semantics.  BORDER: positions with -1.5 < q < dims
// scatter to the stencil cells whose centre-1 indices are >= 0 and centre+1
// indices < dims, never to the centre-centre-centre cell; the reference does
// not check the centre indices, which lie outside the grid for q + 0.5 < 0
// or q + 0.5 >= dims, and those cells are skipped here.
template <int C, int BM>
NGT_DEVICE void scatter_quadratic_reference(float* __restrict__ g, const Dims& d, float qx, float qy, float qz,
                                            const Vec<C>& val, uint32_t* __restrict__ hits) {
  const QuadAxis ax = quad_axis(qx), ay = quad_axis(qy), az = quad_axis(qz);
  if constexpr (BM == BORDER) {
    if (!(qx > -1.5f && qy > -1.5f && qz > -1.5f && qx < static_cast<float>(d.x) && qy < static_cast<float>(d.y) &&
          qz < static_cast<float>(d.z)))
      return;
  }
#pragma unroll
  for (int tz = 0; tz < 3; ++tz)
#pragma unroll
    for (int ty = 0; ty < 3; ++ty)
#pragma unroll
      for (int tx = 0; tx < 3; ++tx) {
        int ix = ax.i[tx], iy = ay.i[ty], iz = az.i[tz];
        if constexpr (BM == BORDER) {
          if (tx == 1 && ty == 1 && tz == 1) continue;
          if (ix < 0 || ix >= d.x || iy < 0 || iy >= d.y || iz < 0 || iz >= d.z) continue;
        } else {
          ix = quad_fix<BM>(ix, d.x); iy = quad_fix<BM>(iy, d.y); iz = quad_fix<BM>(iz, d.z);
        }
        const float w = ax.w[tx] * ay.w[ty] * az.w[tz];
        const int64_t o = cell_offset(d, ix, iy, iz);
#pragma unroll
        for (int i = 0; i < C; ++i) atomic_add(g + o * C + i, val.v[i] * w);
        atomicAdd(hits + o, 1u);
      }
}

template <int C, int FM, int BM>
NGT_DEVICE Vec<C> read_filtered(const float* __restrict__ g, const Dims& d, float qx, float qy, float qz) {
  if constexpr (FM == LINEAR) return read_linear<C, BM>(g, d, qx, qy, qz);
  else if constexpr (FM == QUADRATIC) return read_quadratic<C, BM>(g, d, qx, qy, qz);
  else if constexpr (FM == NEAREST) return read_nearest<C, BM>(g, d, qx, qy, qz);
  else if constexpr (FM == FMIN) return read_minmax<C, BM, false>(g, d, qx, qy, qz);
  else return read_minmax<C, BM, true>(g, d, qx, qy, qz);
}

// d read_linear / d q per channel (sampling.hpp:470-565).  Coinciding
// floor/ceil corners make the derivative along that axis 0 at integer q.
template <int C, int BM>
NGT_DEVICE void read_gradient(const float* __restrict__ g, const Dims& d, float qx, float qy, float qz,
                              Vec<C>& dx, Vec<C>& dy, Vec<C>& dz) {
  const Corners c = make_corners<BM>(d, qx, qy, qz);
  const Corner8<C> k = fetch8<C, BM>(g, d, c);
  dx = vlerp(vlerp(vsub(k.v[1], k.v[0]), vsub(k.v[3], k.v[2]), c.wy),
             vlerp(vsub(k.v[5], k.v[4]), vsub(k.v[7], k.v[6]), c.wy), c.wz);
  dy = vlerp(vlerp(vsub(k.v[2], k.v[0]), vsub(k.v[3], k.v[1]), c.wx),
             vlerp(vsub(k.v[6], k.v[4]), vsub(k.v[7], k.v[5]), c.wx), c.wz);
  dz = vlerp(vlerp(vsub(k.v[4], k.v[0]), vsub(k.v[5], k.v[1]), c.wx),
             vlerp(vsub(k.v[6], k.v[2]), vsub(k.v[7], k.v[3]), c.wx), c.wy);
  if constexpr (BM == BORDER) {
#pragma unroll
    for (int i = 0; i < C; ++i) {
      dx.v[i] = c.inside ? dx.v[i] : 0.f;
      dy.v[i] = c.inside ? dy.v[i] : 0.f;
      dz.v[i] = c.inside ? dz.v[i] : 0.f;
    }
  }
}

// read_linear and read_gradient of one position from one fetch of the 8
// corners (identical results).
template <int C, int BM>
NGT_DEVICE Vec<C> read_linear_gradient(const float* __restrict__ g, const Dims& d, float qx, float qy, float qz,
                                       Vec<C>& dx, Vec<C>& dy, Vec<C>& dz) {
  const Corners c = make_corners<BM>(d, qx, qy, qz);
  Vec<C> k[8];
  fetch8_raw<C, BM>(g, d, c, k);
  mask8<C, BM>(c, k);
  dx = vlerp(vlerp(vsub(k[1], k[0]), vsub(k[3], k[2]), c.wy), vlerp(vsub(k[5], k[4]), vsub(k[7], k[6]), c.wy), c.wz);
  dy = vlerp(vlerp(vsub(k[2], k[0]), vsub(k[3], k[1]), c.wx), vlerp(vsub(k[6], k[4]), vsub(k[7], k[5]), c.wx), c.wz);
  dz = vlerp(vlerp(vsub(k[4], k[0]), vsub(k[5], k[1]), c.wx), vlerp(vsub(k[6], k[2]), vsub(k[7], k[3]), c.wx), c.wy);
  if constexpr (BM == BORDER) {
#pragma unroll
    for (int i = 0; i < C; ++i) {
      dx.v[i] = c.inside ? dx.v[i] : 0.f;
      dy.v[i] = c.inside ? dy.v[i] : 0.f;
      dz.v[i] = c.inside ? dz.v[i] : 0.f;
    }
  }
  return lerp8<C, BM>(k, c);
}

// Scatter ``val`` with trilinear weights to the 8 corners.
//  REF_MASK = false: exact adjoint (BORDER forward validity).
//  REF_MASK = true : the reference's acceptance -1 < q < dims
//                    (resample_grid.cu.cc:1082, sampling_v2.hpp:475).
//  hits  (nullable): +1 per in-grid corner (count normalisation).
//  wsum  (nullable): +weight per in-grid corner (weight normalisation).
template <int C, int BM, bool REF_MASK>
NGT_DEVICE void scatter_linear(float* __restrict__ g, const Dims& d, float qx, float qy, float qz, const Vec<C>& val,
                               uint32_t* __restrict__ hits = nullptr, float* __restrict__ wsum = nullptr) {
  const Corners c = make_corners<BM>(d, qx, qy, qz);
  if constexpr (BM == BORDER) {
    if constexpr (REF_MASK) {
      if (!(qx > -1.f && qx < static_cast<float>(d.x) && qy > -1.f && qy < static_cast<float>(d.y) &&
            qz > -1.f && qz < static_cast<float>(d.z)))
        return;
    } else {
      if (!c.inside) return;
    }
  }
  const float fw[2][3] = {{1.f - c.wx, 1.f - c.wy, 1.f - c.wz}, {c.wx, c.wy, c.wz}};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int bx = k & 1, by = (k >> 1) & 1, bz = (k >> 2) & 1;
    if constexpr (BM == BORDER) {
      const bool ok = (bx ? c.vx1 : c.vx0) && (by ? c.vy1 : c.vy0) && (bz ? c.vz1 : c.vz0);
      if (!ok) continue;
    }
    const int ix = bx ? c.x1 : c.x0, iy = by ? c.y1 : c.y0, iz = bz ? c.z1 : c.z0;
    const float w = fw[bx][0] * fw[by][1] * fw[bz][2];
    const int64_t o = cell_offset(d, ix, iy, iz);
#pragma unroll
    for (int i = 0; i < C; ++i) atomic_add(g + o * C + i, val.v[i] * w);
    if (hits) atomicAdd(hits + o, 1u);
    if (wsum) atomic_add(wsum + o, w);
  }
}

template <int C, int BM>
NGT_DEVICE void scatter_nearest(float* __restrict__ g, const Dims& d, float qx, float qy, float qz, const Vec<C>& val) {
  int ix, iy, iz;
  if (!nearest_cell<BM>(d, qx, qy, qz, ix, iy, iz)) return;
  const int64_t o = cell_offset(d, ix, iy, iz);
#pragma unroll
  for (int i = 0; i < C; ++i) atomic_add(g + o * C + i, val.v[i]);
}

template <int C, int BM, bool IS_MAX>
NGT_DEVICE void scatter_minmax(const float* __restrict__ src, float* __restrict__ g, const Dims& d, float qx, float qy,
                               float qz, const Vec<C>& val) {
  int sel[C];
  Corners c;
  read_minmax<C, BM, IS_MAX>(src, d, qx, qy, qz, sel, &c);
  if constexpr (BM == BORDER) {
    if (!c.inside) return;
  }
#pragma unroll
  for (int i = 0; i < C; ++i) {
    const int k = sel[i];
    if (k < 0) continue;
    const int ix = (k & 1) ? c.x1 : c.x0, iy = (k & 2) ? c.y1 : c.y0, iz = (k & 4) ? c.z1 : c.z0;
    atomic_add(g + cell_offset(d, ix, iy, iz) * C + i, val.v[i]);
  }
}

// Exact adjoint of read_filtered<FM>.
template <int C, int FM, int BM>
NGT_DEVICE void scatter_filtered(const float* __restrict__ src, float* __restrict__ g, const Dims& d, float qx,
                                 float qy, float qz, const Vec<C>& val) {
  if constexpr (FM == LINEAR) scatter_linear<C, BM, false>(g, d, qx, qy, qz, val);
  else if constexpr (FM == QUADRATIC) scatter_quadratic<C, BM>(g, d, qx, qy, qz, val);
  else if constexpr (FM == NEAREST) scatter_nearest<C, BM>(g, d, qx, qy, qz, val);
  else if constexpr (FM == FMIN) scatter_minmax<C, BM, false>(src, g, d, qx, qy, qz, val);
  else scatter_minmax<C, BM, true>(src, g, d, qx, qy, qz, val);
}

// ------------------------------------------------------------- cameras ------

// One (model, camera) pair, matrices row-major in math order, prepared on the
// host in FP64 and rounded to FP32 (the reference inverts in FP32 on the host,
// transformations.hpp:48-68).
struct CamParams {
  float mv[16];
  float mv_inv[16];
  float p[16];
  float p_inv[16];
  float near_, far_;
  int data_index;    // batch entry n
  int out_index;     // output slot n*V + v (or n)
};

constexpr int kMaxCamsPerLaunch = 8;

struct CamBatch {
  CamParams cam[kMaxCamsPerLaunch];
  int count;
};

NGT_DEVICE void mat4_apply(const float* __restrict__ m, float x, float y, float z, float w, float& ox, float& oy,
                           float& oz, float& ow) {
  ox = m[0] * x + m[1] * y + m[2] * z + m[3] * w;
  oy = m[4] * x + m[5] * y + m[6] * z + m[7] * w;
  oz = m[8] * x + m[9] * y + m[10] * z + m[11] * w;
  ow = m[12] * x + m[13] * y + m[14] * z + m[15] * w;
}

// NDC -> VS (divide by w) -> OS (transformations.hpp:124-138).
NGT_DEVICE void ndc_to_os(const CamParams& cp, float nx, float ny, float nz, float& ox, float& oy, float& oz) {
  float vx, vy, vz, vw;
  mat4_apply(cp.p_inv, nx, ny, nz, 1.f, vx, vy, vz, vw);
  const float iw = 1.f / vw;
  vx *= iw; vy *= iw; vz *= iw;
  float ow;
  mat4_apply(cp.mv_inv, vx, vy, vz, 1.f, ox, oy, oz, ow);
}

NGT_DEVICE void idx_to_os(const CamParams& cp, float ix, float iy, float iz, float dx_inv, float dy_inv, float dz_inv,
                          float& ox, float& oy, float& oz) {
  ndc_to_os(cp, ix * dx_inv * 2.f - 1.f, iy * dy_inv * 2.f - 1.f, iz * dz_inv * 2.f - 1.f, ox, oy, oz);
}

// Linear view-space depth slices (transformations.hpp:139-149).
NGT_DEVICE void idx_to_os_lindepth(const CamParams& cp, float ix, float iy, float iz, float dx_inv, float dy_inv,
                                   float dz_inv, float& ox, float& oy, float& oz) {
  const float t = iz * dz_inv;
  const float zv = -fmaf(cp.far_ - cp.near_, t, cp.near_);
  const float nz = -cp.p[10] - cp.p[11] / zv;
  ndc_to_os(cp, ix * dx_inv * 2.f - 1.f, iy * dy_inv * 2.f - 1.f, nz, ox, oy, oz);
}

NGT_DEVICE void os_to_ndc(const CamParams& cp, float x, float y, float z, float& nx, float& ny, float& nz) {
  float vx, vy, vz, vw;
  mat4_apply(cp.mv, x, y, z, 1.f, vx, vy, vz, vw);
  float cx, cy, cz, cw;
  mat4_apply(cp.p, vx, vy, vz, vw, cx, cy, cz, cw);
  const float iw = 1.f / cw;
  nx = cx * iw; ny = cy * iw; nz = cz * iw;
}

// OStoIDX (transformations.hpp:100-108).
NGT_DEVICE void os_to_idx(const CamParams& cp, float x, float y, float z, float dx, float dy, float dz, float& ix,
                          float& iy, float& iz) {
  float nx, ny, nz;
  os_to_ndc(cp, x, y, z, nx, ny, nz);
  ix = (nx * 0.5f + 0.5f) * dx; iy = (ny * 0.5f + 0.5f) * dy; iz = (nz * 0.5f + 0.5f) * dz;
}

// OStoIDXlinearDepth (transformations.hpp:110-121).
NGT_DEVICE void os_to_idx_lindepth(const CamParams& cp, float x, float y, float z, float dx, float dy, float dz,
                                   float& ix, float& iy, float& iz) {
  float nx, ny, nz;
  os_to_ndc(cp, x, y, z, nx, ny, nz);
  ix = (nx * 0.5f + 0.5f) * dx; iy = (ny * 0.5f + 0.5f) * dy;
  const float zv = -cp.p[11] / (nz + cp.p[10]);
  const float t = -(zv + cp.near_) / (cp.far_ - cp.near_);
  iz = t * dz;
}

// ------------------------------------------------------------- blending -----

// e^-d (d >= 0, an optical depth) as 2^(-d log2 e) on the hardware exp2
// (v_exp_f32, about 1 ulp): rounding the product adds a relative error of
// about d * 2^-24, i.e. < 1e-6 for the depths whose transmittance still
// matters at the 1e-5 forward tolerance (e^-16 ~ 1e-7).  Forward and adjoint
// walks share it, so the adjoint differentiates the forward actually computed.
NGT_DEVICE float exp_neg(float d) { return __builtin_amdgcn_exp2f(d * -1.44269504088896341f); }

// Forward blend functors (blending.hpp:9-243).  C == 1 and ADDITIVE are plain
// sums; for C in {2,4} the last channel is density / alpha.
template <int C, int MODE>
NGT_DEVICE void blend_step(Vec<C>& acc, const Vec<C>& cell) {
  if constexpr (C == 1 || MODE == ADDITIVE) {
#pragma unroll
    for (int i = 0; i < C; ++i) acc.v[i] += cell.v[i];
  } else if constexpr (MODE == BEER_LAMBERT) {
    const float dd = fmaxf(acc.v[C - 1] + cell.v[C - 1], 0.f);
    const float t = exp_neg(dd);
    // one explicit rounding per channel, so that every walk that blends
    // (image, recording, adjoint) forms the same image bits
#pragma unroll
    for (int i = 0; i < C - 1; ++i) acc.v[i] = fmaf(t, cell.v[i], acc.v[i]);
    acc.v[C - 1] = dd;
  } else if constexpr (MODE == ALPHA) {
    const float t = 1.f - acc.v[C - 1];
#pragma unroll
    for (int i = 0; i < C - 1; ++i) acc.v[i] += t * cell.v[i];
    acc.v[C - 1] += t * cell.v[C - 1];
  } else {  // ALPHA_ADDITIVE
    const float t = 1.f - cell.v[C - 1];
#pragma unroll
    for (int i = 0; i < C - 1; ++i) acc.v[i] += t * cell.v[i];
    acc.v[C - 1] += cell.v[C - 1];
  }
}

// Reference backward functor verbatim in behaviour (blending.hpp:40-54,
// 139-153, 188-202): updates ``grads`` to the gradient w.r.t. the previous
// accumulator, reconstructs ``cell_out`` to the previous accumulator and
// returns the gradient w.r.t. the cell.
template <int C, int MODE>
NGT_DEVICE Vec<C> blend_grad_reference(Vec<C>& grads, const Vec<C>& cell, Vec<C>& cell_out) {
  if constexpr (C == 1 || MODE == ADDITIVE) {
    return grads;
  } else {
    float lg = 0.f;
#pragma unroll
    for (int i = 0; i < C - 1; ++i) lg += cell.v[i] * grads.v[i];
    Vec<C> gc;
    if constexpr (MODE == BEER_LAMBERT) {
      const float e = exp_neg(cell_out.v[C - 1]);
      cell_out.v[C - 1] = fmaxf(cell_out.v[C - 1] - cell.v[C - 1], 0.f);
      const float dd = grads.v[C - 1] - e * lg;
#pragma unroll
      for (int i = 0; i < C - 1; ++i) gc.v[i] = e * grads.v[i];
      gc.v[C - 1] = dd;
      grads.v[C - 1] = dd;
    } else if constexpr (MODE == ALPHA) {
      cell_out.v[C - 1] = (1.f - cell_out.v[C - 1]) / (1.f - cell.v[C - 1]);
      const float t = 1.f - cell_out.v[C - 1];
      const float dacc = (1.f - cell.v[C - 1]) * grads.v[C - 1] - lg;
#pragma unroll
      for (int i = 0; i < C - 1; ++i) gc.v[i] = t * grads.v[i];
      gc.v[C - 1] = t * grads.v[C - 1];
      grads.v[C - 1] = dacc;
    } else {  // ALPHA_ADDITIVE
      cell_out.v[C - 1] = cell_out.v[C - 1] - cell.v[C - 1];
      const float t = 1.f - cell.v[C - 1];
#pragma unroll
      for (int i = 0; i < C - 1; ++i) gc.v[i] = t * grads.v[i];
      gc.v[C - 1] = grads.v[C - 1] - lg;
    }
    return gc;
  }
}

// ------------------------------------------------------- misc kernels -------

__global__ void fill_kernel(float* __restrict__ p, int64_t n, float value);
__global__ void fill_u32_kernel(uint32_t* __restrict__ p, int64_t n, uint32_t value);

}  // namespace ngt
