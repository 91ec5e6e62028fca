// SampleGridLut / SampleGridTransform / LodTransform and their adjoints on
// gfx950, with mip-mapped reads.
//
// Forward kernels: one work-item per output cell, x fastest, 64x4 work-groups
// so a wavefront reads 64 consecutive LUT float4s (coalesced) and writes 64
// consecutive output cells; the 8-tap gathers hit L2/MALL (the whole input of
// the configurations here is cache-resident).  All (batch, camera) pairs go in
// one launch (grid.z = depth x pairs) instead of the reference's per-camera
// launch + device sync loop (resample_grid.cu.cc:712-794).
//
// Backward: the LUT / transform positions are arbitrary, so the input
// gradient is a scatter with hardware fp32 atomics; the LUT gradient is a
// pure gather (data gradient . dy) accumulated over the data batch in
// registers.
//
// Mips: the pyramid of every batch entry is built once per call in the
// workspace (one launch per level, 2x2x2 means in the reference's pairwise
// order z, y, x: resample_grid.cu.cc:58-84).  The exact adjoint scatters into
// per-level gradient buffers and collapses them coarse to fine, each fine
// cell adding 1/8 of its parent (no atomics in the collapse).
#include "ngt_host.hpp"

namespace ngt {
namespace {

constexpr int kBX = 64, kBY = 4;
constexpr int kMaxMips = 16;

// ----------------------------------------------------------------- mips -----

NGT_HOST_DEVICE_INLINE Dims level_dims(const Dims& d, int m) { return Dims{d.x >> m, d.y >> m, d.z >> m}; }
NGT_HOST_DEVICE_INLINE int64_t dims_cells(const Dims& d) {
  return static_cast<int64_t>(d.x) * d.y * d.z;
}

// A pyramid of n_data grids: level 0 is the caller's buffer (`base`, N
// slices); levels 1..L follow each other in `levels`, level m holding
// [N][dims >> m][C].
struct MipSet {
  float* base;
  float* levels;
  Dims d0;
  int n_data;
  int num_levels;  // L
};

template <int C>
NGT_HOST_DEVICE_INLINE float* level_ptr(const MipSet& s, int n, int m, Dims& dm) {
  dm = level_dims(s.d0, m);
  if (m == 0) return s.base + n * dims_cells(s.d0) * C;
  int64_t off = 0;
  for (int k = 1; k < m; ++k) off += dims_cells(level_dims(s.d0, k));
  return s.levels + (off * s.n_data + n * dims_cells(dm)) * C;
}

// floats of levels 1..L of N grids
int64_t mip_floats(const Dims& d0, int64_t n, int64_t c, int levels) {
  int64_t f = 0;
  for (int m = 1; m <= levels; ++m) f += dims_cells(level_dims(d0, m));
  return f * n * c;
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Level choice of one sample (Sampler2::sample3D, sampling.hpp:939-1010).
struct MipPick {
  int lf, lc;
  float wc;     // weight of level lc (0 unless two levels are read)
  bool two;     // LINEAR and lc > lf
  bool inside;  // raw + bias strictly inside (0, L): d lod / d raw = 1
};

template <int MM>
NGT_DEVICE MipPick mip_pick(float lod_raw, float bias, int L) {
  const float raw = lod_raw + bias;
  const float lod = fminf(fmaxf(raw, 0.f), static_cast<float>(L));
  MipPick p;
  if constexpr (MM == NGT_MIP_NEAREST) {
    p.lf = min(static_cast<int>(lod + 0.5f), L);
    p.lc = p.lf;
    p.two = false;
    p.wc = 0.f;
  } else {
    p.lf = min(static_cast<int>(lod), L);
    p.lc = min(static_cast<int>(ceilf(lod)), L);
    p.two = p.lc > p.lf;
    p.wc = p.two ? lod - floorf(lod) : 0.f;
  }
  p.inside = raw > 0.f && raw < static_cast<float>(L);
  return p;
}

NGT_DEVICE float level_scale(int m) { return 1.f / static_cast<float>(1 << m); }

template <int C, int FM, int BM>
NGT_DEVICE Vec<C> read_level(const MipSet& s, int n, int m, float px, float py, float pz, float cco) {
  Dims dm;
  const float* g = level_ptr<C>(s, n, m, dm);
  const float sc = level_scale(m);
  return read_filtered<C, FM, BM>(g, dm, px * sc - cco, py * sc - cco, pz * sc - cco);
}

// Mip-mapped read of position p (before the cell-centre offset).
template <int C, int FM, int BM, int MM>
NGT_DEVICE Vec<C> read_mip(const MipSet& s, int n, float px, float py, float pz, float lod_raw, float bias,
                           float cco) {
  const MipPick k = mip_pick<MM>(lod_raw, bias, s.num_levels);
  Vec<C> r = read_level<C, FM, BM>(s, n, k.lf, px, py, pz, cco);
  if constexpr (MM == NGT_MIP_LINEAR) {
    if (k.two) r = vlerp(r, read_level<C, FM, BM>(s, n, k.lc, px, py, pz, cco), k.wc);
  }
  return r;
}

template <int C, int FM, int BM>
NGT_DEVICE void scatter_level(const MipSet& data, const MipSet& grad, int n, int m, float px, float py, float pz,
                              float cco, const Vec<C>& val) {
  Dims dm;
  const float* src = level_ptr<C>(data, n, m, dm);
  float* g = level_ptr<C>(grad, n, m, dm);
  const float sc = level_scale(m);
  scatter_filtered<C, FM, BM>(src, g, dm, px * sc - cco, py * sc - cco, pz * sc - cco, val);
}

// Level m+1 of n_data grids from level m: each cell the mean of its 2x2x2
// children, averaged along z, then y, then x.
template <int C>
__global__ __launch_bounds__(256) void mip_down_kernel(const float* __restrict__ fine, Dims df,
                                                       float* __restrict__ coarse, Dims dc, int n_data) {
  const int64_t cells = dims_cells(dc);
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= cells * n_data) return;
  const int n = static_cast<int>(i / cells);
  int64_t r = i - n * cells;
  const int x = static_cast<int>(r % dc.x);
  r /= dc.x;
  const int y = static_cast<int>(r % dc.y);
  const int z = static_cast<int>(r / dc.y);
  const float* f = fine + n * dims_cells(df) * C;
  Vec<C> v[2][2][2];
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int a = 0; a < 2; ++a) v[c][b][a] = vload<C>(f + cell_offset(df, 2 * x + a, 2 * y + b, 2 * z + c) * C);
  Vec<C> out;
#pragma unroll
  for (int k = 0; k < C; ++k) {
    float yx[2][2];
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int a = 0; a < 2; ++a) yx[b][a] = (v[0][b][a].v[k] + v[1][b][a].v[k]) * 0.5f;
    const float x0 = (yx[0][0] + yx[1][0]) * 0.5f;
    const float x1 = (yx[0][1] + yx[1][1]) * 0.5f;
    out.v[k] = (x0 + x1) * 0.5f;
  }
  vstore<C>(coarse + (n * cells + cell_offset(dc, x, y, z)) * C, out);
}

// Adjoint of one mip_down: every fine cell of the 2*dims_coarse region adds
// 1/8 of its parent's gradient.
template <int C>
__global__ __launch_bounds__(256) void mip_collapse_kernel(const float* __restrict__ coarse, Dims dc,
                                                           float* __restrict__ fine, Dims df, int n_data) {
  const Dims dr{2 * dc.x, 2 * dc.y, 2 * dc.z};
  const int64_t cells = dims_cells(dr);
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= cells * n_data) return;
  const int n = static_cast<int>(i / cells);
  int64_t r = i - n * cells;
  const int x = static_cast<int>(r % dr.x);
  r /= dr.x;
  const int y = static_cast<int>(r % dr.y);
  const int z = static_cast<int>(r / dr.y);
  const Vec<C> p = vload<C>(coarse + (n * dims_cells(dc) + cell_offset(dc, x >> 1, y >> 1, z >> 1)) * C);
  float* o = fine + (n * dims_cells(df) + cell_offset(df, x, y, z)) * C;
  Vec<C> v = vload<C>(o);
#pragma unroll
  for (int k = 0; k < C; ++k) v.v[k] += 0.125f * p.v[k];
  vstore<C>(o, v);
}

template <int C>
void launch_mip_generate(const MipSet& s, hipStream_t st) {
  for (int m = 0; m < s.num_levels; ++m) {
    Dims df, dc;
    const float* f = level_ptr<C>(s, 0, m, df);
    float* c = level_ptr<C>(s, 0, m + 1, dc);
    const int64_t n = dims_cells(dc) * s.n_data;
    hipLaunchKernelGGL((mip_down_kernel<C>), dim3(blocks_for(n, 256)), dim3(256), 0, st, f, df, c, dc, s.n_data);
  }
}

template <int C>
void launch_mip_collapse(const MipSet& g, hipStream_t st) {
  for (int m = g.num_levels; m >= 1; --m) {
    Dims df, dc;
    float* f = level_ptr<C>(g, 0, m - 1, df);
    const float* c = level_ptr<C>(g, 0, m, dc);
    const int64_t n = 8 * dims_cells(dc) * g.n_data;
    hipLaunchKernelGGL((mip_collapse_kernel<C>), dim3(blocks_for(n, 256)), dim3(256), 0, st, c, dc, f, df, g.n_data);
  }
}

// ------------------------------------------------------------------ LUT -----

struct LutArgs {
  const float* input;
  Dims din;
  const float* lut;
  Dims dout;
  int n_data, n_cams, separate;
  float cco;
  int relative, normalized;
  float mip_bias;
};

// Sample position (before the cell-centre offset) and raw LoD of one LUT cell
// (resample_grid.cu.cc:610-618).
NGT_DEVICE void lut_position(const LutArgs& a, int v, int x, int y, int z, float& px, float& py, float& pz,
                             float& lod) {
  const float4 l = *reinterpret_cast<const float4*>(a.lut + (cell_offset(a.dout, x, y, z) + static_cast<int64_t>(v) * a.dout.x * a.dout.y * a.dout.z) * 4);
  px = l.x; py = l.y; pz = l.z; lod = l.w;
  if (a.normalized) {
    px *= static_cast<float>(a.dout.x); py *= static_cast<float>(a.dout.y); pz *= static_cast<float>(a.dout.z);
  }
  if (a.relative) {
    px += static_cast<float>(x); py += static_cast<float>(y); pz += static_cast<float>(z);
  }
}

template <int C, int FM, int BM, int MM>
__global__ __launch_bounds__(256) void lut_fwd_kernel(LutArgs a, MipSet mips, float* __restrict__ out) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  const int z = blockIdx.z % a.dout.z;
  const int pair = blockIdx.z / a.dout.z;
  if (x >= a.dout.x || y >= a.dout.y) return;
  int n, v;
  if (a.separate) { n = pair / a.n_cams; v = pair % a.n_cams; }
  else { n = pair; v = pair; }
  float px, py, pz, lod;
  lut_position(a, v, x, y, z, px, py, pz, lod);
  Vec<C> r;
  if constexpr (MM == NGT_MIP_NONE) {
    const int64_t in_slice = static_cast<int64_t>(a.din.x) * a.din.y * a.din.z * C;
    r = read_filtered<C, FM, BM>(a.input + n * in_slice, a.din, px - a.cco, py - a.cco, pz - a.cco);
  } else {
    r = read_mip<C, FM, BM, MM>(mips, n, px, py, pz, lod, a.mip_bias, a.cco);
  }
  const int64_t out_cells = static_cast<int64_t>(a.dout.x) * a.dout.y * a.dout.z;
  vstore<C>(out + (static_cast<int64_t>(pair) * out_cells + cell_offset(a.dout, x, y, z)) * C, r);
}

// One work-item per (camera slot, output cell); loops over the data entries
// that use this camera so the LUT gradient is summed in registers.
template <int C, int FM, int BM, bool REF>
__global__ __launch_bounds__(256) void lut_bwd_kernel(LutArgs a, const float* __restrict__ out_grad,
                                                      float* __restrict__ in_grad, float* __restrict__ lut_grad,
                                                      uint32_t* __restrict__ hits) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  const int z = blockIdx.z % a.dout.z;
  const int v = blockIdx.z / a.dout.z;
  if (x >= a.dout.x || y >= a.dout.y) return;
  float px, py, pz, lod;
  lut_position(a, v, x, y, z, px, py, pz, lod);
  px -= a.cco; py -= a.cco; pz -= a.cco;
  const int64_t in_cells = static_cast<int64_t>(a.din.x) * a.din.y * a.din.z;
  const int64_t out_cells = static_cast<int64_t>(a.dout.x) * a.dout.y * a.dout.z;
  const int64_t cell = cell_offset(a.dout, x, y, z);
  float g0 = 0.f, g1 = 0.f, g2 = 0.f;
  const int n_begin = a.separate ? 0 : v;
  const int n_end = a.separate ? a.n_data : v + 1;
  for (int n = n_begin; n < n_end; ++n) {
    const int slot = a.separate ? n * a.n_cams + v : n;
    const Vec<C> dy = vload<C>(out_grad + (static_cast<int64_t>(slot) * out_cells + cell) * C);
    const float* src = a.input + n * in_cells * C;
    if constexpr (FM == LINEAR) {
      Vec<C> dx, dyy, dz;
      read_gradient<C, BM>(src, a.din, px, py, pz, dx, dyy, dz);
      if constexpr (REF) {  // written per camera, the last batch entry wins
        g0 = vdot(dy, dx); g1 = vdot(dy, dyy); g2 = vdot(dy, dz);
      } else {
        g0 += vdot(dy, dx); g1 += vdot(dy, dyy); g2 += vdot(dy, dz);
      }
    }
    if constexpr (REF) {
      scatter_linear<C, BM, true>(in_grad + n * in_cells * C, a.din, px, py, pz, dy, hits + n * in_cells);
    } else {
      scatter_filtered<C, FM, BM>(src, in_grad + n * in_cells * C, a.din, px, py, pz, dy);
    }
  }
  if (!REF && a.normalized) {
    g0 *= static_cast<float>(a.dout.x); g1 *= static_cast<float>(a.dout.y); g2 *= static_cast<float>(a.dout.z);
  }
  *reinterpret_cast<float4*>(lut_grad + (static_cast<int64_t>(v) * out_cells + cell) * 4) = make_float4(g0, g1, g2, 0.f);
}

// Exact adjoint of the mip-mapped LUT read: scatter into the levels read;
// LUT gradient = level-weighted data gradients (LINEAR filter) and, in w, the
// derivative of the level lerp.
template <int C, int FM, int BM, int MM>
__global__ __launch_bounds__(256) void lut_bwd_mip_kernel(LutArgs a, MipSet data, MipSet grad,
                                                          const float* __restrict__ out_grad,
                                                          float* __restrict__ lut_grad) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  const int z = blockIdx.z % a.dout.z;
  const int v = blockIdx.z / a.dout.z;
  if (x >= a.dout.x || y >= a.dout.y) return;
  float px, py, pz, lod_raw;
  lut_position(a, v, x, y, z, px, py, pz, lod_raw);
  const MipPick k = mip_pick<MM>(lod_raw, a.mip_bias, data.num_levels);
  const float wf = 1.f - k.wc;
  const int64_t out_cells = static_cast<int64_t>(a.dout.x) * a.dout.y * a.dout.z;
  const int64_t cell = cell_offset(a.dout, x, y, z);
  float g0 = 0.f, g1 = 0.f, g2 = 0.f, gl = 0.f;
  const int n_begin = a.separate ? 0 : v;
  const int n_end = a.separate ? a.n_data : v + 1;
  for (int n = n_begin; n < n_end; ++n) {
    const int slot = a.separate ? n * a.n_cams + v : n;
    const Vec<C> dy = vload<C>(out_grad + (static_cast<int64_t>(slot) * out_cells + cell) * C);
    scatter_level<C, FM, BM>(data, grad, n, k.lf, px, py, pz, a.cco, vscale(dy, wf));
    if (k.two) scatter_level<C, FM, BM>(data, grad, n, k.lc, px, py, pz, a.cco, vscale(dy, k.wc));
    if constexpr (FM == LINEAR) {
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int m = t ? k.lc : k.lf;
        const float w = t ? k.wc : wf;
        if (t == 0 || k.two) {
          Dims dm;
          const float* g = level_ptr<C>(data, n, m, dm);
          const float sc = level_scale(m);
          Vec<C> dx, dyy, dz;
          read_gradient<C, BM>(g, dm, px * sc - a.cco, py * sc - a.cco, pz * sc - a.cco, dx, dyy, dz);
          g0 += w * sc * vdot(dy, dx); g1 += w * sc * vdot(dy, dyy); g2 += w * sc * vdot(dy, dz);
        }
      }
    }
    if constexpr (MM == NGT_MIP_LINEAR) {
      if (k.two && k.inside) {
        const Vec<C> vf = read_level<C, FM, BM>(data, n, k.lf, px, py, pz, a.cco);
        const Vec<C> vc = read_level<C, FM, BM>(data, n, k.lc, px, py, pz, a.cco);
        gl += vdot(dy, vsub(vc, vf));
      }
    }
  }
  if (a.normalized) {
    g0 *= static_cast<float>(a.dout.x); g1 *= static_cast<float>(a.dout.y); g2 *= static_cast<float>(a.dout.z);
  }
  *reinterpret_cast<float4*>(lut_grad + (static_cast<int64_t>(v) * out_cells + cell) * 4) = make_float4(g0, g1, g2, gl);
}

// grad /= hits where hits > 0 (resample_grid.cu.cc:1133-1151).
template <int C>
__global__ __launch_bounds__(256) void normalize_by_count_kernel(float* __restrict__ g, const uint32_t* __restrict__ hits,
                                                                 int64_t cells) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= cells) return;
  const uint32_t h = hits[i];
  if (h > 0u) {
    const float w = 1.0f / static_cast<float>(h);
    Vec<C> d = vload<C>(g + i * C);
    vstore<C>(g + i * C, vscale(d, w));
  }
}

// ----------------------------------------------------------- transform ------

struct XformArgs {
  const float* input;
  Dims din;
  Dims dout;
  float out_inv_x, out_inv_y, out_inv_z;
  int coord_mode;
  float mip_bias;
};

// samplePos3D<CM, false> at index coordinates (cell index + 0.5).
NGT_DEVICE void xform_pos(const XformArgs& a, const CamParams& cp, float ix, float iy, float iz, float& px, float& py,
                          float& pz) {
  switch (a.coord_mode) {
    case NGT_COORD_TRANSFORM:
      idx_to_os(cp, ix, iy, iz, a.out_inv_x, a.out_inv_y, a.out_inv_z, px, py, pz);
      break;
    case NGT_COORD_TRANSFORM_REVERSE:
      os_to_idx(cp, ix, iy, iz, static_cast<float>(a.din.x), static_cast<float>(a.din.y), static_cast<float>(a.din.z),
                px, py, pz);
      break;
    case NGT_COORD_TRANSFORM_LINDEPTH_REVERSE:
      os_to_idx_lindepth(cp, ix, iy, iz, static_cast<float>(a.din.x), static_cast<float>(a.din.y),
                         static_cast<float>(a.din.z), px, py, pz);
      break;
    default:
      idx_to_os_lindepth(cp, ix, iy, iz, a.out_inv_x, a.out_inv_y, a.out_inv_z, px, py, pz);
      break;
  }
}

NGT_DEVICE bool forward_mode(int cm) { return cm == NGT_COORD_TRANSFORM || cm == NGT_COORD_TRANSFORM_LINDEPTH; }

NGT_DEVICE float log2_lod(float l0, float l1, float l2) { return fmaxf(0.f, log2f(fmaxf(l0, fmaxf(l1, l2)))); }

// Raw LoD of samplePos3D<CM, true> (sampling.hpp:158-228): forward
// differences p - p(idx + e); per object-space axis lengths in the forward
// modes, per difference vector in the reverse modes.
NGT_DEVICE float xform_lod(const XformArgs& a, const CamParams& cp, float ix, float iy, float iz, float px, float py,
                           float pz) {
  float d[3][3];
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    float qx, qy, qz;
    xform_pos(a, cp, ix + (e == 0 ? 1.f : 0.f), iy + (e == 1 ? 1.f : 0.f), iz + (e == 2 ? 1.f : 0.f), qx, qy, qz);
    d[e][0] = px - qx; d[e][1] = py - qy; d[e][2] = pz - qz;
  }
  float l[3];
  if (forward_mode(a.coord_mode)) {
#pragma unroll
    for (int k = 0; k < 3; ++k) l[k] = sqrtf(d[0][k] * d[0][k] + d[1][k] * d[1][k] + d[2][k] * d[2][k]);
  } else {
#pragma unroll
    for (int e = 0; e < 3; ++e) l[e] = sqrtf(d[e][0] * d[e][0] + d[e][1] * d[e][1] + d[e][2] * d[e][2]);
  }
  return log2_lod(l[0], l[1], l[2]);
}

template <int C, int FM, int BM, int MM>
__global__ __launch_bounds__(256) void xform_fwd_kernel(XformArgs a, CamBatch cb, MipSet mips,
                                                        float* __restrict__ out) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  const int z = blockIdx.z % a.dout.z;
  const int ci = blockIdx.z / a.dout.z;
  if (x >= a.dout.x || y >= a.dout.y) return;
  const CamParams& cp = cb.cam[ci];
  const float ix = static_cast<float>(x) + 0.5f, iy = static_cast<float>(y) + 0.5f, iz = static_cast<float>(z) + 0.5f;
  float px, py, pz;
  xform_pos(a, cp, ix, iy, iz, px, py, pz);
  const int64_t in_slice = static_cast<int64_t>(a.din.x) * a.din.y * a.din.z * C;
  Vec<C> r;
  if constexpr (BM == WRAP) {
    r = vzero<C>();  // Sampler2 has no WRAP path (sampling.hpp:802-826)
  } else if constexpr (MM != NGT_MIP_NONE) {
    const float lod = xform_lod(a, cp, ix, iy, iz, px, py, pz);
    r = read_mip<C, FM, BM, MM>(mips, cp.data_index, px, py, pz, lod, a.mip_bias, 0.5f);
  } else if constexpr (FM == LINEAR) {
    r = read_linear<C, BM>(a.input + cp.data_index * in_slice, a.din, px - 0.5f, py - 0.5f, pz - 0.5f);
  } else if constexpr (FM == QUADRATIC) {
    r = read_quadratic<C, BM>(a.input + cp.data_index * in_slice, a.din, px - 0.5f, py - 0.5f, pz - 0.5f);
  } else {
    r = read_nearest<C, BM>(a.input + cp.data_index * in_slice, a.din, px - 0.5f, py - 0.5f, pz - 0.5f);
  }
  const int64_t out_cells = static_cast<int64_t>(a.dout.x) * a.dout.y * a.dout.z;
  vstore<C>(out + (static_cast<int64_t>(cp.out_index) * out_cells + cell_offset(a.dout, x, y, z)) * C, r);
}

template <int C, int FM, int BM, bool REF, int MM>
__global__ __launch_bounds__(256) void xform_bwd_kernel(XformArgs a, CamBatch cb, MipSet grad,
                                                        const float* __restrict__ out_grad,
                                                        float* __restrict__ in_grad, uint32_t* __restrict__ hits) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  const int z = blockIdx.z % a.dout.z;
  const int ci = blockIdx.z / a.dout.z;
  if (x >= a.dout.x || y >= a.dout.y) return;
  const CamParams& cp = cb.cam[ci];
  const float ix = static_cast<float>(x) + 0.5f, iy = static_cast<float>(y) + 0.5f, iz = static_cast<float>(z) + 0.5f;
  float px, py, pz;
  xform_pos(a, cp, ix, iy, iz, px, py, pz);
  const int64_t in_cells = static_cast<int64_t>(a.din.x) * a.din.y * a.din.z;
  const int64_t out_cells = static_cast<int64_t>(a.dout.x) * a.dout.y * a.dout.z;
  const Vec<C> dy = vload<C>(out_grad + (static_cast<int64_t>(cp.out_index) * out_cells + cell_offset(a.dout, x, y, z)) * C);
  float* g = in_grad + cp.data_index * in_cells * C;
  if constexpr (REF) {
    // kScatter3DGradients: the quadratic scatter for QUADRATIC, the trilinear
    // one for every other filter (resample_grid.cu.cc:1160-1185)
    if constexpr (FM == QUADRATIC)
      scatter_quadratic_reference<C, BM>(g, a.din, px - 0.5f, py - 0.5f, pz - 0.5f, dy,
                                         hits + cp.data_index * in_cells);
    else
      scatter_linear<C, BM, true>(g, a.din, px - 0.5f, py - 0.5f, pz - 0.5f, dy, hits + cp.data_index * in_cells);
  } else if constexpr (BM != WRAP && MM != NGT_MIP_NONE) {
    const float lod = xform_lod(a, cp, ix, iy, iz, px, py, pz);
    const MipPick k = mip_pick<MM>(lod, a.mip_bias, grad.num_levels);
    // the transform filters need no level data for their scatter (LINEAR, NEAREST, QUADRATIC)
    scatter_level<C, FM, BM>(grad, grad, cp.data_index, k.lf, px, py, pz, 0.5f, vscale(dy, 1.f - k.wc));
    if (k.two) scatter_level<C, FM, BM>(grad, grad, cp.data_index, k.lc, px, py, pz, 0.5f, vscale(dy, k.wc));
  } else if constexpr (BM != WRAP) {
    if constexpr (FM == LINEAR) scatter_linear<C, BM, false>(g, a.din, px - 0.5f, py - 0.5f, pz - 0.5f, dy);
    else if constexpr (FM == QUADRATIC) scatter_quadratic<C, BM>(g, a.din, px - 0.5f, py - 0.5f, pz - 0.5f, dy);
    else scatter_nearest<C, BM>(g, a.din, px - 0.5f, py - 0.5f, pz - 0.5f, dy);
  }
}

// kComputeLoD3D (resample_grid.cu.cc:1706-1720): position and central
// difference LoD (calcLoD<CM>, sampling.hpp:72-139) per output cell.
__global__ __launch_bounds__(256) void lod_transform_kernel(XformArgs a, CamBatch cb, int relative,
                                                            float* __restrict__ out) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  const int z = blockIdx.z % a.dout.z;
  const int ci = blockIdx.z / a.dout.z;
  if (x >= a.dout.x || y >= a.dout.y) return;
  const CamParams& cp = cb.cam[ci];
  const float ix = static_cast<float>(x) + 0.5f, iy = static_cast<float>(y) + 0.5f, iz = static_cast<float>(z) + 0.5f;
  float d[3][3];
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const float ex = e == 0 ? 1.f : 0.f, ey = e == 1 ? 1.f : 0.f, ez = e == 2 ? 1.f : 0.f;
    float mx, my, mz, qx, qy, qz;
    xform_pos(a, cp, ix - ex, iy - ey, iz - ez, mx, my, mz);
    xform_pos(a, cp, ix + ex, iy + ey, iz + ez, qx, qy, qz);
    d[e][0] = (mx - qx) * 0.5f; d[e][1] = (my - qy) * 0.5f; d[e][2] = (mz - qz) * 0.5f;
  }
  float l[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) l[k] = sqrtf(d[0][k] * d[0][k] + d[1][k] * d[1][k] + d[2][k] * d[2][k]);
  float px, py, pz;
  xform_pos(a, cp, ix, iy, iz, px, py, pz);
  if (relative) {
    px -= static_cast<float>(x); py -= static_cast<float>(y); pz -= static_cast<float>(z);
  }
  const int64_t out_cells = static_cast<int64_t>(a.dout.x) * a.dout.y * a.dout.z;
  *reinterpret_cast<float4*>(out + (static_cast<int64_t>(cp.out_index) * out_cells + cell_offset(a.dout, x, y, z)) * 4) =
      make_float4(px, py, pz, log2_lod(l[0], l[1], l[2]));
}

// ------------------------------------------------------------- dispatch -----

#define NGT_SWITCH_C(C_, ...)                                 \
  switch (C_) {                                               \
    case 1: { constexpr int CC = 1; __VA_ARGS__; } break;     \
    case 2: { constexpr int CC = 2; __VA_ARGS__; } break;     \
    default: { constexpr int CC = 4; __VA_ARGS__; } break;    \
  }
#define NGT_SWITCH_BM(B_, ...)                                      \
  switch (B_) {                                                     \
    case BORDER: { constexpr int BB = BORDER; __VA_ARGS__; } break; \
    case CLAMP: { constexpr int BB = CLAMP; __VA_ARGS__; } break;   \
    default: { constexpr int BB = WRAP; __VA_ARGS__; } break;       \
  }
#define NGT_SWITCH_FM(F_, ...)                                       \
  switch (F_) {                                                      \
    case LINEAR: { constexpr int FF = LINEAR; __VA_ARGS__; } break;  \
    case NEAREST: { constexpr int FF = NEAREST; __VA_ARGS__; } break; \
    case FMIN: { constexpr int FF = FMIN; __VA_ARGS__; } break;      \
    default: { constexpr int FF = FMAX; __VA_ARGS__; } break;        \
  }
#define NGT_SWITCH_FM2(F_, ...)                                      \
  switch (F_) {                                                      \
    case LINEAR: { constexpr int FF = LINEAR; __VA_ARGS__; } break;  \
    case QUADRATIC: { constexpr int FF = QUADRATIC; __VA_ARGS__; } break; \
    default: { constexpr int FF = NEAREST; __VA_ARGS__; } break;     \
  }
#define NGT_SWITCH_MM(M_, ...)                                                   \
  switch (M_) {                                                                  \
    case NGT_MIP_NONE: { constexpr int MMM = NGT_MIP_NONE; __VA_ARGS__; } break; \
    case NGT_MIP_NEAREST: { constexpr int MMM = NGT_MIP_NEAREST; __VA_ARGS__; } break; \
    default: { constexpr int MMM = NGT_MIP_LINEAR; __VA_ARGS__; } break;         \
  }
#define NGT_SWITCH_MM2(M_, ...)                                                          \
  switch (M_) {                                                                          \
    case NGT_MIP_NEAREST: { constexpr int MMM = NGT_MIP_NEAREST; __VA_ARGS__; } break; \
    default: { constexpr int MMM = NGT_MIP_LINEAR; __VA_ARGS__; } break;                 \
  }

int check_filter_boundary(int interpolation, int boundary) {
  NGT_REQUIRE(interpolation >= 0 && interpolation <= 3, NGT_ERR_INVALID_ARGUMENT, "invalid interpolation");
  NGT_REQUIRE(boundary >= 0 && boundary <= 2, NGT_ERR_INVALID_ARGUMENT, "invalid boundary");
  return NGT_OK;
}

// Mip arguments against an input shape (resample_grid.cc:338-347 plus the
// pyramid limits).
int check_mips(const int64_t* in_shape, int mipmapping, int num_mipmaps) {
  NGT_REQUIRE(mipmapping >= NGT_MIP_NONE && mipmapping <= NGT_MIP_LINEAR, NGT_ERR_INVALID_ARGUMENT,
              "invalid mipmapping");
  NGT_REQUIRE(mipmapping == NGT_MIP_NONE || num_mipmaps > 0, NGT_ERR_INVALID_ARGUMENT,
              "when using mipmaps num_mipmaps must be larger than 0.");
  if (mipmapping == NGT_MIP_NONE) return NGT_OK;
  NGT_REQUIRE(num_mipmaps <= kMaxMips, NGT_ERR_INVALID_ARGUMENT, "num_mipmaps must be at most 16");
  NGT_REQUIRE((in_shape[1] >> num_mipmaps) >= 1 && (in_shape[2] >> num_mipmaps) >= 1 &&
                  (in_shape[3] >> num_mipmaps) >= 1,
              NGT_ERR_INVALID_ARGUMENT,
              "num_mipmaps too large for the input grid (every mip level needs at least one cell per axis)");
  return NGT_OK;
}

int check_mip_grad(int mipmapping, int gradient_semantics) {
  NGT_REQUIRE(mipmapping == NGT_MIP_NONE || gradient_semantics != NGT_GRAD_REFERENCE, NGT_ERR_UNIMPLEMENTED,
              "the reference gradient supports mipmapping 'NONE' only (use exact gradient semantics for mips)");
  return NGT_OK;
}

size_t mip_bytes(const int64_t* in_shape, int mipmapping, int num_mipmaps) {
  if (in_shape == nullptr || mipmapping == NGT_MIP_NONE || num_mipmaps <= 0 || num_mipmaps > kMaxMips) return 0;
  const Shape5 si = shape5(in_shape);
  return align256(static_cast<size_t>(mip_floats(si.dims(), si.n, si.c, num_mipmaps)) * sizeof(float));
}

MipSet make_mips(const float* base, const int64_t* in_shape, void* levels, int num_mipmaps) {
  const Shape5 si = shape5(in_shape);
  return MipSet{const_cast<float*>(base), static_cast<float*>(levels), si.dims(), static_cast<int>(si.n),
                num_mipmaps};
}

int check_lut(const int64_t* in_shape, const int64_t* lut_shape, int mipmapping, int num_mipmaps,
              float cco, int separate, int normalized) {
  int rc = check_grid_shape(in_shape, "input");
  if (rc) return rc;
  NGT_REQUIRE(lut_shape && lut_shape[4] == 4 && lut_shape[0] > 0 && lut_shape[1] > 0 && lut_shape[2] > 0 &&
                  lut_shape[3] > 0,
              NGT_ERR_INVALID_ARGUMENT, "Invalid lut shape (NDHWC) with C=4");
  rc = check_mips(in_shape, mipmapping, num_mipmaps);
  if (rc) return rc;
  NGT_REQUIRE(!normalized || cco == 0.5f, NGT_ERR_INVALID_ARGUMENT,
              "Invalid cell center offset must be 0.5 when using normalized coordinates.");
  if (!separate)
    NGT_REQUIRE(lut_shape[0] == in_shape[0], NGT_ERR_INVALID_ARGUMENT,
                "camera batch must match data batch when not using global sampling.");
  return NGT_OK;
}

LutArgs make_lut_args(const float* input, const int64_t* in_shape, const float* lookup, const int64_t* lut_shape,
                      float cco, int separate, int relative, int normalized, float mip_bias) {
  LutArgs a;
  const Shape5 si = shape5(in_shape);
  a.input = input;
  a.din = si.dims();
  a.lut = lookup;
  a.dout = Dims{static_cast<int>(lut_shape[3]), static_cast<int>(lut_shape[2]), static_cast<int>(lut_shape[1])};
  a.n_data = static_cast<int>(si.n);
  a.n_cams = static_cast<int>(lut_shape[0]);
  a.separate = separate ? 1 : 0;
  a.cco = cco;
  a.relative = relative ? 1 : 0;
  a.normalized = normalized ? 1 : 0;
  a.mip_bias = mip_bias;
  return a;
}

int check_cams(const int64_t* in_shape, const float* M, const float* V, const float* P, const float* F,
               int num_cameras, const int32_t* out_dhw, int separate, std::vector<Pair>& pairs) {
  int rc = check_grid_shape(in_shape, "input");
  if (rc) return rc;
  NGT_REQUIRE(M && V && P && F, NGT_ERR_INVALID_ARGUMENT, "matrices and frustum must be given");
  NGT_REQUIRE(num_cameras > 0, NGT_ERR_INVALID_ARGUMENT, "camera batch size mismatch");
  NGT_REQUIRE(out_dhw && out_dhw[0] > 0 && out_dhw[1] > 0 && out_dhw[2] > 0, NGT_ERR_INVALID_ARGUMENT,
              "Invalid output_shape");
  return make_pairs(static_cast<int>(in_shape[0]), num_cameras, separate != 0, pairs);
}

int check_xform_attrs(int interpolation, int boundary, int coordinate_mode, float cell_center_offset,
                      const char* op) {
  NGT_REQUIRE(interpolation == LINEAR || interpolation == NEAREST || interpolation == QUADRATIC,
              NGT_ERR_INVALID_ARGUMENT, std::string(op) + " supports NEAREST, LINEAR and QUADRATIC interpolation");
  NGT_REQUIRE(boundary == BORDER || boundary == CLAMP, NGT_ERR_INVALID_ARGUMENT, "invalid boundary");
  NGT_REQUIRE(cell_center_offset == 0.5f, NGT_ERR_INVALID_ARGUMENT,
              "Invalid cell center offset for given coordinate mode.");
  NGT_REQUIRE(coordinate_mode == 0 || coordinate_mode == 1 || coordinate_mode == 3 || coordinate_mode == 4,
              NGT_ERR_INVALID_ARGUMENT, "Invalid coordinate_mode");
  return NGT_OK;
}

XformArgs make_xform_args(const float* input, const int64_t* in_shape, const int32_t* out_dhw, int coordinate_mode,
                          float mip_bias) {
  const Shape5 si = shape5(in_shape);
  XformArgs a;
  a.input = input;
  a.din = si.dims();
  a.dout = Dims{out_dhw[2], out_dhw[1], out_dhw[0]};
  a.out_inv_x = 1.f / a.dout.x; a.out_inv_y = 1.f / a.dout.y; a.out_inv_z = 1.f / a.dout.z;
  a.coord_mode = coordinate_mode;
  a.mip_bias = mip_bias;
  return a;
}

CamBatch cam_batch(const std::vector<Pair>& pairs, size_t b, const float* M, const float* V, const float* P,
                   const float* F) {
  CamBatch cb;
  cb.count = static_cast<int>(std::min<size_t>(kMaxCamsPerLaunch, pairs.size() - b));
  for (int i = 0; i < cb.count; ++i) {
    const Pair& p = pairs[b + i];
    cb.cam[i] = make_cam(M + 16 * p.n, V + 16 * p.v, P + 16 * p.v, F + 6 * p.v, p.n, p.out);
  }
  return cb;
}

}  // namespace
}  // namespace ngt

using namespace ngt;

extern "C" {

size_t ngt_sample_grid_lut_fwd_workspace_size(const int64_t* input_shape, int mipmapping, int num_mipmaps) {
  return mip_bytes(input_shape, mipmapping, num_mipmaps);
}

int ngt_sample_grid_lut_fwd(const float* input, const int64_t* input_shape, const float* lookup,
                            const int64_t* lookup_shape, int interpolation, int boundary, int mipmapping,
                            int num_mipmaps, float mip_bias, float cell_center_offset, int separate_camera_batch,
                            int relative_coords, int normalized_coords, float* output, void* workspace,
                            size_t workspace_bytes, ngt_stream_t stream) {
  int rc = check_lut(input_shape, lookup_shape, mipmapping, num_mipmaps, cell_center_offset, separate_camera_batch,
                     normalized_coords);
  if (rc) return rc;
  rc = check_filter_boundary(interpolation, boundary);
  if (rc) return rc;
  const size_t need = ngt_sample_grid_lut_fwd_workspace_size(input_shape, mipmapping, num_mipmaps);
  NGT_REQUIRE(workspace_bytes >= need && (need == 0 || workspace != nullptr), NGT_ERR_WORKSPACE,
              "workspace too small for ngt_sample_grid_lut_fwd");
  const LutArgs a = make_lut_args(input, input_shape, lookup, lookup_shape, cell_center_offset, separate_camera_batch,
                                  relative_coords, normalized_coords, mip_bias);
  const MipSet mips = make_mips(input, input_shape, workspace, mipmapping == NGT_MIP_NONE ? 0 : num_mipmaps);
  const int pairs = separate_camera_batch ? a.n_data * a.n_cams : a.n_data;
  const dim3 grid(blocks_for(a.dout.x, kBX), blocks_for(a.dout.y, kBY), static_cast<unsigned>(a.dout.z * pairs));
  const dim3 block(kBX, kBY);
  hipStream_t s = as_stream(stream);
  if (mipmapping != NGT_MIP_NONE) NGT_SWITCH_C(input_shape[4], launch_mip_generate<CC>(mips, s));
  NGT_SWITCH_C(input_shape[4], NGT_SWITCH_FM(interpolation, NGT_SWITCH_BM(boundary, NGT_SWITCH_MM(mipmapping,
      hipLaunchKernelGGL((lut_fwd_kernel<CC, FF, BB, MMM>), grid, block, 0, s, a, mips, output)))));
  return check_launch("ngt_sample_grid_lut_fwd");
}

size_t ngt_sample_grid_lut_bwd_workspace_size(const int64_t* input_shape, const int64_t* lookup_shape,
                                              int mipmapping, int num_mipmaps, int gradient_semantics) {
  (void)lookup_shape;
  if (input_shape == nullptr) return 0;
  if (gradient_semantics == NGT_GRAD_REFERENCE)
    return static_cast<size_t>(input_shape[0] * input_shape[1] * input_shape[2] * input_shape[3]) * sizeof(uint32_t);
  return 2 * mip_bytes(input_shape, mipmapping, num_mipmaps);  // level data + level gradients
}

int ngt_sample_grid_lut_bwd(const float* input, const int64_t* input_shape, const float* output_grad,
                            const float* lookup, const int64_t* lookup_shape, int interpolation, int boundary,
                            int mipmapping, int num_mipmaps, float mip_bias, float cell_center_offset,
                            int separate_camera_batch, int relative_coords, int normalized_coords,
                            int gradient_semantics, float* input_grad, float* lookup_grad, void* workspace,
                            size_t workspace_bytes, ngt_stream_t stream) {
  int rc = check_lut(input_shape, lookup_shape, mipmapping, num_mipmaps, cell_center_offset, separate_camera_batch,
                     normalized_coords);
  if (rc) return rc;
  rc = check_filter_boundary(interpolation, boundary);
  if (rc) return rc;
  rc = check_mip_grad(mipmapping, gradient_semantics);
  if (rc) return rc;
  const bool ref = gradient_semantics == NGT_GRAD_REFERENCE;
  const size_t need =
      ngt_sample_grid_lut_bwd_workspace_size(input_shape, lookup_shape, mipmapping, num_mipmaps, gradient_semantics);
  NGT_REQUIRE(workspace_bytes >= need && (need == 0 || workspace != nullptr), NGT_ERR_WORKSPACE,
              "workspace too small for ngt_sample_grid_lut_bwd");
  const LutArgs a = make_lut_args(input, input_shape, lookup, lookup_shape, cell_center_offset, separate_camera_batch,
                                  relative_coords, normalized_coords, mip_bias);
  const Shape5 si = shape5(input_shape);
  hipStream_t s = as_stream(stream);
  fill_zero(input_grad, si.elems(), s);
  const int slots = separate_camera_batch ? a.n_cams : a.n_data;
  const dim3 grid(blocks_for(a.dout.x, kBX), blocks_for(a.dout.y, kBY), static_cast<unsigned>(a.dout.z * slots));
  const dim3 block(kBX, kBY);
  if (mipmapping != NGT_MIP_NONE) {
    const size_t half = mip_bytes(input_shape, mipmapping, num_mipmaps);
    char* ws = static_cast<char*>(workspace);
    const MipSet data = make_mips(input, input_shape, ws, num_mipmaps);
    const MipSet grad = make_mips(input_grad, input_shape, ws + half, num_mipmaps);
    fill_zero(grad.levels, static_cast<int64_t>(mip_floats(si.dims(), si.n, si.c, num_mipmaps)), s);
    NGT_SWITCH_C(si.c, launch_mip_generate<CC>(data, s));
    NGT_SWITCH_C(si.c, NGT_SWITCH_FM(interpolation, NGT_SWITCH_BM(boundary, NGT_SWITCH_MM2(mipmapping,
        hipLaunchKernelGGL((lut_bwd_mip_kernel<CC, FF, BB, MMM>), grid, block, 0, s, a, data, grad, output_grad,
                           lookup_grad)))));
    NGT_SWITCH_C(si.c, launch_mip_collapse<CC>(grad, s));
    return check_launch("ngt_sample_grid_lut_bwd");
  }
  uint32_t* hits = static_cast<uint32_t*>(workspace);
  if (ref) fill_zero_u32(hits, si.n * si.cells(), s);
  if (ref) {
    // the reference's input gradient is a trilinear scatter for every filter
    // mode; only LINEAR produces a LUT gradient
    if (interpolation == LINEAR) {
      NGT_SWITCH_C(si.c, NGT_SWITCH_BM(boundary,
          hipLaunchKernelGGL((lut_bwd_kernel<CC, LINEAR, BB, true>), grid, block, 0, s, a, output_grad, input_grad,
                             lookup_grad, hits)));
    } else {
      NGT_SWITCH_C(si.c, NGT_SWITCH_BM(boundary,
          hipLaunchKernelGGL((lut_bwd_kernel<CC, NEAREST, BB, true>), grid, block, 0, s, a, output_grad, input_grad,
                             lookup_grad, hits)));
    }
    NGT_SWITCH_C(si.c,
        hipLaunchKernelGGL((normalize_by_count_kernel<CC>), dim3(blocks_for(si.n * si.cells(), 256)), dim3(256), 0,
                           s, input_grad, hits, si.n * si.cells()));
  } else {
    NGT_SWITCH_C(si.c, NGT_SWITCH_FM(interpolation, NGT_SWITCH_BM(boundary,
        hipLaunchKernelGGL((lut_bwd_kernel<CC, FF, BB, false>), grid, block, 0, s, a, output_grad, input_grad,
                           lookup_grad, nullptr))));
  }
  return check_launch("ngt_sample_grid_lut_bwd");
}

size_t ngt_sample_grid_transform_fwd_workspace_size(const int64_t* input_shape, int mipmapping, int num_mipmaps) {
  return mip_bytes(input_shape, mipmapping, num_mipmaps);
}

int ngt_sample_grid_transform_fwd(const float* input, const int64_t* input_shape, const float* matrix_m,
                                  const float* matrix_v, const float* matrix_p, const float* frustum,
                                  int num_cameras, const int32_t* output_dhw, int interpolation, int boundary,
                                  int mipmapping, int num_mipmaps, float mip_bias, int coordinate_mode,
                                  float cell_center_offset, int separate_camera_batch, float* output,
                                  void* workspace, size_t workspace_bytes, ngt_stream_t stream) {
  std::vector<Pair> pairs;
  int rc = check_cams(input_shape, matrix_m, matrix_v, matrix_p, frustum, num_cameras, output_dhw,
                      separate_camera_batch, pairs);
  if (rc) return rc;
  rc = check_xform_attrs(interpolation, boundary, coordinate_mode, cell_center_offset, "SampleGridTransform");
  if (rc) return rc;
  rc = check_mips(input_shape, mipmapping, num_mipmaps);
  if (rc) return rc;
  const size_t need = ngt_sample_grid_transform_fwd_workspace_size(input_shape, mipmapping, num_mipmaps);
  NGT_REQUIRE(workspace_bytes >= need && (need == 0 || workspace != nullptr), NGT_ERR_WORKSPACE,
              "workspace too small for ngt_sample_grid_transform_fwd");
  const Shape5 si = shape5(input_shape);
  const XformArgs a = make_xform_args(input, input_shape, output_dhw, coordinate_mode, mip_bias);
  const MipSet mips = make_mips(input, input_shape, workspace, mipmapping == NGT_MIP_NONE ? 0 : num_mipmaps);
  hipStream_t s = as_stream(stream);
  if (mipmapping != NGT_MIP_NONE) NGT_SWITCH_C(si.c, launch_mip_generate<CC>(mips, s));
  for (size_t b = 0; b < pairs.size(); b += kMaxCamsPerLaunch) {
    const CamBatch cb = cam_batch(pairs, b, matrix_m, matrix_v, matrix_p, frustum);
    const dim3 grid(blocks_for(a.dout.x, kBX), blocks_for(a.dout.y, kBY), static_cast<unsigned>(a.dout.z * cb.count));
    const dim3 block(kBX, kBY);
    NGT_SWITCH_C(si.c, NGT_SWITCH_FM2(interpolation, NGT_SWITCH_BM(boundary, NGT_SWITCH_MM(mipmapping,
        hipLaunchKernelGGL((xform_fwd_kernel<CC, FF, BB, MMM>), grid, block, 0, s, a, cb, mips, output)))));
  }
  return check_launch("ngt_sample_grid_transform_fwd");
}

size_t ngt_sample_grid_transform_bwd_workspace_size(const int64_t* input_shape, int mipmapping, int num_mipmaps,
                                                    int gradient_semantics) {
  if (input_shape == nullptr) return 0;
  if (gradient_semantics == NGT_GRAD_REFERENCE)
    return static_cast<size_t>(input_shape[0] * input_shape[1] * input_shape[2] * input_shape[3]) * sizeof(uint32_t);
  return mip_bytes(input_shape, mipmapping, num_mipmaps);  // level gradients
}

int ngt_sample_grid_transform_bwd(const float* input, const int64_t* input_shape, const float* output_grad,
                                  const float* matrix_m, const float* matrix_v, const float* matrix_p,
                                  const float* frustum, int num_cameras, const int32_t* output_dhw,
                                  int interpolation, int boundary, int mipmapping, int num_mipmaps, float mip_bias,
                                  int coordinate_mode, float cell_center_offset, int separate_camera_batch,
                                  int gradient_semantics, float* input_grad, void* workspace, size_t workspace_bytes,
                                  ngt_stream_t stream) {
  std::vector<Pair> pairs;
  int rc = check_cams(input_shape, matrix_m, matrix_v, matrix_p, frustum, num_cameras, output_dhw,
                      separate_camera_batch, pairs);
  if (rc) return rc;
  rc = check_xform_attrs(interpolation, boundary, coordinate_mode, cell_center_offset, "SampleGridTransformGrad");
  if (rc) return rc;
  rc = check_mips(input_shape, mipmapping, num_mipmaps);
  if (rc) return rc;
  rc = check_mip_grad(mipmapping, gradient_semantics);
  if (rc) return rc;
  const bool ref = gradient_semantics == NGT_GRAD_REFERENCE;
  const size_t need =
      ngt_sample_grid_transform_bwd_workspace_size(input_shape, mipmapping, num_mipmaps, gradient_semantics);
  NGT_REQUIRE(workspace_bytes >= need && (need == 0 || workspace != nullptr), NGT_ERR_WORKSPACE,
              "workspace too small for ngt_sample_grid_transform_bwd");
  const Shape5 si = shape5(input_shape);
  const XformArgs a = make_xform_args(input, input_shape, output_dhw, coordinate_mode, mip_bias);
  const bool mip = mipmapping != NGT_MIP_NONE;
  const MipSet grad = make_mips(input_grad, input_shape, mip ? workspace : nullptr, mip ? num_mipmaps : 0);
  hipStream_t s = as_stream(stream);
  fill_zero(input_grad, si.elems(), s);
  uint32_t* hits = ref ? static_cast<uint32_t*>(workspace) : nullptr;
  if (ref) fill_zero_u32(hits, si.n * si.cells(), s);
  if (mip) fill_zero(grad.levels, static_cast<int64_t>(mip_floats(si.dims(), si.n, si.c, num_mipmaps)), s);
  for (size_t b = 0; b < pairs.size(); b += kMaxCamsPerLaunch) {
    const CamBatch cb = cam_batch(pairs, b, matrix_m, matrix_v, matrix_p, frustum);
    const dim3 grid(blocks_for(a.dout.x, kBX), blocks_for(a.dout.y, kBY), static_cast<unsigned>(a.dout.z * cb.count));
    const dim3 block(kBX, kBY);
    if (ref && interpolation == QUADRATIC) {
      NGT_SWITCH_C(si.c, NGT_SWITCH_BM(boundary,
          hipLaunchKernelGGL((xform_bwd_kernel<CC, QUADRATIC, BB, true, NGT_MIP_NONE>), grid, block, 0, s, a, cb,
                             grad, output_grad, input_grad, hits)));
    } else if (ref) {
      NGT_SWITCH_C(si.c, NGT_SWITCH_BM(boundary,
          hipLaunchKernelGGL((xform_bwd_kernel<CC, LINEAR, BB, true, NGT_MIP_NONE>), grid, block, 0, s, a, cb, grad,
                             output_grad, input_grad, hits)));
    } else {
      NGT_SWITCH_C(si.c, NGT_SWITCH_FM2(interpolation, NGT_SWITCH_BM(boundary, NGT_SWITCH_MM(mipmapping,
          hipLaunchKernelGGL((xform_bwd_kernel<CC, FF, BB, false, MMM>), grid, block, 0, s, a, cb, grad, output_grad,
                             input_grad, nullptr)))));
    }
  }
  if (mip) NGT_SWITCH_C(si.c, launch_mip_collapse<CC>(grad, s));
  if (ref) {
    NGT_SWITCH_C(si.c,
        hipLaunchKernelGGL((normalize_by_count_kernel<CC>), dim3(blocks_for(si.n * si.cells(), 256)), dim3(256), 0, s,
                           input_grad, hits, si.n * si.cells()));
  }
  return check_launch("ngt_sample_grid_transform_bwd");
}

int ngt_lod_transform(const int32_t* input_dhw, const float* matrix_mv, const float* matrix_p, const float* frustum,
                      int num_cameras, const int32_t* output_dhw, int inverse_transform, int linearize_depth,
                      int relative_coords, float* output, ngt_stream_t stream) {
  NGT_REQUIRE(input_dhw && input_dhw[0] > 0 && input_dhw[1] > 0 && input_dhw[2] > 0, NGT_ERR_INVALID_ARGUMENT,
              "Invalid input_shape");
  NGT_REQUIRE(output_dhw && output_dhw[0] > 0 && output_dhw[1] > 0 && output_dhw[2] > 0, NGT_ERR_INVALID_ARGUMENT,
              "Invalid output_shape");
  NGT_REQUIRE(matrix_mv && matrix_p && frustum, NGT_ERR_INVALID_ARGUMENT, "matrices and frustum must be given");
  NGT_REQUIRE(num_cameras > 0, NGT_ERR_INVALID_ARGUMENT, "camera batch size mismatch");
  const int cm = linearize_depth ? (inverse_transform ? NGT_COORD_TRANSFORM_LINDEPTH_REVERSE : NGT_COORD_TRANSFORM_LINDEPTH)
                                 : (inverse_transform ? NGT_COORD_TRANSFORM_REVERSE : NGT_COORD_TRANSFORM);
  const int64_t in_shape[5] = {1, input_dhw[0], input_dhw[1], input_dhw[2], 4};
  const XformArgs a = make_xform_args(nullptr, in_shape, output_dhw, cm, 0.f);
  static const float kIdentity[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  hipStream_t s = as_stream(stream);
  for (int b = 0; b < num_cameras; b += kMaxCamsPerLaunch) {
    CamBatch cb;
    cb.count = std::min(kMaxCamsPerLaunch, num_cameras - b);
    for (int i = 0; i < cb.count; ++i)
      cb.cam[i] = make_cam(kIdentity, matrix_mv + 16 * (b + i), matrix_p + 16 * (b + i), frustum + 6 * (b + i), 0,
                           b + i);
    const dim3 grid(blocks_for(a.dout.x, kBX), blocks_for(a.dout.y, kBY), static_cast<unsigned>(a.dout.z * cb.count));
    hipLaunchKernelGGL(lod_transform_kernel, grid, dim3(kBX, kBY), 0, s, a, cb, relative_coords ? 1 : 0, output);
  }
  return check_launch("ngt_lod_transform");
}

}  // extern "C"
