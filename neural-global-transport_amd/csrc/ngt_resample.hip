// SampleGridLut / SampleGridTransform and their adjoints on gfx950.
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
#include "ngt_host.hpp"

namespace ngt {
namespace {

constexpr int kBX = 64, kBY = 4;

struct LutArgs {
  const float* input;
  Dims din;
  const float* lut;
  Dims dout;
  int n_data, n_cams, separate;
  float cco;
  int relative, normalized;
};

NGT_DEVICE void lut_position(const LutArgs& a, int v, int x, int y, int z, float& px, float& py, float& pz) {
  const float4 l = *reinterpret_cast<const float4*>(a.lut + (cell_offset(a.dout, x, y, z) + static_cast<int64_t>(v) * a.dout.x * a.dout.y * a.dout.z) * 4);
  px = l.x; py = l.y; pz = l.z;
  if (a.normalized) {
    px *= static_cast<float>(a.dout.x); py *= static_cast<float>(a.dout.y); pz *= static_cast<float>(a.dout.z);
  }
  if (a.relative) {
    px += static_cast<float>(x); py += static_cast<float>(y); pz += static_cast<float>(z);
  }
  px -= a.cco; py -= a.cco; pz -= a.cco;
}

template <int C, int FM, int BM>
__global__ __launch_bounds__(256) void lut_fwd_kernel(LutArgs a, float* __restrict__ out) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  const int z = blockIdx.z % a.dout.z;
  const int pair = blockIdx.z / a.dout.z;
  if (x >= a.dout.x || y >= a.dout.y) return;
  int n, v;
  if (a.separate) { n = pair / a.n_cams; v = pair % a.n_cams; }
  else { n = pair; v = pair; }
  float px, py, pz;
  lut_position(a, v, x, y, z, px, py, pz);
  const int64_t in_slice = static_cast<int64_t>(a.din.x) * a.din.y * a.din.z * C;
  const Vec<C> r = read_filtered<C, FM, BM>(a.input + n * in_slice, a.din, px, py, pz);
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
  float px, py, pz;
  lut_position(a, v, x, y, z, px, py, pz);
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
  float in_inv_x, in_inv_y, in_inv_z;     // unused placeholders for symmetry
  float out_inv_x, out_inv_y, out_inv_z;
  int coord_mode;
};

NGT_DEVICE void xform_position(const XformArgs& a, const CamParams& cp, int x, int y, int z, float& px, float& py,
                               float& pz) {
  const float ix = static_cast<float>(x) + 0.5f, iy = static_cast<float>(y) + 0.5f, iz = static_cast<float>(z) + 0.5f;
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
  px -= 0.5f; py -= 0.5f; pz -= 0.5f;
}

template <int C, int FM, int BM>
__global__ __launch_bounds__(256) void xform_fwd_kernel(XformArgs a, CamBatch cb, float* __restrict__ out) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  const int z = blockIdx.z % a.dout.z;
  const int ci = blockIdx.z / a.dout.z;
  if (x >= a.dout.x || y >= a.dout.y) return;
  const CamParams& cp = cb.cam[ci];
  float px, py, pz;
  xform_position(a, cp, x, y, z, px, py, pz);
  const int64_t in_slice = static_cast<int64_t>(a.din.x) * a.din.y * a.din.z * C;
  Vec<C> r;
  if constexpr (BM == WRAP) {
    r = vzero<C>();  // Sampler2 has no WRAP path (sampling.hpp:802-826)
  } else if constexpr (FM == LINEAR) {
    r = read_linear<C, BM>(a.input + cp.data_index * in_slice, a.din, px, py, pz);
  } else {
    r = read_nearest<C, BM>(a.input + cp.data_index * in_slice, a.din, px, py, pz);
  }
  const int64_t out_cells = static_cast<int64_t>(a.dout.x) * a.dout.y * a.dout.z;
  vstore<C>(out + (static_cast<int64_t>(cp.out_index) * out_cells + cell_offset(a.dout, x, y, z)) * C, r);
}

template <int C, int FM, int BM, bool REF>
__global__ __launch_bounds__(256) void xform_bwd_kernel(XformArgs a, CamBatch cb, const float* __restrict__ out_grad,
                                                        float* __restrict__ in_grad, uint32_t* __restrict__ hits) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  const int z = blockIdx.z % a.dout.z;
  const int ci = blockIdx.z / a.dout.z;
  if (x >= a.dout.x || y >= a.dout.y) return;
  const CamParams& cp = cb.cam[ci];
  float px, py, pz;
  xform_position(a, cp, x, y, z, px, py, pz);
  const int64_t in_cells = static_cast<int64_t>(a.din.x) * a.din.y * a.din.z;
  const int64_t out_cells = static_cast<int64_t>(a.dout.x) * a.dout.y * a.dout.z;
  const Vec<C> dy = vload<C>(out_grad + (static_cast<int64_t>(cp.out_index) * out_cells + cell_offset(a.dout, x, y, z)) * C);
  float* g = in_grad + cp.data_index * in_cells * C;
  if constexpr (REF) {
    scatter_linear<C, BM, true>(g, a.din, px, py, pz, dy, hits + cp.data_index * in_cells);
  } else if constexpr (BM != WRAP) {
    if constexpr (FM == LINEAR) scatter_linear<C, BM, false>(g, a.din, px, py, pz, dy);
    else scatter_nearest<C, BM>(g, a.din, px, py, pz, dy);
  }
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

int check_filter_boundary(int interpolation, int boundary) {
  NGT_REQUIRE(interpolation >= 0 && interpolation <= 3, NGT_ERR_INVALID_ARGUMENT, "invalid interpolation");
  NGT_REQUIRE(boundary >= 0 && boundary <= 2, NGT_ERR_INVALID_ARGUMENT, "invalid boundary");
  return NGT_OK;
}

int check_lut(const int64_t* in_shape, const int64_t* lut_shape, int mipmapping, int num_mipmaps,
              float cco, int separate, int normalized) {
  int rc = check_grid_shape(in_shape, "input");
  if (rc) return rc;
  NGT_REQUIRE(lut_shape && lut_shape[4] == 4 && lut_shape[0] > 0 && lut_shape[1] > 0 && lut_shape[2] > 0 &&
                  lut_shape[3] > 0,
              NGT_ERR_INVALID_ARGUMENT, "Invalid lut shape (NDHWC) with C=4");
  NGT_REQUIRE(mipmapping == NGT_MIP_NONE || num_mipmaps > 0, NGT_ERR_INVALID_ARGUMENT,
              "when using mipmaps num_mipmaps must be larger than 0.");
  NGT_REQUIRE(mipmapping == NGT_MIP_NONE, NGT_ERR_UNIMPLEMENTED, "mipmapped LUT sampling is not implemented");
  NGT_REQUIRE(!normalized || cco == 0.5f, NGT_ERR_INVALID_ARGUMENT,
              "Invalid cell center offset must be 0.5 when using normalized coordinates.");
  if (!separate)
    NGT_REQUIRE(lut_shape[0] == in_shape[0], NGT_ERR_INVALID_ARGUMENT,
                "camera batch must match data batch when not using global sampling.");
  return NGT_OK;
}

LutArgs make_lut_args(const float* input, const int64_t* in_shape, const float* lookup, const int64_t* lut_shape,
                      float cco, int separate, int relative, int normalized) {
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

}  // namespace
}  // namespace ngt

using namespace ngt;

extern "C" {

int ngt_sample_grid_lut_fwd(const float* input, const int64_t* input_shape, const float* lookup,
                            const int64_t* lookup_shape, int interpolation, int boundary, int mipmapping,
                            int num_mipmaps, float mip_bias, float cell_center_offset, int separate_camera_batch,
                            int relative_coords, int normalized_coords, float* output, ngt_stream_t stream) {
  (void)mip_bias;
  int rc = check_lut(input_shape, lookup_shape, mipmapping, num_mipmaps, cell_center_offset, separate_camera_batch,
                     normalized_coords);
  if (rc) return rc;
  rc = check_filter_boundary(interpolation, boundary);
  if (rc) return rc;
  const LutArgs a = make_lut_args(input, input_shape, lookup, lookup_shape, cell_center_offset, separate_camera_batch,
                                  relative_coords, normalized_coords);
  const int pairs = separate_camera_batch ? a.n_data * a.n_cams : a.n_data;
  const dim3 grid(blocks_for(a.dout.x, kBX), blocks_for(a.dout.y, kBY), static_cast<unsigned>(a.dout.z * pairs));
  const dim3 block(kBX, kBY);
  hipStream_t s = as_stream(stream);
  NGT_SWITCH_C(input_shape[4], NGT_SWITCH_FM(interpolation, NGT_SWITCH_BM(boundary,
      hipLaunchKernelGGL((lut_fwd_kernel<CC, FF, BB>), grid, block, 0, s, a, output))));
  return check_launch("ngt_sample_grid_lut_fwd");
}

size_t ngt_sample_grid_lut_bwd_workspace_size(const int64_t* input_shape, const int64_t* lookup_shape,
                                              int gradient_semantics) {
  (void)lookup_shape;
  if (gradient_semantics != NGT_GRAD_REFERENCE || input_shape == nullptr) return 0;
  return static_cast<size_t>(input_shape[0] * input_shape[1] * input_shape[2] * input_shape[3]) * sizeof(uint32_t);
}

int ngt_sample_grid_lut_bwd(const float* input, const int64_t* input_shape, const float* output_grad,
                            const float* lookup, const int64_t* lookup_shape, int interpolation, int boundary,
                            int mipmapping, int num_mipmaps, float mip_bias, float cell_center_offset,
                            int separate_camera_batch, int relative_coords, int normalized_coords,
                            int gradient_semantics, float* input_grad, float* lookup_grad, void* workspace,
                            size_t workspace_bytes, ngt_stream_t stream) {
  (void)mip_bias;
  int rc = check_lut(input_shape, lookup_shape, mipmapping, num_mipmaps, cell_center_offset, separate_camera_batch,
                     normalized_coords);
  if (rc) return rc;
  rc = check_filter_boundary(interpolation, boundary);
  if (rc) return rc;
  const bool ref = gradient_semantics == NGT_GRAD_REFERENCE;
  const size_t need = ngt_sample_grid_lut_bwd_workspace_size(input_shape, lookup_shape, gradient_semantics);
  NGT_REQUIRE(workspace_bytes >= need && (need == 0 || workspace != nullptr), NGT_ERR_WORKSPACE,
              "workspace too small for ngt_sample_grid_lut_bwd");
  const LutArgs a = make_lut_args(input, input_shape, lookup, lookup_shape, cell_center_offset, separate_camera_batch,
                                  relative_coords, normalized_coords);
  const Shape5 si = shape5(input_shape);
  hipStream_t s = as_stream(stream);
  fill_zero(input_grad, si.elems(), s);
  uint32_t* hits = static_cast<uint32_t*>(workspace);
  if (ref) fill_zero_u32(hits, si.n * si.cells(), s);
  const int slots = separate_camera_batch ? a.n_cams : a.n_data;
  const dim3 grid(blocks_for(a.dout.x, kBX), blocks_for(a.dout.y, kBY), static_cast<unsigned>(a.dout.z * slots));
  const dim3 block(kBX, kBY);
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

int ngt_sample_grid_transform_fwd(const float* input, const int64_t* input_shape, const float* matrix_m,
                                  const float* matrix_v, const float* matrix_p, const float* frustum,
                                  int num_cameras, const int32_t* output_dhw, int interpolation, int boundary,
                                  int mipmapping, int num_mipmaps, float mip_bias, int coordinate_mode,
                                  float cell_center_offset, int separate_camera_batch, float* output,
                                  ngt_stream_t stream) {
  (void)mip_bias;
  std::vector<Pair> pairs;
  int rc = check_cams(input_shape, matrix_m, matrix_v, matrix_p, frustum, num_cameras, output_dhw,
                      separate_camera_batch, pairs);
  if (rc) return rc;
  NGT_REQUIRE(interpolation == LINEAR || interpolation == NEAREST, NGT_ERR_UNIMPLEMENTED,
              "SampleGridTransform supports NEAREST and LINEAR interpolation");
  NGT_REQUIRE(boundary == BORDER || boundary == CLAMP, NGT_ERR_INVALID_ARGUMENT, "invalid boundary");
  NGT_REQUIRE(mipmapping == NGT_MIP_NONE || num_mipmaps > 0, NGT_ERR_INVALID_ARGUMENT,
              "when using mipmaps num_mipmaps must be larger than 0.");
  NGT_REQUIRE(mipmapping == NGT_MIP_NONE, NGT_ERR_UNIMPLEMENTED, "mipmapped transform sampling is not implemented");
  NGT_REQUIRE(cell_center_offset == 0.5f, NGT_ERR_INVALID_ARGUMENT,
              "Invalid cell center offset for given coordinate mode.");
  NGT_REQUIRE(coordinate_mode == 0 || coordinate_mode == 1 || coordinate_mode == 3 || coordinate_mode == 4,
              NGT_ERR_INVALID_ARGUMENT, "Invalid coordinate_mode");
  const Shape5 si = shape5(input_shape);
  XformArgs a;
  a.input = input;
  a.din = si.dims();
  a.dout = Dims{output_dhw[2], output_dhw[1], output_dhw[0]};
  a.in_inv_x = a.in_inv_y = a.in_inv_z = 0.f;
  a.out_inv_x = 1.f / a.dout.x; a.out_inv_y = 1.f / a.dout.y; a.out_inv_z = 1.f / a.dout.z;
  a.coord_mode = coordinate_mode;
  hipStream_t s = as_stream(stream);
  for (size_t b = 0; b < pairs.size(); b += kMaxCamsPerLaunch) {
    CamBatch cb;
    cb.count = static_cast<int>(std::min<size_t>(kMaxCamsPerLaunch, pairs.size() - b));
    for (int i = 0; i < cb.count; ++i) {
      const Pair& p = pairs[b + i];
      cb.cam[i] = make_cam(matrix_m + 16 * p.n, matrix_v + 16 * p.v, matrix_p + 16 * p.v, frustum + 6 * p.v, p.n, p.out);
    }
    const dim3 grid(blocks_for(a.dout.x, kBX), blocks_for(a.dout.y, kBY), static_cast<unsigned>(a.dout.z * cb.count));
    const dim3 block(kBX, kBY);
    if (interpolation == LINEAR) {
      NGT_SWITCH_C(si.c, NGT_SWITCH_BM(boundary,
          hipLaunchKernelGGL((xform_fwd_kernel<CC, LINEAR, BB>), grid, block, 0, s, a, cb, output)));
    } else {
      NGT_SWITCH_C(si.c, NGT_SWITCH_BM(boundary,
          hipLaunchKernelGGL((xform_fwd_kernel<CC, NEAREST, BB>), grid, block, 0, s, a, cb, output)));
    }
  }
  return check_launch("ngt_sample_grid_transform_fwd");
}

size_t ngt_sample_grid_transform_bwd_workspace_size(const int64_t* input_shape, int gradient_semantics) {
  if (gradient_semantics != NGT_GRAD_REFERENCE || input_shape == nullptr) return 0;
  return static_cast<size_t>(input_shape[0] * input_shape[1] * input_shape[2] * input_shape[3]) * sizeof(uint32_t);
}

int ngt_sample_grid_transform_bwd(const float* input, const int64_t* input_shape, const float* output_grad,
                                  const float* matrix_m, const float* matrix_v, const float* matrix_p,
                                  const float* frustum, int num_cameras, const int32_t* output_dhw,
                                  int interpolation, int boundary, int coordinate_mode, float cell_center_offset,
                                  int separate_camera_batch, int gradient_semantics, float* input_grad,
                                  void* workspace, size_t workspace_bytes, ngt_stream_t stream) {
  (void)input;
  std::vector<Pair> pairs;
  int rc = check_cams(input_shape, matrix_m, matrix_v, matrix_p, frustum, num_cameras, output_dhw,
                      separate_camera_batch, pairs);
  if (rc) return rc;
  NGT_REQUIRE(interpolation == LINEAR || interpolation == NEAREST, NGT_ERR_UNIMPLEMENTED,
              "SampleGridTransformGrad supports NEAREST and LINEAR interpolation");
  NGT_REQUIRE(boundary == BORDER || boundary == CLAMP, NGT_ERR_INVALID_ARGUMENT, "invalid boundary");
  NGT_REQUIRE(cell_center_offset == 0.5f, NGT_ERR_INVALID_ARGUMENT,
              "Invalid cell center offset for given coordinate mode.");
  NGT_REQUIRE(coordinate_mode == 0 || coordinate_mode == 1 || coordinate_mode == 3 || coordinate_mode == 4,
              NGT_ERR_INVALID_ARGUMENT, "Invalid coordinate_mode");
  const bool ref = gradient_semantics == NGT_GRAD_REFERENCE;
  const size_t need = ngt_sample_grid_transform_bwd_workspace_size(input_shape, gradient_semantics);
  NGT_REQUIRE(workspace_bytes >= need && (need == 0 || workspace != nullptr), NGT_ERR_WORKSPACE,
              "workspace too small for ngt_sample_grid_transform_bwd");
  const Shape5 si = shape5(input_shape);
  XformArgs a;
  a.input = input;
  a.din = si.dims();
  a.dout = Dims{output_dhw[2], output_dhw[1], output_dhw[0]};
  a.in_inv_x = a.in_inv_y = a.in_inv_z = 0.f;
  a.out_inv_x = 1.f / a.dout.x; a.out_inv_y = 1.f / a.dout.y; a.out_inv_z = 1.f / a.dout.z;
  a.coord_mode = coordinate_mode;
  hipStream_t s = as_stream(stream);
  fill_zero(input_grad, si.elems(), s);
  uint32_t* hits = static_cast<uint32_t*>(workspace);
  if (ref) fill_zero_u32(hits, si.n * si.cells(), s);
  for (size_t b = 0; b < pairs.size(); b += kMaxCamsPerLaunch) {
    CamBatch cb;
    cb.count = static_cast<int>(std::min<size_t>(kMaxCamsPerLaunch, pairs.size() - b));
    for (int i = 0; i < cb.count; ++i) {
      const Pair& p = pairs[b + i];
      cb.cam[i] = make_cam(matrix_m + 16 * p.n, matrix_v + 16 * p.v, matrix_p + 16 * p.v, frustum + 6 * p.v, p.n, p.out);
    }
    const dim3 grid(blocks_for(a.dout.x, kBX), blocks_for(a.dout.y, kBY), static_cast<unsigned>(a.dout.z * cb.count));
    const dim3 block(kBX, kBY);
    if (ref) {
      NGT_SWITCH_C(si.c, NGT_SWITCH_BM(boundary,
          hipLaunchKernelGGL((xform_bwd_kernel<CC, LINEAR, BB, true>), grid, block, 0, s, a, cb, output_grad,
                             input_grad, hits)));
    } else if (interpolation == LINEAR) {
      NGT_SWITCH_C(si.c, NGT_SWITCH_BM(boundary,
          hipLaunchKernelGGL((xform_bwd_kernel<CC, LINEAR, BB, false>), grid, block, 0, s, a, cb, output_grad,
                             input_grad, nullptr)));
    } else {
      NGT_SWITCH_C(si.c, NGT_SWITCH_BM(boundary,
          hipLaunchKernelGGL((xform_bwd_kernel<CC, NEAREST, BB, false>), grid, block, 0, s, a, cb, output_grad,
                             input_grad, nullptr)));
    }
  }
  if (ref) {
    NGT_SWITCH_C(si.c,
        hipLaunchKernelGGL((normalize_by_count_kernel<CC>), dim3(blocks_for(si.n * si.cells(), 256)), dim3(256), 0, s,
                           input_grad, hits, si.n * si.cells()));
  }
  return check_launch("ngt_sample_grid_transform_bwd");
}

}  // extern "C"
