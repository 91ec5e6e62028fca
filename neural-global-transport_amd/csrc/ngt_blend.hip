// ReduceGridBlend (+Grad) on gfx950: one work-item per ray (x, y) walking
// depth; a wavefront covers 64 consecutive x so every depth step is one
// coalesced 64 x C x 4-byte read (reduce_blend.cu.cc:34-80 launches 32x8
// threads per batch entry in a host loop; here all entries go in one launch).
//
// Exact backward without the reference's subtraction reconstruction:
//  * keep_dims: reverse walk reading the stored prefix accumulators;
//  * otherwise, BEER_LAMBERT / ALPHA_ADDITIVE / ADDITIVE run a forward walk
//    using  d sigma_k = g_d - (<L_D, g_L> - sum_{j<k} e_j <l_j, g_L>),
//    so no accumulator is reconstructed; ALPHA walks backward rebuilding
//    a_{k-1} = 1 - (1 - a_k) / (1 - alpha_k).
#include "ngt_host.hpp"

namespace ngt {
namespace {

struct BlendArgs {
  int n, d, h, w;
};

template <int C, int MODE, bool KEEP>
__global__ __launch_bounds__(256) void blend_fwd_kernel(BlendArgs a, const float* __restrict__ in,
                                                        float* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  const int n = blockIdx.z;
  if (x >= a.w) return;
  const int64_t plane = static_cast<int64_t>(a.h) * a.w;
  const int64_t base = static_cast<int64_t>(n) * a.d * plane + static_cast<int64_t>(y) * a.w + x;
  Vec<C> acc = vzero<C>();
  for (int k = 0; k < a.d; ++k) {
    const Vec<C> cell = vload<C>(in + (base + k * plane) * C);
    blend_step<C, MODE>(acc, cell);
    if constexpr (KEEP) vstore<C>(out + (base + k * plane) * C, acc);
  }
  if constexpr (!KEEP) vstore<C>(out + (static_cast<int64_t>(n) * plane + static_cast<int64_t>(y) * a.w + x) * C, acc);
}

// Exact cell / previous-accumulator gradients of one blend step given the
// accumulator before the step.
template <int C, int MODE>
NGT_DEVICE Vec<C> blend_step_grad(const Vec<C>& prev, const Vec<C>& cell, Vec<C>& g) {
  if constexpr (C == 1 || MODE == ADDITIVE) {
    return g;
  } else {
    float lg = 0.f;
#pragma unroll
    for (int i = 0; i < C - 1; ++i) lg += cell.v[i] * g.v[i];
    Vec<C> gc;
    if constexpr (MODE == BEER_LAMBERT) {
      const float e = exp_neg(fmaxf(prev.v[C - 1] + cell.v[C - 1], 0.f));
      const float gd = g.v[C - 1] - e * lg;
#pragma unroll
      for (int i = 0; i < C - 1; ++i) gc.v[i] = e * g.v[i];
      gc.v[C - 1] = gd;
      g.v[C - 1] = gd;
    } else if constexpr (MODE == ALPHA) {
      const float t = 1.f - prev.v[C - 1];
#pragma unroll
      for (int i = 0; i < C - 1; ++i) gc.v[i] = t * g.v[i];
      gc.v[C - 1] = t * g.v[C - 1];
      g.v[C - 1] = (1.f - cell.v[C - 1]) * g.v[C - 1] - lg;
    } else {  // ALPHA_ADDITIVE
      const float t = 1.f - cell.v[C - 1];
#pragma unroll
      for (int i = 0; i < C - 1; ++i) gc.v[i] = t * g.v[i];
      gc.v[C - 1] = g.v[C - 1] - lg;
    }
    return gc;
  }
}

template <int C, int MODE, bool KEEP>
__global__ __launch_bounds__(256) void blend_bwd_exact_kernel(BlendArgs a, const float* __restrict__ out_grad,
                                                              const float* __restrict__ out,
                                                              const float* __restrict__ in,
                                                              float* __restrict__ in_grad) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  const int n = blockIdx.z;
  if (x >= a.w) return;
  const int64_t plane = static_cast<int64_t>(a.h) * a.w;
  const int64_t base = static_cast<int64_t>(n) * a.d * plane + static_cast<int64_t>(y) * a.w + x;
  const int64_t pix = (static_cast<int64_t>(n) * plane + static_cast<int64_t>(y) * a.w + x) * C;
  if constexpr (KEEP) {
    Vec<C> g = vzero<C>();
    for (int k = a.d - 1; k >= 0; --k) {
      const Vec<C> gk = vload<C>(out_grad + (base + k * plane) * C);
#pragma unroll
      for (int i = 0; i < C; ++i) g.v[i] += gk.v[i];
      const Vec<C> prev = k > 0 ? vload<C>(out + (base + (k - 1) * plane) * C) : vzero<C>();
      const Vec<C> cell = vload<C>(in + (base + k * plane) * C);
      vstore<C>(in_grad + (base + k * plane) * C, blend_step_grad<C, MODE>(prev, cell, g));
    }
  } else if constexpr (C == 1 || MODE == ADDITIVE) {
    const Vec<C> g = vload<C>(out_grad + pix);
    for (int k = 0; k < a.d; ++k) vstore<C>(in_grad + (base + k * plane) * C, g);
  } else if constexpr (MODE == ALPHA) {
    Vec<C> g = vload<C>(out_grad + pix);
    Vec<C> acc = vload<C>(out + pix);
    for (int k = a.d - 1; k >= 0; --k) {
      const Vec<C> cell = vload<C>(in + (base + k * plane) * C);
      Vec<C> prev = acc;
      prev.v[C - 1] = 1.f - (1.f - acc.v[C - 1]) / (1.f - cell.v[C - 1]);
      vstore<C>(in_grad + (base + k * plane) * C, blend_step_grad<C, MODE>(prev, cell, g));
      acc = prev;
    }
  } else {  // BEER_LAMBERT, ALPHA_ADDITIVE: forward walk
    const Vec<C> g = vload<C>(out_grad + pix);
    const Vec<C> L = vload<C>(out + pix);
    float total = 0.f;
#pragma unroll
    for (int i = 0; i < C - 1; ++i) total += L.v[i] * g.v[i];
    float prefix = 0.f;
    float dacc = 0.f;
    for (int k = 0; k < a.d; ++k) {
      const Vec<C> cell = vload<C>(in + (base + k * plane) * C);
      float lg = 0.f;
#pragma unroll
      for (int i = 0; i < C - 1; ++i) lg += cell.v[i] * g.v[i];
      Vec<C> gc;
      if constexpr (MODE == BEER_LAMBERT) {
        dacc = fmaxf(dacc + cell.v[C - 1], 0.f);
        const float e = exp_neg(dacc);
#pragma unroll
        for (int i = 0; i < C - 1; ++i) gc.v[i] = e * g.v[i];
        gc.v[C - 1] = g.v[C - 1] - (total - prefix);
        prefix += e * lg;
      } else {
        const float t = 1.f - cell.v[C - 1];
#pragma unroll
        for (int i = 0; i < C - 1; ++i) gc.v[i] = t * g.v[i];
        gc.v[C - 1] = g.v[C - 1] - lg;
      }
      vstore<C>(in_grad + (base + k * plane) * C, gc);
    }
  }
}

// ReduceGridBlendGrad verbatim in behaviour (reduce_blend.cu.cc:55-80).
template <int C, int MODE, bool KEEP>
__global__ __launch_bounds__(256) void blend_bwd_reference_kernel(BlendArgs a, const float* __restrict__ out_grad,
                                                                  const float* __restrict__ out,
                                                                  const float* __restrict__ in,
                                                                  float* __restrict__ in_grad) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  const int n = blockIdx.z;
  if (x >= a.w) return;
  const int64_t plane = static_cast<int64_t>(a.h) * a.w;
  const int64_t base = static_cast<int64_t>(n) * a.d * plane + static_cast<int64_t>(y) * a.w + x;
  const int64_t pix = (static_cast<int64_t>(n) * plane + static_cast<int64_t>(y) * a.w + x) * C;
  Vec<C> grads = vzero<C>(), prev_out = vzero<C>();
  if constexpr (!KEEP) {
    grads = vload<C>(out_grad + pix);
    prev_out = vload<C>(out + pix);
  }
  for (int k = a.d - 1; k >= 0; --k) {
    if constexpr (KEEP) {
      const Vec<C> gk = vload<C>(out_grad + (base + k * plane) * C);
#pragma unroll
      for (int i = 0; i < C; ++i) grads.v[i] += gk.v[i];
      prev_out = vload<C>(out + (base + k * plane) * C);
    }
    const Vec<C> cell = vload<C>(in + (base + k * plane) * C);
    vstore<C>(in_grad + (base + k * plane) * C, blend_grad_reference<C, MODE>(grads, cell, prev_out));
  }
}

#define NGT_SWITCH_C(C_, ...)                              \
  switch (C_) {                                            \
    case 1: { constexpr int CC = 1; __VA_ARGS__; } break;  \
    case 2: { constexpr int CC = 2; __VA_ARGS__; } break;  \
    default: { constexpr int CC = 4; __VA_ARGS__; } break; \
  }
#define NGT_SWITCH_MODE(M_, ...)                                                   \
  switch (M_) {                                                                    \
    case BEER_LAMBERT: { constexpr int MM = BEER_LAMBERT; __VA_ARGS__; } break;     \
    case ALPHA: { constexpr int MM = ALPHA; __VA_ARGS__; } break;                   \
    case ALPHA_ADDITIVE: { constexpr int MM = ALPHA_ADDITIVE; __VA_ARGS__; } break; \
    default: { constexpr int MM = ADDITIVE; __VA_ARGS__; } break;                   \
  }
#define NGT_SWITCH_KEEP(K_, ...)                             \
  if (K_) { constexpr bool KK = true; __VA_ARGS__; }         \
  else { constexpr bool KK = false; __VA_ARGS__; }

int check_mode(int mode) {
  NGT_REQUIRE(mode == BEER_LAMBERT || mode == ALPHA || mode == ALPHA_ADDITIVE || mode == ADDITIVE,
              NGT_ERR_INVALID_ARGUMENT, "invalid blending mode");
  return NGT_OK;
}

}  // namespace
}  // namespace ngt

using namespace ngt;

extern "C" {

int ngt_reduce_grid_blend_fwd(const float* input, const int64_t* input_shape, int blending_mode, int keep_dims,
                              float* output, ngt_stream_t stream) {
  int rc = check_grid_shape(input_shape, "input");
  if (rc) return rc;
  rc = check_mode(blending_mode);
  if (rc) return rc;
  const Shape5 s5 = shape5(input_shape);
  const BlendArgs a{static_cast<int>(s5.n), static_cast<int>(s5.d), static_cast<int>(s5.h), static_cast<int>(s5.w)};
  const dim3 grid(blocks_for(a.w, 64), a.h, a.n);
  hipStream_t s = as_stream(stream);
  NGT_SWITCH_C(s5.c, NGT_SWITCH_MODE(blending_mode, NGT_SWITCH_KEEP(keep_dims,
      hipLaunchKernelGGL((blend_fwd_kernel<CC, MM, KK>), grid, dim3(64), 0, s, a, input, output))));
  return check_launch("ngt_reduce_grid_blend_fwd");
}

int ngt_reduce_grid_blend_bwd(const float* output_grad, const float* output, const float* input,
                              const int64_t* input_shape, int blending_mode, int keep_dims, int gradient_semantics,
                              float* input_grad, ngt_stream_t stream) {
  int rc = check_grid_shape(input_shape, "input");
  if (rc) return rc;
  rc = check_mode(blending_mode);
  if (rc) return rc;
  const Shape5 s5 = shape5(input_shape);
  const BlendArgs a{static_cast<int>(s5.n), static_cast<int>(s5.d), static_cast<int>(s5.h), static_cast<int>(s5.w)};
  const dim3 grid(blocks_for(a.w, 64), a.h, a.n);
  hipStream_t s = as_stream(stream);
  if (gradient_semantics == NGT_GRAD_REFERENCE) {
    NGT_REQUIRE(blending_mode != ALPHA_ADDITIVE, NGT_ERR_INVALID_ARGUMENT,
                "ReduceGridBlendGrad: blending_mode ALPHA_ADDITIVE is not supported");
    NGT_SWITCH_C(s5.c, NGT_SWITCH_MODE(blending_mode, NGT_SWITCH_KEEP(keep_dims,
        hipLaunchKernelGGL((blend_bwd_reference_kernel<CC, MM, KK>), grid, dim3(64), 0, s, a, output_grad, output,
                           input, input_grad))));
  } else {
    NGT_SWITCH_C(s5.c, NGT_SWITCH_MODE(blending_mode, NGT_SWITCH_KEEP(keep_dims,
        hipLaunchKernelGGL((blend_bwd_exact_kernel<CC, MM, KK>), grid, dim3(64), 0, s, a, output_grad, output,
                           input, input_grad))));
  }
  return check_launch("ngt_reduce_grid_blend_bwd");
}

}  // extern "C"
