/*
 * ngt_transport.h - C ABI of the MI355X (gfx950) differentiable
 * volume-transport library (libngt_transport.so).
 *
 * This is the drop-in boundary for the four TF custom-op libraries of
 * Neural-Global-Transport (phitest/render/cuda/src/{resample_grid,
 * raymarch_grid,reduce_blend,advect}.cc).  Each entry point replaces one
 * OpKernel::Compute and takes the op's inputs as plain device pointers plus
 * NDHWC shapes, the op's string attributes as the integer enums below, and
 * the 4x4 matrices / frustum parameters as HOST pointers (the reference
 * registers them as host-memory inputs, resample_grid.cc:1183-1219).
 *
 * Conventions
 *  - Shapes are int64[5] = {N, D, H, W, C}; data is fp32, x fastest.
 *  - Matrices are the flat 16-float arrays the reference ops receive (already
 *    transposed by the Renderer), read column-major; `matrix_m` holds N
 *    matrices, `matrix_v`/`matrix_p` hold V, `frustum` holds V x 6 floats
 *    (near, far, left, right, top, bottom).
 *  - `separate_camera_batch` = the reference attribute of that name: 1 ->
 *    output [N, V, ...], every batch entry seen by every camera; 0 -> V == N,
 *    output [N, 1, ...], entry n seen by camera n.
 *  - `gradient_semantics`: NGT_GRAD_EXACT (true adjoint) or
 *    NGT_GRAD_REFERENCE (the reference's count / weight normalised gradients).
 *  - Every call is stream-ordered on `stream` (a hipStream_t, 0 = default
 *    stream), never synchronises the device, never allocates device memory
 *    (scratch comes in through `workspace`, sized by the *_workspace_size
 *    query), keeps no global mutable state and is hipGraph-capturable.
 *  - Return value: NGT_OK or an NGT_ERR_* code; ngt_last_error() returns the
 *    calling thread's message for the last failure (the reference's
 *    InvalidArgument / Unimplemented texts where one exists).
 */
#ifndef NGT_TRANSPORT_H
#define NGT_TRANSPORT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NGT_OK 0
#define NGT_ERR_INVALID_ARGUMENT 1
#define NGT_ERR_UNIMPLEMENTED 2
#define NGT_ERR_WORKSPACE 3
#define NGT_ERR_HIP 4

/* interpolation */
#define NGT_FILTER_NEAREST 0
#define NGT_FILTER_LINEAR 1
#define NGT_FILTER_MIN 2
#define NGT_FILTER_MAX 3
/* boundary */
#define NGT_BOUNDARY_BORDER 0
#define NGT_BOUNDARY_CLAMP 1
#define NGT_BOUNDARY_WRAP 2
/* mipmapping */
#define NGT_MIP_NONE 0
#define NGT_MIP_NEAREST 1
#define NGT_MIP_LINEAR 2
/* coordinate_mode (sampling_settings.hpp:61-70 numbering) */
#define NGT_COORD_TRANSFORM_LINDEPTH 0
#define NGT_COORD_TRANSFORM_LINDEPTH_REVERSE 1
#define NGT_COORD_TRANSFORM 3
#define NGT_COORD_TRANSFORM_REVERSE 4
/* blending_mode */
#define NGT_BLEND_BEER_LAMBERT 0
#define NGT_BLEND_ALPHA 2
#define NGT_BLEND_ALPHA_ADDITIVE 3
#define NGT_BLEND_ADDITIVE 4
/* gradient semantics */
#define NGT_GRAD_EXACT 0
#define NGT_GRAD_REFERENCE 1
/* advection clamp */
#define NGT_CLAMP_NONE 0
#define NGT_CLAMP_MC_SMOOTH 1
#define NGT_CLAMP_NEGATIVE 2

typedef void* ngt_stream_t; /* hipStream_t */

const char* ngt_last_error(void);
int ngt_version(void);

/* ---- Mip pyramids (resample_grid.cu.cc:58-251, sampling.hpp:939-1010) -----
 * With mipmapping NEAREST or LINEAR the sampling ops read a pyramid of
 * num_mipmaps levels built from the input: level m+1 has floor(dims_m / 2)
 * cells, each the mean of its 2x2x2 children.  A sample at position p (before
 * the cell-centre offset) with level of detail lod reads level m at
 * p / 2^m - cell_center_offset on a grid of dims >> m; lod = clamp(raw lod +
 * mip_bias, 0, num_mipmaps); NEAREST reads level min(int(lod + .5), L),
 * LINEAR lerps levels int(lod) and ceil(lod) by fract(lod).  The pyramid
 * lives in the workspace; every level must keep at least one cell per axis
 * (dims >> num_mipmaps >= 1), at most 16 levels. */

/* ---- SampleGridLut (resample_grid.cc:94-120, 742-899) ----------------------
 * input [N,D,H,W,C] (C in 1,2,4), lookup [V,D',H',W',4] -> output
 * [N,V|1,D',H',W',C].  Position = lookup.xyz (x output dims if
 * normalized_coords) (+ output index if relative_coords); raw lod =
 * lookup.w; read at position - cell_center_offset (per mip level as above).
 * The workspace holds the mip pyramid (0 bytes without mipmapping). */
size_t ngt_sample_grid_lut_fwd_workspace_size(const int64_t* input_shape, int mipmapping, int num_mipmaps);
int ngt_sample_grid_lut_fwd(const float* input, const int64_t* input_shape, const float* lookup,
                            const int64_t* lookup_shape, int interpolation, int boundary, int mipmapping,
                            int num_mipmaps, float mip_bias, float cell_center_offset, int separate_camera_batch,
                            int relative_coords, int normalized_coords, float* output, void* workspace,
                            size_t workspace_bytes, ngt_stream_t stream);

/* ---- SampleGridLutGrad (resample_grid.cc:124-150, 925-1081) ----------------
 * output_grad [N,V|1,D',H',W',C] -> input_grad [N,D,H,W,C],
 * lookup_grad [V,D',H',W',4] (both fully overwritten).  With mipmapping
 * (NGT_GRAD_EXACT only: the reference gradient op accepts mipmapping NONE
 * only, resample_grid.cc:135) the input gradient runs through the pyramid's
 * adjoint and lookup_grad.w holds d/d(raw lod). */
size_t ngt_sample_grid_lut_bwd_workspace_size(const int64_t* input_shape, const int64_t* lookup_shape,
                                              int mipmapping, int num_mipmaps, int gradient_semantics);
int ngt_sample_grid_lut_bwd(const float* input, const int64_t* input_shape, const float* output_grad,
                            const float* lookup, const int64_t* lookup_shape, int interpolation, int boundary,
                            int mipmapping, int num_mipmaps, float mip_bias, float cell_center_offset,
                            int separate_camera_batch, int relative_coords, int normalized_coords,
                            int gradient_semantics, float* input_grad, float* lookup_grad, void* workspace,
                            size_t workspace_bytes, ngt_stream_t stream);

/* ---- SampleGridTransform (resample_grid.cc:23-65, 316-516) -----------------
 * input [N,D,H,W,C] -> output [N,V|1,out_dhw,C].  interpolation NEAREST or
 * LINEAR, boundary BORDER or CLAMP, cell_center_offset must be 0.5.  With
 * mipmapping the raw lod of an output cell is log2 of the largest sample
 * displacement over one output cell (samplePos3D<CM, true>,
 * sampling.hpp:151-228). */
size_t ngt_sample_grid_transform_fwd_workspace_size(const int64_t* input_shape, int mipmapping, int num_mipmaps);
int ngt_sample_grid_transform_fwd(const float* input, const int64_t* input_shape, const float* matrix_m,
                                  const float* matrix_v, const float* matrix_p, const float* frustum,
                                  int num_cameras, const int32_t* output_dhw, int interpolation, int boundary,
                                  int mipmapping, int num_mipmaps, float mip_bias, int coordinate_mode,
                                  float cell_center_offset, int separate_camera_batch, float* output,
                                  void* workspace, size_t workspace_bytes, ngt_stream_t stream);

/* ---- SampleGridTransformGrad (resample_grid.cc:68-92, 538-717) ------------
 * Mipmapping with NGT_GRAD_EXACT only (the reference op accepts NONE only,
 * resample_grid.cc:79). */
size_t ngt_sample_grid_transform_bwd_workspace_size(const int64_t* input_shape, int mipmapping, int num_mipmaps,
                                                    int gradient_semantics);
int ngt_sample_grid_transform_bwd(const float* input, const int64_t* input_shape, const float* output_grad,
                                  const float* matrix_m, const float* matrix_v, const float* matrix_p,
                                  const float* frustum, int num_cameras, const int32_t* output_dhw,
                                  int interpolation, int boundary, int mipmapping, int num_mipmaps, float mip_bias,
                                  int coordinate_mode, float cell_center_offset, int separate_camera_batch,
                                  int gradient_semantics, float* input_grad, void* workspace, size_t workspace_bytes,
                                  ngt_stream_t stream);

/* ---- LodTransform (resample_grid.cc:152-172, 1089-1178) --------------------
 * Lookup tables for SampleGridLut from camera transforms: output
 * [V,D',H',W',4] = (sample position of each output cell centre, minus the
 * integer output index if relative_coords; level of detail from central
 * differences, calcLoD<CM> sampling.hpp:72-139).  matrix_mv / matrix_p /
 * frustum hold V cameras (host pointers); input_dhw is the grid sampled
 * FROM, output_dhw the grid sampled TO; inverse_transform and
 * linearize_depth select the coordinate mode. */
int ngt_lod_transform(const int32_t* input_dhw, const float* matrix_mv, const float* matrix_p, const float* frustum,
                      int num_cameras, const int32_t* output_dhw, int inverse_transform, int linearize_depth,
                      int relative_coords, float* output, ngt_stream_t stream);

/* ---- ReduceGridBlend (+Grad) (reduce_blend.cc:24-55) -----------------------
 * input [N,D,H,W,C] -> output [N,H,W,C] or, keep_dims, [N,D,H,W,C]. */
int ngt_reduce_grid_blend_fwd(const float* input, const int64_t* input_shape, int blending_mode, int keep_dims,
                              float* output, ngt_stream_t stream);
int ngt_reduce_grid_blend_bwd(const float* output_grad, const float* output, const float* input,
                              const int64_t* input_shape, int blending_mode, int keep_dims, int gradient_semantics,
                              float* input_grad, ngt_stream_t stream);

/* ---- RaymarchGridTransform (+Grad) (raymarch_grid.cc:29-92) ----------------
 * input [N,D,H,W,C] -> output [N,V|1,H',W',C] for output_dhw = (D',H',W')
 * (D' = samples per ray before BORDER clipping).  With LINEAR interpolation
 * and at least _fwd_workspace_size bytes of workspace the input is first
 * copied to the haloed brick layout that the ray gathers read (one address
 * per trilinear footprint, fewer cache lines per wavefront; see
 * ngt_bricked_grid); with a null / smaller workspace the rays read the NDHWC
 * input directly.  Both give the same result. */
size_t ngt_raymarch_grid_transform_fwd_workspace_size(const int64_t* input_shape, int interpolation);
int ngt_raymarch_grid_transform_fwd(const float* input, const int64_t* input_shape, const float* matrix_m,
                                    const float* matrix_v, const float* matrix_p, const float* frustum,
                                    int num_cameras, const int32_t* output_dhw, int interpolation, int boundary,
                                    int blending_mode, int separate_camera_batch, float* output, void* workspace,
                                    size_t workspace_bytes, ngt_stream_t stream);
/* Exact gradients with BORDER + LINEAR run atomics-free (a per-sample
 * gradient pass, then one gather per brick of voxels); they need this much
 * workspace (0 otherwise).  With less workspace the call falls back to the
 * fp32-atomic scatter.  Reference semantics need C+1 floats per voxel. */
size_t ngt_raymarch_grid_transform_bwd_workspace_size(const int64_t* input_shape, const float* matrix_m,
                                                      const float* matrix_v, const float* matrix_p,
                                                      const float* frustum, int num_cameras,
                                                      const int32_t* output_dhw, int interpolation, int boundary,
                                                      int separate_camera_batch, int gradient_semantics);
/* Host query, no counterpart in the reference: the per-ray sample bound the
 * gather adjoint sizes its per-sample record with (kcap above the BORDER ray
 * lengths of every pixel of every camera pair: (cell-centre box diagonal + 2)
 * / the exact minimum step length over the image, + slack, rounded up to a
 * multiple of 4); -1 when rays are too long for a record and the scatter
 * adjoint runs instead.  Pure host computation (no device memory is
 * touched). */
int ngt_raymarch_step_cap(const int64_t* input_shape, const float* matrix_m, const float* matrix_v,
                          const float* matrix_p, const float* frustum, int num_cameras, const int32_t* output_dhw,
                          int separate_camera_batch);
int ngt_raymarch_grid_transform_bwd(const float* input, const int64_t* input_shape, const float* output,
                                    const float* output_grad, const float* matrix_m, const float* matrix_v,
                                    const float* matrix_p, const float* frustum, int num_cameras,
                                    const int32_t* output_dhw, int interpolation, int boundary, int blending_mode,
                                    int separate_camera_batch, int gradient_semantics, float* input_grad,
                                    void* workspace, size_t workspace_bytes, ngt_stream_t stream);

/* ---- fused density render (Renderer.render_density, renderer.py:718-827) --
 * The fused raymarch of [light, density] * depth_step with constant lights
 * (Light.grid_lighting, lighting.py:26-34): the 4-channel cell
 * (rho*color*ds, rho*ds) is formed inside the kernel from the 1-channel
 * density, so neither the light grid nor the concatenation is materialised.
 * density [N,D,H,W,1], light_rgb = sum of color*intensity over the lights,
 * depth_step = (far-near)/D' per camera (all cameras equal); output
 * [N,V,H',W',4] = (RGB, optical depth), exactly the raymarch of the
 * reference's concatenated grid.  BEER_LAMBERT / ADDITIVE / ALPHA /
 * ALPHA_ADDITIVE as the raymarch op. */
size_t ngt_render_density_fwd_workspace_size(const int64_t* density_shape, int interpolation);
int ngt_render_density_fwd(const float* density, const int64_t* density_shape, const float* light_rgb,
                           float depth_step, const float* matrix_m, const float* matrix_v, const float* matrix_p,
                           const float* frustum, int num_cameras, const int32_t* output_dhw, int interpolation,
                           int boundary, int blending_mode, float* output, void* workspace, size_t workspace_bytes,
                           ngt_stream_t stream);
size_t ngt_render_density_bwd_workspace_size(const int64_t* density_shape, const float* matrix_m,
                                             const float* matrix_v, const float* matrix_p, const float* frustum,
                                             int num_cameras, const int32_t* output_dhw, int interpolation,
                                             int boundary);
int ngt_render_density_bwd(const float* density, const int64_t* density_shape, const float* light_rgb,
                           float depth_step, const float* output, const float* output_grad, const float* matrix_m,
                           const float* matrix_v, const float* matrix_p, const float* frustum, int num_cameras,
                           const int32_t* output_dhw, int interpolation, int boundary, int blending_mode,
                           float* density_grad, void* workspace, size_t workspace_bytes, ngt_stream_t stream);

/* ---- Haloed brick copy of a grid (no reference counterpart: the layout
 * the LINEAR ray kernels read, exposed so that a caller can make it once and
 * hand it to both the forward and the backward of the fused L2 render) -------
 * 4x4x4-cell bricks stored with a one-cell halo (5x5x5 padded to 128 floats),
 * halo cells holding the boundary mode's values.  _bytes returns 0 when the
 * grid's slices are too large for the layout's 32-bit offsets. */
size_t ngt_bricked_grid_bytes(const int64_t* shape);
int ngt_bricked_grid(const float* grid, const int64_t* shape, int boundary, float* bricked, ngt_stream_t stream);

/* ---- Fused render + L2 image loss (optimization.py:478-526, the density
 * target loss with loss_func = squared error) ---------------------------------
 * _l2_fwd renders like ngt_render_density_fwd into output [N,V,H,W,4] and
 * writes *loss (one device float) = loss_weight * sum over N,V,H,W and RGB of
 * (output - target)^2, target [N,V,H,W,3]; the sum is reduced in a fixed
 * order.  _l2_bwd writes density_grad for the upstream gradient *loss_grad
 * (device float): the image gradient 2 * loss_weight * loss_grad *
 * (output - target) is formed inside the render backward's gradient pass
 * instead of being read from memory.  `bricked` (nullable): the haloed brick
 * copy of density made by ngt_bricked_grid with the same boundary; when
 * given, neither call makes its own copy (the workspace sizes still assume
 * NULL).  `record` (nullable, LINEAR and BORDER only, needs `bricked`): a
 * buffer of ngt_render_density_l2_record_size bytes in which the forward also
 * records each ray and its per-sample gradients for a unit loss gradient
 * (BEER_LAMBERT and ADDITIVE within the same walk, the ALPHA blends in a
 * second walk of each ray right after the first); _l2_bwd with
 * `recorded` = 1, the same `bricked` and that buffer as its workspace then
 * only gathers them (scaled by *loss_grad).  The record size is 0 when the
 * cameras' rays are too long for a record (the callers then take the
 * two-walk backward). */
size_t ngt_render_density_l2_fwd_workspace_size(const int64_t* density_shape, int interpolation, int num_cameras,
                                                const int32_t* output_dhw);
int ngt_render_density_l2_fwd(const float* density, const int64_t* density_shape, const float* light_rgb,
                              float depth_step, const float* matrix_m, const float* matrix_v, const float* matrix_p,
                              const float* frustum, int num_cameras, const int32_t* output_dhw, int interpolation,
                              int boundary, int blending_mode, const float* target, float loss_weight,
                              const float* bricked, void* record, size_t record_bytes, float* output, float* loss,
                              void* workspace, size_t workspace_bytes, ngt_stream_t stream);
/* Allocation-free variant of ngt_render_density_l2_fwd.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid it).
 * Same arguments and results without the workspace: the scratch (the haloed
 * brick copy of the density when `bricked` is null, the loss partials) is a
 * static device array of ngt_static_scratch_bytes() bytes per GPU, held
 * until the process ends; calls needing more are rejected with
 * NGT_ERR_WORKSPACE (a 1-channel grid up to about 176^3).  All calls on one
 * GPU share that array: calls that overlap on one GPU race. */
size_t ngt_static_scratch_bytes(void);
int ngt_render_density_l2_fwd_noalloc(const float* density, const int64_t* density_shape, const float* light_rgb,
                                      float depth_step, const float* matrix_m, const float* matrix_v,
                                      const float* matrix_p, const float* frustum, int num_cameras,
                                      const int32_t* output_dhw, int interpolation, int boundary, int blending_mode,
                                      const float* target, float loss_weight, const float* bricked, void* record,
                                      size_t record_bytes, float* output, float* loss, ngt_stream_t stream);
size_t ngt_render_density_l2_record_size(const int64_t* density_shape, const float* matrix_m, const float* matrix_v,
                                          const float* matrix_p, const float* frustum, int num_cameras,
                                          const int32_t* output_dhw);
size_t ngt_render_density_l2_bwd_workspace_size(const int64_t* density_shape, const float* matrix_m,
                                                const float* matrix_v, const float* matrix_p, const float* frustum,
                                                int num_cameras, const int32_t* output_dhw, int interpolation,
                                                int boundary);
int ngt_render_density_l2_bwd(const float* density, const int64_t* density_shape, const float* light_rgb,
                              float depth_step, const float* output, const float* target, float loss_weight,
                              const float* loss_grad, const float* matrix_m, const float* matrix_v,
                              const float* matrix_p, const float* frustum, int num_cameras, const int32_t* output_dhw,
                              int interpolation, int boundary, int blending_mode, const float* bricked,
                              int recorded, float* density_grad, void* workspace, size_t workspace_bytes,
                              ngt_stream_t stream);

/* ---- AdvectGridSemiLangrange (advect.cc:21-71) with VelocityGrid.warp
 * semantics (data_structures.py:1690-1737) --------------------------------
 * input [N,D,H,W,C]; velocity either centered (vel_y = vel_z = NULL, vel_x =
 * [V,D,H,W,3]) or MAC-staggered (vel_x [V,D,H,W+1,1], vel_y [V,D,H+1,W,1],
 * vel_z [V,D+1,H,W,1], averaged to cell centres inside the kernels).
 * Sample position = index + timestep * v (callers pass timestep = -dt),
 * order 1 (semi-Lagrangian) or 2 (MacCormack), clamp NGT_CLAMP_*.
 * output [N,V|1,D,H,W,C].  `sl_out` (nullable, same shape as output) receives
 * the semi-Lagrangian field of an order-2 step for reuse by the backward. */
size_t ngt_advect_workspace_size(const int64_t* input_shape, int num_velocities, int separate_velocity_batch,
                                 int order, int backward);
int ngt_advect_fwd(const float* input, const int64_t* input_shape, const float* vel_x, const float* vel_y,
                   const float* vel_z, int num_velocities, float timestep, int order, int clamp, int boundary,
                   int separate_velocity_batch, float* output, float* sl_out, void* workspace,
                   size_t workspace_bytes, ngt_stream_t stream);
/* Gradients: input_grad [N,D,H,W,C]; velocity grads in the velocity layout
 * (vel_x_grad [V,D,H,W,3] centered, or the three face arrays).  All outputs
 * are fully overwritten.  `sl` (nullable) = the sl_out of the forward. */
int ngt_advect_bwd(const float* input, const int64_t* input_shape, const float* vel_x, const float* vel_y,
                   const float* vel_z, int num_velocities, float timestep, int order, int clamp, int boundary,
                   int separate_velocity_batch, const float* output_grad, const float* sl, int gradient_semantics,
                   float* input_grad, float* vel_x_grad, float* vel_y_grad, float* vel_z_grad, void* workspace,
                   size_t workspace_bytes, ngt_stream_t stream);

/* ---- Uneven batches (no single reference op: the reference launches every
 * (batch, camera) pair separately, resample_grid.cu.cc:748-790,
 * raymarch_grid.cu.cc:399-430).  A list of instances of different sizes runs
 * as one persistent launch per kernel over a work-list of cell tiles / bricks
 * / ray tiles.  Usage: _plan_bytes -> _plan (fills a host buffer, returns the
 * workspace size) -> copy the plan to device memory -> _fwd / _bwd with both
 * copies (the host copy sizes the launches, the kernels read the device
 * copy).  The plan holds the instances' device pointers: re-plan when a
 * buffer moves.  blocks_per_cu: work-groups per compute unit of the
 * persistent grid (0: one work-group per tile).  Launch-only, hence
 * hipGraph-capturable. */
typedef struct {
  int32_t dhw[3];           /* D, H, W of this instance (1-channel density) */
  float timestep;           /* samples at index + timestep * v (callers pass -dt) */
  const float* input;       /* [D,H,W,1] */
  const float* vel_x;       /* MAC faces [D,H,W+1] / [D,H+1,W] / [D+1,H,W], or */
  const float* vel_y;       /* vel_x = centered [D,H,W,3] with vel_y = vel_z = NULL */
  const float* vel_z;       /* (the same layout for every instance of a batch) */
  float* output;            /* fwd: advected [D,H,W,1] */
  float* sl;                /* fwd, order 2: semi-Lagrangian field, read again by the bwd */
  const float* output_grad; /* bwd */
  float* input_grad;        /* bwd: [D,H,W,1] */
  float* vel_x_grad;        /* bwd: in the velocity's layout */
  float* vel_y_grad;
  float* vel_z_grad;
} ngt_advect_instance;

/* No counterpart in the reference; batched AdvectGridSemiLangrange
 * (advect.cc:21-71) with VelocityGrid.warp semantics, exact gradients, C = 1. */
size_t ngt_advect_batch_plan_bytes(int num_instances);
int ngt_advect_batch_plan(const ngt_advect_instance* instances, int num_instances, int order, int clamp,
                          int boundary, void* plan, size_t plan_bytes, size_t* workspace_bytes);
int ngt_advect_batch_fwd(const void* plan_host, const void* plan_device, int blocks_per_cu, ngt_stream_t stream);
int ngt_advect_batch_bwd(const void* plan_host, const void* plan_device, void* workspace, size_t workspace_bytes,
                         int blocks_per_cu, ngt_stream_t stream);

typedef struct {
  int32_t dhw[3];           /* D, H, W of the density */
  const float* density;     /* [D,H,W,1] */
  int32_t num_views;        /* 1 .. 8 */
  const float* matrix_m;    /* host: 16 floats (op layout, as ngt_render_density_fwd) */
  const float* matrix_v;    /* host: 16 per view */
  const float* matrix_p;    /* host: 16 per view */
  const float* frustum;     /* host: 6 per view */
  int32_t image_dhw[3];     /* samples per ray D', image H', W' (all views of the instance) */
  float light_rgb[3];
  float depth_step;
  float* images;            /* fwd / l2: [V,H',W',4] (RGB, optical depth) */
  const float* images_grad; /* bwd */
  float* density_grad;      /* bwd / l2: [D,H,W,1] */
  const float* l2_target;   /* l2: [V,H',W',3] target images */
  float* loss;              /* l2: one float, sum over pixels and RGB of (image - target)^2 */
} ngt_render_instance;

/* No counterpart in the reference; batched fused density render
 * (Renderer.render_density, renderer.py:718-827), LINEAR + BORDER, exact
 * gradients by the brick-gather adjoint.  The plan holds the FP64-inverted
 * camera matrices.  record_budget: bytes for the gather adjoint's per-sample
 * records (0: 4 GiB); consecutive instances are grouped into chunks whose
 * records fit it (an instance whose records alone exceed it is a chunk of its
 * own), and the backward runs chunk by chunk over one record region, so the
 * workspace is the haloed density copies, one float per 16x16 ray tile and
 * the largest chunk's records.  ngt_render_batch_plan_info reports the chunk
 * count and the record region's size.
 * ngt_render_batch_fwd: images.  ngt_render_batch_bwd: density_grad for the
 * given images_grad; reads the images and the haloed copies of the last
 * ngt_render_batch_fwd on this workspace (the densities unchanged since).
 * ngt_render_batch_l2: images, loss and density_grad = d loss / d density in
 * one call (each chunk's recording walk is followed by its gather).
 * gather_brick: the gather's voxel brick, 0 (8x8x8), 1 (16x8x8) or
 * 2 (16x16x8); the results of different bricks agree to the fixed-point
 * rounding (2^-30 of the brick's largest sample gradient per sample). */
size_t ngt_render_batch_plan_bytes(int num_instances);
int ngt_render_batch_plan(const ngt_render_instance* instances, int num_instances, int blending_mode,
                          size_t record_budget, void* plan, size_t plan_bytes, size_t* workspace_bytes);
int ngt_render_batch_plan_info(const void* plan_host, int* num_chunks, size_t* record_bytes);
int ngt_render_batch_fwd(const void* plan_host, const void* plan_device, void* workspace, size_t workspace_bytes,
                         int blocks_per_cu, ngt_stream_t stream);
int ngt_render_batch_bwd(const void* plan_host, const void* plan_device, void* workspace, size_t workspace_bytes,
                         int blocks_per_cu, int gather_brick, ngt_stream_t stream);
int ngt_render_batch_l2(const void* plan_host, const void* plan_device, void* workspace, size_t workspace_bytes,
                        int blocks_per_cu, int gather_brick, ngt_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* NGT_TRANSPORT_H */
