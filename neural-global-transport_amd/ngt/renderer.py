"""Renderer: the call surface of phitest/render/renderer.py on torch + the HIP ops.

Same constructor and methods as the reference ``Renderer`` for the render and
warp paths: ``render_density`` (renderer.py:718-827), ``_render_cameras``
(:669-682), ``raymarch_camera`` / ``_raymarch_camera_transform`` (:372-555),
``sample_camera`` / ``_sample_camera_transform`` (:290-504), ``_sample_LuT``
(:559-574), ``_sample_transform`` / ``resample_grid3D_aligned`` (:222-286,
843-905), the lookup-table helpers (:123-196, 325-370), ``unproject``
(:952-976), ``_blend_grid`` (:577-640), ``_build_light_grid`` (:687-716),
``resample_grid3D_offset`` (:922-936).  Data are NDHWC torch tensors on the
HIP device; images are returned as a list of NHWC tensors (one per camera).

MI355X-specific: when rendering is fused, every light is a constant
``Light`` and no custom op hooks the grid, ``render_density`` runs the fused
density renderer that forms ``[rho*color, rho]*depth_step`` inside the
raymarch kernel (no light grid, no 4-channel concatenation in HBM, a
1-channel gradient scatter); the result equals the reference's raymarch of the
concatenated grid.
"""
import logging
from collections.abc import Iterable

import numpy as np
import torch

from . import ops
from .camera import Camera
from .lighting import Light
from .transform import GridTransform


class _NullProfiler:
    class _S:
        def __enter__(self):
            return self

        def __exit__(self, *a):
            return False

    def sample(self, name):
        return self._S()


def _mat_list(mats):
    return np.stack([np.asarray(m, dtype=np.float32).reshape(16) for m in mats])


class Renderer(object):
    def __init__(self, profiler=None, filter_mode="LINEAR", boundary_mode="BORDER", mipmapping="LINEAR", num_mips=2,
                 mip_bias=0.0, sample_gradients=False, fast_gradient_mip_bias_add=0.0, blend_mode="BEER_LAMBERT",
                 name="Renderer", **kwargs):
        self.log = logging.getLogger(name)
        self.profiler = profiler if profiler is not None else _NullProfiler()
        self.filter_mode = filter_mode
        self.boundary_mode = boundary_mode
        self.mip_mode = mipmapping
        self.num_mips = num_mips
        self.mip_bias = mip_bias
        self.gradient_mip_bias_add = fast_gradient_mip_bias_add
        self.sample_gradients = sample_gradients
        self.blend_mode = blend_mode
        self.luma = kwargs.get("luma", [0.2126, 0.7152, 0.0722])
        self.allow_fused = kwargs.get("fused", False)
        self.fused_lighting = kwargs.get("fused_lighting", True)
        self.scattering_ratio = kwargs.get("scattering_ratio", 1.0)

    @property
    def can_render_fused(self):
        return self.allow_fused and (self.blend_mode in ["BEER_LAMBERT", "ADDITIVE", "ALPHA_ADDITIVE", "ALPHA"])

    def _sort_cameras(self, cameras):
        def split_by_size(cams):
            sizes = []
            for cam in cams:
                if cam.transform.grid_size not in sizes:
                    sizes.append(cam.transform.grid_size)
            return [(size, [cam for cam in cams if cam.transform.grid_size == size]) for size in sizes]

        static_cams = split_by_size([c for c in cameras if c.static is not None])
        non_static_cams = split_by_size([c for c in cameras if c.static is None])
        return non_static_cams, static_cams

    def _get_camera_params_batch(self, cameras):
        M_V = [camera.view_matrix().transpose() for camera in cameras]
        M_P = [camera.projection_matrix().transpose() for camera in cameras]
        F = [camera.frustum() for camera in cameras]
        return M_V, M_P, F

    def _transform_params(self, transformations, cameras):
        M = [t.get_transform_matrix().transpose() for t in transformations]
        V, P, F = self._get_camera_params_batch(cameras)
        return _mat_list(M), _mat_list(V), _mat_list(P), np.stack([np.asarray(f, np.float32) for f in F])

    # --- sampling ------------------------------------------------------------------

    def _sample_camera_transform(self, data, transformations, cameras, inverse=False):
        CM = {False: "TRANSFORM_LINDEPTH", True: "TRANSFORM_LINDEPTH_REVERSE"}
        M, V, P, F = self._transform_params(transformations, cameras)
        out_shape = transformations[0].grid_size if inverse else cameras[0].transform.grid_size
        kw = dict(interpolation=self.filter_mode, boundary=self.boundary_mode, mipmapping=self.mip_mode,
                  num_mipmaps=self.num_mips, mip_bias=self.mip_bias)
        if self.sample_gradients:
            return _GatherGradSample.apply(data, (M, V, P, F), tuple(out_shape), CM[inverse], CM[not inverse], kw,
                                           self.gradient_mip_bias_add)
        return ops.sample_grid_transform(data, M, V, P, F, out_shape, coordinate_mode=CM[inverse], **kw)

    # --- lookup tables (LodTransform) ------------------------------------------------

    def get_camera_LuT(self, grid_transform, cam, inverse=False):
        """[1, D, H, W, 4] lookup table (position, LoD) of one camera
        (renderer.py:123-135)."""
        M_model = grid_transform.get_transform_matrix().transpose()
        V, P, F = self._get_camera_params_batch([cam])
        MV = [M_model @ v for v in V]
        in_shape, out_shape = ((cam.transform.grid_size, grid_transform.grid_size) if inverse
                               else (grid_transform.grid_size, cam.transform.grid_size))
        return ops.lod_transform(in_shape, _mat_list(MV), _mat_list(P), np.stack(F), out_shape,
                                 inverse_transform=inverse)

    def get_transform_LuT(self, from_transformations, to_transformations, relative=False, inverse=False,
                          fix_scale_center=False):
        """Lookup tables between grid transforms, no perspective
        (renderer.py:137-161)."""
        M = [t.get_transform_matrix().transpose() for t in from_transformations]
        if fix_scale_center:
            V = [(GridTransform(t.grid_size, scale=[2, 2, 2], center=True, normalize="ALL").get_transform_matrix()
                  @ t.get_inverse_transform()).transpose() for t in to_transformations]
        else:
            V = [t.get_inverse_transform().transpose() for t in to_transformations]
        MV = [m @ v for m, v in zip(M, V)]
        P = [np.eye(4, dtype=np.float32) for _ in to_transformations]
        F = [np.asarray([-1, 1, -1, 1, 1, -1], dtype=np.float32) for _ in to_transformations]
        in_shape, out_shape = ((to_transformations[0].grid_size, from_transformations[0].grid_size) if inverse
                               else (from_transformations[0].grid_size, to_transformations[0].grid_size))
        return ops.lod_transform(in_shape, _mat_list(MV), _mat_list(P), np.stack(F), out_shape,
                                 linearize_depth=False, relative_coords=relative, inverse_transform=inverse)

    def _setup_static_camera_LuT(self, grid_transform, sizes_cams, inverse=False):
        """Builds and stores the lookup tables of static cameras
        (renderer.py:163-181)."""
        size, cams = sizes_cams
        M_model = grid_transform.get_transform_matrix().transpose()
        V, P, F = self._get_camera_params_batch(cams)
        MV = [M_model @ v for v in V]
        in_shape, out_shape = (size, grid_transform.grid_size) if inverse else (grid_transform.grid_size, size)
        luts = ops.lod_transform(in_shape, _mat_list(MV), _mat_list(P), np.stack(F), out_shape,
                                 inverse_transform=inverse)
        for lut, cam in zip(luts.unbind(0), cams):
            if inverse:
                cam.inverseLuT = lut
            else:
                cam.LuT = lut

    def _get_cam_luts(self, transformation, size, cams, inverse):
        pending = [c for c in cams if (c.inverseLuT if inverse else c.LuT) is None]
        if pending:
            self._setup_static_camera_LuT(transformation, (size, pending), inverse)
        return torch.stack([c.inverseLuT if inverse else c.LuT for c in cams])

    def _sample_camera_LuT(self, data, transformation, cameras, inverse=False):
        """Camera sampling through cached lookup tables (renderer.py:325-370)."""
        kw = dict(interpolation=self.filter_mode, boundary=self.boundary_mode, mipmapping=self.mip_mode,
                  num_mipmaps=self.num_mips, mip_bias=self.mip_bias, coordinate_mode="LOOKUP",
                  relative_coords=False, normalized_coords=False)
        size = cameras[0].transform.grid_size
        luts = self._get_cam_luts(transformation, size, cameras, inverse)
        if self.sample_gradients:
            if data.shape[0] != 1:
                raise NotImplementedError("camera lut rendering does not support data batches.")
            inv = self._get_cam_luts(transformation, size, cameras, not inverse)
            return _GatherGradLut.apply(data, luts, inv, kw, self.gradient_mip_bias_add)
        return ops.sample_grid_lut(data, luts, **kw)

    # --- orthographic resampling between grids --------------------------------------

    def _sample_transform(self, data, from_transformations, to_transformations, inverse=False,
                          fix_scale_center=False, **kernelargs):
        """Resampling between grid transforms through an orthographic identity
        camera over the [-1, 1] cube of the target (renderer.py:222-286);
        returns NVDHWC."""
        CM = {False: "TRANSFORM", True: "TRANSFORM_REVERSE"}
        M = [t.get_transform_matrix().transpose() for t in from_transformations]
        if fix_scale_center:
            V = [(GridTransform(t.grid_size, scale=[2, 2, 2], center=True, normalize="ALL").get_transform_matrix()
                  @ t.get_inverse_transform()).transpose() for t in to_transformations]
        else:
            V = [t.get_inverse_transform().transpose() for t in to_transformations]
        P = [np.eye(4, dtype=np.float32) for _ in to_transformations]
        F = np.asarray([[-1, 1, -1, 1, 1, -1]] * len(to_transformations), dtype=np.float32)
        cams = (_mat_list(M), _mat_list(V), _mat_list(P), F)
        out_shape = tuple(int(v) for v in (from_transformations[0].grid_size if inverse
                                           else to_transformations[0].grid_size))
        kw = dict(interpolation=kernelargs.get("interpolation", self.filter_mode),
                  boundary=kernelargs.get("boundary", self.boundary_mode),
                  mipmapping=kernelargs.get("mipmapping", self.mip_mode), num_mipmaps=self.num_mips,
                  mip_bias=self.mip_bias)
        if self.sample_gradients:
            return _GatherGradSample.apply(data, cams, out_shape, CM[inverse], CM[not inverse], kw,
                                           self.gradient_mip_bias_add)
        return ops.sample_grid_transform(data, *cams, out_shape, coordinate_mode=CM[inverse], **kw)

    @staticmethod
    def _aligned_scale(align, in_size, out_size):
        a = align.upper()
        if a == "BORDER":
            return 2.0 / in_size
        if a == "CENTER":  # outer cell centres meet
            return 2.0 / out_size * (out_size - 1.0) / (in_size - 1.0)
        if a == "STAGGER_INPUT":  # input centres on output borders
            return 2.0 / (in_size - 1.0)
        if a == "STAGGER_OUTPUT":  # input borders on output centres
            return 2.0 / (in_size + 1.0)
        raise ValueError("unknown alignment {}".format(align))

    def resample_grid3D_aligned(self, data, target_shape, align_x="BORDER", align_y="BORDER", align_z="BORDER",
                                allow_split_channels=False):
        """[N, D, H, W, C] -> [N, target D, H, W, C] with per-axis alignment
        BORDER / CENTER / STAGGER_INPUT / STAGGER_OUTPUT (renderer.py:843-905).
        Channel counts other than 1, 2, 4 need ``allow_split_channels`` (the
        channels are then resampled as separate 1-channel grids)."""
        if data.dim() != 5:
            raise ValueError("input shape must be NDHWC, is : {}".format(list(data.shape)))
        if len(target_shape) != 3:
            raise ValueError("target_shape must have 3 entries")
        N, D, H, W, C = data.shape
        if not allow_split_channels and C not in (1, 2, 4):
            raise ValueError("resampling supports 1, 2 or 4 channels, is {}".format(C))
        tz, ty, tx = (int(v) for v in target_shape)
        scale = [self._aligned_scale(align_x, W, tx), self._aligned_scale(align_y, H, ty),
                 self._aligned_scale(align_z, D, tz)]
        in_transform = GridTransform([D, H, W], scale=scale, center=True)
        target_transform = GridTransform([tz, ty, tx])
        if C not in (1, 2, 4):
            x = data.permute(0, 4, 1, 2, 3).reshape(N * C, D, H, W, 1)
            y = self._sample_transform(x, [in_transform] * (N * C), [target_transform]).squeeze(1)
            return y.reshape(N, C, tz, ty, tx).permute(0, 2, 3, 4, 1).contiguous()
        return self._sample_transform(data, [in_transform] * N, [target_transform]).squeeze(1)

    def sample_camera(self, data, transformations, cameras, inverse=False, allow_static=True, force_static=False,
                      use_step_channel=None, squeeze_batch=None):
        if data.dim() != 5:
            raise ValueError("data must be 5D (NDHWC), is {}".format(data.dim()))
        n = data.shape[0]
        if not isinstance(transformations, Iterable):
            transformations = [transformations]
            if n != 1:
                transformations = transformations * n
            elif squeeze_batch is None:
                squeeze_batch = True
        elif len(transformations) != n:
            raise ValueError("transformation and data batch size mismatch: {} - {}".format(len(transformations), n))
        if squeeze_batch is None:
            squeeze_batch = False
        if not isinstance(cameras, Iterable):
            cameras = [cameras]
        if len(cameras) > 1 and not all(c.transform.grid_size == cameras[0].transform.grid_size for c in cameras[1:]):
            raise ValueError("all cameras must have the same resolution (DHW). (use Renderer._sort_cameras() for batching.)")
        sample_lut = False
        if allow_static and any(c.static is not None for c in cameras):
            # static cameras sample through their cached lookup tables when they belong to this grid
            # transform and there is no data batch (renderer.py:455-465)
            if not all(c.static == transformations[0] for c in cameras) or n != 1:
                if force_static:
                    raise ValueError("Camera static setup does not match transformation or data is batched")
                self.log.warning("Incorrect static camera setup, falling back to transform rendering for static "
                                 "cameras.")
            else:
                sample_lut = True
        apply_step = None
        if use_step_channel is not None and use_step_channel != []:
            data, apply_step = self._premultiply_step(data, cameras, use_step_channel, allow_post=True)
        if sample_lut:
            sampled = self._sample_camera_LuT(data, transformations[0], cameras, inverse)
        else:
            sampled = self._sample_camera_transform(data, transformations, cameras, inverse)
        if apply_step is not None:
            sampled = sampled * apply_step
        return sampled.squeeze(0) if (squeeze_batch and n == 1) else sampled

    def _premultiply_step(self, data, cameras, use_step_channel, allow_post=False):
        C = data.shape[-1]
        if np.isscalar(use_step_channel):
            return data * use_step_channel, None
        step_channel = sorted(c % C for c in use_step_channel if -C <= c < C)
        depth_steps = [cam.depth_step for cam in cameras]
        if all(s == depth_steps[0] for s in depth_steps):
            if step_channel == list(range(C)):
                return data * depth_steps[0], None
            scale = torch.tensor([depth_steps[0] if c in step_channel else 1.0 for c in range(C)],
                                 dtype=torch.float32, device=data.device)
            return data * scale, None
        if not allow_post:
            raise ValueError("All cameras must have the same step size for batched rendering:\n%s" % (depth_steps,))
        step = torch.tensor([[s if c in step_channel else 1.0 for c in range(C)] for s in depth_steps],
                            dtype=torch.float32, device=data.device).reshape(1, len(cameras), 1, 1, 1, C)
        return data, step

    def _raymarch_camera_transform(self, data, transformations, cameras, inverse=False, **kwargs):
        M, V, P, F = self._transform_params(transformations, cameras)
        kw = dict(interpolation=kwargs.get("filter_mode", self.filter_mode),
                  boundary=kwargs.get("boundary_mode", self.boundary_mode),
                  blending_mode=kwargs.get("blend_mode", self.blend_mode),
                  separate_camera_batch=kwargs.get("global_sampling", True))
        if inverse:
            if kw["blending_mode"] not in ["ADDITIVE"]:
                raise ValueError("Inverse raymarching not possible for blend mode %s" % kw["blending_mode"])
            grid_size = transformations[0].grid_size
            dummy = torch.zeros((data.shape[0],) + tuple(grid_size) + (data.shape[-1],), dtype=data.dtype,
                                device=data.device)
            out_shape = [cameras[0].transform.grid_size[0], data.shape[2], data.shape[3]]
            return ops.raymarch_grid_transform_grad(dummy, data, data, M, V, P, F, out_shape, **kw)
        return ops.raymarch_grid_transform(data, M, V, P, F, cameras[0].transform.grid_size, **kw)

    def unproject(self, grid_transform, targets, cameras, blend_func=torch.minimum):
        """Volume estimate from per-camera target images [V, H, W, C]: each
        target, extruded along its frustum depth, sampled back into the grid
        and combined with ``blend_func`` (renderer.py:952-976, non-fused
        branch)."""
        C = targets.shape[-1]
        out = torch.ones((1,) + tuple(int(v) for v in grid_transform.grid_size) + (C,), dtype=torch.float32,
                         device=targets.device)
        for i, cam in enumerate(cameras):
            t = targets[i].reshape((1, 1) + tuple(targets[i].shape))
            t = t.expand((1, int(cam.transform.grid_size[0])) + tuple(targets[i].shape)).contiguous()
            out = blend_func(out, self.sample_camera(t, grid_transform, cam, inverse=True))
        return out

    def raymarch_camera(self, data, transformations, cameras, use_step_channel=None, squeeze_batch=None,
                        **raymarch_kwargs):
        if data.dim() != 5:
            raise ValueError("data must be 5D (NDHWC for forward or NVHWC for inverse sampling), is {}".format(data.dim()))
        inverse = raymarch_kwargs.get("inverse", False)
        n = data.shape[0]
        if not isinstance(transformations, Iterable):
            transformations = [transformations]
            if n != 1:
                transformations = transformations * n
            elif squeeze_batch is None:
                squeeze_batch = True
        if squeeze_batch is None:
            squeeze_batch = False
        if not isinstance(cameras, Iterable):
            cameras = [cameras]
        if not inverse and len(transformations) != n:
            raise ValueError("transformation and data batch size mismatch: {} - {}".format(len(transformations), n))
        if len(cameras) > 1 and not all(c.transform.grid_size == cameras[0].transform.grid_size for c in cameras[1:]):
            raise ValueError("all cameras must have the same resolution (DHW). (use Renderer._sort_cameras() for batching.)")
        if use_step_channel is not None and use_step_channel != []:
            data, _ = self._premultiply_step(data, cameras, use_step_channel)
        sampled = self._raymarch_camera_transform(data, transformations, cameras, **raymarch_kwargs)
        return sampled.squeeze(0) if (squeeze_batch and n == 1) else sampled

    def _sample_LuT(self, data, luts, combined_batch=False, relative=False, normalized=False, cell_center_offset=0.0):
        if luts.shape[-1] == 3:
            luts = torch.nn.functional.pad(luts, (0, 1))
        if luts.shape[-1] != 4:
            raise ValueError("LuT must have 3 or 4 channels")
        res = ops.sample_grid_lut(data, luts, interpolation=self.filter_mode, boundary=self.boundary_mode,
                                  mipmapping=self.mip_mode if self.mip_mode else "NONE", num_mipmaps=self.num_mips,
                                  mip_bias=self.mip_bias, coordinate_mode="LOOKUP",
                                  separate_camera_batch=not combined_batch, cell_center_offset=cell_center_offset,
                                  relative_coords=relative, normalized_coords=normalized)
        return res.squeeze(1) if combined_batch else res

    def _blend_grid(self, data, blend_mode=None, keep_dims=False):
        """Blend along depth (axis -4).  MAX / MEAN / MIN are torch reductions;
        BEER_LAMBERT (which the reference composes from cumsum / exp,
        renderer.py:605-614) and the other modes use the ReduceGridBlend kernel
        (identical for non-negative densities; the kernel clamps the optical
        depth at 0 as the reference op does)."""
        blend_mode = (blend_mode or self.blend_mode).upper()
        if blend_mode == "MAX":
            return data.amax(dim=-4, keepdim=keep_dims)
        if blend_mode == "MEAN":
            return data.mean(dim=-4, keepdim=keep_dims)
        if blend_mode == "MIN":
            return data.amin(dim=-4, keepdim=keep_dims)
        shp = data.shape
        flat = data.reshape((-1,) + tuple(shp[-4:]))
        out = ops.reduce_grid_blend(flat, blend_mode, keep_dims)
        return out.reshape(tuple(shp[:-4]) + tuple(out.shape[1:]))

    def _apply_custom_ops(self, tensor, custom_ops, name):
        result = tensor
        if custom_ops is not None and name in custom_ops:
            fns = custom_ops[name]
            for fn in (fns if isinstance(fns, Iterable) else [fns]):
                result = fn(result)
        return result

    def _render_cameras(self, light_density, grid_transform, cameras, camera_size, custom_ops=None):
        light_density = self._apply_custom_ops(light_density, custom_ops, "GRID")
        step = [0, 1, 2, 3] if self.blend_mode not in ["MAX", "MEAN", "MIN"] else None
        if self.can_render_fused and (custom_ops is None or "FRUSTUM" not in custom_ops):
            image_hdr = self.raymarch_camera(light_density, grid_transform, cameras, use_step_channel=step)
        else:
            frustum = self.sample_camera(light_density, grid_transform, cameras, False, use_step_channel=step)
            frustum = self._apply_custom_ops(frustum, custom_ops, "FRUSTUM")
            image_hdr = self._blend_grid(frustum, self.blend_mode)
        return self._apply_custom_ops(image_hdr, custom_ops, "IMAGE")

    def _build_light_grid(self, density_data, density_transforms, light_list, monochrome=False):
        shape = list(density_data.shape)
        shape[-1] = 1 if monochrome else 3
        light = torch.zeros(shape, dtype=torch.float32, device=density_data.device)
        for lt in light_list:
            if isinstance(lt, Light):
                g = lt.grid_lighting(density_data, density_transforms, self)
                if monochrome and not lt.monochrome:
                    g = (g * torch.tensor(self.luma, device=g.device, dtype=g.dtype)).sum(-1, keepdim=True)
                if not monochrome and lt.monochrome:
                    g = g.expand(shape)
            elif isinstance(lt, (torch.Tensor, np.ndarray)):
                g = torch.as_tensor(lt, dtype=torch.float32, device=density_data.device)
                if g.shape[-1] == 3 and monochrome:
                    g = (g * torch.tensor(self.luma, device=g.device)).sum(-1, keepdim=True)
                g = g.expand(shape)
            else:
                raise ValueError("Type of light '{}' is not supported.".format(type(lt).__name__))
            light = light + g
        return light

    def _fused_light_rgb(self, light_list, monochrome, custom_ops):
        """color*intensity summed over the lights when the fused density
        renderer applies, else None."""
        if not (self.fused_lighting and self.can_render_fused and not monochrome and light_list):
            return None
        if custom_ops is not None and any(k in custom_ops for k in ("DENSITY", "GRID", "FRUSTUM")):
            return None
        if not all(type(lt) is Light for lt in light_list):
            return None
        return np.sum([lt.rgb() for lt in light_list], axis=0)

    def render_density(self, density_transform, light_list, camera_list, cut_alpha=True, background=None,
                       monochrome=False, split_cameras=False, custom_ops=None, tonemapping="NONE"):
        """Render density (GridTransform with NDHW1 data, or a list of them)
        with the lights into every camera; returns a list of NHWC images in
        camera_list order (renderer.py:718-827)."""
        if not isinstance(density_transform, Iterable):
            density_transforms = [density_transform.copy_no_data() for _ in range(density_transform.get_batch_size())]
            density_data = density_transform.data
        else:
            datas, density_transforms = [], []
            size = density_transform[0].get_grid_size()
            for dt in density_transform:
                if np.any(np.not_equal(size, dt.get_grid_size())):
                    raise ValueError("All density grids must have the same spatial shape for batched rendering.")
                datas.append(dt.data)
                density_transforms.extend(dt.copy_no_data() for _ in range(dt.get_batch_size()))
            density_data = torch.cat(datas, 0)
        density_data = self._apply_custom_ops(density_data, custom_ops, "DENSITY")
        fused_rgb = self._fused_light_rgb(light_list, monochrome, custom_ops)
        data = None
        if fused_rgb is None:
            light = self._build_light_grid(density_data, density_transforms, light_list, monochrome)
            data = torch.cat([light, density_data], dim=-1)
        cam_images = [None] * len(camera_list)
        non_static, static = self._sort_cameras(camera_list)
        for cam_size, cams in non_static + static:
            groups = [[c] for c in cams] if split_cameras else [cams]
            for group in groups:
                if fused_rgb is not None:
                    images = self._render_fused(density_data, fused_rgb, density_transforms, group)
                else:
                    images = self._render_cameras(data, density_transforms, group, cam_size, custom_ops=custom_ops)
                ts = torch.exp(-images[..., -1:])
                if cut_alpha:
                    images = images[..., :-1]
                for i, cam in enumerate(group):
                    image = images[:, i]
                    t = ts[:, i]
                    if background is not None:
                        bkg = torch.as_tensor(background[camera_list.index(cam)], dtype=torch.float32,
                                              device=image.device)
                        image = image + bkg.expand(image.shape) * t
                    image = self._tonemap(image, tonemapping)
                    cam_images[camera_list.index(cam)] = image
        return cam_images

    def _render_fused(self, density, light_rgb, transformations, cameras):
        depth_steps = [c.depth_step for c in cameras]
        if not all(s == depth_steps[0] for s in depth_steps):
            raise ValueError("All cameras must have the same step size for batched rendering:\n%s" % (depth_steps,))
        M, V, P, F = self._transform_params(transformations, cameras)
        return ops.render_density_fused(density, light_rgb, depth_steps[0], M, V, P, F,
                                        cameras[0].transform.grid_size, interpolation=self.filter_mode,
                                        boundary=self.boundary_mode, blending_mode=self.blend_mode)

    def _tonemap(self, image, mode="NONE"):
        mode = mode.upper()
        if mode == "NONE":
            return image
        if mode == "CLIP_NEGATIVE":
            return torch.clamp(image, min=0)
        if mode == "SATURATE":
            return torch.clamp(image, 0, 1)
        if mode == "NORMALIZE":
            mn = image.min()
            if mn < 0:
                image = image - mn
            mx = image.max()
            return image / mx if mx > 0 else image
        raise ValueError(mode)

    def resample_grid3D_offset(self, data, offsets, target_shape):
        """Sample ``data`` at ``index + offset`` for each offset row (N,3) into
        ``target_shape`` (renderer.py:922-936)."""
        offsets = torch.as_tensor(np.asarray(offsets, dtype=np.float32), device=data.device)
        if offsets.dim() != 2 or offsets.shape[1] != 3:
            raise ValueError("Shape of offsets must be (N,3), is %s" % (list(offsets.shape),))
        if len(target_shape) != 3:
            raise ValueError("target_shape must be (3,), is %s" % (target_shape,))
        lut = torch.nn.functional.pad(offsets, (0, 1)).reshape(offsets.shape[0], 1, 1, 1, 4)
        lut = lut.expand((offsets.shape[0],) + tuple(int(v) for v in target_shape) + (4,)).contiguous()
        return self._sample_LuT(data, lut, relative=True)


class _GatherGradSample(torch.autograd.Function):
    """``sample_gradients=True`` (renderer.py:296-323): the backward of camera
    sampling is the inverse-transform gather of dy, summed over the views,
    instead of the scatter transpose."""

    @staticmethod
    def forward(ctx, data, cams, out_shape, cm_fwd, cm_bwd, kw, bias_add):
        M, V, P, F = cams
        y = ops.sample_grid_transform(data, M, V, P, F, out_shape, coordinate_mode=cm_fwd, **kw)
        ctx.cams, ctx.cm_bwd, ctx.kw, ctx.bias_add = cams, cm_bwd, kw, bias_add
        ctx.in_shape = tuple(data.shape[1:4])
        return y

    @staticmethod
    def backward(ctx, dy):
        M, V, P, F = ctx.cams
        kw = dict(ctx.kw)
        kw["mip_bias"] = kw["mip_bias"] + ctx.bias_add
        views = V.shape[0]
        dx = []
        for i in range(dy.shape[0]):
            dxv = ops.sample_grid_transform(dy[i].contiguous(), np.repeat(M[i:i + 1], views, 0), V, P, F, ctx.in_shape,
                                            coordinate_mode=ctx.cm_bwd, separate_camera_batch=False, **kw)
            dx.append(dxv.sum(0))
        return torch.cat(dx, 0), None, None, None, None, None, None


class _GatherGradLut(torch.autograd.Function):
    """``sample_gradients=True`` through lookup tables (renderer.py:330-352):
    the backward gathers dy through the inverse lookup tables and sums the
    views."""

    @staticmethod
    def forward(ctx, data, luts, inv_luts, kw, bias_add):
        ctx.inv_luts, ctx.kw, ctx.bias_add = inv_luts, kw, bias_add
        return ops.sample_grid_lut(data, luts, **kw)

    @staticmethod
    def backward(ctx, dy):
        kw = dict(ctx.kw)
        kw["mip_bias"] = kw["mip_bias"] + ctx.bias_add
        dx = ops.sample_grid_lut(dy[0].contiguous(), ctx.inv_luts, separate_camera_batch=False, **kw)
        return dx.sum(0), None, None, None, None
