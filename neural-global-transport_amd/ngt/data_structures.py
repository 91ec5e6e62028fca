"""Grid data model for the transport path: DensityGrid, VelocityGrid, State.

Mirror of the warp / render parts of phitest/render/data_structures.py:
``VelocityGrid.component_shapes`` (:765-773), ``centered`` (:1634-1653),
``centered_lut_grid`` (:1684-1688), ``warp`` (:1690-1737),
``_sampled_to_component_shape`` (:1655-1682), ``_warp_vel_component`` /
``warped`` / ``copy_warped`` (:1081-1177), ``DensityGrid.warped`` (:614-617),
``State.render_density`` / ``density_advected`` / ``velocity_advected``
(:2260-2330), and the resolution changes ``DensityGrid.scaled`` (:548-564),
``VelocityGrid.scaled`` / ``resample_velocity`` / ``_centered_to_staggered``
/ ``_scalar_centered_to_staggered`` / ``_sampled_to_shape`` (:975-1058,
1341-1434).  Tensors are torch NDHWC on the HIP device.

``VelocityGrid.warp`` runs the fused MacCormack kernels (one SL kernel + one
correction/clamp kernel, velocity centred on the fly from the MAC faces).
``fused=False`` reproduces the reference's composition instead: 2-4
SampleGridLut launches on an explicit 4-channel LUT plus elementwise torch
ops - the per-op "plain" path.
"""
import copy
import logging

import numpy as np
import torch

from . import ops
from .renderer import Renderer
from .transform import GridTransform


def _default_warp_renderer(boundary="BORDER"):
    return Renderer(filter_mode="LINEAR", boundary_mode=boundary, mipmapping="NONE", blend_mode="ADDITIVE",
                    sample_gradients=False)


_LOG = logging.getLogger("data_structures")

# resampling between resolutions: LINEAR, no mips (the reference's scale renderer, reconstruct_sequence.py:602-609)
_default_scale_renderer = _default_warp_renderer


def _as_tensor(a, device):
    if a is None or isinstance(a, torch.Tensor):
        return a
    return torch.as_tensor(np.asarray(a, dtype=np.float32), device=device)


def _to_numpy(t):
    return t.detach().cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t)


class DensityGrid:
    """Density [1, D, H, W, 1] with optional hull and inflow.  The inflow is a
    sub-grid placed at ``inflow_offset`` (z, y, x) and optionally masked;
    ``with_inflow`` = max(d + padded inflow, 0) (data_structures.py:380-462)."""

    def __init__(self, shape, d=None, scale_renderer=None, hull=None, inflow=None, inflow_offset=None,
                 inflow_mask=None, device=None, **kwargs):
        self.shape = [int(v) for v in shape]
        dev = device or (d.device if isinstance(d, torch.Tensor) else "cuda")
        if d is None:
            d = torch.zeros([1] + self.shape + [1], dtype=torch.float32, device=dev)
        self._d = _as_tensor(d, dev)
        self.hull = _as_tensor(hull, dev)
        self._inflow = _as_tensor(inflow, dev)
        self.inflow_mask = _as_tensor(inflow_mask, dev) if inflow is not None else None
        self.inflow_offset = None
        if self._inflow is not None:
            self.inflow_offset = [int(v) for v in (inflow_offset if inflow_offset is not None else (0, 0, 0))]
            ishape = list(self._inflow.shape[-4:-1])
            for a in range(3):
                if self.inflow_offset[a] < 0 or self.inflow_offset[a] + ishape[a] > self.shape[a]:
                    raise ValueError("inflow of shape %s at offset %s does not fit the grid %s"
                                     % (ishape, self.inflow_offset, self.shape))
        self.scale_renderer = scale_renderer if scale_renderer is not None else _default_scale_renderer()

    @property
    def d(self):
        return self._d

    @property
    def inflow(self):
        """The (masked) inflow padded to the grid shape (zeros without inflow)."""
        if self._inflow is None:
            return torch.zeros_like(self._d)
        f = self._inflow * self.inflow_mask if self.inflow_mask is not None else self._inflow
        ish = f.shape[-4:-1]
        off = self.inflow_offset
        pad = []
        for a in (2, 1, 0):  # F.pad takes the last axes first: x, y, z
            pad += [off[a], self.shape[a] - off[a] - ish[a]]
        return torch.nn.functional.pad(f, [0, 0] + pad)

    def with_inflow(self):
        return self._d if self._inflow is None else torch.clamp(self._d + self.inflow, min=0.0)

    @classmethod
    def from_file(cls, path, scale_renderer=None, hull=None, inflow=None, inflow_offset=None, inflow_mask=None,
                  device=None, **kwargs):
        """Density npz (data_structures.py:465-485): 'arr_0' [1, D, H, W, 1] and
        optionally 'hull', 'inflow', 'inflow_mask', 'inflow_offset'.  Read
        without pickle; None (and a warning) if the file cannot be read."""
        try:
            with np.load(path, allow_pickle=False) as f:
                d = f["arr_0"]
                if "hull" in f and hull is None:
                    hull = f["hull"]
                if "inflow" in f and inflow is None:
                    inflow = f["inflow"]
                    if "inflow_mask" in f and inflow_mask is None:
                        inflow_mask = f["inflow_mask"]
                    if "inflow_offset" in f and inflow_offset is None:
                        inflow_offset = f["inflow_offset"].tolist()
                return cls(list(d.shape[-4:-1]), d=d, scale_renderer=scale_renderer, hull=hull, inflow=inflow,
                           inflow_offset=inflow_offset, inflow_mask=inflow_mask, device=device)
        except Exception:
            _LOG.warning("Failed to load density from '%s':", path, exc_info=True)
            return None

    def save(self, path):
        """npz written as the reference writes it (data_structures.py:704-726)."""
        extra = {}
        if self.hull is not None:
            extra["hull"] = _to_numpy(self.hull)
        if self._inflow is not None:
            extra["inflow"] = _to_numpy(self._inflow)
            if self.inflow_mask is not None:
                extra["inflow_mask"] = _to_numpy(self.inflow_mask)
            extra["inflow_offset"] = np.asarray(self.inflow_offset)
        np.savez_compressed(path, _to_numpy(self._d), **extra)

    def assign(self, d):
        self._d = d

    def warped(self, vel_grid, order=1, dt=1.0, clamp="NONE"):
        if not isinstance(vel_grid, VelocityGrid):
            raise ValueError("Invalid velocity grid")
        return vel_grid.warp(self.with_inflow(), order=order, dt=dt, clamp=clamp)

    def copy_warped(self, vel_grid, order=1, dt=1.0, clamp="NONE", **kwargs):
        return DensityGrid(self.shape, d=self.warped(vel_grid, order, dt, clamp), scale_renderer=self.scale_renderer)

    def scaled(self, new_shape, with_inflow=False):
        """The density resampled to ``new_shape`` (BORDER-aligned,
        data_structures.py:548-564)."""
        if not (isinstance(new_shape, (list, tuple)) and len(new_shape) == 3):
            raise ValueError("Invalid shape")
        density = self.with_inflow() if with_inflow else self.d
        if list(new_shape) == self.shape:
            return density
        return self.scale_renderer.resample_grid3D_aligned(density, list(new_shape))

    def copy_scaled(self, new_shape, **kwargs):
        hull = None if self.hull is None else self.scale_renderer.resample_grid3D_aligned(self.hull, list(new_shape))
        return DensityGrid(new_shape, d=self.scaled(new_shape), scale_renderer=self.scale_renderer, hull=hull)


class VelocityGrid:
    @staticmethod
    def component_shapes(centered_shape):
        assert len(centered_shape) == 3
        x_shape = list(centered_shape)
        x_shape[2] += 1
        y_shape = list(centered_shape)
        y_shape[1] += 1
        z_shape = list(centered_shape)
        z_shape[0] += 1
        return x_shape, y_shape, z_shape

    def __init__(self, centered_shape, std=0.1, x=None, y=None, z=None, boundary=None, scale_renderer=None,
                 warp_renderer=None, device=None, **kwargs):
        self.centered_shape = [int(v) for v in centered_shape]
        self.x_shape, self.y_shape, self.z_shape = self.component_shapes(self.centered_shape)
        dev = device or "cuda"
        self.x = _as_tensor(x, dev) if x is not None else torch.randn([1] + self.x_shape + [1], device=dev) * std
        self.y = _as_tensor(y, dev) if y is not None else torch.randn([1] + self.y_shape + [1], device=dev) * std
        self.z = _as_tensor(z, dev) if z is not None else torch.randn([1] + self.z_shape + [1], device=dev) * std
        self.boundary = boundary
        self.scale_renderer = scale_renderer if scale_renderer is not None else _default_scale_renderer()
        self.warp_renderer = warp_renderer if warp_renderer is not None else _default_warp_renderer()

    @staticmethod
    def shape_centered_to_staggered(shape):
        return [v + 1 for v in shape]

    @property
    def staggered_shape(self):
        return self.shape_centered_to_staggered(self.centered_shape)

    # --- resolution changes (data_structures.py:975-1058, 1341-1382, 1413-1434) ------

    def scaled(self, centered_shape, scale_magnitude=True):
        """MAC components resampled to a new centred resolution, each aligned
        at the cell centres along its own axis; magnitudes scaled by the
        per-axis resolution ratio."""
        if not (isinstance(centered_shape, (list, tuple)) and len(centered_shape) == 3):
            raise ValueError("Invalid shape")
        centered_shape = [int(v) for v in centered_shape]
        if centered_shape == self.centered_shape:
            return self.x, self.y, self.z
        xs, ys, zs = self.component_shapes(centered_shape)
        r = self.scale_renderer
        x = r.resample_grid3D_aligned(self.x, xs, align_x="CENTER")
        y = r.resample_grid3D_aligned(self.y, ys, align_y="CENTER")
        z = r.resample_grid3D_aligned(self.z, zs, align_z="CENTER")
        if scale_magnitude:
            s = [o / i for i, o in zip(self.centered_shape, centered_shape)]  # z, y, x
            x, y, z = x * s[2], y * s[1], z * s[0]
        return x, y, z

    def copy_scaled(self, centered_shape, scale_magnitude=True, **kwargs):
        x, y, z = self.scaled(centered_shape, scale_magnitude)
        return VelocityGrid(centered_shape, x=x, y=y, z=z, boundary=self.boundary, scale_renderer=self.scale_renderer,
                            warp_renderer=self.warp_renderer)

    @staticmethod
    def resample_velocity(scale_renderer, vel, scale=None, shape=None, is_staggered=True, scale_magnitude=True):
        """Velocity (MAC component triple, or centred [N,D,H,W,3]) at a new
        resolution given by ``scale`` or ``shape`` (data_structures.py:1000-1058)."""
        if (scale is None) == (shape is None):
            raise ValueError("You must provide 1 of scale or shape.")
        if isinstance(vel, tuple):
            centered = list(vel[0].shape)
            centered[-2] -= 1
        elif isinstance(vel, torch.Tensor):
            centered = list(vel.shape)
        else:
            raise ValueError("Unknown velcity format '%s'" % (type(vel).__name__,))
        if len(centered) != 5:
            raise ValueError("velocity must be 5D")
        if scale is not None:
            shape = [int(round(v * scale)) for v in centered[1:-1]]
        else:
            if not (isinstance(shape, (list, tuple)) and len(shape) == 3):
                raise ValueError("invalid shape")
            scale = float(np.mean([float(t) / float(s) for s, t in zip(centered[1:-1], shape)]))
        r = scale_renderer
        if is_staggered:
            if not isinstance(vel, tuple):
                raise NotImplementedError("combined staggered velocity tensors are not supported")
            xs, ys, zs = VelocityGrid.component_shapes(shape)
            x, y, z = vel
            comps = (r.resample_grid3D_aligned(x, xs, align_x="CENTER", align_y="BORDER", align_z="BORDER"),
                     r.resample_grid3D_aligned(y, ys, align_x="BORDER", align_y="CENTER", align_z="BORDER"),
                     r.resample_grid3D_aligned(z, zs, align_x="BORDER", align_y="BORDER", align_z="CENTER"))
            return tuple(c * scale for c in comps) if scale_magnitude else comps
        parts = torch.split(vel, [2, 1], dim=-1)
        out = torch.cat([r.resample_grid3D_aligned(p.contiguous(), shape) for p in parts], dim=-1)
        return out * scale if scale_magnitude else out

    def _centered_to_staggered(self, centered, output_centered_shape=None, scale_magnitude=True):
        """Centred [N,D,H,W,3] velocity -> MAC components of
        ``output_centered_shape`` (data_structures.py:1341-1364)."""
        if centered.dim() != 5 or centered.shape[-1] != 3:
            raise ValueError("centered velocity must be [N,D,H,W,3]")
        in_shape = [int(v) for v in centered.shape[1:4]]
        if output_centered_shape is None:
            output_centered_shape, factors = in_shape, None
        else:
            factors = [o / i for o, i in zip(output_centered_shape, in_shape)]
        xs, ys, zs = self.component_shapes(list(output_centered_shape))
        cx, cy, cz = (c.contiguous() for c in torch.split(centered, 1, dim=-1))
        r = self.scale_renderer
        x = r.resample_grid3D_aligned(cx, xs, align_x="CENTER", align_y="BORDER", align_z="BORDER")
        y = r.resample_grid3D_aligned(cy, ys, align_x="BORDER", align_y="CENTER", align_z="BORDER")
        z = r.resample_grid3D_aligned(cz, zs, align_x="BORDER", align_y="BORDER", align_z="CENTER")
        if scale_magnitude and factors is not None:
            x, y, z = x * factors[2], y * factors[1], z * factors[0]
        return x, y, z

    def _scalar_centered_to_staggered(self, centered, allow_split_channels=False):
        """A centred scalar field on the staggered (+1) grid, its cell centres
        on the input's cell borders (data_structures.py:1366-1382)."""
        if centered.dim() != 5:
            raise ValueError("centered field must be 5D")
        staggered = self.shape_centered_to_staggered([int(v) for v in centered.shape[1:4]])
        return self.scale_renderer.resample_grid3D_aligned(centered, staggered, align_x="STAGGER_OUTPUT",
                                                           align_y="STAGGER_OUTPUT", align_z="STAGGER_OUTPUT",
                                                           allow_split_channels=allow_split_channels)

    def _sampled_to_shape(self, shape):
        """The MAC components resampled onto a grid of ``shape`` with uniform
        scaling (data_structures.py:1413-1434)."""
        shape = [int(v) for v in shape]
        scale = [2.0 / v for v in shape[::-1]]
        target = GridTransform(shape)
        out = []
        for comp, cs in ((self.x, self.x_shape), (self.y, self.y_shape), (self.z, self.z_shape)):
            if shape == cs:
                out.append(comp)
            else:
                t = GridTransform(cs, scale=scale, center=True)
                out.append(self.scale_renderer._sample_transform(comp, [t], [target]).squeeze(1))
        return out

    def assign(self, x, y, z):
        self.x, self.y, self.z = x, y, z

    @classmethod
    def from_file(cls, path, boundary=None, scale_renderer=None, warp_renderer=None, device=None, **kwargs):
        """Velocity npz (data_structures.py:909-925): 'vel_x', 'vel_y',
        'vel_z' MAC components and 'centered_shape' (derived from vel_x in
        legacy files).  Read without pickle; None (and a warning) if the file
        cannot be read."""
        try:
            with np.load(path, allow_pickle=False) as v:
                if "centered_shape" in v:
                    shape = [int(s) for s in v["centered_shape"].tolist()]
                else:
                    shape = list(v["vel_x"].shape)
                    shape[-2] -= 1
                    shape = shape[1:-1]
                dev = device or "cuda"
                comps = [_as_tensor(v[k].astype(np.float32), dev) for k in ("vel_x", "vel_y", "vel_z")]
            return cls(shape, x=comps[0], y=comps[1], z=comps[2], boundary=boundary, scale_renderer=scale_renderer,
                       warp_renderer=warp_renderer, device=device)
        except Exception:
            _LOG.warning("Failed to load velocity from '%s':", path, exc_info=True)
            return None

    def save(self, path):
        """npz written as the reference writes it (data_structures.py:1197-1198)."""
        np.savez_compressed(path, centered_shape=self.centered_shape, vel_x=_to_numpy(self.x), vel_y=_to_numpy(self.y),
                            vel_z=_to_numpy(self.z))

    def centered(self, pad_lod=False, concat=True):
        h = 0.5
        c = [(self.x[:, :, :, 1:] + self.x[:, :, :, :-1]) * h,
             (self.y[:, :, 1:] + self.y[:, :, :-1]) * h,
             (self.z[:, 1:] + self.z[:, :-1]) * h]
        if pad_lod:
            c.append(torch.zeros_like(c[0]))
        return torch.cat(c, dim=-1) if concat else c

    def centered_lut_grid(self, dt=1.0, centered_velocity=None):
        vel_centered = self.centered() if centered_velocity is None else centered_velocity
        return vel_centered * (-dt)

    def _can_fuse(self):
        return self.warp_renderer.filter_mode == "LINEAR" and self.warp_renderer.mip_mode in ("NONE", None)

    def warp(self, data, order=1, dt=1.0, clamp="NONE", centered_velocity=None, fused=True):
        """Advect ``data`` [N,D,H,W,C] with this velocity (semi-Lagrangian, or
        MacCormack for order 2, with the MC_SMOOTH revert clamp)."""
        clamp = clamp.upper()
        if list(data.shape[1:4]) != self.centered_shape:
            raise ValueError("Shape mismatch in centered warp: data {}, velocity {}".format(
                list(data.shape[1:4]), self.centered_shape))
        if order not in (1, 2):
            raise ValueError("Unsupported warp order '{}'".format(order))
        if clamp == "MC":
            raise NotImplementedError("MIM and MAX warp sampling have wrong gradients.")
        if fused and self._can_fuse():
            vel = centered_velocity if centered_velocity is not None else (self.x, self.y, self.z)
            out = ops.advect_grid_semi_langrange(data, vel, timestep=-dt, order=order,
                                                 clamp_extrema=clamp if clamp in ("MC_SMOOTH", "NEGATIVE") else "NONE",
                                                 boundary=self.warp_renderer.boundary_mode,
                                                 separate_velocity_batch=False)
            return out.squeeze(1)  # a view: its backward launches nothing
        return self._warp_plain(data, order, dt, clamp, centered_velocity)

    def _warp_plain(self, data, order, dt, clamp, centered_velocity=None):
        r = self.warp_renderer
        v = self.centered_lut_grid(dt=dt, centered_velocity=centered_velocity)
        warped = r._sample_LuT(data, v, True, relative=True)
        if order == 2:
            back = r._sample_LuT(warped, -v, True, relative=True)
            corr = warped + 0.5 * (data - back)
            if clamp == "MC_SMOOTH":
                fm = r.filter_mode
                r.filter_mode = "MIN"
                dmin = r._sample_LuT(data, v, True, relative=True)
                r.filter_mode = "MAX"
                dmax = r._sample_LuT(data, v, True, relative=True)
                r.filter_mode = fm
                oob = torch.logical_or(corr < dmin, corr > dmax)
                corr = torch.where(oob, warped, corr)
            warped = corr
        if clamp == "NEGATIVE":
            warped = torch.clamp(warped, min=0)
        return warped

    # --- velocity self-advection --------------------------------------------------

    def _sampled_to_component_shape(self, component, pad_lod=False, concat=True):
        component = component.upper()
        off_from, off_to = 0.5, -0.5
        r = self.warp_renderer
        out = []
        out.append(self.x if component == "X" else r.resample_grid3D_offset(
            self.x, [[off_from, off_to, 0.0] if component == "Y" else [off_from, 0.0, off_to]],
            self.y_shape if component == "Y" else self.z_shape).squeeze(1))
        out.append(self.y if component == "Y" else r.resample_grid3D_offset(
            self.y, [[off_to, off_from, 0.0] if component == "X" else [0.0, off_from, off_to]],
            self.x_shape if component == "X" else self.z_shape).squeeze(1))
        out.append(self.z if component == "Z" else r.resample_grid3D_offset(
            self.z, [[off_to, 0.0, off_from] if component == "X" else [0.0, off_to, off_from]],
            self.x_shape if component == "X" else self.y_shape).squeeze(1))
        if pad_lod:
            out.append(torch.zeros_like(out[0]))
        return torch.cat(out, dim=-1) if concat else out

    def _warp_vel_component(self, data, lut, order=1, dt=1.0, clamp="NONE"):
        """Warp one MAC component with a component-shaped relative LUT."""
        if order < 1 or order > 2:
            raise ValueError("Unsupported warp order '{}'".format(order))
        clamp = clamp.upper()
        if clamp == "MC":
            raise NotImplementedError("MIM and MAX warp sampling have wrong gradients.")
        if self._can_fuse():
            # sample position = index + 1 * lut, exactly the relative LUT read
            out = ops.advect_grid_semi_langrange(data, lut.contiguous(), timestep=1.0, order=order,
                                                 clamp_extrema=clamp if clamp in ("MC_SMOOTH", "NEGATIVE") else "NONE",
                                                 boundary=self.warp_renderer.boundary_mode,
                                                 separate_velocity_batch=False)
            return out.squeeze(1)  # a view: its backward launches nothing
        r = self.warp_renderer
        warped = r._sample_LuT(data, lut, True, relative=True)
        if order == 2:
            back = r._sample_LuT(warped, -lut, True, relative=True)
            corr = warped + 0.5 * (data - back)
            if clamp == "MC_SMOOTH":
                fm = r.filter_mode
                r.filter_mode = "MIN"
                dmin = r._sample_LuT(data, lut, True, relative=True)
                r.filter_mode = "MAX"
                dmax = r._sample_LuT(data, lut, True, relative=True)
                r.filter_mode = fm
                corr = torch.where(torch.logical_or(corr < dmin, corr > dmax), warped, corr)
            warped = corr
        return warped

    def warped(self, vel_grid=None, order=1, dt=1.0, clamp="NONE"):
        """Self-advect the three MAC components (vel_grid None) or advect them
        with another velocity grid of the same centered shape."""
        res = []
        for comp, data in (("X", self.x), ("Y", self.y), ("Z", self.z)):
            if vel_grid is None:
                lut = self._sampled_to_component_shape(comp) * (-dt)
            else:
                lut = vel_grid._sampled_to_component_shape_of(self, comp) * (-dt)
            res.append(self._warp_vel_component(data, lut, order=order, dt=dt, clamp=clamp))
        return tuple(res)

    def _sampled_to_component_shape_of(self, other, comp):
        if other.centered_shape != self.centered_shape:
            raise NotImplementedError("advecting a velocity grid of another resolution")
        return self._sampled_to_component_shape(comp)

    def copy_warped(self, vel_grid=None, order=1, dt=1.0, clamp="NONE", **kwargs):
        x, y, z = self.warped(vel_grid, order, dt, clamp=clamp)
        return VelocityGrid(self.centered_shape, x=x, y=y, z=z, boundary=self.boundary,
                            scale_renderer=self.scale_renderer, warp_renderer=self.warp_renderer)


class RenderingContext:
    """Cameras + lights + renderer (renderer.py:35-59, the fields the render
    path uses)."""

    def __init__(self, cameras, lights, dens_renderer, vel_renderer=None, monochrome=False, render_SDF=False):
        self.cameras = cameras
        self.lights = lights
        self.dens_renderer = dens_renderer
        self.vel_renderer = vel_renderer
        self.monochrome = monochrome
        self.render_SDF = render_SDF


class State:
    def __init__(self, density, velocity, frame=0, transform=None, target_cameras=None, **kwargs):
        self.density = density
        self.velocity = velocity
        self.frame = frame
        self.transform = transform
        self.target_cameras = target_cameras
        self.images = None
        self.t = None

    def get_density_transform(self):
        return self.transform.copy_new_data(self.density.d)

    def render_density(self, render_ctx, custom_ops=None, **kwargs):
        cams = self.target_cameras if self.target_cameras is not None else render_ctx.cameras
        imgs = torch.stack(render_ctx.dens_renderer.render_density(
            self.get_density_transform(), light_list=render_ctx.lights, camera_list=cams, cut_alpha=False,
            monochrome=render_ctx.monochrome, custom_ops=custom_ops), dim=1)
        imgs, d = imgs[..., :3], imgs[..., 3:]
        self.images = imgs
        self.t = torch.exp(-d)
        return imgs, self.t

    def density_advected(self, dt=1.0, order=1, clamp="NONE"):
        return self.density.warped(self.velocity, order=order, dt=dt, clamp=clamp)

    def velocity_advected(self, dt=1.0, order=1, clamp="NONE"):
        return self.velocity.copy_warped(order=order, dt=dt, clamp=clamp)
