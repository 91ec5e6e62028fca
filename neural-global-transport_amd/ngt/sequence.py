"""Multi-step transport with periodic multi-view rendering, differentiated
with step-wise recomputation checkpointing.

One step advances the density and the centered velocity by MacCormack
advection (the density with ``VelocityGrid.warp`` semantics,
data_structures.py:1690-1737; the velocity advected by itself, the centered
analogue of ``VelocityGrid.warped``, data_structures.py:1081-1164).  Every
``render_every``-th step the density is rendered into the cameras (fused
density renderer, renderer.py:718-827) and an L2 image loss is accumulated.

With ``checkpoint=True`` only the (density, velocity) pair entering each
segment of ``render_every`` steps is kept for the backward pass; a segment's
steps and render are recomputed when its gradient is needed
(``torch.utils.checkpoint``).  Device memory then grows with the number of
segments times two grids instead of with every intermediate field, sample
buffer and image of the chain.
"""
import torch
from torch.utils.checkpoint import checkpoint as _checkpoint

from . import ops


def transport_step(rho, vel, dt=1.0, clamp="MC_SMOOTH", boundary="BORDER"):
    """One MacCormack step of the density [1,D,H,W,1] and the centered
    velocity [1,D,H,W,3] (both advected by the velocity at the start of the
    step)."""
    rho1 = ops.advect_grid_semi_langrange(rho, vel, -dt, 2, clamp, boundary, False).squeeze(1)
    vel1 = ops.advect_grid_semi_langrange(vel, vel, -dt, 2, clamp, boundary, False).squeeze(1)
    return rho1, vel1


class RenderSpec:
    """Fused-render arguments of the chain: op-format camera matrices,
    light color*intensity, depth step and output (D', H', W')."""

    def __init__(self, matrix_m, matrix_v, matrix_p, frustum, depth_step, output_dhw, light_rgb=(1.0, 1.0, 1.0),
                 boundary="BORDER", blending_mode="BEER_LAMBERT"):
        self.M, self.V, self.P, self.F = matrix_m, matrix_v, matrix_p, frustum
        self.depth_step = float(depth_step)
        self.output_dhw = tuple(int(v) for v in output_dhw)
        self.light_rgb = tuple(float(c) for c in light_rgb)
        self.boundary = boundary
        self.blending_mode = blending_mode

    def render(self, rho):
        """[N,V,H,W,4] images (RGB, optical depth)."""
        return ops.render_density_fused(rho, self.light_rgb, self.depth_step, self.M, self.V, self.P, self.F,
                                        self.output_dhw, boundary=self.boundary, blending_mode=self.blending_mode)

    def l2_loss(self, rho, target):
        """sum((render(rho)[..., :3] - target)^2), rendered and reduced in one
        fused op whose adjoint forms the image gradient in-kernel."""
        loss, _ = ops.render_density_fused_l2(rho, target, self.light_rgb, self.depth_step, self.M, self.V, self.P,
                                              self.F, self.output_dhw, boundary=self.boundary,
                                              blending_mode=self.blending_mode)
        return loss

    @classmethod
    def from_cameras(cls, grid_transform, cameras, light_rgb=(1.0, 1.0, 1.0), boundary="BORDER"):
        from .renderer import Renderer
        M, V, P, F = Renderer()._transform_params([grid_transform], cameras)
        return cls(M, V, P, F, cameras[0].depth_step, cameras[0].transform.grid_size, light_rgb, boundary)


def _segment(rho, vel, steps, spec, target, dt, clamp, boundary):
    for _ in range(steps):
        rho, vel = transport_step(rho, vel, dt, clamp, boundary)
    return rho, vel, spec.l2_loss(rho, target)


def chain_loss(rho0, vel0, spec, targets, render_every=5, dt=1.0, clamp="MC_SMOOTH", boundary="BORDER",
               checkpoint=True):
    """Sum over k of || render(rho_{(k+1)*render_every}) - targets[k] ||^2 for
    len(targets) segments of ``render_every`` steps; differentiable w.r.t.
    rho0 and vel0.  ``targets[k]`` are [N,V,H,W,3] images."""
    rho, vel = rho0, vel0
    total = None
    for target in targets:
        if checkpoint:
            rho, vel, loss = _checkpoint(_segment, rho, vel, render_every, spec, target, dt, clamp, boundary,
                                         use_reentrant=False)
        else:
            rho, vel, loss = _segment(rho, vel, render_every, spec, target, dt, clamp, boundary)
        total = loss if total is None else total + loss
    return total


def render_targets(rho0, vel0, spec, segments, render_every=5, dt=1.0, clamp="MC_SMOOTH", boundary="BORDER"):
    """The chain's images without gradients (e.g. targets from a perturbed
    initial state)."""
    out = []
    with torch.no_grad():
        rho, vel = rho0, vel0
        for _ in range(segments):
            for _ in range(render_every):
                rho, vel = transport_step(rho, vel, dt, clamp, boundary)
            out.append(spec.render(rho)[..., :3].contiguous())
    return out
