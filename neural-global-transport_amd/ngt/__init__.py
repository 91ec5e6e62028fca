"""MI355X-native differentiable volume transport (drop-in for the render and
advect ops of Neural-Global-Transport's phitest.render package)."""
from . import ops  # noqa: F401
from .ops import (sample_grid_lut, sample_grid_lut_grad, sample_grid_transform, sample_grid_transform_grad,  # noqa: F401
                  reduce_grid_blend, reduce_grid_blend_grad, raymarch_grid_transform, raymarch_grid_transform_grad,
                  advect_grid_semi_langrange, sampling_ops, blending_ops, raymarching_ops, advection_ops)

__version__ = "0.1.0"
