"""Lights.  Mirror of phitest/render/lighting.py:12-40 (``Light``): constant
light whose per-cell contribution is ``density * color * intensity``
(``OLD_LIGHTING`` path, lighting.py:26-34)."""
import numpy as np


class Light(object):
    def __init__(self, color=(1.0, 1.0, 1.0), intensity=1.0, reflection=1.0):
        self.c = np.asarray(color, dtype=np.float64)
        self.i = intensity
        self.r = reflection

    @property
    def monochrome(self):
        return np.isscalar(self.c) or np.asarray(self.c).shape in ((), (1,))

    def grid_lighting(self, density_data, density_transforms=None, renderer=None, scattering_func=None):
        import torch
        c = torch.as_tensor(np.asarray(self.c, dtype=np.float32).reshape(-1), device=density_data.device)
        return density_data * c * float(self.i)

    def rgb(self):
        """color * intensity as 3 floats (used by the fused renderer)."""
        c = np.broadcast_to(np.asarray(self.c, dtype=np.float64).reshape(-1), (3,))
        return c * float(self.i)

    def to_dict(self):
        return {"intensity": float(self.i), "color": list(np.asarray(self.c).tolist())}
