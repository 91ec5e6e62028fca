"""Writes the golden fixtures of this directory: small seeded inputs and the
FP64 oracle's outputs for every op of the hot path (inputs rounded to fp32,
as the device sees them).  Run from the repository root:
``python tests/golden/make_golden.py``.  The fixtures are data; the oracle
(oracle/) is the program that produced them."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import BORDER, CLAMP, LINEAR, MIN  # noqa: E402
from oracle import geometry as G  # noqa: E402
from oracle import resample as R  # noqa: E402
from oracle import blend as B  # noqa: E402
from oracle import raymarch as RM  # noqa: E402
from oracle import advect as A  # noqa: E402


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def ring(dims, out_dhw, views=2):
    t = (0.0123, -0.0217, 0.0311)
    M = G.grid_transform(dims, translation_=t, center=True, scale_=[1.0 / dims[2]] * 3)
    cams = G.look_at_ring_cameras(views, 3.0, 40.0, out_dhw[2] / out_dhw[1], 2.0, 4.0)
    Ms = np.stack([G.op_matrix(M)]).astype(np.float32)
    Vs = np.stack([G.op_matrix(c[0]) for c in cams]).astype(np.float32)
    Ps = np.stack([G.op_matrix(c[1]) for c in cams]).astype(np.float32)
    Fs = np.stack([c[2] for c in cams]).astype(np.float32)
    return Ms, Vs, Ps, Fs


def smooth(rng, shape, amp):
    from scipy.ndimage import gaussian_filter
    f = gaussian_filter(rng.standard_normal(shape), 1.2)
    return amp * f / np.max(np.abs(f))


def main():
    out = {}
    rng = np.random.default_rng(2024)
    # SampleGridLut, relative LUT, LINEAR / MIN, BORDER / CLAMP, with the gradient
    x = f32(rng.standard_normal((1, 5, 6, 7, 2)))
    lut = f32(np.concatenate([rng.uniform(-1.3, 1.3, (1, 4, 5, 6, 3)), np.zeros((1, 4, 5, 6, 1))], -1))
    g = f32(rng.standard_normal((1, 1, 4, 5, 6, 2)))
    for name, fm, bm in (("lut_linear_border", LINEAR, BORDER), ("lut_min_clamp", MIN, CLAMP)):
        y = R.sample_grid_lut(x, lut, fm, bm, relative=True)
        gi, gl = R.sample_grid_lut_grad(x, lut, g, fm, bm, relative=True)
        out[name] = dict(input=x, lut=lut, out_grad=g, output=y, input_grad=gi, lut_grad=gl)
    # SampleGridTransform (LINDEPTH) and its gradient
    dims = (6, 7, 8)
    Ms, Vs, Ps, Fs = ring(dims, (5, 6, 7))
    x = f32(rng.standard_normal((1,) + dims + (1,)))
    y = R.sample_grid_transform(x, Ms, Vs, Ps, Fs, (5, 6, 7), LINEAR, BORDER)
    g = f32(rng.standard_normal(y.shape))
    gi = R.sample_grid_transform_grad(x, g, Ms, Vs, Ps, Fs, LINEAR, BORDER)
    out["transform_lindepth"] = dict(input=x, M=Ms, V=Vs, P=Ps, F=Fs, out_grad=g, output=y, input_grad=gi)
    # ReduceGridBlend BEER_LAMBERT (4 channels) and its gradient
    x = f32(rng.uniform(0, 0.4, (2, 6, 3, 4, 4)))
    y = B.reduce_grid_blend(x, B.BEER_LAMBERT)
    g = f32(rng.standard_normal(y.shape))
    gi = B.reduce_grid_blend_grad(g, y, x, B.BEER_LAMBERT)
    out["blend_beer_lambert"] = dict(input=x, out_grad=g, output=y, input_grad=gi)
    # RaymarchGridTransform BEER_LAMBERT (4 channels) and its exact gradient
    dims = (7, 6, 8)
    Ms, Vs, Ps, Fs = ring(dims, (9, 5, 6))
    x = f32(rng.uniform(0, 0.2, (1,) + dims + (4,)))
    y = RM.raymarch_grid_transform(x, Ms, Vs, Ps, Fs, (9, 5, 6), LINEAR, BORDER, B.BEER_LAMBERT)
    g = f32(rng.standard_normal(y.shape))
    gi = RM.raymarch_grid_transform_grad(x, y, g, Ms, Vs, Ps, Fs, (9, 5, 6), LINEAR, BORDER, B.BEER_LAMBERT)
    out["raymarch_beer_lambert"] = dict(input=x, M=Ms, V=Vs, P=Ps, F=Fs, out_grad=g, output=y, input_grad=gi)
    # MacCormack + MC_SMOOTH advection (centered velocity) and its gradient
    D, H, W = 6, 7, 8
    rho = f32(np.clip(smooth(rng, (D, H, W), 1.0), 0, 1))[..., None]
    vc = f32(np.stack([smooth(rng, (D, H, W), 1.4) for _ in range(3)], -1))
    y = A.warp_forward(rho, vc, 1.0, 2, "MC_SMOOTH", BORDER)
    g = f32(rng.standard_normal(y.shape))
    gr, gv = A.warp_backward(rho, vc, g, 1.0, 2, "MC_SMOOTH", BORDER)
    out["advect_maccormack"] = dict(input=rho, velocity=vc, out_grad=g, output=y, input_grad=gr, velocity_grad=gv)
    for name, arrays in out.items():
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **{k: np.asarray(v) for k, v in arrays.items()})
        print(name, {k: np.asarray(v).shape for k, v in arrays.items()})


if __name__ == "__main__":
    main()
