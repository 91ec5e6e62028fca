"""The FP32 OpenMP CPU restatement (oracle/cpu_port.cpp, the timed CPU
baseline of bench.py) against the FP64 NumPy oracle: MacCormack + MC_SMOOTH
on MAC velocity and the fused render + L2 loss, forward and exact adjoint.

Tolerances: forward 1e-5, gradients 1e-4, normwise relative (FP32 vs FP64).
"""
import numpy as np
import pytest

from oracle import BORDER, LINEAR
from oracle import advect as A
from oracle import blend as B
from oracle import geometry as G
from oracle import raymarch as RM


@pytest.fixture(scope="module")
def port():
    from oracle import build_cpu_port, cpu_port
    build_cpu_port.build()
    return cpu_port


def rel(got, exp):
    return float(np.max(np.abs(got - exp))) / max(float(np.max(np.abs(exp))), 1e-30)


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def _smooth(rng, shape, amp):
    from scipy.ndimage import gaussian_filter
    f = gaussian_filter(rng.standard_normal(shape), 1.5)
    return amp * f / np.max(np.abs(f))


def test_advect_mc_forward_and_adjoint(port):
    rng = np.random.default_rng(3)
    D, H, W = 10, 12, 14
    rho = f32(np.clip(_smooth(rng, (D, H, W), 1.0), 0, 1))
    faces = [f32(_smooth(rng, s, 1.6)) for s in ((D, H, W + 1), (D, H + 1, W), (D + 1, H, W))]
    vc = A.centered_from_staggered(*(f[..., None] for f in faces))
    out, sl = port.advect_mc_fwd(rho, *faces, -1.0)
    exp = A.warp_forward(rho[..., None], vc, 1.0, 2, "MC_SMOOTH", BORDER)[..., 0]
    assert rel(out, exp) < 1e-5
    g = f32(rng.standard_normal((D, H, W)))
    g_rho, gx, gy, gz = port.advect_mc_bwd(rho, *faces, -1.0, sl, g)
    eg_rho, eg_v = A.warp_backward(rho[..., None], vc, g[..., None], 1.0, 2, "MC_SMOOTH", BORDER)
    assert rel(g_rho, eg_rho[..., 0]) < 1e-4
    for got, e in zip((gx, gy, gz), A.centered_from_staggered_adjoint(eg_v)):
        assert rel(got, e[..., 0]) < 1e-4


def test_render_l2_forward_and_adjoint(port):
    rng = np.random.default_rng(4)
    dims, hw, depth = (12, 11, 13), (10, 14), 16
    M = G.grid_transform(dims, translation_=(0.0123, -0.0217, 0.0311), center=True, scale_=[1.0 / dims[2]] * 3)
    cams = G.look_at_ring_cameras(3, 3.0, 40.0, hw[1] / hw[0], 2.0, 4.0)
    Ms = [G.op_matrix(M).astype(np.float32)]
    Vs = [G.op_matrix(c[0]).astype(np.float32) for c in cams]
    Ps = [G.op_matrix(c[1]).astype(np.float32) for c in cams]
    Fs = [c[2].astype(np.float32) for c in cams]
    rgb = np.array([0.9, 0.7, 0.5])
    ds = 2.0 / depth
    rho = f32(rng.uniform(0.0, 0.5, dims))
    out_dhw = (depth,) + hw
    img = port.render_fwd(rho, Ms[0], Vs, Ps, out_dhw, rgb, ds)
    cat = np.concatenate([rho[..., None] * rgb, rho[..., None]], -1)[None] * ds
    eimg = RM.raymarch_grid_transform(cat, Ms, Vs, Ps, Fs, out_dhw, LINEAR, BORDER, B.BEER_LAMBERT)[0]
    assert rel(img, eimg) < 1e-5
    tgt = f32(rng.uniform(0.0, 0.3, (3,) + hw + (3,)))
    w = 0.37
    loss = port.l2_loss(img, tgt, w)
    assert abs(loss - w * np.sum((eimg[..., :3] - tgt) ** 2)) < 1e-5 * abs(loss)
    g = port.render_l2_bwd(rho, Ms[0], Vs, Ps, out_dhw, rgb, ds, img, tgt, w, 2.0)
    gimg = np.zeros_like(eimg)
    gimg[..., :3] = 2.0 * w * 2.0 * (eimg[..., :3] - tgt)
    g_cat = RM.raymarch_grid_transform_grad(cat, eimg[None], gimg[None], Ms, Vs, Ps, Fs, out_dhw, LINEAR, BORDER,
                                            B.BEER_LAMBERT)[0]
    eg = ds * (g_cat[..., :3] @ rgb + g_cat[..., 3])
    assert rel(g, eg) < 1e-4
