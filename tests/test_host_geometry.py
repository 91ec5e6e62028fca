"""The product's host-side matrices (ngt.transform / ngt.camera /
Renderer._transform_params) equal the oracle's restatement of the reference's
camera.py / transform.py, and the synthetic generators meet their specs."""
import numpy as np
import pytest

from oracle import geometry as G


def test_grid_transform_matches_oracle():
    from ngt.transform import GridTransform
    for kw in (dict(center=True, scale=[1 / 64.0] * 3), dict(translation=[0.1, -0.2, 0.3], rotation_deg=[10, 20, 30]),
               dict(center=True, normalize="MAX", scale=[1, 2, 3])):
        t = GridTransform([32, 48, 64], **kw)
        ref = G.grid_transform([32, 48, 64], translation_=kw.get("translation", (0, 0, 0)),
                               rotation_deg=kw.get("rotation_deg", (0, 0, 0)), scale_=kw.get("scale", (1, 1, 1)),
                               center=kw.get("center", False), normalize=kw.get("normalize", "NONE"))
        np.testing.assert_allclose(t.get_transform_matrix(), ref, atol=1e-12)


def test_ring_cameras_and_renderer_params_match_oracle():
    from ngt.synth import ring_setup
    from ngt.renderer import Renderer
    grid, cams = ring_setup((128, 128, 128), (144, 256), 128, 5)
    M, V, P, F = Renderer(fused=True)._transform_params([grid], cams)
    near, far = 3 - np.sqrt(3) / 2, 3 + np.sqrt(3) / 2
    ref = G.look_at_ring_cameras(5, 3.0, 40.0, 256 / 144, near, far)
    np.testing.assert_allclose(M[0], G.op_matrix(G.grid_transform((128,) * 3, center=True, scale_=[1 / 128.0] * 3)),
                               atol=1e-7)
    for i, (Vr, Pr, Fr) in enumerate(ref):
        np.testing.assert_allclose(V[i], G.op_matrix(Vr), atol=1e-6)
        np.testing.assert_allclose(P[i], G.op_matrix(Pr), rtol=1e-6, atol=1e-6)
        np.testing.assert_allclose(F[i], Fr, rtol=1e-6)
    assert cams[0].depth_step == pytest.approx((far - near) / 128)


def test_synthetic_generators():
    from ngt import synth
    rho = synth.blobs((32, 32, 32), 8, seed=1)
    assert rho.shape == (32, 32, 32, 1) and rho.min() >= 0 and rho.max() <= 1 and rho.max() > 0.2
    v = synth.curl_noise_centered((24, 24, 24), seed=2, vmax=2.0)
    assert np.max(np.linalg.norm(v, axis=-1)) == pytest.approx(2.0, rel=1e-5)
    vx, vy, vz = synth.curl_noise_staggered((16, 18, 20), seed=3, vmax=2.0)
    assert vx.shape == (16, 18, 21, 1) and vy.shape == (16, 19, 20, 1) and vz.shape == (17, 18, 20, 1)
    div = (vx[:, :, 1:] - vx[:, :, :-1]) + (vy[:, 1:] - vy[:, :-1]) + (vz[1:] - vz[:-1])
    assert np.max(np.abs(div)) < 1e-5 * np.max(np.abs(vx))
