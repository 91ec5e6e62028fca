"""The per-ray sample bound of the render / raymarch gather adjoint
(ngt_raymarch_step_cap, a host-only query of the C ABI) against the oracle's
FP64 ray setup of every pixel (raymarch_grid.cu.cc:173-200): no BORDER ray
has more samples than the bound, for ring cameras outside and inside the
grid, off-axis principal points, wide fields of view, rotated and
non-uniformly scaled grids, orthographic cameras and a projection whose
inverse mixes pixel coordinates into w (the exhaustive branch).  A ray with
more samples than the bound would have its recorded gradient cut short."""
import ctypes
import os

import numpy as np
import pytest

from oracle import BORDER
from oracle import geometry as G
from oracle import raymarch as RM


@pytest.fixture(scope="module")
def lib():
    from ngt import _lib
    if not os.path.exists(_lib.lib_path()):
        pytest.skip("library not built")
    return _lib.load()


def step_cap(lib, dims, Ms, Vs, Ps, Fs, out_dhw):
    from ngt import _lib
    shape = (ctypes.c_int64 * 5)(1, *dims, 1)
    # (pointer, count, keep-alive array) per argument; the arrays outlive the call
    m, v, p, fr = (_lib.host_floats(x, n) for x, n in ((Ms, 16), (Vs, 16), (Ps, 16), (Fs, 6)))
    odhw = (ctypes.c_int32 * 3)(*out_dhw)
    return lib.ngt_raymarch_step_cap(shape, m[0], v[0], p[0], fr[0], len(Vs), odhw, 1)


def max_steps(dims, M, V, P, F, out_dhw):
    most = 0
    for v in range(len(V)):
        cx = G.CameraXform.from_op_args(M[0], V[v], P[v], F[v])
        _, _, _, steps, hit = RM.ray_setup(cx, out_dhw, dims, BORDER)
        if np.any(hit):
            most = max(most, int(steps[hit].max()))
    return most


def op_args(M, cams):
    return ([G.op_matrix(M).astype(np.float32)], [G.op_matrix(c[0]).astype(np.float32) for c in cams],
            [G.op_matrix(c[1]).astype(np.float32) for c in cams], [np.asarray(c[2], np.float32) for c in cams])


def ring(views, dist, fov, aspect, near, far, height=0.0, start=0.0):
    return G.look_at_ring_cameras(views, dist, fov, aspect, near, far, height=height, start_deg=start)


def off_axis(dist, trbl, near, far, yaw=0.0, pitch=0.0):
    P = G.projection(near, far, trbl)
    F = G.frustum_params(near, far, trbl)
    C = G.rotation_xyz_deg([pitch, yaw, 0.0]) @ G.translation([0.0, 0.0, dist])
    return (np.linalg.inv(C), P, F)


CASES = {
    "ring_far": ((32, 24, 40), (48, 30, 50), lambda a: ring(5, 3.0, 40.0, a, 2.0, 4.0)),
    "ring_inside": ((20, 20, 20), (40, 24, 30), lambda a: ring(4, 1.1, 60.0, a, 0.2, 2.5, height=0.1, start=9.0)),
    "wide_fov": ((16, 28, 20), (24, 40, 40), lambda a: ring(3, 1.6, 120.0, a, 0.3, 3.0, height=-0.2)),
    "off_axis": ((24, 24, 24), (32, 36, 44), lambda a: [off_axis(2.5, [0.9, 0.2, 0.1, -0.6], 1.5, 3.5, yaw=17.0),
                                                      off_axis(2.0, [-0.1, 1.1, -0.9, 0.3], 1.0, 3.0, pitch=-12.0),
                                                      off_axis(2.2, [0.05, 0.6, -0.7, -0.1], 0.8, 3.2, 40.0, 25.0)]),
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("grid_kind", ["uniform", "rotated_anisotropic"])
def test_step_cap_bounds_every_ray(lib, name, grid_kind):
    dims, out_dhw, make = CASES[name]
    if grid_kind == "uniform":
        M = G.grid_transform(dims, center=True, scale_=[1.0 / max(dims)] * 3)
    else:
        M = G.grid_transform(dims, translation_=(0.03, -0.05, 0.02), rotation_deg=(11.0, -23.0, 7.0), center=True,
                             scale_=[1.3 / dims[2], 0.8 / dims[1], 1.0 / dims[0]])
    cams = make(out_dhw[2] / out_dhw[1])
    args = op_args(M, cams)
    cap = step_cap(lib, dims, *args, out_dhw)
    most = max_steps(dims, *args, out_dhw)
    print("%s/%s: longest ray %d samples, bound %d" % (name, grid_kind, most, cap))
    assert most > 0
    assert 0 < most <= cap
    assert cap % 4 == 0  # the record stores samples in quads of 4


def test_step_cap_orthographic_and_general_projection(lib):
    dims, out_dhw = (12, 14, 10), (20, 16, 18)
    M = G.grid_transform(dims, center=True, scale_=[1.0 / max(dims)] * 3)
    # orthographic: every pixel has the same step
    P = G.projection(1.0, 3.0, [0.6, 0.6, -0.6, -0.6], perspective=False)
    cam_o = (np.linalg.inv(G.rotation_xyz_deg([5.0, 30.0, 0.0]) @ G.translation([0.0, 0.0, 2.0])), P,
             G.frustum_params(1.0, 3.0, [0.6, 0.6, -0.6, -0.6]))
    # a projective map whose inverse has x / y terms in its last row: the
    # bound takes the exhaustive branch over all pixels
    V, Pp, F = ring(1, 2.5, 50.0, out_dhw[2] / out_dhw[1], 1.5, 3.5)[0]
    Pg = Pp.copy()
    Pg[3, 0], Pg[3, 1] = 0.07, -0.05
    assert abs(np.linalg.inv(Pg)[3, 0]) > 1e-3
    for cam in (cam_o, (V, Pg, F)):
        args = op_args(M, [cam])
        cap = step_cap(lib, dims, *args, out_dhw)
        most = max_steps(dims, *args, out_dhw)
        print("longest ray %d samples, bound %d" % (most, cap))
        assert 0 < most <= cap


def test_step_cap_of_the_benchmark_views(lib):
    """128^3, 5 ring views of 256x144 with 128 slices: the bound and how far
    the longest ray is below it (the record is sized by the bound, which
    takes the grid's full diagonal where the ring's rays cross at most a
    face diagonal)."""
    from ngt import synth
    from ngt.renderer import Renderer
    grid_t, cams = synth.ring_setup((128, 128, 128), (144, 256), 128, 5)
    M, V, P, F = Renderer()._transform_params([grid_t], cams)
    cap = step_cap(lib, (128, 128, 128), M, V, P, F, (128, 144, 256))
    most = max_steps((128, 128, 128), M, V, P, F, (128, 144, 256))
    print("benchmark views: longest ray %d samples, bound %d" % (most, cap))
    assert 0 < most <= cap <= 1.5 * most + 8


def test_record_size_of_the_benchmark_views(lib):
    """The fused render + L2 forward's record (ngt_render_density_l2_record_size)
    holds a 48-byte ray record per pixel, a 4-byte gradient bound per 8x8
    pixel tile and one float per recorded sample (kcap per ray, pixels in
    groups of 8), each region 256-byte aligned; the two-walk backward's workspace adds the haloed
    brick copy to it."""
    from ngt import _lib, synth
    from ngt.renderer import Renderer
    grid_t, cams = synth.ring_setup((128, 128, 128), (144, 256), 128, 5)
    M, V, P, F = Renderer()._transform_params([grid_t], cams)
    shape = (ctypes.c_int64 * 5)(1, 128, 128, 128, 1)
    m, v, p, fr = (_lib.host_floats(x, n) for x, n in ((M, 16), (V, 16), (P, 16), (F, 6)))
    odhw = (ctypes.c_int32 * 3)(128, 144, 256)
    record = lib.ngt_render_density_l2_record_size(shape, m[0], v[0], p[0], fr[0], 5, odhw)
    two_walk = lib.ngt_render_density_l2_bwd_workspace_size(shape, m[0], v[0], p[0], fr[0], 5, odhw, 1, 0)
    cap = step_cap(lib, (128, 128, 128), M, V, P, F, (128, 144, 256))
    up = lambda n: (n + 255) // 256 * 256
    npix = 144 * 256
    assert record == up(5 * npix * 48) + up(5 * (256 // 8) * (144 // 8) * 4) + up(5 * cap * ((npix + 7) // 8 * 8) * 4)
    assert two_walk == record + up(lib.ngt_bricked_grid_bytes(shape))
    print("record %.1f MB, two-walk workspace %.1f MB" % (record / 1e6, two_walk / 1e6))
