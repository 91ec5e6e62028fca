"""GPU parity of the MacCormack advection and its brick-tiled adjoint on grids
that span several adjoint bricks per axis, against the FP64 oracle
(oracle/advect.py: VelocityGrid.warp semantics, data_structures.py:1690-1737).

The exact backward of the HIP library works on bricks of 32 x 8 x 16 cells
(x, y, z) with LDS tiles grown by a 2-cell halo; tile partials of
neighbouring bricks are gathered back per cell, and every work-item walks a
z column holding the shared corner plane of consecutive cells.  The grids
here have at least two bricks along every axis, partial bricks at the high
ends, and so cross every kind of brick boundary (x, y, z and their edges and
corners) in the tile gather and in the held-plane walks.

Also here: the 64^3 advection workload (8 seeded Gaussian blobs, centered
curl-noise velocity with max |v| = 2 cells, dt = 1, BORDER), SL and MC +
MC_SMOOTH, forward and adjoint; and the MC_SMOOTH revert decisions compared
cell by cell with the oracle's.

Tolerances: forward 1e-5 and gradients 1e-4, normwise relative
(max |gpu - oracle| / max |oracle|); revert masks exact outside an fp32
rounding margin.  Gradients are compared everywhere except at the
non-differentiable points that fp32 and fp64 may resolve differently
(tests/advect_masks.py: velocity components whose fp32 sample position
rounds onto an integer, and the influence of MC_SMOOTH ties whose device
choice is not visible in its outputs); each test prints how many cells that
leaves out.
"""
import numpy as np
import pytest
import torch

from oracle import BORDER, CLAMP, WRAP
from oracle import advect as A

import advect_masks as AM

pytestmark = pytest.mark.gpu

FWD_TOL = 1e-5
GRAD_TOL = 1e-4
BNAME = {BORDER: "BORDER", CLAMP: "CLAMP", WRAP: "WRAP"}

# (D, H, W): >= 2 bricks along each axis, ragged last bricks
SHAPES = [(40, 20, 70), (34, 17, 37), (33, 9, 12)]


@pytest.fixture(scope="module")
def ngt_ops():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from ngt import ops, _lib
    _lib.load()
    return ops


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float32))).to("cuda:0")


def host(t):
    return t.detach().cpu().double().numpy()


def rel(got, exp):
    scale = max(float(np.max(np.abs(exp))), 1e-30)
    return float(np.max(np.abs(got - exp))) / scale


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def smooth_field(rng, shape, amp):
    from scipy.ndimage import gaussian_filter
    f = gaussian_filter(rng.standard_normal(shape), 2.0)
    return amp * f / np.max(np.abs(f))


def _case_inputs(rng, shape, C, staggered, vmax):
    D, H, W = shape
    # smooth, never flat (no clipping): branch ties stay isolated
    rho = f32(np.stack([0.4 + 0.6 * smooth_field(rng, shape, 1.0) for _ in range(C)], -1))
    if staggered:
        faces = [f32(smooth_field(rng, (D, H, W + 1), vmax))[..., None],
                 f32(smooth_field(rng, (D, H + 1, W), vmax))[..., None],
                 f32(smooth_field(rng, (D + 1, H, W), vmax))[..., None]]
        vc = A.centered_from_staggered(*faces)
    else:
        faces = None
        vc = f32(np.stack([smooth_field(rng, shape, vmax) for _ in range(3)], -1))
    return rho, faces, vc


def _vel_leaves(faces, vc):
    if faces is not None:
        return tuple(dev(f[None]).requires_grad_() for f in faces)
    return dev(vc[None]).requires_grad_()


def _gpu_parts(ngt_ops, rho_np, vel_np, order, clamp, bound, out):
    """The device's SL and unclamped MacCormack results for the same inputs
    (its branch choices are read off them), with its tested result."""
    if clamp == "NONE":
        return None
    args = dict(timestep=-1.0, boundary=BNAME[bound], separate_velocity_batch=False)
    with torch.no_grad():
        sl = ngt_ops.advect_grid_semi_langrange(rho_np, vel_np, order=1, clamp_extrema="NONE", **args)
        corr = ngt_ops.advect_grid_semi_langrange(rho_np, vel_np, order=2, clamp_extrema="NONE", **args)
    return {"sl": host(sl)[0, 0], "corr": host(corr)[0, 0], "out": host(out)[0, 0]}


def _check_grads(rt, vel, faces, rho, vc, g, order, clamp, bound, gpu, label=""):
    eg_rho, eg_v, rho_keep, vel_keep, info = AM.oracle_adjoint(rho, vc, g, order, clamp, bound, gpu)
    print("%s left out: %s" % (label, info), flush=True)
    assert info["rho_left_out"] <= 0.005 * rho_keep.size and info["vel_left_out"] <= 0.005 * vel_keep.size
    assert AM.rel_on(host(rt.grad)[0], eg_rho, rho_keep[..., None]) < GRAD_TOL
    if faces is not None:
        for t, e, m in zip(vel, A.centered_from_staggered_adjoint(eg_v), AM.faces_of_cells(~vel_keep)):
            assert AM.rel_on(host(t.grad)[0], e, ~m) < GRAD_TOL
    else:
        assert AM.rel_on(host(vel.grad)[0], eg_v, vel_keep) < GRAD_TOL


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("order,clamp", [(1, "NONE"), (2, "NONE"), (2, "MC_SMOOTH"), (1, "NEGATIVE"),
                                         (2, "NEGATIVE")])
@pytest.mark.parametrize("bound", [BORDER, CLAMP, WRAP])
@pytest.mark.parametrize("staggered", [False, True])
def test_advect_across_bricks(ngt_ops, shape, order, clamp, bound, staggered):
    """Forward and adjoint (rho and every velocity component) across brick
    boundaries; 1 or 2 channels (alternating over the cases)."""
    key = SHAPES.index(shape) * 97 + order * 13 + len(clamp) * 5 + bound * 3 + staggered
    C = 1 + key % 2
    rng = np.random.default_rng(1000 + key)
    rho, faces, vc = _case_inputs(rng, shape, C, staggered, vmax=1.9)
    print("case shape=%s C=%d order=%d clamp=%s bound=%s staggered=%s" % (shape, C, order, clamp, BNAME[bound],
                                                                          staggered), flush=True)
    exp = A.warp_forward(rho, vc, 1.0, order, clamp, bound)
    rt = dev(rho[None]).requires_grad_()
    vel = _vel_leaves(faces, vc)
    out = ngt_ops.advect_grid_semi_langrange(rt, vel, timestep=-1.0, order=order, clamp_extrema=clamp,
                                             boundary=BNAME[bound], separate_velocity_batch=False)
    assert rel(host(out)[0, 0], exp) < FWD_TOL
    g = f32(rng.standard_normal(exp.shape))
    out.backward(dev(g[None, None]))
    vel_np = tuple(dev(f[None]) for f in faces) if staggered else dev(vc[None])
    gpu = _gpu_parts(ngt_ops, dev(rho[None]), vel_np, order, clamp, bound, out)
    _check_grads(rt, vel, faces, rho, vc, g, order, clamp, bound, gpu)


@pytest.mark.parametrize("C,separate,vmax", [(1, True, 1.9), (2, True, 3.5), (2, False, 1.0)])
def test_advect_across_bricks_batches(ngt_ops, C, separate, vmax):
    """Several (data, velocity) pairs on a multi-brick MAC grid: separate
    velocity batches (N = 2 data x V = 2 velocities) where both data entries
    share each velocity's gradient, and displacements beyond the tile halo
    (vmax 3.5) whose far corners take the overflow path."""
    rng = np.random.default_rng(int(vmax * 100) + C)
    shape = (36, 18, 40)
    D, H, W = shape
    N, V = 2, 2
    rho = f32(np.stack([np.stack([np.clip(smooth_field(rng, shape, 1.0), 0, 1) for _ in range(C)], -1)
                        for _ in range(N)]))
    faces = [f32(np.stack([smooth_field(rng, (D, H, W + 1), vmax) for _ in range(V)]))[..., None],
             f32(np.stack([smooth_field(rng, (D, H + 1, W), vmax) for _ in range(V)]))[..., None],
             f32(np.stack([smooth_field(rng, (D + 1, H, W), vmax) for _ in range(V)]))[..., None]]
    vel = tuple(dev(c).requires_grad_() for c in faces)
    rt = dev(rho).requires_grad_()
    out = ngt_ops.advect_grid_semi_langrange(rt, vel, timestep=-1.0, order=2, clamp_extrema="MC_SMOOTH",
                                             boundary="BORDER", separate_velocity_batch=separate)
    pairs = [(n, v, v) for n in range(N) for v in range(V)] if separate else [(n, n, 0) for n in range(N)]
    g = f32(rng.standard_normal(out.shape))
    out.backward(dev(g))
    args = dict(timestep=-1.0, boundary="BORDER", separate_velocity_batch=separate)
    with torch.no_grad():
        vel_np = tuple(dev(c) for c in faces)
        sl = host(ngt_ops.advect_grid_semi_langrange(dev(rho), vel_np, order=1, clamp_extrema="NONE", **args))
        corr = host(ngt_ops.advect_grid_semi_langrange(dev(rho), vel_np, order=2, clamp_extrema="NONE", **args))
    eg_rho = np.zeros_like(rho)
    eg_v = [np.zeros_like(c) for c in faces]
    rho_keep = np.ones(rho.shape[:4], dtype=bool)
    face_keep = [np.ones(c.shape, dtype=bool) for c in faces]
    for n, v, slot in pairs:
        vc = A.centered_from_staggered(faces[0][v], faces[1][v], faces[2][v])
        exp = A.warp_forward(rho[n], vc, 1.0, 2, "MC_SMOOTH", BORDER)
        assert rel(host(out)[n, slot], exp) < FWD_TOL
        gpu = {"sl": sl[n, slot], "corr": corr[n, slot], "out": host(out)[n, slot]}
        gr, gv, rk, vk, info = AM.oracle_adjoint(rho[n], vc, g[n, slot], 2, "MC_SMOOTH", BORDER, gpu)
        print("pair (%d, %d) left out: %s" % (n, v, info), flush=True)
        eg_rho[n] += gr
        rho_keep[n] &= rk
        for k, (e, m) in enumerate(zip(A.centered_from_staggered_adjoint(gv), AM.faces_of_cells(~vk))):
            eg_v[k][v] += e
            face_keep[k][v] &= ~m
    assert AM.rel_on(host(rt.grad), eg_rho, rho_keep[..., None]) < GRAD_TOL
    for t, e, m in zip(vel, eg_v, face_keep):
        assert AM.rel_on(host(t.grad), e, m) < GRAD_TOL


def _check_revert_mask(ngt_ops, rho, vel, vc, bound):
    """The device's MC_SMOOTH decisions cell by cell against the oracle's.
    The device's decision is read off its outputs: its unclamped MacCormack
    value against the min / max of the sample's 8 corners (exact in either
    precision), and the MC_SMOOTH result must equal bit for bit the SL value
    where that decision reverts and the unclamped value where it keeps."""
    _, parts = A.warp_forward(rho, vc, 1.0, 2, "MC_SMOOTH", bound, return_parts=True)
    oob = parts["oob"]
    args = dict(timestep=-1.0, boundary=BNAME[bound], separate_velocity_batch=False)
    r = dev(rho[None])
    with torch.no_grad():
        sl = host(ngt_ops.advect_grid_semi_langrange(r, vel, order=1, clamp_extrema="NONE", **args))[0, 0]
        corr = host(ngt_ops.advect_grid_semi_langrange(r, vel, order=2, clamp_extrema="NONE", **args))[0, 0]
        mc = host(ngt_ops.advect_grid_semi_langrange(r, vel, order=2, clamp_extrema="MC_SMOOTH", **args))[0, 0]
    reverted = (corr < parts["mn"]) | (corr > parts["mx"])
    np.testing.assert_array_equal(mc, np.where(reverted, sl, corr))
    # decisions closer to a bound than fp32 rounding may legitimately differ
    ties = AM.branch_ties(rho, vc, parts, 2, "MC_SMOOTH", bound)
    n_ties = int(np.sum(ties))
    print("revert mask: %d cells, %d reverted (oracle), %d within the fp32 margin, %d shown by the output (SL != corr)"
          % (oob.size, int(np.sum(oob)), n_ties, int(np.sum(sl != corr))), flush=True)
    assert int(np.sum(oob & ~ties)) > 0
    np.testing.assert_array_equal(reverted[~ties], oob[~ties])
    return n_ties


def test_advect_64_workload(ngt_ops):
    """64^3: 8 Gaussian blobs (seed 1), centered curl-noise velocity (seed 5,
    max |v| = 2 cells), dt = 1, trilinear, BORDER; SL, then MacCormack with
    the MC_SMOOTH clamp; forward, adjoint and the exact revert mask."""
    from ngt import synth
    shape = (64, 64, 64)
    rho = f32(synth.blobs(shape, 8, seed=1))
    vc = f32(synth.curl_noise_centered(shape, seed=5, vmax=2.0))
    rng = np.random.default_rng(64)
    g = f32(rng.standard_normal(rho.shape))
    for order, clamp in ((1, "NONE"), (2, "MC_SMOOTH")):
        print("case 64^3 order=%d clamp=%s" % (order, clamp), flush=True)
        exp = A.warp_forward(rho, vc, 1.0, order, clamp, BORDER)
        rt = dev(rho[None]).requires_grad_()
        vt = dev(vc[None]).requires_grad_()
        out = ngt_ops.advect_grid_semi_langrange(rt, vt, timestep=-1.0, order=order, clamp_extrema=clamp,
                                                 boundary="BORDER", separate_velocity_batch=False)
        assert rel(host(out)[0, 0], exp) < FWD_TOL
        # zero boundary: the oracle's exact zeros (cells whose sample leaves
        # the grid, or reads only zero cells) are zeros on the GPU too
        assert np.all(host(out)[0, 0][exp == 0] == 0)
        out.backward(dev(g[None, None]))
        gpu = _gpu_parts(ngt_ops, dev(rho[None]), dev(vc[None]), order, clamp, BORDER, out)
        _check_grads(rt, vt, None, rho, vc, g, order, clamp, BORDER, gpu, "64^3 order %d" % order)
    _check_revert_mask(ngt_ops, rho, dev(vc[None]), vc, BORDER)


@pytest.mark.parametrize("bound", [BORDER, CLAMP, WRAP])
@pytest.mark.parametrize("staggered", [False, True])
def test_revert_mask_across_bricks(ngt_ops, bound, staggered):
    """MC_SMOOTH revert decisions on a multi-brick grid equal the oracle's
    cell by cell (outside the fp32 margin), centered and MAC velocity."""
    rng = np.random.default_rng(2000 + 3 * bound + staggered)
    rho, faces, vc = _case_inputs(rng, SHAPES[0], 1, staggered, vmax=1.9)
    vel = tuple(dev(f[None]) for f in faces) if staggered else dev(vc[None])
    _check_revert_mask(ngt_ops, rho, vel, vc, bound)
