"""Full-size (128^3, 5 x 256x144x128) GPU checks through size-independent
properties: atomics-free gather adjoint == atomic scatter adjoint, adjoint
identities <A x, y> = <x, A^T y> of the linear ops, exact shifts for integer
velocities, identity for dt = 0, and FD spot checks of the full step
gradient."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GRID = (128, 128, 128)
IMG = (144, 256)
DEPTH = 128


@pytest.fixture(scope="module")
def setup():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from ngt import ops, synth, _lib
    from ngt.renderer import Renderer
    _lib.load()
    grid, cams = synth.ring_setup(GRID, IMG, DEPTH, 5)
    M, V, P, F = Renderer(fused=True)._transform_params([grid], cams)
    rho = torch.from_numpy(synth.blobs(GRID, 16, seed=2)[None]).cuda()
    vel = tuple(torch.from_numpy(c[None]).cuda() for c in synth.curl_noise_staggered(GRID, seed=4, vmax=2.0))
    return dict(ops=ops, M=M, V=V, P=P, F=F, ds=float(cams[0].depth_step), rho=rho, vel=vel)


def rel(a, b):
    return float((a - b).abs().max() / b.abs().max().clamp_min(1e-30))


def test_render_gather_adjoint_equals_atomic_scatter(setup):
    ops = setup["ops"]
    rho = setup["rho"].clone().requires_grad_()
    args = ((1.0, 0.9, 0.8), setup["ds"], setup["M"], setup["V"], setup["P"], setup["F"], (DEPTH,) + IMG)
    img = ops.render_density_fused(rho, *args)
    g = torch.randn_like(img)
    old = ops.use_gather_adjoint
    try:
        ops.use_gather_adjoint = True
        (ga,) = torch.autograd.grad(img, rho, g, retain_graph=True)
        ops.use_gather_adjoint = False
        (gs,) = torch.autograd.grad(img, rho, g)
    finally:
        ops.use_gather_adjoint = old
    assert rel(ga, gs) < 1e-4
    # linearity: <d img / d rho . delta, g> is the directional derivative; for
    # ADDITIVE blending the render is linear, so <R x, y> == <x, R^T y>
    x = setup["rho"]
    img_add = ops.render_density_fused(x.clone().requires_grad_(), args[0], args[1], *args[2:],
                                       blending_mode="ADDITIVE")
    y = torch.randn_like(img_add)
    xr = x.clone().requires_grad_()
    out = ops.render_density_fused(xr, args[0], args[1], *args[2:], blending_mode="ADDITIVE")
    (gt,) = torch.autograd.grad(out, xr, y)
    lhs = float((out.double() * y.double()).sum())
    rhs = float((x.double() * gt.double()).sum())
    # the fp32 rounding of the image and of the gradient sums (taken in no
    # fixed order) scales with the magnitudes of the terms, not with the
    # cancelling inner product itself
    l1 = float((out.double() * y.double()).abs().sum())
    assert abs(lhs - rhs) <= 1e-7 * l1, (lhs, rhs, l1)


def test_raymarch_gather_equals_scatter_multichannel(setup):
    ops = setup["ops"]
    torch.manual_seed(0)
    x = (torch.rand((1, 64, 64, 64, 4), device="cuda") * 0.05).requires_grad_()
    from ngt import synth
    from ngt.renderer import Renderer
    grid, cams = synth.ring_setup((64, 64, 64), (72, 128), 64, 3)
    M, V, P, F = Renderer(fused=True)._transform_params([grid], cams)
    out = ops.raymarch_grid_transform(x, M, V, P, F, (64, 72, 128))
    g = torch.randn_like(out)
    old = ops.use_gather_adjoint
    try:
        ops.use_gather_adjoint = True
        (ga,) = torch.autograd.grad(out, x, g, retain_graph=True)
        ops.use_gather_adjoint = False
        (gs,) = torch.autograd.grad(out, x, g)
    finally:
        ops.use_gather_adjoint = old
    assert rel(ga, gs) < 1e-4


def test_advect_integer_shift_and_identity(setup):
    ops = setup["ops"]
    rho = setup["rho"]
    v = torch.zeros(rho.shape[:4] + (3,), device="cuda")
    v[..., 0] = 1.0
    v[..., 2] = -2.0
    out = ops.advect_grid_semi_langrange(rho, v, -1.0, 2, "MC_SMOOTH", "WRAP", False)[:, 0]
    exp = torch.roll(torch.roll(rho, 1, dims=3), -2, dims=1)
    assert torch.equal(out, exp)
    out0 = ops.advect_grid_semi_langrange(rho, setup["vel"], 0.0, 2, "MC_SMOOTH", "BORDER", False)[:, 0]
    assert torch.equal(out0, rho)


def test_advect_adjoint_identity_full_size(setup):
    """Order-1 advection is linear in rho: <A rho, y> == <rho, A^T y>."""
    ops = setup["ops"]
    rho = setup["rho"].clone().requires_grad_()
    out = ops.advect_grid_semi_langrange(rho, setup["vel"], -1.0, 1, "NONE", "BORDER", False)
    y = torch.randn_like(out)
    (g,) = torch.autograd.grad(out, rho, y)
    lhs = float((out.double() * y.double()).sum())
    rhs = float((setup["rho"].double() * g.double()).sum())
    # fp32 rounding (atomics in no fixed order) scales with the magnitudes of
    # the terms, not with the cancelling inner product
    l1 = float((out.double() * y.double()).abs().sum())
    assert abs(lhs - rhs) <= 1e-7 * l1, (lhs, rhs, l1)


def test_full_step_gradient_fd(setup):
    """Central finite differences (fp64 accumulation of fp32 evaluations) of
    the semi-Lagrangian step + render loss w.r.t. a few density voxels and x
    faces inside the blobs (order 1: no clamp switching inside the FD
    stencil)."""
    ops = setup["ops"]
    args = ((1.0, 1.0, 1.0), setup["ds"], setup["M"], setup["V"], setup["P"], setup["F"], (DEPTH,) + IMG)
    target = ops.render_density_fused(setup["rho"] * 0.9, *args)[..., :3].detach()

    def loss(rho, vel):
        adv = ops.advect_grid_semi_langrange(rho, vel, -1.0, 1, "NONE", "BORDER", False)[:, 0]
        img = ops.render_density_fused(adv, *args)
        return ((img[..., :3] - target) ** 2).double().sum()

    rho = setup["rho"].clone().requires_grad_()
    vel = tuple(v.clone().requires_grad_() for v in setup["vel"])
    L = loss(rho, vel)
    grads = torch.autograd.grad(L, (rho,) + vel)
    d = setup["rho"][0, ..., 0]
    idx = torch.nonzero(d > 0.3)[::997][:6].tolist()
    h = 1e-3
    for (z, y, x) in idx:
        for t, gt, name in ((rho, grads[0], "rho"), (vel[0], grads[1], "vx")):
            base = t.detach()
            p = base.clone(); p[0, z, y, x, 0] += h
            m = base.clone(); m[0, z, y, x, 0] -= h
            if name == "rho":
                fd = (loss(p, setup["vel"]) - loss(m, setup["vel"])) / (2 * h)
            else:
                fd = (loss(setup["rho"], (p,) + setup["vel"][1:]) - loss(setup["rho"], (m,) + setup["vel"][1:])) / (2 * h)
            an = float(gt[0, z, y, x, 0])
            assert abs(an - float(fd)) <= 3e-2 * max(abs(float(fd)), 1e-4 * float(L)), (name, z, y, x, an, float(fd))
