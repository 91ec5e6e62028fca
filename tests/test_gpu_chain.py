"""Checkpointed multi-step transport + periodic render (ngt.sequence):
the recomputation-checkpointed gradient equals the stored-activation
gradient, uses less device memory, and matches a directional finite
difference of the chain loss."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GRID = (20, 24, 16)   # D, H, W
IMG = (18, 24)
DEPTH = 20


@pytest.fixture(scope="module")
def chain():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from ngt import synth, sequence, _lib
    _lib.load()
    grid, cams = synth.ring_setup(GRID, IMG, DEPTH, 3)
    spec = sequence.RenderSpec.from_cameras(grid, cams, light_rgb=(0.9, 0.8, 0.7))
    rho = torch.from_numpy(synth.blobs(GRID, 6, seed=11, sigma_range=(2.0, 4.0))[None]).cuda()
    vel = torch.from_numpy(synth.curl_noise_centered(GRID, seed=12, vmax=1.5, sigma=3.0)[None]).cuda()
    rho_t = torch.from_numpy(synth.blobs(GRID, 6, seed=13, sigma_range=(2.0, 4.0))[None]).cuda()
    return dict(seq=sequence, spec=spec, rho=rho, vel=vel, rho_t=rho_t)


def rel(a, b):
    return float((a - b).abs().max() / b.abs().max().clamp_min(1e-30))


def _grads(c, targets, checkpoint, clamp="MC_SMOOTH", render_every=3):
    rho = c["rho"].clone().requires_grad_()
    vel = c["vel"].clone().requires_grad_()
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats()
    base = torch.cuda.memory_allocated()
    loss = c["seq"].chain_loss(rho, vel, c["spec"], targets, render_every=render_every, clamp=clamp,
                               checkpoint=checkpoint)
    loss.backward()
    torch.cuda.synchronize()
    peak = torch.cuda.max_memory_allocated() - base
    return loss.detach(), rho.grad, vel.grad, peak


def test_checkpointed_gradient_equals_stored(chain):
    seq = chain["seq"]
    targets = seq.render_targets(chain["rho_t"], chain["vel"], chain["spec"], 4, render_every=3)
    l0, gr0, gv0, peak0 = _grads(chain, targets, checkpoint=False)
    l1, gr1, gv1, peak1 = _grads(chain, targets, checkpoint=True)
    assert torch.equal(l0, l1)
    assert rel(gr1, gr0) < 1e-5
    assert rel(gv1, gv0) < 1e-5
    assert float(gr0.abs().max()) > 0 and float(gv0.abs().max()) > 0
    assert peak1 < peak0, (peak1, peak0)


def test_chain_directional_derivative(chain):
    """<grad L, d> against a central difference (fp64 loss accumulation) of L
    along the smooth direction d = (rho0, vel0) / 10 (order-2 advection without
    the revert clamp, so L is smooth)."""
    seq = chain["seq"]
    targets = seq.render_targets(chain["rho_t"], chain["vel"], chain["spec"], 2, render_every=2, clamp="NONE")
    _, gr, gv, _ = _grads(chain, targets, checkpoint=True, clamp="NONE", render_every=2)
    dr = 0.1 * chain["rho"]
    dv = 0.1 * chain["vel"]
    an = float((gr.double() * dr.double()).sum() + (gv.double() * dv.double()).sum())

    def L(eps):
        with torch.no_grad():
            rho, vel = chain["rho"] + eps * dr, chain["vel"] + eps * dv
            total = 0.0
            for target in targets:
                for _ in range(2):
                    rho, vel = seq.transport_step(rho, vel, clamp="NONE")
                img = chain["spec"].render(rho)
                total += float(((img[..., :3].double() - target.double()) ** 2).sum())
            return total
    h = 1e-2
    fd = (L(h) - L(-h)) / (2 * h)
    # central differences of an fp32 chain whose trilinear weights have kinks
    # at cell faces: a few percent
    assert abs(an - fd) <= 3e-2 * abs(fd), (an, fd)


def test_checkpointed_memory_grows_by_the_segment_states(chain):
    """With checkpointing, each further segment keeps only its entering
    (density, velocity): no brick copy, recorded gradients or images of the
    segment's render stay alive until the backward."""
    seq = chain["seq"]
    targets = seq.render_targets(chain["rho_t"], chain["vel"], chain["spec"], 6, render_every=3)
    p3 = _grads(chain, targets[:3], checkpoint=True)[3]
    p6 = _grads(chain, targets, checkpoint=True)[3]
    state = (chain["rho"].numel() + chain["vel"].numel()) * 4
    per_segment = (p6 - p3) / 3.0
    print("checkpointed peak: 3 segments %d B, 6 segments %d B, per segment %.0f B, state %d B"
          % (p3, p6, per_segment, state), flush=True)
    assert per_segment <= 1.25 * state + 4096, (per_segment, state)


def _oracle_chain(rho0, vel0, targets, spec_args, light, ds, render_every, steps):
    """The chain's loss and gradients composed from the oracle's pinned
    pieces: MacCormack + MC_SMOOTH warps of the density and of the velocity
    by itself, the raymarch of [rho * light, rho] * ds after every
    `render_every` steps, L2 against the targets; reverse accumulation."""
    from oracle import advect as A, raymarch as RM, BORDER, LINEAR
    from oracle import blend as B
    M, V, P, F, out_dhw = spec_args
    rho, vel = [rho0], [vel0]
    for _ in range(steps):
        rho.append(A.warp_forward(rho[-1], vel[-1], 1.0, 2, "MC_SMOOTH", BORDER))
        vel.append(A.warp_forward(vel[-1], vel[-1], 1.0, 2, "MC_SMOOTH", BORDER))
    L = np.asarray(light, np.float64)
    loss = 0.0
    g_rho = np.zeros_like(rho[-1])
    g_vel = np.zeros_like(vel[-1])
    for k in range(steps, 0, -1):
        if k % render_every == 0:
            t = targets[k // render_every - 1]
            cat = np.concatenate([rho[k] * L[0], rho[k] * L[1], rho[k] * L[2], rho[k]], -1)[None] * ds
            img = RM.raymarch_grid_transform(cat, M, V, P, F, out_dhw, LINEAR, BORDER, B.BEER_LAMBERT)
            loss += float(np.sum((img[0, ..., :3] - t) ** 2))
            g_img = np.zeros_like(img)
            g_img[..., :3] = 2.0 * (img[..., :3] - t[None])
            gc = RM.raymarch_grid_transform_grad(cat, img, g_img, M, V, P, F, out_dhw, LINEAR, BORDER,
                                                 B.BEER_LAMBERT)[0]
            g_rho = g_rho + ds * (gc[..., :3] @ L + gc[..., 3])[..., None]
        gr_prev, gv_from_rho = A.warp_backward(rho[k - 1], vel[k - 1], g_rho, 1.0, 2, "MC_SMOOTH", BORDER)
        g_in, g_v = A.warp_backward(vel[k - 1], vel[k - 1], g_vel, 1.0, 2, "MC_SMOOTH", BORDER)
        g_rho, g_vel = gr_prev, g_in + g_v + gv_from_rho
    return loss, g_rho, g_vel


def test_chain_three_steps_against_oracle():
    """3 MacCormack steps of density and velocity at 16^3, a 2-view render
    after every step: loss (1e-5) and gradients w.r.t. the initial density
    and velocity (1e-4) against the oracle composition."""
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from ngt import synth, sequence, _lib
    _lib.load()
    grid = (16, 16, 16)
    grid_t, cams = synth.ring_setup(grid, (12, 14), 16, 2)
    spec = sequence.RenderSpec.from_cameras(grid_t, cams, light_rgb=(0.9, 0.8, 0.7))
    rho0 = synth.blobs(grid, 4, seed=31, sigma_range=(2.0, 3.0))
    vel0 = synth.curl_noise_centered(grid, seed=32, vmax=1.5, sigma=3.0)
    rng = np.random.default_rng(33)
    targets = [rng.uniform(0.0, 0.2, (2, 12, 14, 3)).astype(np.float32) for _ in range(3)]
    r = torch.from_numpy(rho0[None]).cuda().requires_grad_()
    v = torch.from_numpy(vel0[None]).cuda().requires_grad_()
    loss = sequence.chain_loss(r, v, spec, [torch.from_numpy(t[None]).cuda() for t in targets], render_every=1,
                               checkpoint=True)
    loss.backward()
    f64 = lambda a: np.asarray(a, np.float64)
    e_loss, e_gr, e_gv = _oracle_chain(f64(rho0), f64(vel0), [f64(t) for t in targets],
                                       (spec.M, spec.V, spec.P, spec.F, spec.output_dhw), spec.light_rgb,
                                       spec.depth_step, 1, 3)
    assert abs(float(loss) - e_loss) <= 1e-5 * e_loss
    assert rel(r.grad[0].cpu().double(), torch.from_numpy(e_gr)) < 1e-4
    assert rel(v.grad[0].cpu().double(), torch.from_numpy(e_gv)) < 1e-4


def test_chain_30_steps_96x160x96():
    """The long-horizon workload: 30 MacCormack steps at 96x160x96 (DHW),
    5-view 240x135x96 render after every 5th step.  The checkpointed
    gradient equals the stored-activation one; both peaks stay inside 8 GB;
    a directional central difference of the loss (order 2 without the revert
    clamp, so the loss is smooth) matches <grad, d>."""
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from ngt import synth, sequence, _lib
    _lib.load()
    grid, img, depth = (96, 160, 96), (135, 240), 96
    grid_t, cams = synth.ring_setup(grid, img, depth, 5)
    spec = sequence.RenderSpec.from_cameras(grid_t, cams)
    c = dict(seq=sequence, spec=spec, rho=torch.from_numpy(synth.blobs(grid, 16, seed=2)[None]).cuda(),
             vel=torch.from_numpy(synth.curl_noise_centered(grid, seed=4, vmax=2.0)[None]).cuda())
    rho_t = torch.from_numpy(synth.blobs(grid, 16, seed=3)[None]).cuda()
    targets = sequence.render_targets(rho_t, c["vel"], spec, 6, render_every=5)
    print("case: chain 30 steps 96x160x96, MC_SMOOTH", flush=True)
    l0, gr0, gv0, peak0 = _grads(c, targets, checkpoint=False, render_every=5)
    l1, gr1, gv1, peak1 = _grads(c, targets, checkpoint=True, render_every=5)
    print("peak above inputs: stored %.1f MB, checkpointed %.1f MB" % (peak0 / 1e6, peak1 / 1e6), flush=True)
    assert torch.equal(l0, l1)
    assert rel(gr1, gr0) < 1e-5 and rel(gv1, gv0) < 1e-5
    assert peak1 < peak0 < 8e9
    print("case: chain 30 steps 96x160x96, directional difference", flush=True)
    targets = sequence.render_targets(rho_t, c["vel"], spec, 6, render_every=5, clamp="NONE")
    _, gr, gv, _ = _grads(c, targets, checkpoint=True, clamp="NONE", render_every=5)
    dr, dv = 0.1 * c["rho"], 0.1 * c["vel"]
    an = float((gr.double() * dr.double()).sum() + (gv.double() * dv.double()).sum())

    def L(eps):
        with torch.no_grad():
            rho, vel = c["rho"] + eps * dr, c["vel"] + eps * dv
            total = 0.0
            for target in targets:
                for _ in range(5):
                    rho, vel = sequence.transport_step(rho, vel, clamp="NONE")
                out = spec.render(rho)
                total += float(((out[..., :3].double() - target.double()) ** 2).sum())
            return total
    h = 1e-2
    fd = (L(h) - L(-h)) / (2 * h)
    print("directional derivative %.6e, central difference %.6e" % (an, fd), flush=True)
    assert abs(an - fd) <= 3e-2 * abs(fd), (an, fd)
