"""The fused step (MacCormack + MC_SMOOTH advection of a staggered-velocity
density, fused render + L2 loss, backward) captured as one HIP graph: every
replay, with eager steps and synchronisations of the program in between,
gives the eager step's loss and gradients.  The gradients accumulate with
atomics in no fixed order, so they agree to rounding, not bit for bit."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GRID = (32, 40, 48)
IMG = (36, 64)
DEPTH = 32


@pytest.fixture(scope="module")
def problem():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from ngt import ops, synth, _lib
    from ngt.renderer import Renderer
    _lib.load()
    grid, cams = synth.ring_setup(GRID, IMG, DEPTH, 3)
    M, V, P, F = Renderer(fused=True)._transform_params([grid], cams)
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    rho = t(synth.blobs(GRID, 6, seed=11)[None])
    vel = tuple(t(c[None]) for c in synth.curl_noise_staggered(GRID, seed=12, vmax=2.0))
    light = (1.0, 0.9, 0.8)
    ds = float(cams[0].depth_step)
    with torch.no_grad():
        target = ops.render_density_fused(t(synth.blobs(GRID, 6, seed=13)[None]), light, ds, M, V, P, F,
                                          (DEPTH,) + IMG)[..., :3].contiguous()
    return dict(ops=ops, rho=rho, vel=vel, light=light, ds=ds, M=M, V=V, P=P, F=F, target=target)


def step(p, leaves):
    rho, vx, vy, vz = leaves
    adv = p["ops"].advect_grid_semi_langrange(rho, (vx, vy, vz), timestep=-1.0, order=2, clamp_extrema="MC_SMOOTH",
                                              boundary="BORDER", separate_velocity_batch=False).squeeze(1)
    loss, _ = p["ops"].render_density_fused_l2(adv, p["target"], p["light"], p["ds"], p["M"], p["V"], p["P"],
                                               p["F"], (DEPTH,) + IMG)
    loss.backward()
    return loss


def leaves_of(p):
    return [p["rho"].clone().requires_grad_()] + [v.clone().requires_grad_() for v in p["vel"]]


def rel(a, b):
    return float((a - b).abs().max() / b.abs().max().clamp_min(1e-30))


def test_fused_step_graph_replays_match_eager(problem):
    p = problem
    eager = leaves_of(p)
    eager_loss = float(step(p, eager))
    torch.cuda.synchronize()

    static = leaves_of(p)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(2):
            for l in static:
                l.grad = None
            step(p, static)
    torch.cuda.current_stream().wait_stream(s)
    for l in static:
        l.grad = None
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        static_loss = step(p, static)

    for replay in range(4):
        g.replay()
        torch.cuda.synchronize()
        assert abs(float(static_loss) - eager_loss) <= 1e-5 * abs(eager_loss), replay
        for a, b in zip(static, eager):
            assert rel(a.grad, b.grad) < 1e-5, replay
        # eager work of the program between replays
        other = leaves_of(p)
        step(p, other)
        torch.cuda.synchronize()
        for a, b in zip(other, eager):
            assert rel(a.grad, b.grad) < 1e-5, replay
