"""Uneven batches (ngt.batch): every instance of a batched launch equals the
per-instance op on the same inputs; the batched step replays correctly from
a captured HIP graph."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

# (grid D,H,W), views, image (H, W), depth slices
SPECS = [((12, 14, 16), 1, (10, 14), 12), ((20, 16, 18), 3, (16, 24), 20), ((9, 9, 9), 2, (8, 8), 9),
         ((24, 20, 33), 4, (18, 30), 24), ((16, 16, 16), 8, (12, 16), 16)]


@pytest.fixture(scope="module")
def batch():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from ngt import synth, batch as B, ops, _lib
    from ngt.renderer import Renderer
    _lib.load()
    dens, vels, cams, tgts = [], [], [], []
    for i, (dhw, nv, hw, depth) in enumerate(SPECS):
        dens.append(torch.from_numpy(synth.blobs(dhw, 4, seed=20 + i, sigma_range=(1.5, 3.0))).cuda())
        vels.append(tuple(torch.from_numpy(c).cuda() for c in
                          synth.curl_noise_staggered(dhw, seed=40 + i, vmax=1.5, sigma=2.5)))
        grid, cs = synth.ring_setup(dhw, hw, depth, nv)
        grid.set_translation([0.0123, -0.0217, 0.0311])
        M, V, P, F = Renderer()._transform_params([grid], cs)
        cams.append((M, V, P, F, cs[0].depth_step, (depth,) + hw))
        tgts.append(torch.rand((nv,) + hw + (3,), device="cuda", generator=torch.Generator("cuda").manual_seed(i)))
    return dict(B=B, ops=ops, dens=dens, vels=vels, cams=cams, tgts=tgts)


def rel(a, b):
    return float((a - b).abs().max() / b.abs().max().clamp_min(1e-30))


def test_advect_batch_matches_per_instance(batch):
    B, ops = batch["B"], batch["ops"]
    ab = B.AdvectBatch(batch["dens"], batch["vels"], -1.0, 2, "MC_SMOOTH", "BORDER", tuner=B.Autotuner())
    outs = ab.forward()
    g = torch.Generator("cuda").manual_seed(3)
    for og in ab.out_grads:
        og.copy_(torch.randn(og.shape, device="cuda", generator=g))
    gin, gvel = ab.backward()
    torch.cuda.synchronize()
    for i in range(len(SPECS)):
        d = batch["dens"][i][None].clone().requires_grad_()
        v = tuple(c[None].clone().requires_grad_() for c in batch["vels"][i])
        ref = ops.advect_grid_semi_langrange(d, v, -1.0, 2, "MC_SMOOTH", "BORDER", False)[0, 0]
        assert torch.equal(outs[i], ref), i
        ref.backward(ab.out_grads[i])
        assert rel(gin[i], d.grad[0]) < 1e-6, i
        for a, b in zip(gvel[i], v):
            assert rel(a, b.grad[0]) < 1e-6, i


@pytest.mark.parametrize("brick", [0, 1, 2])
def test_render_batch_matches_per_instance(batch, brick):
    """Every gather brick (8x8x8, 16x8x8, 16x16x8), each instance in a
    record chunk of its own (record_budget 1 byte)."""
    B, ops = batch["B"], batch["ops"]
    rb = B.RenderBatch(batch["dens"], batch["cams"], (0.9, 0.8, 0.7), tuner=B.Autotuner(), record_budget=1)
    assert rb.num_chunks == len(SPECS)
    imgs = rb.forward()
    g = torch.Generator("cuda").manual_seed(4)
    for ig in rb.image_grads:
        ig.copy_(torch.randn(ig.shape, device="cuda", generator=g))
    dg = rb.backward(shape=(0, brick))
    torch.cuda.synchronize()
    for i in range(len(SPECS)):
        M, V, P, F, ds, out_dhw = batch["cams"][i]
        d = batch["dens"][i][None].clone().requires_grad_()
        ref = ops.render_density_fused(d, (0.9, 0.8, 0.7), ds, M, V, P, F, out_dhw)[0]
        assert torch.equal(imgs[i], ref), i
        ref.backward(rb.image_grads[i])
        assert rel(dg[i], d.grad[0]) < 1e-5, i


@pytest.mark.parametrize("blend", ["BEER_LAMBERT", "ALPHA"])
def test_render_batch_l2_step_matches_fused_l2(batch, blend):
    """The batched L2 step (recording walk + gather per chunk) against the
    per-instance fused render + L2 op; one chunk and one chunk per instance
    give the same bits."""
    B, ops = batch["B"], batch["ops"]
    light = (0.9, 0.8, 0.7)
    one = B.RenderBatch(batch["dens"], batch["cams"], light, blend, tuner=None, targets=batch["tgts"])
    imgs, losses, dg = one.l2_step(shape=(0, 2))
    torch.cuda.synchronize()
    assert one.num_chunks == 1
    got = ([t.clone() for t in imgs], losses.clone(), [t.clone() for t in dg])
    many = B.RenderBatch(batch["dens"], batch["cams"], light, blend, tuner=None, targets=batch["tgts"],
                         record_budget=1)
    imgs2, losses2, dg2 = many.l2_step(shape=(4, 2))
    torch.cuda.synchronize()
    assert many.num_chunks == len(SPECS)
    assert torch.equal(losses2, got[1])
    for i in range(len(SPECS)):
        assert torch.equal(imgs2[i], got[0][i]) and torch.equal(dg2[i], got[2][i]), i
        M, V, P, F, ds, out_dhw = batch["cams"][i]
        d = batch["dens"][i][None].clone().requires_grad_()
        loss, img = ops.render_density_fused_l2(d, batch["tgts"][i][None], light, ds, M, V, P, F, out_dhw,
                                                blending_mode=blend)
        loss.backward()
        assert torch.equal(got[0][i], img.detach()[0]), i
        assert abs(float(got[1][i]) - float(loss)) <= 1e-6 * float(loss), i
        assert rel(got[2][i], d.grad[0]) < 1e-5, i


def _per_instance_step(batch):
    ops = batch["ops"]
    grads = []
    for i in range(len(SPECS)):
        M, V, P, F, ds, out_dhw = batch["cams"][i]
        d = batch["dens"][i][None].clone().requires_grad_()
        v = tuple(c[None].clone().requires_grad_() for c in batch["vels"][i])
        adv = ops.advect_grid_semi_langrange(d, v, -1.0, 2, "MC_SMOOTH", "BORDER", False)[:, 0]
        img = ops.render_density_fused(adv, (1.0, 1.0, 1.0), ds, M, V, P, F, out_dhw)[0]
        ((img[..., :3] - batch["tgts"][i]) ** 2).sum().backward()
        grads.append((d.grad[0], tuple(c.grad[0] for c in v)))
    return grads


def test_transport_render_step_and_graph_replay(batch):
    B = batch["B"]
    step = B.TransportRenderBatch(batch["dens"], batch["vels"], batch["cams"], batch["tgts"], tuner=B.Autotuner())
    gin, gvel = step.step()
    torch.cuda.synchronize()
    eager = [(a.clone(), tuple(c.clone() for c in b)) for a, b in zip(gin, gvel)]
    ref = _per_instance_step(batch)
    for (a, va), (b, vb) in zip(eager, ref):
        assert rel(a, b) < 1e-5
        for x, y in zip(va, vb):
            assert rel(x, y) < 1e-5
    # capture the whole batched step and replay it, with eager work between replays
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        step.step()
    torch.cuda.current_stream().wait_stream(s)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step.step()
    for rep in range(3):
        for t in step.adv.in_grads:
            t.fill_(float("nan"))
        graph.replay()
        torch.cuda.synchronize()
        for (a, va), b, vb in zip(eager, step.adv.in_grads, step.adv.vel_grads):
            assert rel(b, a) < 1e-5, rep
            for x, y in zip(vb, va):
                assert rel(x, y) < 1e-5, rep
        _ = torch.randn(1000, device="cuda").sum().item()
