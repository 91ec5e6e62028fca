"""The uneven-batch workload at full size: 32 sequences x 4 frames
(ngt.synth.uneven_batch_layout(5, 32): grid edges 48-160, 1-8 views,
128x72 / 256x144 / 480x270 images, edge-many depth slices), one MacCormack
step + fused render + L2 loss, forward and backward, as batched launches.

* Every one of the 128 instances against the library's per-instance ops
  (advection, then the fused render + L2 op) on the same inputs.
* The batch with a 1 GiB record budget (many chunks) against the default
  budget: the render's results bit for bit.
* Sequence 6, frame 0 (48^3, 8 views of 128x72; inputs from the numpy
  generators) against the FP64 oracle run on the same inputs (about 15 s
  of CPU): advected density, images and loss within 1e-5 of the field's
  max |.| (of the loss), every gradient within 1e-4 of its max |g|.
"""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SLOT = (6, 0)  # sequence, frame
FWD_TOL = 1e-5
GRAD_TOL = 1e-4


@pytest.fixture(scope="module")
def workload():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from ngt import synth, batch as B, ops, _lib
    from ngt.renderer import Renderer
    _lib.load()
    dens, vels, cams, tgts = [], [], [], []
    for q, (e, nv, hw) in enumerate(synth.uneven_batch_layout(5, 32)):
        grid_t, cs = synth.ring_setup((e, e, e), hw, e, nv)
        M, V, P, F = Renderer()._transform_params([grid_t], cs)
        if q == ORACLE_SLOT[0]:
            vel = tuple(torch.from_numpy(c).cuda() for c in
                        synth.curl_noise_staggered((e, e, e), seed=1000 + q, vmax=2.0))
        else:
            vel = synth.curl_noise_staggered_torch((e, e, e), seed=1000 + q, vmax=2.0)
        for f in range(4):
            if (q, f) == ORACLE_SLOT:
                dens.append(torch.from_numpy(synth.blobs((e, e, e), 16, seed=100 * q + f)).cuda())
                tgts.append(torch.from_numpy(np.random.default_rng(7).uniform(0.0, 1.0, (nv,) + hw + (3,))
                                             .astype(np.float32)).cuda())
            else:
                dens.append(synth.blobs_torch((e, e, e), 16, seed=100 * q + f))
                tgts.append(torch.rand((nv,) + hw + (3,), device="cuda",
                                       generator=torch.Generator("cuda").manual_seed(1000 * q + f)))
            vels.append(vel)
            cams.append((M, V, P, F, cs[0].depth_step, (e,) + hw))
    torch.cuda.synchronize()
    print("uneven batch: %d instances" % len(dens), flush=True)
    return dict(B=B, ops=ops, dens=dens, vels=vels, cams=cams, tgts=tgts)


def _batched(w, record_budget=0):
    B = w["B"]
    step = B.TransportRenderBatch(w["dens"], w["vels"], w["cams"], w["tgts"], tuner=None,
                                  record_budget=record_budget)
    gin, gvel = step.step()
    torch.cuda.synchronize()
    return step, gin, gvel


def rel(a, b):
    return float((a - b).abs().max() / b.abs().max().clamp_min(1e-30))


def test_uneven_batch_matches_per_instance(workload):
    w = workload
    ops = w["ops"]
    step, gin, gvel = _batched(w)
    print("batched: record chunks %d, record region %.2f GB, render workspace %.2f GB" % (
        step.ren.num_chunks, step.ren.record_bytes / 1e9, step.ren.workspace.numel() / 1e9), flush=True)
    losses = step.losses().cpu()
    worst = [0.0, 0.0, 0.0]
    for i in range(len(w["dens"])):
        M, V, P, F, ds, out_dhw = w["cams"][i]
        d = w["dens"][i][None].clone().requires_grad_()
        v = tuple(c[None].clone().requires_grad_() for c in w["vels"][i])
        adv = ops.advect_grid_semi_langrange(d, v, -1.0, 2, "MC_SMOOTH", "BORDER", False).squeeze(1)
        loss, img = ops.render_density_fused_l2(adv, w["tgts"][i][None], (1.0, 1.0, 1.0), ds, M, V, P, F, out_dhw)
        loss.backward()
        assert torch.equal(step.adv.outputs[i], adv.detach()[0]), i
        assert torch.equal(step.ren.images[i], img.detach()[0]), i
        lv = float(loss.detach())
        worst[0] = max(worst[0], abs(float(losses[i]) - lv) / lv)
        worst[1] = max(worst[1], rel(gin[i], d.grad[0]))
        worst[2] = max(worst[2], max(rel(a, b.grad[0]) for a, b in zip(gvel[i], v)))
        assert worst[0] <= 1e-6 and worst[1] <= 1e-5 and worst[2] <= 1e-5, (i, worst)
    print("128 instances vs per-instance ops: loss %.1e, dL/drho %.1e, dL/dv %.1e (relative)" % tuple(worst),
          flush=True)


def test_record_chunks_are_bitwise_equal(workload):
    """The render's results (images, losses, dL/d advected density) do not
    depend on the chunking, bit for bit (the gather's fixed-point sums are
    independent of order); the advection adjoint adds with LDS float
    atomics in no fixed order, so its outputs agree to rounding."""
    w = workload
    step, gin, gvel = _batched(w)
    ref = ([g.clone() for g in step.ren.density_grads], [t.clone() for t in step.ren.images], step.losses().clone(),
           [g.clone() for g in gin], [tuple(c.clone() for c in v) for v in gvel])
    del step, gin, gvel
    # every instance's records fit 1 GiB alone, so no chunk exceeds it
    small, gin2, gvel2 = _batched(w, record_budget=1 << 30)
    print("record budget 1 GiB: %d chunks, record region %.2f GB" % (small.ren.num_chunks,
                                                                    small.ren.record_bytes / 1e9), flush=True)
    assert small.ren.num_chunks >= 4
    assert small.ren.record_bytes <= 1 << 30
    assert torch.equal(small.losses(), ref[2])
    for i in range(len(w["dens"])):
        assert torch.equal(small.ren.density_grads[i], ref[0][i]), i
        assert torch.equal(small.ren.images[i], ref[1][i]), i
        assert rel(gin2[i], ref[3][i]) <= 1e-6, i
        for a, b in zip(gvel2[i], ref[4][i]):
            assert rel(a, b) <= 1e-6, i


def test_oracle_instance(workload):
    """The batched results of one instance against the FP64 oracle run on
    the same float32 inputs, with the device's MC_SMOOTH revert choices at
    ties (tests/advect_masks.py)."""
    import sys
    sys.path.insert(0, os.path.dirname(HERE))
    sys.path.insert(0, HERE)
    from oracle import BORDER, LINEAR
    from oracle import advect as A
    from oracle import blend as OB
    from oracle import raymarch as RM
    import advect_masks as AM
    w = workload
    ops = w["ops"]
    i = 4 * ORACLE_SLOT[0] + ORACLE_SLOT[1]
    M, V, P, F, ds, out_dhw = w["cams"][i]
    print("case: uneven batch instance %d (48^3, 8 views of 128x72) vs the oracle" % i, flush=True)
    step, gin, gvel = _batched(w)
    host = lambda t: t.detach().cpu().double().numpy()
    rho = host(w["dens"][i])
    faces = [host(c) for c in w["vels"][i]]
    tgt = host(w["tgts"][i])
    ds = float(np.float32(ds))
    # the device's SL and unclamped MacCormack values: its revert choices at ties
    args = dict(timestep=-1.0, boundary="BORDER", separate_velocity_batch=False)
    with torch.no_grad():
        d1, v1 = w["dens"][i][None], tuple(c[None] for c in w["vels"][i])
        sl = host(ops.advect_grid_semi_langrange(d1, v1, order=1, clamp_extrema="NONE", **args))[0, 0]
        corr = host(ops.advect_grid_semi_langrange(d1, v1, order=2, clamp_extrema="NONE", **args))[0, 0]
    adv_gpu = host(step.adv.outputs[i])
    vc = A.centered_from_staggered(*faces)
    _, parts = A.warp_forward(rho, vc, 1.0, 2, "MC_SMOOTH", BORDER, return_parts=True)
    ties = AM.branch_ties(rho, vc, parts, 2, "MC_SMOOTH", BORDER)
    oob = np.where(ties, (corr < parts["mn"]) | (corr > parts["mx"]), parts["oob"])
    adv = np.where(oob, parts["fwd"], parts["corr"])
    e = rel(torch.from_numpy(adv_gpu), torch.from_numpy(adv))
    print("advected density: %.2e of max (%d ties, device choices)" % (e, int(np.sum(ties))), flush=True)
    assert e <= FWD_TOL
    cells = np.concatenate([adv, adv, adv, adv], axis=-1)[None] * ds
    img = RM.raymarch_grid_transform(cells, M, V, P, F, out_dhw, LINEAR, BORDER, OB.BEER_LAMBERT)
    e = rel(torch.from_numpy(host(step.ren.images[i])), torch.from_numpy(img[0]))
    print("images: %.2e of max" % e, flush=True)
    assert e <= FWD_TOL
    loss = float(np.sum((img[0, ..., :3] - tgt) ** 2))
    lv = float(step.losses()[i])
    print("loss %.9e (oracle %.9e)" % (lv, loss), flush=True)
    assert abs(lv - loss) <= FWD_TOL * loss
    g_img = np.zeros_like(img)
    g_img[..., :3] = 2.0 * (img[..., :3] - tgt[None])
    g_cells = RM.raymarch_grid_transform_grad(cells, img, g_img, M, V, P, F, out_dhw, LINEAR, BORDER,
                                              OB.BEER_LAMBERT)[0]
    g_adv = ds * np.sum(g_cells, axis=-1, keepdims=True)  # light (1, 1, 1)
    e = rel(torch.from_numpy(host(step.ren.density_grads[i])), torch.from_numpy(g_adv))
    print("dL/d advected density: %.2e of max" % e, flush=True)
    assert e <= GRAD_TOL
    gpu = {"sl": sl, "corr": corr, "out": adv_gpu}
    g_rho, g_vc, rho_keep, vel_keep, info = AM.oracle_adjoint(rho, vc, g_adv, 2, "MC_SMOOTH", BORDER, gpu)
    print("left out: %s" % info, flush=True)
    assert info["undecided"] == 0 and info["vel_left_out"] <= 0.005 * vel_keep.size
    e = AM.rel_on(host(gin[i]), g_rho, rho_keep[..., None])
    print("dL/drho: %.2e of max" % e, flush=True)
    assert e <= GRAD_TOL
    for k, (t, ex, m) in enumerate(zip(gvel[i], A.centered_from_staggered_adjoint(g_vc),
                                       AM.faces_of_cells(~vel_keep))):
        e = AM.rel_on(host(t), ex, ~m)
        print("dL/dv%s: %.2e of max" % ("xyz"[k], e), flush=True)
        assert e <= GRAD_TOL
