#!/usr/bin/env python3
"""Benchmark: differentiable single step at 128^3 on one MI355X.

Workload (one "step"): density 128^3 fp32 (16 seeded Gaussian blobs) and a
MAC-staggered curl-noise velocity (129x128x128 / 128x129x128 / 128x128x129,
max |v| = 2 voxels/step); one MacCormack step (dt=1, trilinear, MC_SMOOTH
revert clamp, BORDER) followed by the 5-view render (ring of 5 pinhole
cameras, 256x144 pixels, 128 depth slices, BEER_LAMBERT, constant light);
L2 loss against target images rendered from a perturbed density (reduced,
and its image gradient formed, inside the render op); full backward to the
density and the three face arrays.  Inputs are resident in
HBM before timing starts.

Prints one JSON line.  value = voxel-updates/s (128^3 cells advected per
fwd+bwd step / time); ray_samples_per_s = 5*256*144*128 / time.  vs_plain =
speed-up over the per-op "plain" path (the reference's composition: an explicit
4-channel LUT, four SampleGridLut launches and elementwise ops for the warp,
a materialised light grid, channel concatenation and one raymarch launch per
camera for the render).
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "neural-global-transport_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
GRID = (128, 128, 128)
IMAGE_HW = (144, 256)
DEPTH = 128
VIEWS = 5


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--no-plain", action="store_true")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--graph", action="store_true",
                   help="replay the fused step as a captured HIP graph (the default; kept for old command lines)")
    p.add_argument("--eager", action="store_true", help="time the eagerly launched step only, no HIP graph")
    p.add_argument("--chain", action="store_true",
                   help="instead of the single step: the 30-step checkpointed chain at 96x160x96 (render every 5th "
                        "step, 5 views 240x135x96); prints its own JSON line (time, peak device memory)")
    p.add_argument("--batch", action="store_true",
                   help="instead of the single step: the uneven batch (32 sequences x 4 frames, grid edges "
                        "48-160, 1-8 views, 3 image sizes) as batched launches (eager and HIP-graph replay) against "
                        "per-instance launches; prints its own JSON line")
    p.add_argument("--traffic-probe", action="store_true",
                   help="run a 256 MiB copy and each op of the step once (for rocprofv3 --pmc FETCH_SIZE / "
                        "WRITE_SIZE runs) and exit")
    p.add_argument("--traffic-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                   help="turn the two counter_collection.csv files of the --traffic-probe runs into the "
                        "per-kernel / per-op HBM traffic JSON (stdout) and exit")
    return p.parse_args()


# kernels launched by each op of the step (substring of the kernel name)
OP_KERNELS = {
    "advect_fwd": ["advect_sl_kernel", "advect_mc_kernel"],
    "advect_bwd": ["advect_mc_bwd_brick_kernel", "advect_sl_bwd_brick_kernel", "advect_rho_faces_kernel",
                   "advect_rho_gather_kernel", "faces_from_centered_grad_kernel", "fill_kernel"],
    "render_fwd": ["to_bricked_kernel", "raymarch_fwd_record_kernel", "l2_finish_kernel"],
    "render_bwd": ["raymarch_bwd_brick_kernel"],
}
COPY_BYTES = 256 << 20
# the op whose single kernel is the step's largest (see main)
ROOFLINE_OP = "render_bwd"
assert len(OP_KERNELS[ROOFLINE_OP]) == 1


def traffic_probe():
    """One 256 MiB device copy (the calibration kernel: reads and writes
    COPY_BYTES), then every op of the step twice (the second call of each is
    the measured dispatch)."""
    import torch
    from ngt import _lib
    _lib.load()
    dev = torch.device("cuda:0")
    x = torch.rand(COPY_BYTES // 4, device=dev)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    torch.mul(x, 1.0, out=y)  # an elementwise kernel reading and writing COPY_BYTES
    torch.cuda.synchronize()
    del x, y
    prob = build_problem(torch, dev)
    calls = op_calls(torch, prob)
    for name in ("advect_fwd", "advect_bwd", "render_fwd", "render_bwd"):
        for _ in range(2):
            calls[name]()
        torch.cuda.synchronize()
        print("probe: %s done" % name, file=sys.stderr, flush=True)


def _read_counters(path, counter):
    """{kernel name: [value per dispatch]} of one counter from a rocprofv3
    counter_collection.csv."""
    import csv
    out = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            if r.get("Counter_Name") != counter:
                continue
            out.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    return out


def traffic_summary(fetch_csv, write_csv):
    fetch = _read_counters(fetch_csv, "FETCH_SIZE")
    write = _read_counters(write_csv, "WRITE_SIZE")
    copy_f = [v for k, vs in fetch.items() if "elementwise" in k for v in vs]
    copy_w = [v for k, vs in write.items() if "elementwise" in k for v in vs]
    # the calibration copy is the largest elementwise dispatch
    scale_f = COPY_BYTES / (max(copy_f) * 1024.0)
    scale_w = COPY_BYTES / (max(copy_w) * 1024.0)
    kernels = {}
    for k in fetch:
        if k not in write:
            continue
        # the last dispatch of each kernel is the warm one
        kernels[k if len(k) <= 160 else k[:157] + "..."] = {"fetch_kib": fetch[k][-1], "write_kib": write[k][-1],
                      "hbm_bytes": fetch[k][-1] * 1024.0 * scale_f + write[k][-1] * 1024.0 * scale_w}
    per_op = {}
    for op, names in OP_KERNELS.items():
        tot = 0.0
        for k, v in kernels.items():
            if any(n in k for n in names) and "raymarch_bwd_exact" not in k:
                tot += v["hbm_bytes"] * (2 if (op == "advect_bwd" and "fill_kernel" in k) else 1)
        per_op[op] = tot
    cmd = ("export TMPDIR=/tmp; rocprofv3 --pmc FETCH_SIZE --output-format csv -d gpurun_out/pmc_fetch -o f -- "
           "python -u bench.py --traffic-probe; rocprofv3 --pmc WRITE_SIZE --output-format csv -d gpurun_out/pmc_write "
           "-o w -- python -u bench.py --traffic-probe; python bench.py --traffic-summary "
           "gpurun_out/pmc_fetch/f_counter_collection.csv gpurun_out/pmc_write/w_counter_collection.csv")
    return {"command": cmd,
            "note": "FETCH_SIZE / WRITE_SIZE (KiB per dispatch) of the warm dispatch of each kernel, scaled by "
                    "COPY_BYTES / counter of a 256 MiB elementwise copy; bytes fetched into the L2s (each XCD's "
                    "L2 fetches separately; MALL hits count as fetched)",
            "calibration": {"copy_bytes": COPY_BYTES, "fetch_scale": scale_f, "write_scale": scale_w},
            "kernels": kernels, "hbm_bytes_per_launch": per_op}


def build_problem(torch, dev):
    from ngt import synth, ops
    from ngt.lighting import Light
    rho = synth.blobs(GRID, 16, seed=2)
    target_rho = synth.blobs(GRID, 16, seed=3)
    vx, vy, vz = synth.curl_noise_staggered(GRID, seed=4, vmax=2.0)
    grid_t, cams = synth.ring_setup(GRID, IMAGE_HW, DEPTH, VIEWS)
    from ngt.renderer import Renderer
    r = Renderer(filter_mode="LINEAR", boundary_mode="BORDER", mipmapping="NONE", blend_mode="BEER_LAMBERT",
                 fused=True)
    M, V, P, F = r._transform_params([grid_t], cams)
    light = Light(color=[1.0, 1.0, 1.0], intensity=1.0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    prob = dict(rho=t(rho[None]), vx=t(vx[None]), vy=t(vy[None]), vz=t(vz[None]), M=M, V=V, P=P, F=F,
                ds=float(cams[0].depth_step), light=tuple(light.rgb()), cams=cams, grid_t=grid_t, renderer=r)
    with torch.no_grad():
        timg = ops.render_density_fused(t(target_rho[None]), prob["light"], prob["ds"], M, V, P, F,
                                        (DEPTH,) + IMAGE_HW)
    prob["target"] = timg[..., :3].contiguous()
    return prob


def fused_step(torch, ops, prob, leaves):
    rho, vx, vy, vz = leaves
    adv = ops.advect_grid_semi_langrange(rho, (vx, vy, vz), timestep=-1.0, order=2, clamp_extrema="MC_SMOOTH",
                                         boundary="BORDER", separate_velocity_batch=False).squeeze(1)
    loss, _ = ops.render_density_fused_l2(adv, prob["target"], prob["light"], prob["ds"], prob["M"], prob["V"],
                                          prob["P"], prob["F"], (DEPTH,) + IMAGE_HW)
    loss.backward()
    return loss


def plain_step(torch, ops, prob, leaves):
    """The reference's composition on the same library's per-op entry points,
    with the reference's scatter-form (fp32 atomic) raymarch adjoint."""
    old = ops.use_gather_adjoint
    ops.use_gather_adjoint = False
    try:
        return _plain_step(torch, ops, prob, leaves)
    finally:
        ops.use_gather_adjoint = old


def _plain_step(torch, ops, prob, leaves):
    rho, vx, vy, vz = leaves
    r = prob["renderer"]
    from ngt.data_structures import VelocityGrid
    vg = VelocityGrid(list(GRID), x=vx, y=vy, z=vz, warp_renderer=r)
    adv = vg.warp(rho, order=2, dt=1.0, clamp="MC_SMOOTH", fused=False)
    light = adv * torch.tensor(prob["light"], dtype=torch.float32, device=adv.device)
    data = torch.cat([light, adv], dim=-1) * prob["ds"]
    imgs = []
    for v in range(VIEWS):
        imgs.append(ops.raymarch_grid_transform(data, prob["M"], prob["V"][v:v + 1], prob["P"][v:v + 1],
                                                prob["F"][v:v + 1], (DEPTH,) + IMAGE_HW))
    img = torch.cat(imgs, dim=1)
    loss = ((img[..., :3] - prob["target"]) ** 2).sum()
    loss.backward()
    return loss


def make_leaves(torch, prob):
    return [prob[k].clone().requires_grad_() for k in ("rho", "vx", "vy", "vz")]


def time_steps(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        fn()
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / steps  # ms


def time_call(torch, call, iters):
    """Average duration of one library call (HIP events on the stream the
    kernels are launched on)."""
    s = torch.cuda.current_stream()
    for _ in range(3):
        call()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(iters):
        call()
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / iters


def op_breakdown(torch, prob, iters):
    """Per-op call durations of the fused step (direct C-ABI calls)."""
    calls = op_calls(torch, prob)
    calls["render_fwd"]()
    return {k: time_call(torch, c, iters) for k, c in calls.items()}


def op_calls(torch, prob):
    """The four ops of the fused step as direct C-ABI calls on preallocated
    buffers."""
    from ngt import _lib, ops
    lib = _lib.load()
    st = lambda: ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    cams = ops._Cams(prob["M"], prob["V"], prob["P"], prob["F"])
    lp, _, keep = _lib.host_floats(prob["light"], 3)
    out_dhw = _lib.dhw((DEPTH,) + IMAGE_HW)
    rho, vx, vy, vz = prob["rho"], prob["vx"], prob["vy"], prob["vz"]
    s_rho = _lib.shape5(rho)
    adv = torch.empty((1, 1) + tuple(rho.shape[1:]), device=rho.device)
    sl = torch.empty_like(adv)
    g_in, gx, gy, gz = (torch.empty_like(t) for t in (rho, vx, vy, vz))
    wsb = lib.ngt_advect_workspace_size(s_rho, 1, 0, 2, 1)
    ws = torch.empty(wsb, dtype=torch.uint8, device=rho.device)
    g_out = torch.randn_like(adv)  # gradient of the advected field, [1,1,D,H,W,1]
    img = torch.empty((1, VIEWS) + IMAGE_HW + (4,), device=rho.device)
    tgt = prob["target"]
    loss = torch.empty((), device=rho.device)
    loss_grad = torch.ones((), device=rho.device)
    g_adv = torch.empty_like(rho)
    rws_b = lib.ngt_render_density_l2_record_size(s_rho, cams.m, cams.v, cams.p, cams.f, cams.nv, out_dhw)
    rws = torch.empty(max(rws_b, 1), dtype=torch.uint8, device=rho.device)
    fws_b = lib.ngt_render_density_l2_fwd_workspace_size(s_rho, 1, cams.nv, out_dhw)
    fws = torch.empty(max(fws_b, 1), dtype=torch.uint8, device=rho.device)
    brk_b = lib.ngt_bricked_grid_bytes(s_rho)
    brk = torch.empty(brk_b, dtype=torch.uint8, device=rho.device)  # the forward's brick copy, reused backward
    P = _lib.ptr

    def render_fwd():  # brick copy, render + loss, and the recorded per-sample gradients
        _lib.check(lib.ngt_bricked_grid(P(rho), s_rho, 0, P(brk), st()))
        _lib.check(lib.ngt_render_density_l2_fwd(
            P(rho), s_rho, lp, prob["ds"], cams.m, cams.v, cams.p, cams.f, cams.nv, out_dhw, 1, 0, 0, P(tgt), 1.0,
            P(brk), P(rws), rws_b, P(img), P(loss), P(fws), fws_b, st()))

    calls = {
        "advect_fwd": lambda: _lib.check(lib.ngt_advect_fwd(
            P(rho), s_rho, P(vx), P(vy), P(vz), 1, -1.0, 2, 1, 0, 0, P(adv), P(sl), None, 0, st())),
        "advect_bwd": lambda: _lib.check(lib.ngt_advect_bwd(
            P(rho), s_rho, P(vx), P(vy), P(vz), 1, -1.0, 2, 1, 0, 0, P(g_out), P(sl), 0, P(g_in), P(gx), P(gy),
            P(gz), P(ws), wsb, st())),
        "render_fwd": render_fwd,
        "render_bwd": lambda: _lib.check(lib.ngt_render_density_l2_bwd(
            P(rho), s_rho, lp, prob["ds"], P(img), P(tgt), 1.0, P(loss_grad), cams.m, cams.v, cams.p, cams.f,
            cams.nv, out_dhw, 1, 0, 0, P(brk), 1, P(g_adv), P(rws), rws_b, st())),
    }
    # every raw buffer handed to the C ABI has the element count its entry
    # point reads (the ABI takes plain pointers and cannot check them)
    cells = int(np.prod(GRID))
    for t, n in ((adv, cells), (sl, cells), (g_out, cells), (g_in, cells), (g_adv, cells),
                 (img, VIEWS * IMAGE_HW[0] * IMAGE_HW[1] * 4), (tgt, VIEWS * IMAGE_HW[0] * IMAGE_HW[1] * 3),
                 (loss, 1), (loss_grad, 1)):
        assert t.numel() == n, (tuple(t.shape), n)
    calls["render_fwd"]()
    return calls


def _oracle_fp64_estimate():
    """FP64 NumPy oracle (single thread) on a bounded sample of the step: the
    MacCormack step fwd+bwd at 48^3 and the fused render fwd+bwd of one view at
    64x36 pixels (128 slices) on the 128^3 grid; seconds per full step, scaled
    per voxel-update and per ray-sample."""
    from oracle import advect as A, raymarch as RM, geometry as G, BORDER, LINEAR
    from oracle import blend as B
    from ngt import synth
    n = 48
    rho = synth.blobs((n, n, n), 8, seed=1).astype(np.float64)
    v = synth.curl_noise_centered((n, n, n), seed=5, vmax=2.0).astype(np.float64)
    t0 = time.perf_counter()
    out = A.warp_forward(rho, v, 1.0, 2, "MC_SMOOTH", BORDER)
    A.warp_backward(rho, v, out, 1.0, 2, "MC_SMOOTH", BORDER)
    t_adv = (time.perf_counter() - t0) / n ** 3
    grid_t, cams = synth.ring_setup(GRID, (36, 64), DEPTH, 1)
    M = G.op_matrix(grid_t.get_transform_matrix())
    V = G.op_matrix(cams[0].view_matrix())
    P = G.op_matrix(cams[0].projection_matrix())
    F = cams[0].frustum().astype(np.float64)
    dens = synth.blobs(GRID, 16, seed=2).astype(np.float64)
    cat = np.concatenate([dens, dens, dens, dens], -1)[None] * cams[0].depth_step
    t0 = time.perf_counter()
    img = RM.raymarch_grid_transform(cat, [M], [V], [P], [F], (DEPTH, 36, 64), LINEAR, BORDER, B.BEER_LAMBERT)
    RM.raymarch_grid_transform_grad(cat, img, img, [M], [V], [P], [F], (DEPTH, 36, 64), LINEAR, BORDER, B.BEER_LAMBERT)
    t_ren = (time.perf_counter() - t0) / (36 * 64 * DEPTH)
    return float(t_adv * np.prod(GRID) + t_ren * VIEWS * IMAGE_HW[0] * IMAGE_HW[1] * DEPTH)


def cpu_baseline(reps=10, warmup=2):
    """The FP32 OpenMP CPU restatement (oracle/cpu_port.cpp) timed on this
    host's cores on the full workload: the same 128^3 MacCormack step and
    5-view 256x144x128 render + L2 loss, forward and backward, the median of
    ``reps`` timed steps after ``warmup`` untimed ones.  The FP64 NumPy
    oracle's single-thread estimate is reported beside it."""
    from oracle import cpu_port
    from ngt import synth
    rho = synth.blobs(GRID, 16, seed=2)[..., 0]
    vx, vy, vz = (a[..., 0] for a in synth.curl_noise_staggered(GRID, seed=4, vmax=2.0))
    grid_t, cams = synth.ring_setup(GRID, IMAGE_HW, DEPTH, VIEWS)
    from ngt.renderer import Renderer
    M, V, P, F = Renderer()._transform_params([grid_t], cams)
    ds = float(cams[0].depth_step)
    light = np.ones(3, np.float32)
    target = cpu_port.render_fwd(synth.blobs(GRID, 16, seed=3)[..., 0], M[0], V, P, (DEPTH,) + IMAGE_HW, light,
                                 ds)[..., :3].copy()

    def step():
        adv, sl = cpu_port.advect_mc_fwd(rho, vx, vy, vz, -1.0)
        img = cpu_port.render_fwd(adv, M[0], V, P, (DEPTH,) + IMAGE_HW, light, ds)
        cpu_port.l2_loss(img, target)
        g_adv = cpu_port.render_l2_bwd(adv, M[0], V, P, (DEPTH,) + IMAGE_HW, light, ds, img, target)
        cpu_port.advect_mc_bwd(rho, vx, vy, vz, -1.0, sl, g_adv)

    for _ in range(warmup):
        step()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        step()
        times.append(time.perf_counter() - t0)
    t_step = float(np.median(times))
    return {"value": float(np.prod(GRID) / t_step), "unit": "voxel-updates/s", "cores": cpu_port.threads(),
            "kind": "port",
            "sample": "FP32 OpenMP C++ restatement (oracle/cpu_port.cpp, g++ -O3 -march=x86-64-v3 -fopenmp) of the "
                      "full step: 128^3 MacCormack+MC_SMOOTH and 5-view 256x144x128 render + L2, fwd+bwd; median "
                      "of %d steps after %d warm-ups" % (reps, warmup),
            "seconds_per_step": t_step, "seconds_min_max": [float(min(times)), float(max(times))],
            "cpu_model": _cpu_model(), "nproc": os.cpu_count(),
            "cpus_available": len(os.sched_getaffinity(0)), "openmp_threads": cpu_port.threads(),
            "oracle_fp64_1thread_seconds_per_step_est": _oracle_fp64_estimate()}


def _cpu_model():
    """The first processor's /proc/cpuinfo model name, with its family /
    model / stepping numbers (a virtualised host may report a short name)."""
    info = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if not line.strip():
                    break
                k, _, v = line.partition(":")
                info[k.strip()] = v.strip()
    except OSError:
        return None
    name = info.get("model name")
    ids = ", ".join("%s %s" % (k, info[k]) for k in ("cpu family", "model", "stepping") if k in info)
    return "%s (%s)" % (name, ids) if name and ids else name


def chain_bench(steps, warmup):
    """30 MacCormack steps of density + centered velocity at 96x160x96 (DHW),
    5-view render (240x135 pixels, 96 slices) after every 5th step, L2 loss
    against targets from a perturbed initial density; full backward to the
    initial density and velocity.  Timed with and without recomputation
    checkpointing; peak device memory above the inputs is reported."""
    import torch
    from ngt import synth, sequence, _lib
    _lib.load()
    grid_dhw, img_hw, depth = (96, 160, 96), (135, 240), 96
    grid_t, cams = synth.ring_setup(grid_dhw, img_hw, depth, VIEWS)
    spec = sequence.RenderSpec.from_cameras(grid_t, cams)
    rho0 = torch.from_numpy(synth.blobs(grid_dhw, 16, seed=2)[None]).cuda()
    vel0 = torch.from_numpy(synth.curl_noise_centered(grid_dhw, seed=4, vmax=2.0)[None]).cuda()
    rho_t = torch.from_numpy(synth.blobs(grid_dhw, 16, seed=3)[None]).cuda()
    targets = sequence.render_targets(rho_t, vel0, spec, 6, render_every=5)
    res = {}
    for ckpt in (True, False):
        def run():
            r = rho0.clone().requires_grad_()
            v = vel0.clone().requires_grad_()
            loss = sequence.chain_loss(r, v, spec, targets, render_every=5, checkpoint=ckpt)
            loss.backward()
            return loss
        for _ in range(warmup):
            run()
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats()
        base = torch.cuda.memory_allocated()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            run()
        e1.record()
        torch.cuda.synchronize()
        res["checkpoint" if ckpt else "stored"] = {
            "ms_per_chain": e0.elapsed_time(e1) / steps,
            "peak_device_bytes": int(torch.cuda.max_memory_allocated() - base)}
    cells = int(np.prod(grid_dhw))
    return {"workload": "30-step MacCormack chain (density + centered velocity, MC_SMOOTH, BORDER) at 96x160x96, "
                        "5-view 240x135x96 render after every 5th step, L2 loss, full backward",
            "chains": steps, "warmup": warmup, "results": res,
            "voxel_updates_per_s_checkpoint": 30 * cells / (res["checkpoint"]["ms_per_chain"] * 1e-3),
            "peak_budget_bytes": 8 << 30,
            "within_budget": res["checkpoint"]["peak_device_bytes"] < (8 << 30)}


def batch_bench(steps, warmup):
    """32 sequences (seed 5), each with a cubic grid edge drawn from {48, 64,
    96, 128, 160}, 4 frames (densities), 1-8 ring views with an image size
    drawn from {128x72, 256x144, 480x270} and edge-many depth slices.  Every
    frame runs one MacCormack step (MC_SMOOTH, BORDER, MAC velocity) and the
    fused render with L2 loss, forward and backward.  Batched: one persistent
    launch per kernel over all 128 instances (eager and captured in a HIP
    graph); per-instance: the library's per-instance ops (advection, then
    the fused render + L2 op) launched instance by instance; per-instance
    plain kernels: the reference's composition (LUT-sampled MacCormack warp,
    one raymarch per camera, scatter-form adjoints) instance by instance."""
    import torch
    from ngt import synth, batch as B, ops, _lib
    from ngt.renderer import Renderer
    _lib.load()
    dens, vels, cams, tgts = [], [], [], []
    cells = ray_samples = 0
    for q, (e, nv, hw) in enumerate(synth.uneven_batch_layout(5, 32)):
        grid_t, cs = synth.ring_setup((e, e, e), hw, e, nv)
        M, V, P, F = Renderer()._transform_params([grid_t], cs)
        vel = synth.curl_noise_staggered_torch((e, e, e), seed=1000 + q, vmax=2.0)
        for f in range(4):
            dens.append(synth.blobs_torch((e, e, e), 16, seed=100 * q + f))
            vels.append(vel)
            cams.append((M, V, P, F, cs[0].depth_step, (e,) + hw))
            tgts.append(torch.rand((nv,) + hw + (3,), device="cuda"))
            cells += e ** 3
            ray_samples += nv * hw[0] * hw[1] * e
    torch.cuda.synchronize()
    print("batch: %d instances generated" % len(dens), file=sys.stderr, flush=True)
    step = B.TransportRenderBatch(dens, vels, cams, tgts)
    print("batch: planned, workspace %.2f GB" % ((step.adv.workspace.numel() + step.ren.workspace.numel()) / 1e9),
          file=sys.stderr, flush=True)
    for _ in range(max(1, warmup)):
        step.step()   # the first call autotunes the persistent grids
    torch.cuda.synchronize()

    def timed(fn, n):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(n):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n

    print("batch: autotuned", file=sys.stderr, flush=True)
    ms_eager = timed(step.step, steps)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        step.step()
    torch.cuda.current_stream().wait_stream(s)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step.step()
    graph.replay()
    torch.cuda.synchronize()
    ms_graph = timed(graph.replay, steps)

    def per_instance():
        # the fastest per-instance composition: the fused render + L2 op
        for i in range(len(dens)):
            M, V, P, F, ds, out_dhw = cams[i]
            d = dens[i][None].requires_grad_()
            v = tuple(c[None].requires_grad_() for c in vels[i])
            adv = ops.advect_grid_semi_langrange(d, v, -1.0, 2, "MC_SMOOTH", "BORDER", False).squeeze(1)
            loss, _ = ops.render_density_fused_l2(adv, tgts[i][None], (1.0, 1.0, 1.0), ds, M, V, P, F, out_dhw)
            loss.backward()

    def per_instance_plain():
        # the reference's composition per instance on the same library's per-op
        # entry points (LUT-sampled warp, one raymarch per camera, scatter-form
        # adjoints), as the single-step bench's plain path
        from ngt.data_structures import VelocityGrid
        from ngt.renderer import Renderer
        old = ops.use_gather_adjoint
        ops.use_gather_adjoint = False
        try:
            for i in range(len(dens)):
                M, V, P, F, ds, out_dhw = cams[i]
                e = int(dens[i].shape[0])
                r = Renderer(filter_mode="LINEAR", boundary_mode="BORDER", mipmapping="NONE",
                             blend_mode="BEER_LAMBERT", fused=True)
                d = dens[i][None].requires_grad_()
                v = [c[None].requires_grad_() for c in vels[i]]
                vg = VelocityGrid([e, e, e], x=v[0], y=v[1], z=v[2], warp_renderer=r)
                adv = vg.warp(d, order=2, dt=1.0, clamp="MC_SMOOTH", fused=False)
                data = torch.cat([adv, adv, adv, adv], dim=-1) * ds
                imgs = [ops.raymarch_grid_transform(data, M, V[k:k + 1], P[k:k + 1], F[k:k + 1], out_dhw)
                        for k in range(len(V))]
                ((torch.cat(imgs, dim=1)[..., :3] - tgts[i][None]) ** 2).sum().backward()
        finally:
            ops.use_gather_adjoint = old

    print("batch: batched %.2f ms eager, %.2f ms graph" % (ms_eager, ms_graph), file=sys.stderr, flush=True)
    per_instance()
    torch.cuda.synchronize()
    ms_plain = timed(per_instance, max(1, steps // 4))
    print("batch: per-instance fused ops %.2f ms" % ms_plain, file=sys.stderr, flush=True)
    per_instance_plain()
    torch.cuda.synchronize()
    ms_plain_kernels = timed(per_instance_plain, 1)
    print("batch: per-instance plain kernels %.2f ms" % ms_plain_kernels, file=sys.stderr, flush=True)
    best = min(ms_eager, ms_graph)
    tuner = B.DEFAULT_TUNER
    return {"workload": "uneven batch: 32 sequences x 4 frames (grid edge 48-160, 1-8 views, 128x72 / 256x144 / "
                        "480x270 images, edge depth slices), MacCormack + fused render + L2, fwd+bwd",
            "instances": len(dens), "voxel_updates_per_step": cells, "ray_samples_per_step": ray_samples,
            "ms_per_step_batched_eager": ms_eager, "ms_per_step_batched_graph": ms_graph,
            "ms_per_step_per_instance": ms_plain, "speedup_vs_per_instance": ms_plain / best,
            "ms_per_step_per_instance_plain_kernels": ms_plain_kernels,
            "speedup_vs_per_instance_plain_kernels": ms_plain_kernels / best,
            "voxel_updates_per_s": cells / (best * 1e-3), "ray_samples_per_s": ray_samples / (best * 1e-3),
            "workspace_bytes": int(step.adv.workspace.numel() + step.ren.workspace.numel()),
            "render_record_chunks": step.ren.num_chunks, "render_record_bytes": step.ren.record_bytes,
            "autotuned": {"/".join(str(x) for x in (k[0], k[-1])): (
                v if k[0] == "advect" or k[-1] == "fwd" else {"blocks_per_cu": v[0],
                                                              "gather_brick": B.GATHER_BRICKS[v[1]]})
                for k, v in tuner.best.items()},
            "autotune_ms": {"/".join(str(x) for x in (k[0], k[-1])): {str(c): round(t, 3) for c, t in v.items()}
                            for k, v in tuner.timings.items()}}


def main():
    args = parse()
    if args.batch:
        print(json.dumps(batch_bench(max(1, args.steps), max(1, args.warmup))), flush=True)
        return
    if args.chain:
        print(json.dumps(chain_bench(max(1, args.steps // 5), min(args.warmup, 1))), flush=True)
        return
    if args.traffic_summary:
        print(json.dumps(traffic_summary(*args.traffic_summary), indent=1))
        return
    if args.traffic_probe:
        traffic_probe()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only; sharding the step across GPUs is not implemented. "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import torch
    from ngt import ops, _lib
    _lib.load()
    dev = torch.device("cuda:0")
    print("bench: building inputs", file=sys.stderr, flush=True)
    prob = build_problem(torch, dev)
    leaves = make_leaves(torch, prob)

    def zero():
        for l in leaves:
            l.grad = None

    def step():
        zero()
        fused_step(torch, ops, prob, leaves)

    print("bench: fused step", file=sys.stderr, flush=True)
    ms = time_steps(torch, step, args.steps, args.warmup)
    ms_graph = None
    graph_check = None
    if not args.eager:
        g = torch.cuda.CUDAGraph()
        static = make_leaves(torch, prob)

        def gstep():
            for l in static:
                l.grad = None
            fused_step(torch, ops, prob, static)

        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for _ in range(3):
                gstep()
        torch.cuda.current_stream().wait_stream(s)
        with torch.cuda.graph(g):
            gstep()
        ms_graph = time_steps(torch, g.replay, args.steps, args.warmup)
        # the replayed step computes what the eager step computes (gradients
        # to rounding: their sums use atomics in no fixed order)
        zero()
        step()
        torch.cuda.synchronize()
        graph_check = max(float((a.grad - b.grad).abs().max() / b.grad.abs().max().clamp_min(1e-30))
                          for a, b in zip(static, leaves))
        if not graph_check < 1e-4:
            print("bench: HIP-graph replay disagrees with the eager step (max rel %.3g)" % graph_check,
                  file=sys.stderr, flush=True)
            sys.exit(1)

    ms_plain = None
    if not args.no_plain:
        print("bench: plain step", file=sys.stderr, flush=True)
        pleaves = make_leaves(torch, prob)

        def pstep():
            for l in pleaves:
                l.grad = None
            plain_step(torch, ops, prob, pleaves)

        ms_plain = time_steps(torch, pstep, max(3, args.steps // 4), 2)

    ops_ms = op_breakdown(torch, prob, max(5, args.steps))
    cells = int(np.prod(GRID))
    pixels = VIEWS * IMAGE_HW[0] * IMAGE_HW[1]
    # algorithmic bytes per call (compulsory interface traffic):
    #   advect fwd 20 B/voxel, advect bwd 36 B/voxel (staggered faces ~12 B/cell),
    #   render + L2 fwd 4 B/voxel + 28 B/pixel (image out, RGB target in),
    #   render + L2 bwd 8 B/voxel + 28 B/pixel (image and RGB target in)
    algo = {"advect_fwd": cells * 20, "advect_bwd": cells * 36, "render_fwd": cells * 4 + pixels * 28,
            "render_bwd": cells * 8 + pixels * 28}
    # the roofline names the step's dominant kernel (rocprofv3 kernel-trace
    # summary of this workload): the render adjoint's brick gather,
    # raymarch_bwd_brick_kernel, which the render + L2 backward launches alone,
    # so that call's HIP-event duration is the kernel's
    dominant = ROOFLINE_OP
    ms_kernel = ops_ms[dominant]
    kbytes = algo[dominant]
    achieved = kbytes / (ms_kernel * 1e-3) / 1e9
    # whole-step compulsory bytes: advect fwd 20 + bwd 36 per voxel, render + L2
    # fwd 4/voxel + 28/pixel, bwd 8/voxel + 28/pixel
    step_bytes = cells * (20 + 36 + 4 + 8) + pixels * (28 + 28)
    best_ms = ms_graph if ms_graph is not None else ms
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_step128.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f)["hbm_bytes_per_launch"].get(dominant)
    rec = {
        "metric": "fwd+bwd voxel-updates/s + ray-samples/s at 128^3 (1 GPU), % HBM peak, vs plain",
        "value": cells / (best_ms * 1e-3),
        "unit": "voxel-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": best_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded blobs + curl noise)",
        "config": {"workload": "MacCormack step (MC_SMOOTH, BORDER) of 128^3 density with MAC curl-noise velocity "
                               "+ 5-view 256x144x128 fused render, L2 loss, full backward",
                   "grid": list(GRID), "views": VIEWS, "image_hw": list(IMAGE_HW), "depth_slices": DEPTH,
                   "graph": not args.eager},
        "ray_samples_per_s": pixels * DEPTH / (best_ms * 1e-3),
        "hbm_frac_step": step_bytes / (best_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "ms_per_step_eager": ms,
        "ms_per_step_graph": ms_graph,
        "graph_check_max_rel": graph_check,
        "ms_per_step_plain": ms_plain,
        "vs_plain": (ms_plain / best_ms) if ms_plain else None,
        "roofline": {"kernel": " + ".join(OP_KERNELS[dominant]) + " (the whole ngt_" +
                               dominant.replace("render", "render_density_l2") + " call)",
                     "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "ms_per_launch": ms_kernel,
                     "algorithmic_bytes_per_launch": kbytes},
        "op_ms": ops_ms,
        "op_hbm_frac": {k: algo[k] / (ops_ms[k] * 1e-3) / 1e9 / HBM_PEAK_GBS for k in ops_ms},
    }
    if not args.no_cpu:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        rec["cpu_baseline"] = cpu_baseline()
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
