"""Uneven batches: many (density, velocity, cameras) instances of different
grid and image sizes advected and rendered with one persistent launch per
kernel (ngt_advect_batch_* / ngt_render_batch_* in include/ngt_transport.h).

A plan is built once per batch layout (instance sizes and buffers) on the
host, copied to device memory and reused; the launches only read it, so a
whole batched forward + backward step can be captured in a HIP graph
(``torch.cuda.CUDAGraph``) and replayed.  ``Autotuner`` times candidate
launch shapes on first use and caches the fastest per batch layout: the
persistent grid size of each kernel family (work-groups per compute unit,
0 = one work-group per tile) and, for the render's gather adjoint, the voxel
brick of a work-group (8x8x8, 16x8x8 or 16x16x8).

The render's per-sample records (the gather adjoint's only large buffer) are
bounded by a byte budget: instances are grouped into chunks whose records
fit it and the backward runs chunk by chunk over one record region.

The reference processes every instance and camera with its own launches
(resample_grid.cu.cc:748-790, raymarch_grid.cu.cc:399-430); the batched path
computes the same values per instance (checked against the per-instance ops
in the GPU tests).
"""
import ctypes

import numpy as np
import torch

from . import _lib, ops

CANDIDATE_BLOCKS_PER_CU = (0, 2, 4, 8, 16)
GATHER_BRICKS = ("8x8x8", "16x8x8", "16x16x8")
# (work-groups per compute unit, gather brick index)
CANDIDATE_RENDER = tuple((b, g) for b in CANDIDATE_BLOCKS_PER_CU for g in range(len(GATHER_BRICKS)))
DEFAULT_GATHER_BRICK = 2


class Autotuner:
    """Times each candidate of a launch parameter on the current stream and
    keeps the fastest per key (the batch layout)."""

    def __init__(self, candidates=CANDIDATE_BLOCKS_PER_CU, reps=3):
        self.candidates = tuple(candidates)
        self.reps = reps
        self.best = {}
        self.timings = {}

    def pick(self, key, run, candidates=None):
        if key in self.best:
            return self.best[key]
        s = torch.cuda.current_stream()
        times = {}
        for c in (self.candidates if candidates is None else candidates):
            run(c)  # warm
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            for _ in range(self.reps):
                run(c)
            e1.record(s)
            torch.cuda.synchronize()
            times[c] = e0.elapsed_time(e1) / self.reps
        best = min(times, key=times.get)
        self.best[key] = best
        self.timings[key] = times
        return best


DEFAULT_TUNER = Autotuner()


def _ptr(t):
    return t.data_ptr() if t is not None else None


def _upload(blob):
    host = np.frombuffer(blob, dtype=np.uint8).copy()
    dev = torch.from_numpy(host).cuda()
    return host, dev


class AdvectBatch:
    """MacCormack / semi-Lagrangian advection of N 1-channel densities, each
    with its own velocity (MAC face triple or centered [D,H,W,3])."""

    def __init__(self, densities, velocities, timestep=-1.0, order=2, clamp="MC_SMOOTH", boundary="BORDER",
                 tuner=DEFAULT_TUNER):
        lib = _lib.load()
        self.n = len(densities)
        if self.n == 0 or len(velocities) != self.n:
            raise ValueError("one velocity per density")
        self.order = int(order)
        self.clamp = ops._enum(ops.CLAMP, clamp, "clamp")
        self.boundary = ops._enum(ops.BOUNDARY, boundary, "boundary")
        self.tuner = tuner
        self.densities = [_lib.device_tensor(d, "density").reshape(d.shape[-4:-1] + (1,)) for d in densities]
        self.velocities = []
        for v in velocities:
            if isinstance(v, (tuple, list)):
                self.velocities.append(tuple(_lib.device_tensor(c, "velocity") for c in v))
            else:
                self.velocities.append((_lib.device_tensor(v, "velocity"),))
        self.outputs = [torch.empty_like(d) for d in self.densities]
        self.sl = [torch.empty_like(d) for d in self.densities] if self.order == 2 else [None] * self.n
        self.out_grads = [torch.zeros_like(d) for d in self.densities]
        self.in_grads = [torch.empty_like(d) for d in self.densities]
        self.vel_grads = [tuple(torch.empty_like(c) for c in v) for v in self.velocities]
        arr = (_lib.AdvectInstance * self.n)()
        for i in range(self.n):
            q = arr[i]
            D, H, W = self.densities[i].shape[:3]
            q.dhw[0], q.dhw[1], q.dhw[2] = D, H, W
            q.timestep = float(timestep)
            q.input = _ptr(self.densities[i])
            v = self.velocities[i]
            q.vel_x = _ptr(v[0])
            q.vel_y = _ptr(v[1]) if len(v) == 3 else None
            q.vel_z = _ptr(v[2]) if len(v) == 3 else None
            q.output = _ptr(self.outputs[i])
            q.sl = _ptr(self.sl[i])
            q.output_grad = _ptr(self.out_grads[i])
            q.input_grad = _ptr(self.in_grads[i])
            g = self.vel_grads[i]
            q.vel_x_grad = _ptr(g[0])
            q.vel_y_grad = _ptr(g[1]) if len(g) == 3 else None
            q.vel_z_grad = _ptr(g[2]) if len(g) == 3 else None
        nbytes = lib.ngt_advect_batch_plan_bytes(self.n)
        blob = (ctypes.c_uint8 * nbytes)()
        ws = ctypes.c_size_t(0)
        _lib.check(lib.ngt_advect_batch_plan(arr, self.n, self.order, self.clamp, self.boundary, blob, nbytes,
                                             ctypes.byref(ws)))
        self.plan_host, self.plan_dev = _upload(bytes(blob))
        self.workspace = torch.empty(max(int(ws.value), 1), dtype=torch.uint8, device="cuda")
        self.key = ("advect", tuple(tuple(d.shape) for d in self.densities), self.order, self.clamp)

    def _fwd(self, bpc):
        lib = _lib.load()
        _lib.check(lib.ngt_advect_batch_fwd(self.plan_host.ctypes.data, self.plan_dev.data_ptr(), bpc,
                                            _lib.stream(torch.device("cuda"))))

    def _bwd(self, bpc):
        lib = _lib.load()
        _lib.check(lib.ngt_advect_batch_bwd(self.plan_host.ctypes.data, self.plan_dev.data_ptr(),
                                            self.workspace.data_ptr(), self.workspace.numel(), bpc,
                                            _lib.stream(torch.device("cuda"))))

    def forward(self):
        """Fills ``outputs`` (and ``sl``); returns ``outputs``."""
        self._fwd(self.tuner.pick(self.key + ("fwd",), self._fwd) if self.tuner else 0)
        return self.outputs

    def backward(self):
        """Reads ``out_grads``; fills ``in_grads`` and ``vel_grads``."""
        self._bwd(self.tuner.pick(self.key + ("bwd",), self._bwd) if self.tuner else 0)
        return self.in_grads, self.vel_grads


class RenderBatch:
    """Fused density render of N instances, each a 1-channel density with its
    own grid transform and 1-8 cameras of one image size."""

    def __init__(self, densities, cameras, light_rgb=(1.0, 1.0, 1.0), blending_mode="BEER_LAMBERT",
                 tuner=DEFAULT_TUNER, density_grads=None, targets=None, record_budget=0):
        """densities: list of [D,H,W,1] (or [1,D,H,W,1]) device tensors;
        cameras: list of (matrix_m, matrix_v, matrix_p, frustum, depth_step,
        output_dhw) in the op layout of ``ngt.ops.render_density_fused``;
        targets: optional [V,H',W',3] target images per instance (for
        ``l2_step``); record_budget: bytes for the gather adjoint's records
        (0: the library's default, 4 GiB)."""
        lib = _lib.load()
        self.n = len(densities)
        if self.n == 0 or len(cameras) != self.n:
            raise ValueError("one camera set per density")
        if targets is not None and len(targets) != self.n:
            raise ValueError("one target per density")
        self.tuner = tuner
        self.mode = ops._enum(ops.BLEND, blending_mode, "blend mode")
        self.densities = [_lib.device_tensor(d, "density").reshape(d.shape[-4:-1] + (1,)) for d in densities]
        self._keep = []
        arr = (_lib.RenderInstance * self.n)()
        self.images, self.image_grads, self.density_grads, self.targets = [], [], [], []
        self.losses = torch.zeros(self.n, device="cuda")
        for i, (M, V, P, F, ds, out_dhw) in enumerate(cameras):
            q = arr[i]
            D, H, W = self.densities[i].shape[:3]
            q.dhw[0], q.dhw[1], q.dhw[2] = D, H, W
            q.density = _ptr(self.densities[i])
            mp, _, km = _lib.host_floats(M, 16)
            vp, nv, kv = _lib.host_floats(V, 16)
            pp, _, kp = _lib.host_floats(P, 16)
            fp, _, kf = _lib.host_floats(F, 6)
            self._keep += [km, kv, kp, kf]
            q.num_views = nv
            q.matrix_m, q.matrix_v, q.matrix_p, q.frustum = mp, vp, pp, fp
            q.image_dhw[0], q.image_dhw[1], q.image_dhw[2] = (int(v) for v in out_dhw)
            for c in range(3):
                q.light_rgb[c] = float(light_rgb[c])
            q.depth_step = float(ds)
            hw = (int(out_dhw[1]), int(out_dhw[2]))
            img = torch.empty((nv,) + hw + (4,), device="cuda")
            self.images.append(img)
            self.image_grads.append(torch.zeros_like(img))
            self.density_grads.append(torch.empty_like(self.densities[i]) if density_grads is None
                                      else _lib.device_tensor(density_grads[i], "density_grad"))
            q.images = _ptr(img)
            q.images_grad = _ptr(self.image_grads[-1])
            q.density_grad = _ptr(self.density_grads[-1])
            if targets is not None:
                tgt = _lib.device_tensor(targets[i], "target").reshape((nv,) + hw + (3,))
                self.targets.append(tgt)
                q.l2_target = _ptr(tgt)
                q.loss = self.losses.data_ptr() + 4 * i
        nbytes = lib.ngt_render_batch_plan_bytes(self.n)
        blob = (ctypes.c_uint8 * nbytes)()
        ws = ctypes.c_size_t(0)
        _lib.check(lib.ngt_render_batch_plan(arr, self.n, self.mode, int(record_budget), blob, nbytes,
                                             ctypes.byref(ws)))
        chunks, rbytes = ctypes.c_int(0), ctypes.c_size_t(0)
        _lib.check(lib.ngt_render_batch_plan_info(blob, ctypes.byref(chunks), ctypes.byref(rbytes)))
        self.num_chunks, self.record_bytes = int(chunks.value), int(rbytes.value)
        self.plan_host, self.plan_dev = _upload(bytes(blob))
        self.workspace = torch.empty(max(int(ws.value), 1), dtype=torch.uint8, device="cuda")
        self.key = ("render", tuple(tuple(d.shape) for d in self.densities),
                    tuple(tuple(i.shape) for i in self.images), self.mode, self.num_chunks)

    def _fwd(self, bpc):
        lib = _lib.load()
        _lib.check(lib.ngt_render_batch_fwd(self.plan_host.ctypes.data, self.plan_dev.data_ptr(),
                                            self.workspace.data_ptr(), self.workspace.numel(), bpc,
                                            _lib.stream(torch.device("cuda"))))

    def _bwd(self, shape):
        bpc, brick = shape
        lib = _lib.load()
        _lib.check(lib.ngt_render_batch_bwd(self.plan_host.ctypes.data, self.plan_dev.data_ptr(),
                                            self.workspace.data_ptr(), self.workspace.numel(), bpc, brick,
                                            _lib.stream(torch.device("cuda"))))

    def _l2(self, shape):
        bpc, brick = shape
        lib = _lib.load()
        _lib.check(lib.ngt_render_batch_l2(self.plan_host.ctypes.data, self.plan_dev.data_ptr(),
                                           self.workspace.data_ptr(), self.workspace.numel(), bpc, brick,
                                           _lib.stream(torch.device("cuda"))))

    def forward(self):
        self._fwd(self.tuner.pick(self.key + ("fwd",), self._fwd) if self.tuner else 0)
        return self.images

    def backward(self, shape=None):
        """Reads ``image_grads`` and the images of the last ``forward``;
        fills ``density_grads``.  shape: (work-groups per CU, gather brick
        index) instead of the tuned one."""
        if shape is None:
            shape = (self.tuner.pick(self.key + ("bwd",), self._bwd, CANDIDATE_RENDER) if self.tuner
                     else (0, DEFAULT_GATHER_BRICK))
        self._bwd(shape)
        return self.density_grads

    def l2_step(self, shape=None):
        """Images, per-instance losses sum((image - target)^2 over RGB) in
        ``losses`` and their density gradients in ``density_grads``."""
        if not self.targets:
            raise ValueError("l2_step needs targets")
        if shape is None:
            shape = (self.tuner.pick(self.key + ("l2",), self._l2, CANDIDATE_RENDER) if self.tuner
                     else (0, DEFAULT_GATHER_BRICK))
        self._l2(shape)
        return self.images, self.losses, self.density_grads


class TransportRenderBatch:
    """One differentiable step for every instance: MacCormack advection of
    the density, fused render of the advected density, L2 loss against
    per-instance target images, and the full backward to the initial density
    and velocity: the advection forward, the render's L2 step (images, loss
    and density gradient) and the advection backward."""

    def __init__(self, densities, velocities, cameras, targets, light_rgb=(1.0, 1.0, 1.0), timestep=-1.0,
                 clamp="MC_SMOOTH", boundary="BORDER", tuner=DEFAULT_TUNER, record_budget=0):
        self.adv = AdvectBatch(densities, velocities, timestep, 2, clamp, boundary, tuner)
        # the render's density gradient is written straight into the advection's output gradient
        self.ren = RenderBatch(self.adv.outputs, cameras, light_rgb, tuner=tuner, density_grads=self.adv.out_grads,
                               targets=targets, record_budget=record_budget)
        self.targets = self.ren.targets

    def step(self):
        self.adv.forward()
        self.ren.l2_step()
        return self.adv.backward()

    def losses(self):
        """Per-instance losses of the last step (device tensor [N])."""
        return self.ren.losses

    def loss(self):
        return float(self.ren.losses.double().sum())
