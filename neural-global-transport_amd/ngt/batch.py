"""Uneven batches: many (density, velocity, cameras) instances of different
grid and image sizes advected and rendered with one persistent launch per
kernel (ngt_advect_batch_* / ngt_render_batch_* in include/ngt_transport.h).

A plan is built once per batch layout (instance sizes and buffers) on the
host, copied to device memory and reused; the launches only read it, so a
whole batched forward + backward step can be captured in a HIP graph
(``torch.cuda.CUDAGraph``) and replayed.  The persistent grid size of each
kernel family (work-groups per compute unit) is chosen by ``Autotuner`` from
timed candidates on first use and cached per batch layout.

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


class Autotuner:
    """Times each candidate of a launch parameter on the current stream and
    keeps the fastest per key (the batch layout)."""

    def __init__(self, candidates=CANDIDATE_BLOCKS_PER_CU, reps=3):
        self.candidates = tuple(candidates)
        self.reps = reps
        self.best = {}
        self.timings = {}

    def pick(self, key, run):
        if key in self.best:
            return self.best[key]
        s = torch.cuda.current_stream()
        times = {}
        for c in self.candidates:
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
                 tuner=DEFAULT_TUNER, density_grads=None):
        """densities: list of [D,H,W,1] (or [1,D,H,W,1]) device tensors;
        cameras: list of (matrix_m, matrix_v, matrix_p, frustum, depth_step,
        output_dhw) in the op layout of ``ngt.ops.render_density_fused``."""
        lib = _lib.load()
        self.n = len(densities)
        if self.n == 0 or len(cameras) != self.n:
            raise ValueError("one camera set per density")
        self.tuner = tuner
        self.mode = ops._enum(ops.BLEND, blending_mode, "blend mode")
        self.densities = [_lib.device_tensor(d, "density").reshape(d.shape[-4:-1] + (1,)) for d in densities]
        self._keep = []
        arr = (_lib.RenderInstance * self.n)()
        self.images, self.image_grads, self.density_grads = [], [], []
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
            img = torch.empty((nv, int(out_dhw[1]), int(out_dhw[2]), 4), device="cuda")
            self.images.append(img)
            self.image_grads.append(torch.zeros_like(img))
            self.density_grads.append(torch.empty_like(self.densities[i]) if density_grads is None
                                      else _lib.device_tensor(density_grads[i], "density_grad"))
            q.images = _ptr(img)
            q.images_grad = _ptr(self.image_grads[-1])
            q.density_grad = _ptr(self.density_grads[-1])
        nbytes = lib.ngt_render_batch_plan_bytes(self.n)
        blob = (ctypes.c_uint8 * nbytes)()
        ws = ctypes.c_size_t(0)
        _lib.check(lib.ngt_render_batch_plan(arr, self.n, self.mode, blob, nbytes, ctypes.byref(ws)))
        self.plan_host, self.plan_dev = _upload(bytes(blob))
        self.workspace = torch.empty(max(int(ws.value), 1), dtype=torch.uint8, device="cuda")
        self.key = ("render", tuple(tuple(d.shape) for d in self.densities),
                    tuple(tuple(i.shape) for i in self.images), self.mode)

    def _fwd(self, bpc):
        lib = _lib.load()
        _lib.check(lib.ngt_render_batch_fwd(self.plan_host.ctypes.data, self.plan_dev.data_ptr(),
                                            self.workspace.data_ptr(), self.workspace.numel(), bpc,
                                            _lib.stream(torch.device("cuda"))))

    def _bwd(self, bpc):
        lib = _lib.load()
        _lib.check(lib.ngt_render_batch_bwd(self.plan_host.ctypes.data, self.plan_dev.data_ptr(),
                                            self.workspace.data_ptr(), self.workspace.numel(), bpc,
                                            _lib.stream(torch.device("cuda"))))

    def forward(self):
        self._fwd(self.tuner.pick(self.key + ("fwd",), self._fwd) if self.tuner else 0)
        return self.images

    def backward(self):
        """Reads ``image_grads``; fills ``density_grads``."""
        self._bwd(self.tuner.pick(self.key + ("bwd",), self._bwd) if self.tuner else 0)
        return self.density_grads


class TransportRenderBatch:
    """One differentiable step for every instance: MacCormack advection of
    the density, fused render of the advected density, L2 loss against
    per-instance target images, and the full backward to the initial density
    and velocity, as two batched ops forward and two backward."""

    def __init__(self, densities, velocities, cameras, targets, light_rgb=(1.0, 1.0, 1.0), timestep=-1.0,
                 clamp="MC_SMOOTH", boundary="BORDER", tuner=DEFAULT_TUNER):
        self.adv = AdvectBatch(densities, velocities, timestep, 2, clamp, boundary, tuner)
        # the render's density gradient is written straight into the advection's output gradient
        self.ren = RenderBatch(self.adv.outputs, cameras, light_rgb, tuner=tuner, density_grads=self.adv.out_grads)
        self.targets = [t.reshape(img.shape[:3] + (3,)) for t, img in zip(targets, self.ren.images)]

    def step(self):
        self.adv.forward()
        images = self.ren.forward()
        for img, tgt, g in zip(images, self.targets, self.ren.image_grads):
            torch.mul(img[..., :3] - tgt, 2.0, out=g[..., :3])
        self.ren.backward()
        return self.adv.backward()

    def loss(self):
        return sum(float(((img[..., :3] - t) ** 2).double().sum()) for img, t in zip(self.ren.images, self.targets))
