"""ctypes wrapper of oracle/build/libngt_cpu_port.so (oracle/cpu_port.cpp):
the FP32 OpenMP restatement of the benchmark step (MacCormack advection with
MC_SMOOTH on MAC velocity, fused BEER_LAMBERT render with the L2 image loss,
forward and exact adjoint).

Test infrastructure only (see ``oracle/__init__.py``): used by the tests as a
second checker and by bench.py's ``cpu_baseline`` leg as the timed CPU path.
Arrays are float32 numpy, single instance: density [D, H, W] (or
[D, H, W, 1]), faces [D, H, W+1], [D, H+1, W], [D+1, H, W]; matrices in the op
layout (16 floats, column-major); images [V, H', W', 4]; targets
[V, H', W', 3].
"""
import ctypes
import os

import numpy as np

from . import build_cpu_port

_LIB = None
_fp = ctypes.POINTER(ctypes.c_float)


def load():
    global _LIB
    if _LIB is None:
        if not os.path.exists(build_cpu_port.LIB):
            raise RuntimeError("CPU port not built: run python -m oracle.build_cpu_port (or __graft_entry__.build())")
        lib = ctypes.CDLL(build_cpu_port.LIB)
        i, f = ctypes.c_int, ctypes.c_float
        lib.ngtcpu_threads.restype = i
        lib.ngtcpu_advect_mc_fwd.argtypes = [_fp, i, i, i, _fp, _fp, _fp, f, _fp, _fp]
        lib.ngtcpu_advect_mc_bwd.argtypes = [_fp, i, i, i, _fp, _fp, _fp, f, _fp, _fp, _fp, _fp, _fp, _fp, _fp]
        lib.ngtcpu_render_fwd.argtypes = [_fp, i, i, i, _fp, _fp, _fp, i, i, i, i, _fp, f, _fp]
        lib.ngtcpu_l2_loss.argtypes = [_fp, _fp, ctypes.c_int64, f]
        lib.ngtcpu_l2_loss.restype = ctypes.c_double
        lib.ngtcpu_render_l2_bwd.argtypes = [_fp, i, i, i, _fp, _fp, _fp, i, i, i, i, _fp, f, _fp, _fp, f, f, _fp]
        _LIB = lib
    return _LIB


def _f(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _p(a):
    return a.ctypes.data_as(_fp)


def threads():
    return load().ngtcpu_threads()


def advect_mc_fwd(rho, vx, vy, vz, ts):
    """Returns (out, sl), each [D, H, W]."""
    lib = load()
    rho = _f(rho).reshape(np.shape(rho)[:3])
    D, H, W = rho.shape
    vx, vy, vz = (_f(a) for a in (vx, vy, vz))
    sl = np.empty_like(rho)
    out = np.empty_like(rho)
    lib.ngtcpu_advect_mc_fwd(_p(rho), D, H, W, _p(vx), _p(vy), _p(vz), float(ts), _p(sl), _p(out))
    return out, sl


def advect_mc_bwd(rho, vx, vy, vz, ts, sl, g_out):
    """Returns (g_rho [D,H,W], g_vx, g_vy, g_vz) for the output gradient g_out."""
    lib = load()
    rho = _f(rho).reshape(np.shape(rho)[:3])
    D, H, W = rho.shape
    vx, vy, vz, sl, g_out = (_f(a) for a in (vx, vy, vz, sl, g_out))
    g_rho = np.empty_like(rho)
    gx, gy, gz = (np.empty(np.shape(a), np.float32) for a in (vx, vy, vz))
    work = np.empty(4 * D * H * W, np.float32)
    lib.ngtcpu_advect_mc_bwd(_p(rho), D, H, W, _p(vx), _p(vy), _p(vz), float(ts), _p(sl), _p(g_out), _p(g_rho),
                             _p(gx), _p(gy), _p(gz), _p(work))
    return g_rho, gx, gy, gz


def render_fwd(rho, M, V, P, out_dhw, light_rgb, depth_step):
    """images [V, H', W', 4]; M one op matrix, V / P per view."""
    lib = load()
    rho = _f(rho).reshape(np.shape(rho)[:3])
    D, H, W = rho.shape
    M, V, P = _f(M).reshape(16), _f(V).reshape(-1, 16), _f(P).reshape(-1, 16)
    nv = V.shape[0]
    OD, OH, OW = (int(v) for v in out_dhw)
    img = np.empty((nv, OH, OW, 4), np.float32)
    light = _f(light_rgb).reshape(3)
    lib.ngtcpu_render_fwd(_p(rho), D, H, W, _p(M), _p(V), _p(P), nv, OD, OH, OW, _p(light), float(depth_step),
                          _p(img))
    return img


def l2_loss(images, target, loss_weight=1.0):
    images, target = _f(images), _f(target)
    return load().ngtcpu_l2_loss(_p(images), _p(target), images.size // 4, float(loss_weight))


def render_l2_bwd(rho, M, V, P, out_dhw, light_rgb, depth_step, images, target, loss_weight=1.0, loss_grad=1.0):
    """Density gradient [D, H, W] of loss_grad * loss_weight * sum((img - target)^2)."""
    lib = load()
    rho = _f(rho).reshape(np.shape(rho)[:3])
    D, H, W = rho.shape
    M, V, P = _f(M).reshape(16), _f(V).reshape(-1, 16), _f(P).reshape(-1, 16)
    nv = V.shape[0]
    OD, OH, OW = (int(v) for v in out_dhw)
    images, target = _f(images), _f(target)
    light = _f(light_rgb).reshape(3)
    g = np.empty_like(rho)
    lib.ngtcpu_render_l2_bwd(_p(rho), D, H, W, _p(M), _p(V), _p(P), nv, OD, OH, OW, _p(light), float(depth_step),
                             _p(images), _p(target), float(loss_weight), float(loss_grad), _p(g))
    return g
