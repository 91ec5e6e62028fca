"""ctypes binding of libngt_transport.so (the C ABI in include/ngt_transport.h).

The library is the only compute path: if it is missing or a tensor is not on
a HIP device the call raises; there is no CPU fallback.
"""
import ctypes
import os
import threading

import numpy as np
import torch

from . import _build

_LIB = None
_LOCK = threading.Lock()

_c_i64p = ctypes.POINTER(ctypes.c_int64)
_c_i32p = ctypes.POINTER(ctypes.c_int32)
_c_fp = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p
_i = ctypes.c_int
_f = ctypes.c_float
_sz = ctypes.c_size_t

# name -> (restype, argtypes)
_SIGNATURES = {
    "ngt_last_error": (ctypes.c_char_p, []),
    "ngt_version": (_i, []),
    "ngt_sample_grid_lut_fwd_workspace_size": (_sz, [_c_i64p, _i, _i]),
    "ngt_sample_grid_lut_fwd": (_i, [_vp, _c_i64p, _vp, _c_i64p, _i, _i, _i, _i, _f, _f, _i, _i, _i, _vp, _vp, _sz,
                                     _vp]),
    "ngt_sample_grid_lut_bwd_workspace_size": (_sz, [_c_i64p, _c_i64p, _i, _i, _i]),
    "ngt_sample_grid_lut_bwd": (_i, [_vp, _c_i64p, _vp, _vp, _c_i64p, _i, _i, _i, _i, _f, _f, _i, _i, _i, _i, _vp, _vp,
                                     _vp, _sz, _vp]),
    "ngt_sample_grid_transform_fwd_workspace_size": (_sz, [_c_i64p, _i, _i]),
    "ngt_sample_grid_transform_fwd": (_i, [_vp, _c_i64p, _c_fp, _c_fp, _c_fp, _c_fp, _i, _c_i32p, _i, _i, _i, _i, _f,
                                           _i, _f, _i, _vp, _vp, _sz, _vp]),
    "ngt_sample_grid_transform_bwd_workspace_size": (_sz, [_c_i64p, _i, _i, _i]),
    "ngt_sample_grid_transform_bwd": (_i, [_vp, _c_i64p, _vp, _c_fp, _c_fp, _c_fp, _c_fp, _i, _c_i32p, _i, _i, _i, _i,
                                           _f, _i, _f, _i, _i, _vp, _vp, _sz, _vp]),
    "ngt_lod_transform": (_i, [_c_i32p, _c_fp, _c_fp, _c_fp, _i, _c_i32p, _i, _i, _i, _vp, _vp]),
    "ngt_reduce_grid_blend_fwd": (_i, [_vp, _c_i64p, _i, _i, _vp, _vp]),
    "ngt_reduce_grid_blend_bwd": (_i, [_vp, _vp, _vp, _c_i64p, _i, _i, _i, _vp, _vp]),
    "ngt_raymarch_grid_transform_fwd_workspace_size": (_sz, [_c_i64p, _i]),
    "ngt_raymarch_grid_transform_fwd": (_i, [_vp, _c_i64p, _c_fp, _c_fp, _c_fp, _c_fp, _i, _c_i32p, _i, _i, _i, _i,
                                             _vp, _vp, _sz, _vp]),
    "ngt_raymarch_step_cap": (_i, [_c_i64p, _c_fp, _c_fp, _c_fp, _c_fp, _i, _c_i32p, _i]),
    "ngt_raymarch_grid_transform_bwd_workspace_size": (_sz, [_c_i64p, _c_fp, _c_fp, _c_fp, _c_fp, _i, _c_i32p, _i, _i,
                                                            _i, _i]),
    "ngt_raymarch_grid_transform_bwd": (_i, [_vp, _c_i64p, _vp, _vp, _c_fp, _c_fp, _c_fp, _c_fp, _i, _c_i32p, _i, _i,
                                             _i, _i, _i, _vp, _vp, _sz, _vp]),
    "ngt_render_density_fwd_workspace_size": (_sz, [_c_i64p, _i]),
    "ngt_render_density_fwd": (_i, [_vp, _c_i64p, _c_fp, _f, _c_fp, _c_fp, _c_fp, _c_fp, _i, _c_i32p, _i, _i, _i,
                                    _vp, _vp, _sz, _vp]),
    "ngt_render_density_bwd_workspace_size": (_sz, [_c_i64p, _c_fp, _c_fp, _c_fp, _c_fp, _i, _c_i32p, _i, _i]),
    "ngt_render_density_bwd": (_i, [_vp, _c_i64p, _c_fp, _f, _vp, _vp, _c_fp, _c_fp, _c_fp, _c_fp, _i, _c_i32p, _i,
                                    _i, _i, _vp, _vp, _sz, _vp]),
    "ngt_render_density_l2_fwd_workspace_size": (_sz, [_c_i64p, _i, _i, _c_i32p]),
    "ngt_bricked_grid_bytes": (_sz, [_c_i64p]),
    "ngt_bricked_grid": (_i, [_vp, _c_i64p, _i, _vp, _vp]),
    "ngt_render_density_l2_fwd": (_i, [_vp, _c_i64p, _c_fp, _f, _c_fp, _c_fp, _c_fp, _c_fp, _i, _c_i32p, _i, _i, _i,
                                       _vp, _f, _vp, _vp, _sz, _vp, _vp, _vp, _sz, _vp]),
    "ngt_static_scratch_bytes": (_sz, []),
    "ngt_render_density_l2_fwd_noalloc": (_i, [_vp, _c_i64p, _c_fp, _f, _c_fp, _c_fp, _c_fp, _c_fp, _i, _c_i32p, _i, _i,
                                               _i, _vp, _f, _vp, _vp, _sz, _vp, _vp, _vp]),
    "ngt_render_density_l2_bwd_workspace_size": (_sz, [_c_i64p, _c_fp, _c_fp, _c_fp, _c_fp, _i, _c_i32p, _i, _i]),
    "ngt_render_density_l2_record_size": (_sz, [_c_i64p, _c_fp, _c_fp, _c_fp, _c_fp, _i, _c_i32p]),
    "ngt_render_density_l2_bwd": (_i, [_vp, _c_i64p, _c_fp, _f, _vp, _vp, _f, _vp, _c_fp, _c_fp, _c_fp, _c_fp, _i,
                                       _c_i32p, _i, _i, _i, _vp, _i, _vp, _vp, _sz, _vp]),
    "ngt_advect_workspace_size": (_sz, [_c_i64p, _i, _i, _i, _i]),
    "ngt_advect_fwd": (_i, [_vp, _c_i64p, _vp, _vp, _vp, _i, _f, _i, _i, _i, _i, _vp, _vp, _vp, _sz, _vp]),
    "ngt_advect_bwd": (_i, [_vp, _c_i64p, _vp, _vp, _vp, _i, _f, _i, _i, _i, _i, _vp, _vp, _i, _vp, _vp, _vp, _vp,
                            _vp, _sz, _vp]),
}



class AdvectInstance(ctypes.Structure):
    """ngt_advect_instance (include/ngt_transport.h)."""
    _fields_ = [("dhw", ctypes.c_int32 * 3), ("timestep", ctypes.c_float), ("input", _vp), ("vel_x", _vp),
                ("vel_y", _vp), ("vel_z", _vp), ("output", _vp), ("sl", _vp), ("output_grad", _vp),
                ("input_grad", _vp), ("vel_x_grad", _vp), ("vel_y_grad", _vp), ("vel_z_grad", _vp)]


class RenderInstance(ctypes.Structure):
    """ngt_render_instance (include/ngt_transport.h)."""
    _fields_ = [("dhw", ctypes.c_int32 * 3), ("density", _vp), ("num_views", ctypes.c_int32), ("matrix_m", _c_fp),
                ("matrix_v", _c_fp), ("matrix_p", _c_fp), ("frustum", _c_fp), ("image_dhw", ctypes.c_int32 * 3),
                ("light_rgb", ctypes.c_float * 3), ("depth_step", ctypes.c_float), ("images", _vp),
                ("images_grad", _vp), ("density_grad", _vp), ("l2_target", _vp), ("loss", _vp)]


_szp = ctypes.POINTER(ctypes.c_size_t)
_SIGNATURES.update({
    "ngt_advect_batch_plan_bytes": (_sz, [_i]),
    "ngt_advect_batch_plan": (_i, [ctypes.POINTER(AdvectInstance), _i, _i, _i, _i, _vp, _sz, _szp]),
    "ngt_advect_batch_fwd": (_i, [_vp, _vp, _i, _vp]),
    "ngt_advect_batch_bwd": (_i, [_vp, _vp, _vp, _sz, _i, _vp]),
    "ngt_render_batch_plan_bytes": (_sz, [_i]),
    "ngt_render_batch_plan": (_i, [ctypes.POINTER(RenderInstance), _i, _i, _sz, _vp, _sz, _szp]),
    "ngt_render_batch_plan_info": (_i, [_vp, ctypes.POINTER(ctypes.c_int), _szp]),
    "ngt_render_batch_fwd": (_i, [_vp, _vp, _vp, _sz, _i, _vp]),
    "ngt_render_batch_bwd": (_i, [_vp, _vp, _vp, _sz, _i, _i, _vp]),
    "ngt_render_batch_l2": (_i, [_vp, _vp, _vp, _sz, _i, _i, _vp]),
})

SYMBOLS = tuple(_SIGNATURES)


class NGTError(RuntimeError):
    """A non-zero status from the HIP library; ``code`` is the NGT_ERR_* value."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class NGTInvalidArgument(NGTError, ValueError):
    pass


class NGTUnimplemented(NGTError, NotImplementedError):
    pass


def lib_path():
    return os.environ.get("NGT_LIBRARY", _build.LIB_PATH)


def load():
    """Load (once) and return the ctypes library.  Raises if it is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    with _LOCK:
        if _LIB is None:
            path = lib_path()
            if not os.path.exists(path):
                raise RuntimeError("libngt_transport.so is not built (%s); run `python -m ngt._build` "
                                   "or __graft_entry__.build()" % path)
            lib = ctypes.CDLL(path)
            for name, (res, args) in _SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _LIB = lib
    return _LIB


def check(rc):
    if rc == 0:
        return
    msg = load().ngt_last_error().decode("utf-8", "replace")
    if rc == 1:
        raise NGTInvalidArgument(rc, msg)
    if rc == 2:
        raise NGTUnimplemented(rc, msg)
    raise NGTError(rc, msg)


# --- argument helpers ------------------------------------------------------------

def shape5(t):
    s = list(t.shape)
    if len(s) != 5:
        raise ValueError("Invalid input shape (NDHWC): %s" % (s,))
    return (ctypes.c_int64 * 5)(*s)


def dhw(shape):
    s = [int(v) for v in shape]
    if len(s) != 3:
        raise ValueError("Invalid output_shape")
    return (ctypes.c_int32 * 3)(*s)


def host_floats(x, per_item):
    """Matrices / frustum params (list of arrays, array or CPU tensor) ->
    (ctypes float pointer, count, keepalive array)."""
    if isinstance(x, torch.Tensor):
        x = x.detach().cpu().numpy()
    if isinstance(x, (list, tuple)):
        x = np.stack([np.asarray(v.detach().cpu().numpy() if isinstance(v, torch.Tensor) else v, dtype=np.float32)
                      .reshape(-1) for v in x])
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32).reshape(-1, per_item))
    return a.ctypes.data_as(_c_fp), a.shape[0], a


def device_tensor(t, name):
    if not isinstance(t, torch.Tensor):
        raise TypeError("%s must be a torch.Tensor" % name)
    if t.device.type != "cuda":
        raise RuntimeError("%s must be on a HIP device (got %s); the NGT ops have no CPU path" % (name, t.device))
    if t.dtype != torch.float32:
        raise TypeError("%s must be float32, is %s" % (name, t.dtype))
    return t.contiguous()


def empty_grad(t):
    """Uninitialised float32 tensor of ``t``'s shape with standard contiguous
    strides.  (``empty_like`` copies zero strides of size-1 dimensions, and
    autograd clones such a gradient instead of taking it as the leaf's .grad.)"""
    return torch.empty(t.shape, dtype=torch.float32, device=t.device)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def stream(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def workspace(nbytes, device):
    if nbytes == 0:
        return None, 0
    return torch.empty(int(nbytes), dtype=torch.uint8, device=device), int(nbytes)
