"""Drop-in replacement of ``phitest.render.cuda.ops_loader`` on PyTorch-ROCm.

Same op names, keyword arguments, string attributes, shapes (NDHWC) and
error behaviour as the TF custom ops (ops_loader.py:14-118,
resample_grid.cc:23-172, raymarch_grid.cc:29-92, reduce_blend.cc:24-55,
advect.cc:21-71).  Every differentiable op is a ``torch.autograd.Function``
whose backward calls the hand-written HIP adjoint kernels of
libngt_transport.so.

``gradient_semantics`` (module attribute, or the ``semantics()`` context
manager) selects the gradients:

* ``"exact"`` (default): the true adjoint of every forward op;
* ``"reference"``: the reference's registered gradients - input gradients of
  the sampling ops divided by per-voxel corner-hit counts, the raymarch
  gradient divided per camera by the summed corner weights on the reference's
  backward lattice, the reference ALPHA blend backward, the LUT gradient of
  the last batch entry.

Matrices and frustum parameters are host arrays (numpy, lists or CPU
tensors), exactly what the reference ops received.
"""
import contextlib

import torch

from . import _lib

gradient_semantics = "exact"

# Exact raymarch / render gradients with BORDER + LINEAR use the atomics-free
# brick-gather adjoint (per-sample gradient pass + per-brick LDS gather); False selects
# the fp32-atomic scatter kernels instead (same result up to summation order).
use_gather_adjoint = True

# Raymarch / render forwards with LINEAR interpolation hand the library a
# workspace for a haloed 4x4x4-brick copy of the grid that the rays read;
# False makes them read the NDHWC grid directly (same result).
use_bricked_layout = True

INTERPOLATION = {"NEAREST": 0, "LINEAR": 1, "MIN": 2, "MAX": 3, "QUADRATIC": 4}
BOUNDARY = {"BORDER": 0, "CLAMP": 1, "WRAP": 2}
MIPMAPPING = {"NONE": 0, "NEAREST": 1, "LINEAR": 2}
COORDINATE_MODE = {"TRANSFORM_LINDEPTH": 0, "TRANSFORM_LINDEPTH_REVERSE": 1, "TRANSFORM": 3, "TRANSFORM_REVERSE": 4}
BLEND = {"BEER_LAMBERT": 0, "ALPHA": 2, "ALPHA_ADDITIVE": 3, "ADDITIVE": 4}
CLAMP = {"NONE": 0, "MC_SMOOTH": 1, "NEGATIVE": 2}


@contextlib.contextmanager
def semantics(mode):
    global gradient_semantics
    if mode not in ("exact", "reference"):
        raise ValueError("gradient semantics must be 'exact' or 'reference'")
    old = gradient_semantics
    gradient_semantics = mode
    try:
        yield
    finally:
        gradient_semantics = old


def _sem():
    return 1 if gradient_semantics == "reference" else 0


def _enum(table, value, what):
    if isinstance(value, bytes):
        value = value.decode()
    v = str(value).upper()
    if v not in table:
        raise ValueError("invalid %s '%s'" % (what, value))
    return table[v]


# --- SampleGridLut ----------------------------------------------------------------

def _lut_attrs(interpolation, boundary, mipmapping, num_mipmaps, mip_bias, coordinate_mode, cell_center_offset,
               separate_camera_batch, relative_coords, normalized_coords):
    if str(coordinate_mode).upper() not in ("LOOKUP", "LUT"):
        raise ValueError("invalid coordinate_mode.")
    interp = _enum(INTERPOLATION, interpolation, "interpolation")
    if interp == 4:
        raise ValueError("invalid interpolation 'QUADRATIC' for SampleGridLut")
    return dict(interpolation=interp, boundary=_enum(BOUNDARY, boundary, "boundary"),
                mipmapping=_enum(MIPMAPPING, mipmapping, "mipmapping"), num_mipmaps=int(num_mipmaps),
                mip_bias=float(mip_bias), cell_center_offset=float(cell_center_offset),
                separate_camera_batch=bool(separate_camera_batch), relative_coords=bool(relative_coords),
                normalized_coords=bool(normalized_coords))


def _lut_fwd(inp, lut, a):
    lib = _lib.load()
    inp = _lib.device_tensor(inp, "input")
    lut = _lib.device_tensor(lut, "lookup")
    N = inp.shape[0]
    if inp.dim() != 5:
        raise ValueError("Invalid input shape (NDHWC): %s" % (list(inp.shape),))
    if lut.dim() != 5 or lut.shape[-1] != 4:
        raise ValueError("Invalid lut shape (NDHWC) with C=4: %s" % (list(lut.shape),))
    V = lut.shape[0] if a["separate_camera_batch"] else 1
    out = torch.empty((N, V) + tuple(lut.shape[1:4]) + (inp.shape[-1],), dtype=torch.float32, device=inp.device)
    s_in = _lib.shape5(inp)
    ws, nbytes = _lib.workspace(lib.ngt_sample_grid_lut_fwd_workspace_size(s_in, a["mipmapping"], a["num_mipmaps"]),
                                inp.device)
    _lib.check(lib.ngt_sample_grid_lut_fwd(
        _lib.ptr(inp), s_in, _lib.ptr(lut), _lib.shape5(lut), a["interpolation"], a["boundary"],
        a["mipmapping"], a["num_mipmaps"], a["mip_bias"], a["cell_center_offset"], int(a["separate_camera_batch"]),
        int(a["relative_coords"]), int(a["normalized_coords"]), _lib.ptr(out), _lib.ptr(ws), nbytes,
        _lib.stream(inp.device)))
    return out


def _lut_bwd(inp, out_grad, lut, a, sem):
    lib = _lib.load()
    inp = _lib.device_tensor(inp, "input")
    lut = _lib.device_tensor(lut, "lookup")
    out_grad = _lib.device_tensor(out_grad, "output_grad")
    if out_grad.dim() != 6:
        raise ValueError("Invalid output_grads shape (NVDHWC): %s" % (list(out_grad.shape),))
    in_grad = _lib.empty_grad(inp)
    lut_grad = _lib.empty_grad(lut)
    s_in, s_lut = _lib.shape5(inp), _lib.shape5(lut)
    ws, nbytes = _lib.workspace(lib.ngt_sample_grid_lut_bwd_workspace_size(s_in, s_lut, a["mipmapping"],
                                                                           a["num_mipmaps"], sem), inp.device)
    _lib.check(lib.ngt_sample_grid_lut_bwd(
        _lib.ptr(inp), s_in, _lib.ptr(out_grad), _lib.ptr(lut), s_lut, a["interpolation"], a["boundary"],
        a["mipmapping"], a["num_mipmaps"], a["mip_bias"], a["cell_center_offset"], int(a["separate_camera_batch"]),
        int(a["relative_coords"]), int(a["normalized_coords"]), sem, _lib.ptr(in_grad), _lib.ptr(lut_grad),
        _lib.ptr(ws), nbytes, _lib.stream(inp.device)))
    return in_grad, lut_grad


class _SampleGridLut(torch.autograd.Function):
    @staticmethod
    def forward(ctx, inp, lut, attrs):
        out = _lut_fwd(inp, lut, attrs)
        ctx.save_for_backward(inp, lut)
        ctx.attrs = attrs
        ctx.sem = _sem()
        return out

    @staticmethod
    def backward(ctx, grad):
        inp, lut = ctx.saved_tensors
        gi, gl = _lut_bwd(inp, grad, lut, ctx.attrs, ctx.sem)
        return gi, gl, None


def sample_grid_lut(input, lookup, interpolation="LINEAR", boundary="BORDER", mipmapping="NONE", num_mipmaps=0,
                    mip_bias=0.0, coordinate_mode="LOOKUP", cell_center_offset=0.0, separate_camera_batch=True,
                    relative_coords=True, normalized_coords=False):
    """SampleGridLut: NDHWC input, VDHW4 lookup -> NVDHWC (resample_grid.cc:94-120)."""
    a = _lut_attrs(interpolation, boundary, mipmapping, num_mipmaps, mip_bias, coordinate_mode, cell_center_offset,
                   separate_camera_batch, relative_coords, normalized_coords)
    return _SampleGridLut.apply(input, lookup, a)


def sample_grid_lut_grad(input, output_grad, lookup, interpolation="LINEAR", boundary="BORDER", mipmapping="NONE",
                         num_mipmaps=0, mip_bias=0.0, coordinate_mode="LOOKUP", cell_center_offset=0.0,
                         separate_camera_batch=True, relative_coords=True, normalized_coords=False):
    """SampleGridLutGrad -> (input_grad, lookup_grad) (resample_grid.cc:124-150)."""
    a = _lut_attrs(interpolation, boundary, mipmapping, num_mipmaps, mip_bias, coordinate_mode, cell_center_offset,
                   separate_camera_batch, relative_coords, normalized_coords)
    return _lut_bwd(input, output_grad, lookup, a, _sem())


# --- SampleGridTransform ------------------------------------------------------------

class _Cams:
    """Host-side camera arguments kept alive for the duration of a call."""

    def __init__(self, matrix_m, matrix_v, matrix_p, frustum_params):
        self.m, self.nm, self._m = _lib.host_floats(matrix_m, 16)
        self.v, self.nv, self._v = _lib.host_floats(matrix_v, 16)
        self.p, self.np_, self._p = _lib.host_floats(matrix_p, 16)
        self.f, self.nf, self._f = _lib.host_floats(frustum_params, 6)
        if not (self.nv == self.np_ == self.nf):
            raise ValueError("camera batch size mismatch")

    def check_batch(self, n):
        if self.nm != n:
            raise ValueError("model matrix batch size mismatch")


def _xform_attrs(interpolation, boundary, mipmapping, num_mipmaps, mip_bias, coordinate_mode, cell_center_offset,
                 separate_camera_batch):
    interp = _enum(INTERPOLATION, interpolation, "interpolation")
    return dict(interpolation=interp, boundary=_enum(BOUNDARY, boundary, "boundary"),
                mipmapping=_enum(MIPMAPPING, mipmapping, "mipmapping"), num_mipmaps=int(num_mipmaps),
                mip_bias=float(mip_bias), coordinate_mode=_enum(COORDINATE_MODE, coordinate_mode, "coordinate_mode"),
                cell_center_offset=float(cell_center_offset), separate_camera_batch=bool(separate_camera_batch))


def _xform_fwd(inp, cams, out_shape, a):
    lib = _lib.load()
    inp = _lib.device_tensor(inp, "input")
    cams.check_batch(inp.shape[0])
    V = cams.nv if a["separate_camera_batch"] else 1
    od = [int(v) for v in out_shape]
    out = torch.empty((inp.shape[0], V) + tuple(od) + (inp.shape[-1],), dtype=torch.float32, device=inp.device)
    s_in = _lib.shape5(inp)
    ws, nbytes = _lib.workspace(lib.ngt_sample_grid_transform_fwd_workspace_size(s_in, a["mipmapping"],
                                                                                 a["num_mipmaps"]), inp.device)
    _lib.check(lib.ngt_sample_grid_transform_fwd(
        _lib.ptr(inp), s_in, cams.m, cams.v, cams.p, cams.f, cams.nv, _lib.dhw(od), a["interpolation"],
        a["boundary"], a["mipmapping"], a["num_mipmaps"], a["mip_bias"], a["coordinate_mode"],
        a["cell_center_offset"], int(a["separate_camera_batch"]), _lib.ptr(out), _lib.ptr(ws), nbytes,
        _lib.stream(inp.device)))
    return out


def _xform_bwd(inp, out_grad, cams, a, sem):
    lib = _lib.load()
    inp = _lib.device_tensor(inp, "input")
    out_grad = _lib.device_tensor(out_grad, "output_grad")
    cams.check_batch(inp.shape[0])
    in_grad = _lib.empty_grad(inp)
    s_in = _lib.shape5(inp)
    ws, nbytes = _lib.workspace(lib.ngt_sample_grid_transform_bwd_workspace_size(s_in, a["mipmapping"],
                                                                                 a["num_mipmaps"], sem), inp.device)
    _lib.check(lib.ngt_sample_grid_transform_bwd(
        _lib.ptr(inp), s_in, _lib.ptr(out_grad), cams.m, cams.v, cams.p, cams.f, cams.nv,
        _lib.dhw(out_grad.shape[2:5]), a["interpolation"], a["boundary"], a["mipmapping"], a["num_mipmaps"],
        a["mip_bias"], a["coordinate_mode"], a["cell_center_offset"], int(a["separate_camera_batch"]), sem,
        _lib.ptr(in_grad), _lib.ptr(ws), nbytes, _lib.stream(inp.device)))
    return in_grad


class _SampleGridTransform(torch.autograd.Function):
    @staticmethod
    def forward(ctx, inp, cams, out_shape, attrs):
        out = _xform_fwd(inp, cams, out_shape, attrs)
        ctx.save_for_backward(inp)
        ctx.cams, ctx.attrs, ctx.sem = cams, attrs, _sem()
        return out

    @staticmethod
    def backward(ctx, grad):
        (inp,) = ctx.saved_tensors
        return _xform_bwd(inp, grad, ctx.cams, ctx.attrs, ctx.sem), None, None, None


def sample_grid_transform(input, matrix_m, matrix_v, matrix_p, frustum_params, output_shape, interpolation="LINEAR",
                          boundary="BORDER", mipmapping="NONE", num_mipmaps=0, mip_bias=0.0,
                          coordinate_mode="TRANSFORM_LINDEPTH", cell_center_offset=0.5, separate_camera_batch=True):
    """SampleGridTransform: NDHWC -> NVDHWC (resample_grid.cc:23-65)."""
    a = _xform_attrs(interpolation, boundary, mipmapping, num_mipmaps, mip_bias, coordinate_mode, cell_center_offset,
                     separate_camera_batch)
    cams = _Cams(matrix_m, matrix_v, matrix_p, frustum_params)
    return _SampleGridTransform.apply(input, cams, tuple(int(v) for v in output_shape), a)


def sample_grid_transform_grad(input, output_grad, matrix_m, matrix_v, matrix_p, frustum_params,
                               interpolation="LINEAR", boundary="BORDER", mipmapping="NONE", num_mipmaps=0,
                               mip_bias=0.0, coordinate_mode="TRANSFORM_LINDEPTH", cell_center_offset=0.5,
                               separate_camera_batch=True):
    """SampleGridTransformGrad (resample_grid.cc:68-92)."""
    a = _xform_attrs(interpolation, boundary, mipmapping, num_mipmaps, mip_bias, coordinate_mode, cell_center_offset,
                     separate_camera_batch)
    return _xform_bwd(input, output_grad, _Cams(matrix_m, matrix_v, matrix_p, frustum_params), a, _sem())


def lod_transform(input_shape, matrix_mv, matrix_p, frustum_params, output_shape, inverse_transform=False,
                  linearize_depth=True, relative_coords=False, device=None):
    """LodTransform (resample_grid.cc:152-172): [V, D', H', W', 4] lookup
    tables (sample position, level of detail) of V cameras, to be read by
    ``sample_grid_lut`` with mipmapping.  ``input_shape`` is the DHW of the
    grid sampled from, ``output_shape`` the DHW sampled to; matrices are host
    arrays (model-view and projection, op layout)."""
    lib = _lib.load()
    mv, nmv, _kmv = _lib.host_floats(matrix_mv, 16)
    p, npr, _kp = _lib.host_floats(matrix_p, 16)
    f, nf, _kf = _lib.host_floats(frustum_params, 6)
    if not (nmv == npr == nf):
        raise ValueError("camera batch size mismatch")
    ind = [int(v) for v in input_shape][-3:]
    od = [int(v) for v in output_shape][-3:]
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    out = torch.empty((nmv,) + tuple(od) + (4,), dtype=torch.float32, device=dev)
    _lib.check(lib.ngt_lod_transform(_lib.dhw(ind), mv, p, f, nmv, _lib.dhw(od), int(bool(inverse_transform)),
                                     int(bool(linearize_depth)), int(bool(relative_coords)), _lib.ptr(out),
                                     _lib.stream(dev)))
    return out


# --- ReduceGridBlend -------------------------------------------------------------

def _blend_fwd(inp, mode, keep_dims):
    lib = _lib.load()
    inp = _lib.device_tensor(inp, "input")
    if inp.dim() != 5:
        raise ValueError("Invalid input shape %s" % (list(inp.shape),))
    shape = tuple(inp.shape) if keep_dims else (inp.shape[0],) + tuple(inp.shape[2:])
    out = torch.empty(shape, dtype=torch.float32, device=inp.device)
    _lib.check(lib.ngt_reduce_grid_blend_fwd(_lib.ptr(inp), _lib.shape5(inp), mode, int(keep_dims), _lib.ptr(out),
                                             _lib.stream(inp.device)))
    return out


def _blend_bwd(out_grad, out, inp, mode, keep_dims, sem):
    lib = _lib.load()
    inp = _lib.device_tensor(inp, "input")
    out = _lib.device_tensor(out, "output")
    out_grad = _lib.device_tensor(out_grad, "output_grad")
    in_grad = _lib.empty_grad(inp)
    _lib.check(lib.ngt_reduce_grid_blend_bwd(_lib.ptr(out_grad), _lib.ptr(out), _lib.ptr(inp), _lib.shape5(inp), mode,
                                             int(keep_dims), sem, _lib.ptr(in_grad), _lib.stream(inp.device)))
    return in_grad


class _ReduceGridBlend(torch.autograd.Function):
    @staticmethod
    def forward(ctx, inp, mode, keep_dims):
        out = _blend_fwd(inp, mode, keep_dims)
        ctx.save_for_backward(inp, out)
        ctx.mode, ctx.keep, ctx.sem = mode, keep_dims, _sem()
        return out

    @staticmethod
    def backward(ctx, grad):
        inp, out = ctx.saved_tensors
        return _blend_bwd(grad, out, inp, ctx.mode, ctx.keep, ctx.sem), None, None


def reduce_grid_blend(input, blending_mode="BEER_LAMBERT", keep_dims=False):
    """ReduceGridBlend: NDHWC -> NHWC (or NDHWC with keep_dims) (reduce_blend.cc:24-34)."""
    return _ReduceGridBlend.apply(input, _enum(BLEND, blending_mode, "blending mode"), bool(keep_dims))


def reduce_grid_blend_grad(output_grad, output, input, blending_mode="BEER_LAMBERT", keep_dims=False):
    """ReduceGridBlendGrad (reduce_blend.cc:37-55)."""
    return _blend_bwd(output_grad, output, input, _enum(BLEND, blending_mode, "blending mode"), bool(keep_dims), _sem())


# --- RaymarchGridTransform -------------------------------------------------------

def _ray_attrs(interpolation, boundary, blending_mode, keep_dims, separate_camera_batch):
    interp = _enum(INTERPOLATION, interpolation, "filter mode")
    return dict(interpolation=interp, boundary=_enum(BOUNDARY, boundary, "boundary mode"),
                blending_mode=_enum(BLEND, blending_mode, "blend mode"), keep_dims=bool(keep_dims),
                separate_camera_batch=bool(separate_camera_batch))


def _ray_out_shape(inp, cams, out_shape, a):
    V = cams.nv if a["separate_camera_batch"] else 1
    s = (inp.shape[0], V, int(out_shape[1]), int(out_shape[2]), inp.shape[-1])
    return s


def _ray_fwd(inp, cams, out_shape, a):
    lib = _lib.load()
    inp = _lib.device_tensor(inp, "input")
    cams.check_batch(inp.shape[0])
    out = torch.empty(_ray_out_shape(inp, cams, out_shape, a), dtype=torch.float32, device=inp.device)
    s_in = _lib.shape5(inp)
    ws, nbytes = (None, 0)
    if use_bricked_layout:
        ws, nbytes = _lib.workspace(lib.ngt_raymarch_grid_transform_fwd_workspace_size(s_in, a["interpolation"]),
                                    inp.device)
    _lib.check(lib.ngt_raymarch_grid_transform_fwd(
        _lib.ptr(inp), s_in, cams.m, cams.v, cams.p, cams.f, cams.nv, _lib.dhw(out_shape),
        a["interpolation"], a["boundary"], a["blending_mode"], int(a["separate_camera_batch"]), _lib.ptr(out),
        _lib.ptr(ws), nbytes, _lib.stream(inp.device)))
    return out


def _ray_bwd(inp, out, out_grad, cams, out_shape, a, sem):
    lib = _lib.load()
    inp = _lib.device_tensor(inp, "input")
    out_shape5 = _ray_out_shape(inp, cams, out_shape, a)
    out = _lib.device_tensor(out, "output").reshape(out_shape5)
    out_grad = _lib.device_tensor(out_grad, "output_grad").reshape(out_shape5)
    cams.check_batch(inp.shape[0])
    in_grad = _lib.empty_grad(inp)
    s_in = _lib.shape5(inp)
    nb = lib.ngt_raymarch_grid_transform_bwd_workspace_size(
        s_in, cams.m, cams.v, cams.p, cams.f, cams.nv, _lib.dhw(out_shape), a["interpolation"], a["boundary"],
        int(a["separate_camera_batch"]), sem)
    if sem == 0 and not use_gather_adjoint:
        nb = 0
    ws, nbytes = _lib.workspace(nb, inp.device)
    _lib.check(lib.ngt_raymarch_grid_transform_bwd(
        _lib.ptr(inp), s_in, _lib.ptr(out), _lib.ptr(out_grad), cams.m, cams.v, cams.p, cams.f, cams.nv,
        _lib.dhw(out_shape), a["interpolation"], a["boundary"], a["blending_mode"], int(a["separate_camera_batch"]),
        sem, _lib.ptr(in_grad), _lib.ptr(ws), nbytes, _lib.stream(inp.device)))
    return in_grad


def _keep(out, a):
    # keep_dims returns NVDHWC with D = 1 (raymarch_grid.cc:290-295)
    return out.unsqueeze(2) if a["keep_dims"] else out


class _RaymarchGridTransform(torch.autograd.Function):
    @staticmethod
    def forward(ctx, inp, cams, out_shape, attrs):
        out = _ray_fwd(inp, cams, out_shape, attrs)
        ctx.save_for_backward(inp, out)
        ctx.cams, ctx.out_shape, ctx.attrs, ctx.sem = cams, out_shape, attrs, _sem()
        return _keep(out, attrs)

    @staticmethod
    def backward(ctx, grad):
        inp, out = ctx.saved_tensors
        return _ray_bwd(inp, out, grad, ctx.cams, ctx.out_shape, ctx.attrs, ctx.sem), None, None, None


def raymarch_grid_transform(input, matrix_m, matrix_v, matrix_p, frustum_params, output_shape,
                            interpolation="LINEAR", boundary="BORDER", blending_mode="BEER_LAMBERT", keep_dims=False,
                            separate_camera_batch=True):
    """RaymarchGridTransform: NDHWC -> NVHWC (raymarch_grid.cc:29-63)."""
    a = _ray_attrs(interpolation, boundary, blending_mode, keep_dims, separate_camera_batch)
    cams = _Cams(matrix_m, matrix_v, matrix_p, frustum_params)
    return _RaymarchGridTransform.apply(input, cams, tuple(int(v) for v in output_shape), a)


class _RaymarchGridTransformGrad(torch.autograd.Function):
    """The gradient op as a differentiable function of ``output_grad``: for
    ADDITIVE blending it is linear in output_grad and its own gradient is the
    forward raymarch (ops_loader.py:98-118)."""

    @staticmethod
    def forward(ctx, inp, out, out_grad, cams, out_shape, attrs):
        ctx.cams, ctx.out_shape, ctx.attrs, ctx.sem = cams, out_shape, attrs, _sem()
        ctx.grad_shape = out_grad.shape
        return _ray_bwd(inp, out, out_grad, cams, out_shape, attrs, ctx.sem)

    @staticmethod
    def backward(ctx, grad):
        if ctx.attrs["blending_mode"] != BLEND["ADDITIVE"]:
            raise NotImplementedError("Gradient of raymach op gradient only available for ADDITIVE blending")
        gg = _ray_fwd(grad, ctx.cams, ctx.out_shape, ctx.attrs).reshape(ctx.grad_shape)
        return None, None, gg, None, None, None


def raymarch_grid_transform_grad(input, output, output_grad, matrix_m, matrix_v, matrix_p, frustum_params,
                                 output_shape, interpolation="LINEAR", boundary="BORDER",
                                 blending_mode="BEER_LAMBERT", keep_dims=False, separate_camera_batch=True):
    """RaymarchGridTransformGrad (raymarch_grid.cc:67-92); differentiable
    w.r.t. output_grad for ADDITIVE blending (used for inverse raymarching,
    renderer.py:372-398)."""
    a = _ray_attrs(interpolation, boundary, blending_mode, keep_dims, separate_camera_batch)
    cams = _Cams(matrix_m, matrix_v, matrix_p, frustum_params)
    return _RaymarchGridTransformGrad.apply(input, output, output_grad, cams, tuple(int(v) for v in output_shape), a)


# --- fused density render -----------------------------------------------------------

def _render_fwd(density, light_rgb, depth_step, cams, out_shape, interp, boundary, mode):
    lib = _lib.load()
    density = _lib.device_tensor(density, "density")
    cams.check_batch(density.shape[0])
    out = torch.empty((density.shape[0], cams.nv, int(out_shape[1]), int(out_shape[2]), 4), dtype=torch.float32,
                      device=density.device)
    lp, _, keep = _lib.host_floats(light_rgb, 3)
    s_in = _lib.shape5(density)
    ws, nbytes = (None, 0)
    if use_bricked_layout:
        ws, nbytes = _lib.workspace(lib.ngt_render_density_fwd_workspace_size(s_in, interp), density.device)
    _lib.check(lib.ngt_render_density_fwd(
        _lib.ptr(density), s_in, lp, float(depth_step), cams.m, cams.v, cams.p, cams.f, cams.nv,
        _lib.dhw(out_shape), interp, boundary, mode, _lib.ptr(out), _lib.ptr(ws), nbytes, _lib.stream(density.device)))
    return out


def _render_bwd(density, light_rgb, depth_step, out, out_grad, cams, out_shape, interp, boundary, mode):
    lib = _lib.load()
    density = _lib.device_tensor(density, "density")
    out = _lib.device_tensor(out, "output")
    out_grad = _lib.device_tensor(out_grad, "output_grad")
    g = _lib.empty_grad(density)
    lp, _, keep = _lib.host_floats(light_rgb, 3)
    s_d = _lib.shape5(density)
    nb = lib.ngt_render_density_bwd_workspace_size(s_d, cams.m, cams.v, cams.p, cams.f, cams.nv,
                                                   _lib.dhw(out_shape), interp, boundary) if use_gather_adjoint else 0
    ws, nbytes = _lib.workspace(nb, density.device)
    _lib.check(lib.ngt_render_density_bwd(
        _lib.ptr(density), s_d, lp, float(depth_step), _lib.ptr(out), _lib.ptr(out_grad), cams.m,
        cams.v, cams.p, cams.f, cams.nv, _lib.dhw(out_shape), interp, boundary, mode, _lib.ptr(g),
        _lib.ptr(ws), nbytes, _lib.stream(density.device)))
    return g


class _RenderDensity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, density, light_rgb, depth_step, cams, out_shape, interp, boundary, mode):
        out = _render_fwd(density, light_rgb, depth_step, cams, out_shape, interp, boundary, mode)
        ctx.save_for_backward(density, out)
        ctx.args = (light_rgb, depth_step, cams, out_shape, interp, boundary, mode)
        return out

    @staticmethod
    def backward(ctx, grad):
        density, out = ctx.saved_tensors
        light_rgb, depth_step, cams, out_shape, interp, boundary, mode = ctx.args
        g = _render_bwd(density, light_rgb, depth_step, out, grad, cams, out_shape, interp, boundary, mode)
        return g, None, None, None, None, None, None, None


def render_density_fused(density, light_rgb, depth_step, matrix_m, matrix_v, matrix_p, frustum_params, output_shape,
                         interpolation="LINEAR", boundary="BORDER", blending_mode="BEER_LAMBERT"):
    """Raymarch of ``concat([density*light_rgb, density]) * depth_step`` without
    materialising the 4-channel grid; NDHW1 -> NVHW4 (RGB, optical depth)."""
    interp = _enum(INTERPOLATION, interpolation, "filter mode")
    cams = _Cams(matrix_m, matrix_v, matrix_p, frustum_params)
    return _RenderDensity.apply(density, tuple(float(c) for c in light_rgb), float(depth_step), cams,
                                tuple(int(v) for v in output_shape), interp, _enum(BOUNDARY, boundary, "boundary"),
                                _enum(BLEND, blending_mode, "blend mode"))


class _RenderDensityL2(torch.autograd.Function):
    """Fused render + L2 image loss: returns (loss, images); only the loss
    carries a gradient."""

    @staticmethod
    def forward(ctx, density, target, light_rgb, depth_step, cams, out_shape, interp, boundary, mode, weight):
        lib = _lib.load()
        density = _lib.device_tensor(density, "density")
        target = _lib.device_tensor(target, "target")
        cams.check_batch(density.shape[0])
        img_shape = (density.shape[0], cams.nv, int(out_shape[1]), int(out_shape[2]))
        if tuple(target.shape) != img_shape + (3,):
            raise ValueError("target must be [N,V,H,W,3] = %s, is %s" % (list(img_shape + (3,)), list(target.shape)))
        out = torch.empty(img_shape + (4,), dtype=torch.float32, device=density.device)
        loss = torch.empty((), dtype=torch.float32, device=density.device)
        lp, _, keep = _lib.host_floats(light_rgb, 3)
        s_d = _lib.shape5(density)
        # one haloed brick copy of the density serves the forward and the backward
        bricked = None
        if interp == INTERPOLATION["LINEAR"] and use_bricked_layout:
            nb = lib.ngt_bricked_grid_bytes(s_d)
            if nb > 0:
                bricked, _ = _lib.workspace(nb, density.device)
                _lib.check(lib.ngt_bricked_grid(_lib.ptr(density), s_d, boundary, _lib.ptr(bricked),
                                                _lib.stream(density.device)))
        ws, nbytes = _lib.workspace(lib.ngt_render_density_l2_fwd_workspace_size(s_d, interp, cams.nv,
                                                                                 _lib.dhw(out_shape)), density.device)
        # with a gradient to come, the forward also records the gather adjoint's
        # rays and per-sample gradients (LINEAR, BORDER)
        record, rbytes = None, 0
        if (ctx.needs_input_grad[0] and use_gather_adjoint and bricked is not None
                and boundary == BOUNDARY["BORDER"]):
            rb = lib.ngt_render_density_l2_record_size(s_d, cams.m, cams.v, cams.p, cams.f, cams.nv,
                                                       _lib.dhw(out_shape))
            if rb > 0:
                record, rbytes = _lib.workspace(rb, density.device)
        _lib.check(lib.ngt_render_density_l2_fwd(
            _lib.ptr(density), s_d, lp, float(depth_step), cams.m, cams.v, cams.p, cams.f, cams.nv,
            _lib.dhw(out_shape), interp, boundary, mode, _lib.ptr(target), float(weight),
            None if bricked is None else _lib.ptr(bricked), None if record is None else _lib.ptr(record), rbytes,
            _lib.ptr(out), _lib.ptr(loss), _lib.ptr(ws), nbytes, _lib.stream(density.device)))
        # the brick copy and the recorded gradients are saved like the inputs,
        # so that a checkpointed caller drops them with everything else and
        # the recomputed forward makes them again
        ctx.save_for_backward(density, target, out, bricked, record)
        ctx.record_bytes = rbytes
        ctx.args = (light_rgb, depth_step, cams, out_shape, interp, boundary, mode, weight)
        ctx.mark_non_differentiable(out)
        ctx.set_materialize_grads(False)  # no zero-filled gradient for the images
        return loss, out

    @staticmethod
    def backward(ctx, grad_loss, grad_out):
        density, target, out, bricked, record = ctx.saved_tensors
        light_rgb, depth_step, cams, out_shape, interp, boundary, mode, weight = ctx.args
        lib = _lib.load()
        if grad_loss is None:
            return None, None, None, None, None, None, None, None, None, None
        gl = _lib.device_tensor(grad_loss.reshape(()).to(torch.float32), "loss_grad").contiguous()
        g = _lib.empty_grad(density)
        lp, _, keep = _lib.host_floats(light_rgb, 3)
        s_d = _lib.shape5(density)
        rbytes = ctx.record_bytes
        if record is not None:
            ws, nbytes = record, rbytes
        else:
            if use_gather_adjoint:
                nb = lib.ngt_render_density_l2_bwd_workspace_size(s_d, cams.m, cams.v, cams.p, cams.f, cams.nv,
                                                                  _lib.dhw(out_shape), interp, boundary)
            else:
                nb = out.numel() * 4
            ws, nbytes = _lib.workspace(nb, density.device)
        _lib.check(lib.ngt_render_density_l2_bwd(
            _lib.ptr(density), s_d, lp, float(depth_step), _lib.ptr(out), _lib.ptr(target), float(weight),
            _lib.ptr(gl), cams.m, cams.v, cams.p, cams.f, cams.nv, _lib.dhw(out_shape), interp, boundary, mode,
            None if bricked is None else _lib.ptr(bricked), 0 if record is None else 1, _lib.ptr(g), _lib.ptr(ws),
            nbytes, _lib.stream(density.device)))
        return g, None, None, None, None, None, None, None, None, None


def render_density_fused_l2(density, target, light_rgb, depth_step, matrix_m, matrix_v, matrix_p, frustum_params,
                            output_shape, loss_weight=1.0, interpolation="LINEAR", boundary="BORDER",
                            blending_mode="BEER_LAMBERT"):
    """``render_density_fused`` followed by the squared-error image loss
    ``loss_weight * sum((images[..., :3] - target)^2)`` (the density target
    loss, optimization.py:478-526) in one op: returns (loss, images).  The
    backward forms the image gradient inside the render adjoint."""
    interp = _enum(INTERPOLATION, interpolation, "filter mode")
    cams = _Cams(matrix_m, matrix_v, matrix_p, frustum_params)
    return _RenderDensityL2.apply(density, target, tuple(float(c) for c in light_rgb), float(depth_step), cams,
                                  tuple(int(v) for v in output_shape), interp, _enum(BOUNDARY, boundary, "boundary"),
                                  _enum(BLEND, blending_mode, "blend mode"), float(loss_weight))


# --- AdvectGridSemiLangrange -------------------------------------------------------

def _adv_layout(velocity):
    """velocity: centered [V,D,H,W,3] tensor or (x, y, z) staggered tuple."""
    if isinstance(velocity, (tuple, list)):
        vx, vy, vz = (_lib.device_tensor(v, "velocity") for v in velocity)
        return vx, vy, vz, vx.shape[0]
    v = _lib.device_tensor(velocity, "velocity_centered")
    if v.dim() != 5 or v.shape[-1] != 3:
        raise ValueError("velocity_centered must be [V,D,H,W,3], is %s" % (list(v.shape),))
    return v, None, None, v.shape[0]


def _check_vel_shape(inp, vx, vy, vz, order):
    D, H, W = inp.shape[1:4]
    if vy is None:
        if tuple(vx.shape[1:4]) != (D, H, W):
            raise ValueError("Shape mismatch in centered warp: data %s, velocity %s"
                             % (list(inp.shape[1:4]), list(vx.shape[1:4])))
    else:
        if (tuple(vx.shape[1:4]) != (D, H, W + 1) or tuple(vy.shape[1:4]) != (D, H + 1, W)
                or tuple(vz.shape[1:4]) != (D + 1, H, W)):
            raise ValueError("shapes of components do not fit. x: %s, y: %s, z: %s"
                             % (list(vx.shape), list(vy.shape), list(vz.shape)))


def _adv_fwd(inp, vel, ts, order, clamp, boundary, separate, keep_sl):
    lib = _lib.load()
    inp = _lib.device_tensor(inp, "input")
    vx, vy, vz, V = _adv_layout(vel)
    _check_vel_shape(inp, vx, vy, vz, order)
    N = inp.shape[0]
    out = torch.empty((N, V if separate else 1) + tuple(inp.shape[1:]), dtype=torch.float32, device=inp.device)
    sl = torch.empty_like(out) if (order == 2 and keep_sl) else None
    s_in = _lib.shape5(inp)
    ws, nbytes = (None, 0)
    if order == 2 and sl is None:
        ws, nbytes = _lib.workspace(lib.ngt_advect_workspace_size(s_in, V, int(separate), order, 0), inp.device)
    _lib.check(lib.ngt_advect_fwd(
        _lib.ptr(inp), s_in, _lib.ptr(vx), _lib.ptr(vy), _lib.ptr(vz), V, float(ts), order, clamp, boundary,
        int(separate), _lib.ptr(out), _lib.ptr(sl), _lib.ptr(ws), nbytes, _lib.stream(inp.device)))
    return out, sl


def _adv_bwd(inp, vel, g, sl, ts, order, clamp, boundary, separate, sem):
    lib = _lib.load()
    inp = _lib.device_tensor(inp, "input")
    vx, vy, vz, V = _adv_layout(vel)
    g = _lib.device_tensor(g, "output_grad")
    s_in = _lib.shape5(inp)
    gi = _lib.empty_grad(inp)
    gvx = _lib.empty_grad(vx)
    gvy = _lib.empty_grad(vy) if vy is not None else None
    gvz = _lib.empty_grad(vz) if vz is not None else None
    ws, nbytes = _lib.workspace(lib.ngt_advect_workspace_size(s_in, V, int(separate), order, 1), inp.device)
    _lib.check(lib.ngt_advect_bwd(
        _lib.ptr(inp), s_in, _lib.ptr(vx), _lib.ptr(vy), _lib.ptr(vz), V, float(ts), order, clamp, boundary,
        int(separate), _lib.ptr(g), _lib.ptr(sl), sem, _lib.ptr(gi), _lib.ptr(gvx), _lib.ptr(gvy), _lib.ptr(gvz),
        _lib.ptr(ws), nbytes, _lib.stream(inp.device)))
    return gi, gvx, gvy, gvz


class _Advect(torch.autograd.Function):
    @staticmethod
    def forward(ctx, inp, ts, order, clamp, boundary, separate, vx, vy, vz):
        vel = vx if vy is None else (vx, vy, vz)
        out, sl = _adv_fwd(inp, vel, ts, order, clamp, boundary, separate, keep_sl=True)
        ctx.save_for_backward(inp, vx, vy, vz, sl)
        ctx.args = (ts, order, clamp, boundary, separate)
        ctx.sem = _sem()
        return out

    @staticmethod
    def backward(ctx, grad):
        inp, vx, vy, vz, sl = ctx.saved_tensors
        ts, order, clamp, boundary, separate = ctx.args
        vel = vx if vy is None else (vx, vy, vz)
        gi, gvx, gvy, gvz = _adv_bwd(inp, vel, grad, sl, ts, order, clamp, boundary, separate, ctx.sem)
        return gi, None, None, None, None, None, gvx, gvy, gvz


def advect_grid_semi_langrange(input, velocity_centered, timestep=1.0, order=1, clamp_extrema="NONE",
                               boundary="BORDER", separate_velocity_batch=True):
    """AdvectGridSemiLangrange (advect.cc:21-71) with VelocityGrid.warp
    semantics (data_structures.py:1690-1737): output[n,v] = warp of input[n] by
    velocity v sampling at ``index + timestep * v`` (timestep = -dt advects
    forward in time); order 2 = MacCormack; clamp_extrema NONE | MC_SMOOTH
    (revert to the semi-Lagrangian value outside the 8-corner min/max) |
    NEGATIVE.  ``velocity_centered`` may also be a (x, y, z) tuple of MAC
    face arrays.  Output NVDHWC."""
    if int(order) not in (1, 2):
        raise ValueError("Unsupported warp order '{}'".format(order))
    clamp = str(clamp_extrema).upper()
    if clamp in ("TRUE", "1"):
        clamp = "MC_SMOOTH"
    if clamp in ("FALSE", "0"):
        clamp = "NONE"
    if clamp == "MC":
        raise NotImplementedError("MIM and MAX warp sampling have wrong gradients.")
    if isinstance(velocity_centered, (tuple, list)):
        vx, vy, vz = velocity_centered
    else:
        vx, vy, vz = velocity_centered, None, None
    return _Advect.apply(input, float(timestep), int(order), _enum(CLAMP, clamp, "clamp"),
                         _enum(BOUNDARY, boundary, "boundary"), bool(separate_velocity_batch), vx, vy, vz)


# --- namespaces mirroring ops_loader ------------------------------------------------

class _Namespace:
    def __init__(self, **fns):
        self.__dict__.update(fns)


sampling_ops = _Namespace(sample_grid_transform=sample_grid_transform, sample_grid_transform_grad=sample_grid_transform_grad,
                          sample_grid_lut=sample_grid_lut, sample_grid_lut_grad=sample_grid_lut_grad)
blending_ops = _Namespace(reduce_grid_blend=reduce_grid_blend, reduce_grid_blend_grad=reduce_grid_blend_grad)
raymarching_ops = _Namespace(raymarch_grid_transform=raymarch_grid_transform,
                             raymarch_grid_transform_grad=raymarch_grid_transform_grad)
advection_ops = _Namespace(advect_grid_semi_langrange=advect_grid_semi_langrange)
