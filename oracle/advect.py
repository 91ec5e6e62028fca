"""Semi-Lagrangian / MacCormack advection with the MC_SMOOTH revert clamp (FP64).

Test infrastructure only (see ``oracle/__init__.py``).

Semantics of ``VelocityGrid.warp`` (data_structures.py:1690-1737) on top of
``Renderer._sample_LuT(..., combined_batch=True, relative=True)``
(renderer.py:559-574, cell_center_offset 0):

    lut  = -dt * v_centered                      (centered_lut_grid, :1684-1688)
    fwd  = S(rho, i + lut)
    order 2:  back = S(fwd, i - lut)
              corr = fwd + 0.5 * (rho - back)
              MC_SMOOTH: mn = S_MIN(rho, i + lut), mx = S_MAX(rho, i + lut)
                         corr = where(corr < mn or corr > mx, fwd, corr)
    NEGATIVE: out = max(out, 0)

Staggered (MAC) velocity: ``v_centered`` averages the two faces of each cell
(VelocityGrid.centered, :1634-1653); component shapes x [D,H,W+1],
y [D,H+1,W], z [D+1,H,W] (:765-773).  Velocity self-advection
(``VelocityGrid.warped`` with ``vel_grid=None``, :1108-1164) interpolates the
other two components onto each component's faces with constant +-0.5 offsets
(``_sampled_to_component_shape`` :1655-1682, ``resample_grid3D_offset``
renderer.py:922-936) and applies the same warp per component.

Grids here are single instances ``[D, H, W, C]``; velocities ``[D, H, W, 3]``.
"""
import numpy as np

from . import BORDER, LINEAR, MIN, MAX
from . import sampling as S


def _idx(shape_dhw):
    d, h, w = shape_dhw
    z, y, x = np.meshgrid(np.arange(d), np.arange(h), np.arange(w), indexing="ij")
    return np.stack([x, y, z], axis=-1).astype(np.float64)


def centered_from_staggered(vx, vy, vz):
    """[D,H,W+1,1], [D,H+1,W,1], [D+1,H,W,1] -> [D,H,W,3]."""
    cx = 0.5 * (vx[:, :, 1:] + vx[:, :, :-1])
    cy = 0.5 * (vy[:, 1:] + vy[:, :-1])
    cz = 0.5 * (vz[1:] + vz[:-1])
    return np.concatenate([cx, cy, cz], axis=-1)


def centered_from_staggered_adjoint(g):
    """Adjoint of ``centered_from_staggered``: g [D,H,W,3] -> face gradients."""
    D, H, W, _ = g.shape
    gx = np.zeros((D, H, W + 1, 1))
    gy = np.zeros((D, H + 1, W, 1))
    gz = np.zeros((D + 1, H, W, 1))
    gx[:, :, 1:, 0] += 0.5 * g[..., 0]
    gx[:, :, :-1, 0] += 0.5 * g[..., 0]
    gy[:, 1:, :, 0] += 0.5 * g[..., 1]
    gy[:, :-1, :, 0] += 0.5 * g[..., 1]
    gz[1:, :, :, 0] += 0.5 * g[..., 2]
    gz[:-1, :, :, 0] += 0.5 * g[..., 2]
    return gx, gy, gz


def warp_forward(rho, v, dt=1.0, order=1, clamp="NONE", boundary=BORDER, return_parts=False):
    """rho [D,H,W,C], v (centered) [D,H,W,3] -> advected rho."""
    rho = np.asarray(rho, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    if list(rho.shape[:3]) != list(v.shape[:3]):
        raise ValueError("Shape mismatch in centered warp")
    clamp = clamp.upper()
    i = _idx(rho.shape[:3])
    lut = -dt * v
    p_fwd = i + lut
    fwd = S.sample_linear(rho, p_fwd, boundary)
    parts = {"fwd": fwd}
    out = fwd
    if order == 2:
        p_back = i - lut
        back = S.sample_linear(fwd, p_back, boundary)
        corr = fwd + 0.5 * (rho - back)
        parts.update(back=back, corr=corr)
        if clamp in ("MC", "MC_SMOOTH"):
            if clamp == "MC":
                raise NotImplementedError("MIM and MAX warp sampling have wrong gradients.")
            mn = S.sample_minmax(rho, p_fwd, boundary, False)
            mx = S.sample_minmax(rho, p_fwd, boundary, True)
            oob = (corr < mn) | (corr > mx)
            parts.update(mn=mn, mx=mx, oob=oob)
            corr = np.where(oob, fwd, corr)
        out = corr
    elif order != 1:
        raise ValueError("Unsupported warp order '{}'".format(order))
    if clamp == "NEGATIVE":
        parts["neg_pass"] = out >= 0.0
        out = np.maximum(out, 0.0)
    if return_parts:
        return out, parts
    return out


def _lut_grad_exact(grid, p, g, boundary):
    """d<g, S(grid, p)>/dp, summed over channels -> [..., 3]."""
    dx, dy, dz = S.data_gradient(grid, p, boundary)
    return np.stack([np.sum(g * dx, -1), np.sum(g * dy, -1), np.sum(g * dz, -1)], axis=-1)


def _scatter(shape, p, g, boundary, semantics):
    if semantics == "reference":
        cnt = np.zeros(shape[:3], dtype=np.int64)
        out = S.scatter_linear(shape, p, g, boundary, mask="reference", count=cnt)
        nz = cnt > 0
        out[nz] = out[nz] / cnt[nz][:, None]
        return out
    return S.scatter_linear(shape, p, g, boundary, mask="forward")


def warp_backward(rho, v, g_out, dt=1.0, order=1, clamp="NONE", boundary=BORDER, semantics="exact",
                  decisions=None):
    """Gradient of ``warp_forward`` w.r.t. (rho, v_centered).

    semantics="exact": the true adjoint (the where / max masks follow the
    forward's branch choice; MIN / MAX reads carry no gradient because they
    only feed the condition).
    semantics="reference": the same chain with each SampleGridLutGrad input
    gradient divided by its corner-hit count, as TF's autodiff over the
    reference ops computes it.
    decisions: optional {"oob": mask, "neg_pass": mask} replacing the
    forward's branch choices (the MC_SMOOTH revert mask, the NEGATIVE pass
    mask), e.g. another evaluation's choices at ties, where the forward is
    not differentiable and either branch is a valid one-sided derivative.
    """
    rho = np.asarray(rho, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    g = np.asarray(g_out, dtype=np.float64).copy()
    out, parts = warp_forward(rho, v, dt, order, clamp, boundary, return_parts=True)
    for key, mask in (decisions or {}).items():
        if key not in parts:
            raise ValueError("no branch choice '%s' in this warp" % key)
        parts[key] = np.asarray(mask, dtype=bool)
    clamp = clamp.upper()
    i = _idx(rho.shape[:3])
    lut = -dt * v
    g_rho = np.zeros_like(rho)
    g_lut = np.zeros(v.shape)
    if clamp == "NEGATIVE":
        g = np.where(parts["neg_pass"], g, 0.0)
    if order == 2:
        if "oob" in parts:
            g_fwd = np.where(parts["oob"], g, 0.0)
            g_corr = np.where(parts["oob"], 0.0, g)
        else:
            g_fwd = np.zeros_like(g)
            g_corr = g
        g_fwd = g_fwd + g_corr
        g_rho += 0.5 * g_corr
        g_back = -0.5 * g_corr
        p_back = i - lut
        g_fwd = g_fwd + _scatter(rho.shape, p_back, g_back, boundary, semantics)
        g_lut -= _lut_grad_exact(parts["fwd"], p_back, g_back, boundary)
    else:
        g_fwd = g
    p_fwd = i + lut
    g_rho += _scatter(rho.shape, p_fwd, g_fwd, boundary, semantics)
    g_lut += _lut_grad_exact(rho, p_fwd, g_fwd, boundary)
    g_v = -dt * g_lut
    return g_rho, g_v


# --- velocity self-advection (staggered) --------------------------------------

_OFF_FROM, _OFF_TO = 0.5, -0.5


def _component_offsets(target):
    """(source component, xyz offset) pairs for the two transverse components
    interpolated onto ``target``'s faces (data_structures.py:1655-1682)."""
    if target == "X":
        return [("Y", (_OFF_TO, _OFF_FROM, 0.0)), ("Z", (_OFF_TO, 0.0, _OFF_FROM))]
    if target == "Y":
        return [("X", (_OFF_FROM, _OFF_TO, 0.0)), ("Z", (0.0, _OFF_TO, _OFF_FROM))]
    return [("X", (_OFF_FROM, 0.0, _OFF_TO)), ("Y", (0.0, _OFF_FROM, _OFF_TO))]


def _shape_of(comps, name):
    return comps["XYZ".index(name)].shape[:3]


def sampled_to_component_shape(comps, target, boundary):
    """[x@t, y@t, z@t] each [*t_shape, 1] (resample_grid3D_offset uses
    sample_grid_lut with relative constant offsets)."""
    tshape = _shape_of(comps, target)
    i = _idx(tshape)
    res = {}
    res[target] = comps["XYZ".index(target)]
    for src, off in _component_offsets(target):
        res[src] = S.sample_linear(comps["XYZ".index(src)], i + np.asarray(off), boundary)
    return [res["X"], res["Y"], res["Z"]]


def velocity_self_advect_forward(vx, vy, vz, dt=1.0, order=1, clamp="NONE", boundary=BORDER):
    comps = [np.asarray(c, dtype=np.float64) for c in (vx, vy, vz)]
    out = []
    for t in "XYZ":
        vel = np.concatenate(sampled_to_component_shape(comps, t, boundary), axis=-1)
        out.append(warp_forward(comps["XYZ".index(t)], vel, dt, order, clamp, boundary))
    return out


def velocity_self_advect_backward(vx, vy, vz, g_outs, dt=1.0, order=1, clamp="NONE", boundary=BORDER,
                                  semantics="exact"):
    comps = [np.asarray(c, dtype=np.float64) for c in (vx, vy, vz)]
    grads = [np.zeros_like(c) for c in comps]
    for ti, t in enumerate("XYZ"):
        parts = sampled_to_component_shape(comps, t, boundary)
        vel = np.concatenate(parts, axis=-1)
        g_data, g_vel = warp_backward(comps[ti], vel, g_outs[ti], dt, order, clamp, boundary, semantics)
        grads[ti] += g_data
        grads[ti] += g_vel[..., ti:ti + 1]
        tshape = _shape_of(comps, t)
        i = _idx(tshape)
        for src, off in _component_offsets(t):
            si = "XYZ".index(src)
            grads[si] += _scatter(comps[si].shape, i + np.asarray(off), g_vel[..., si:si + 1], boundary, semantics)
    return grads
