"""Writes the full-size fixtures ``fullsize_step128.npz`` and
``fullsize_render128.npz`` with the FP64 oracle.  Run from the repository
root: ``python tests/golden/make_fullsize.py`` (about five minutes, a few GB
of memory).  The fixtures are data; the oracle (oracle/) is the program that
produced them.  The GPU tests regenerate the same seeded inputs with
``ngt.synth`` and compare against these values.

fullsize_step128.npz - the differentiable step at 128^3:
  density = 16 Gaussian blobs (seed 2), MAC curl-noise velocity (seed 4,
  max |v| = 2 cells); one MacCormack step (dt 1, MC_SMOOTH, BORDER); the
  5-view ring render (256x144 pixels, 128 slices, BEER_LAMBERT, light
  (1,1,1)); loss = sum((rgb - target)^2) with the target rendered from the
  seed-3 density.  Stored:
  * the loss; advected-density and image samples with full-field checksums;
  * dL/drho at 64 voxels and dL/dv at 64 faces (22 x, 21 y, 21 z), each as
    the oracle's adjoint and as an FP64 central difference of the loss.  The
    differences are evaluated locally: a perturbation of one voxel (face)
    changes the advected density only within 6 (4) cells, so the step is
    recomputed on a sub-block around it and only the rays whose samples
    touch the changed cells are marched again; both restrictions are
    checked against the full evaluation at the unperturbed point.
    Voxels whose difference would cross a non-smooth point (a revert
    decision of the clamp flipping, a sample crossing a cell boundary, an
    advected value leaving [0, inf) where the blend's max(d + sigma, 0)
    switches) are redrawn;
  * per gradient field: max |g|, ||g||_2 and <g, r> with r a seeded standard
    normal field (seeds 11 rho, 12-14 faces), over the entries whose value
    the arithmetic fixes: left out (and listed, rho_out / face_out) are the
    faces beside a velocity component whose fp32 sample position may round
    onto an integer and the voxels / faces that MC_SMOOTH ties reach
    (tests/advect_masks.py).
fullsize_render128.npz - the forward render at 128^3 (16 blobs, seed 2):
  fused raymarch (raymarch_grid_transform) and the plain path
  (sample_grid_transform TRANSFORM_LINDEPTH into 128 depth slices, then
  reduce_grid_blend), 5 views: 4096 sampled pixels per path, max |img|,
  ||img||_2 and <img, r> (seed 21).
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "neural-global-transport_amd"))
sys.path.insert(0, os.path.dirname(HERE))

from oracle import BORDER, LINEAR  # noqa: E402
from oracle import advect as A  # noqa: E402
from oracle import blend as B  # noqa: E402
from oracle import geometry as G  # noqa: E402
from oracle import raymarch as RM  # noqa: E402
from oracle import resample as R  # noqa: E402
import advect_masks as AM  # noqa: E402

GRID = (128, 128, 128)
IMG = (144, 256)
DEPTH = 128
VIEWS = 5
OUT_DHW = (DEPTH,) + IMG
LIGHT = np.array([1.0, 1.0, 1.0])
H_RHO = 1e-5
H_VEL = 1e-5
R_CHANGE = 7      # cells around the perturbed cell whose advected value may change
R_BLOCK = 14      # sub-block radius: R_CHANGE + the 6-cell read radius of a MacCormack cell, + 1


def log(*a):
    print(*a, flush=True)


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def inputs():
    from ngt import synth
    from ngt.renderer import Renderer
    rho = f32(synth.blobs(GRID, 16, seed=2))
    tgt_rho = f32(synth.blobs(GRID, 16, seed=3))
    faces = [f32(c) for c in synth.curl_noise_staggered(GRID, seed=4, vmax=2.0)]
    grid_t, cams = synth.ring_setup(GRID, IMG, DEPTH, VIEWS)
    M, V, P, F = Renderer()._transform_params([grid_t], cams)
    ds = float(np.float32(cams[0].depth_step))
    return rho, tgt_rho, faces, (M, V, P, F), ds


def cells_of(adv, ds):
    return np.concatenate([adv * LIGHT[0], adv * LIGHT[1], adv * LIGHT[2], adv], axis=-1)[None] * ds


def render(adv, cams, ds):
    M, V, P, F = cams
    return RM.raymarch_grid_transform(cells_of(adv, ds), M, V, P, F, OUT_DHW, LINEAR, BORDER, B.BEER_LAMBERT)


def checksums(field, seed):
    r = np.random.default_rng(seed).standard_normal(field.shape)
    return float(np.max(np.abs(field))), float(np.sqrt(np.sum(field * field))), float(np.sum(field * r))


# --- local finite differences -------------------------------------------------------

class Rays:
    """Per view: every pixel's march (start, step, steps) from the oracle's
    ray setup, and the segment each ray's samples span."""

    def __init__(self, cams):
        M, V, P, F = cams
        self.views = []
        for v in range(VIEWS):
            cx = G.CameraXform.from_op_args(M[0], V[v], P[v], F[v])
            start, _end, step, steps, hit = RM.ray_setup(cx, OUT_DHW, GRID, BORDER)
            steps = np.where(hit, steps, 0)
            last = start + np.maximum(steps - 1, 0)[..., None] * step
            self.views.append((start, step, steps, last))

    def touching(self, v, lo_xyz, hi_xyz):
        """Pixels of view v with a sample whose trilinear footprint may touch
        cells [lo, hi) (xyz), by a slab test of the sample segment against
        the positions p with floor(p - 0.5) in [lo - 1, hi), widened by a
        cell."""
        start, step, steps, last = self.views[v]
        blo = np.asarray(lo_xyz, np.float64) - 1.5
        bhi = np.asarray(hi_xyz, np.float64) + 1.5
        d = last - start
        t0 = np.zeros(steps.shape)
        t1 = np.ones(steps.shape)
        for a in range(3):
            with np.errstate(divide="ignore", invalid="ignore"):
                ta = (blo[a] - start[..., a]) / d[..., a]
                tb = (bhi[a] - start[..., a]) / d[..., a]
            lo_t = np.where(d[..., a] != 0, np.minimum(ta, tb), np.where(
                (start[..., a] >= blo[a]) & (start[..., a] <= bhi[a]), -np.inf, np.inf))
            hi_t = np.where(d[..., a] != 0, np.maximum(ta, tb), np.where(
                (start[..., a] >= blo[a]) & (start[..., a] <= bhi[a]), np.inf, -np.inf))
            t0 = np.maximum(t0, lo_t)
            t1 = np.minimum(t1, hi_t)
        return np.nonzero((t0 <= t1) & (steps > 0))

    def march(self, cells, v, sel):
        start, step, steps, _ = self.views[v]
        s = start[sel][:, None]
        st = step[sel][:, None]
        n = steps[sel][:, None]
        _, mask, samp = RM._ray_samples(cells[0], s, st, n, LINEAR, BORDER)
        scan = B.blend_scan(samp, B.BEER_LAMBERT, mask)
        return scan[:, 0, -1, :]


def sub_block(center_zyx):
    lo = [max(c - R_BLOCK, 0) for c in center_zyx]
    hi = [min(c + R_BLOCK + 1, n) for c, n in zip(center_zyx, GRID)]
    return lo, hi


def local_warp(rho, vc, lo, hi):
    sl = tuple(slice(a, b) for a, b in zip(lo, hi))
    return A.warp_forward(rho[sl], vc[sl], 1.0, 2, "MC_SMOOTH", BORDER, return_parts=True)


def fd_point(state, kind, idx):
    """Central difference of the loss for one voxel (kind "rho", idx zyx) or
    one face (kind 0/1/2, idx zyx of the face).  Returns None when the
    perturbation crosses a non-smooth point."""
    rho, faces, vc, adv, cells, tgt, rays, ds = state
    if kind == "rho":
        center = idx
        h = H_RHO
    else:
        center = idx
        h = H_VEL
    lo, hi = sub_block(center)
    clo = [max(c - R_CHANGE, 0) for c in center]
    chi = [min(c + R_CHANGE + 1, n) for c, n in zip(center, GRID)]
    inner = tuple(slice(a - b, c - b) for a, b, c in zip(clo, lo, chi))
    region = tuple(slice(a, c) for a, c in zip(clo, chi))
    out0, parts0 = local_warp(rho, vc, lo, hi)
    assert np.array_equal(out0[inner], adv[region]), "sub-block advection differs from the full one"
    res = []
    for sgn in (1.0, -1.0):
        r2, v2 = rho, vc
        if kind == "rho":
            r2 = rho.copy()
            r2[idx] += sgn * h
        else:
            k = kind
            f2 = faces[k].copy()
            f2[idx] += sgn * h
            comps = [faces[0], faces[1], faces[2]]
            comps[k] = f2
            # only the two cells beside the face change; recompute their centering
            v2 = vc.copy()
            z, y, x = idx
            cells_k = [(z, y, x)]
            prev = [z, y, x]
            prev[2 - k] -= 1
            cells_k.append(tuple(prev))
            for c in cells_k:
                if all(0 <= c[a] < GRID[a] for a in range(3)):
                    cz, cy, cx_ = c
                    if k == 0:
                        v2[cz, cy, cx_, 0] = 0.5 * (f2[cz, cy, cx_ + 1, 0] + f2[cz, cy, cx_, 0])
                    elif k == 1:
                        v2[cz, cy, cx_, 1] = 0.5 * (f2[cz, cy + 1, cx_, 0] + f2[cz, cy, cx_, 0])
                    else:
                        v2[cz, cy, cx_, 2] = 0.5 * (f2[cz + 1, cy, cx_, 0] + f2[cz, cy, cx_, 0])
        out, parts = local_warp(r2, v2, lo, hi)
        # the change stays inside the inner region
        diff = out != out0
        diff[inner] = False
        assert not np.any(diff), "perturbation reached beyond the change radius"
        if not np.array_equal(parts["oob"][inner], parts0["oob"][inner]):
            return None
        if np.min(out[inner]) < 0.0:
            return None
        if kind != "rho":
            sl = tuple(slice(a, b) for a, b in zip(lo, hi))
            i = A._idx(out.shape[:3])
            if not (np.array_equal(np.floor(i - v2[sl]), np.floor(i - vc[sl]))
                    and np.array_equal(np.floor(i + v2[sl]), np.floor(i + vc[sl]))):
                return None
        res.append(out[inner])
    # march the affected rays with the + and - advected fields
    dl = 0.0
    xyz_lo = (clo[2], clo[1], clo[0])
    xyz_hi = (chi[2], chi[1], chi[0])
    saved = cells[(0,) + region].copy()
    try:
        for v in range(VIEWS):
            sel = rays.touching(v, xyz_lo, xyz_hi)
            if sel[0].size == 0:
                continue
            imgs = []
            for a in res:
                cells[(0,) + region] = np.concatenate([a * LIGHT[0], a * LIGHT[1], a * LIGHT[2], a], -1) * ds
                imgs.append(rays.march(cells, v, sel))
            cells[(0,) + region] = saved
            ip, im = imgs
            t = tgt[v][sel]
            dl += float(np.sum((ip[:, :3] - im[:, :3]) * (ip[:, :3] + im[:, :3] - 2.0 * t)))
    finally:
        cells[(0,) + region] = saved
    return dl / (2.0 * h)


def pick(rng, gfield, count, taken):
    """Half uniform over the field, half among entries with |g| above 1 % of
    the maximum."""
    flat = np.abs(gfield).ravel()
    support = np.nonzero(flat > 0.01 * flat.max())[0]
    out = []
    while len(out) < count:
        pool = support if len(out) % 2 else None
        i = int(rng.choice(pool)) if pool is not None else int(rng.integers(flat.size))
        if i in taken:
            continue
        taken.add(i)
        out.append(np.unravel_index(i, gfield.shape[:3]))
    return out


def make_step():
    t0 = time.time()
    rho, tgt_rho, faces, cams, ds = inputs()
    M, V, P, F = cams
    vc = A.centered_from_staggered(*faces)
    adv = A.warp_forward(rho, vc, 1.0, 2, "MC_SMOOTH", BORDER)
    log("advect %.1fs" % (time.time() - t0))
    tgt = f32(render(tgt_rho, cams, ds)[0, ..., :3])
    img = render(adv, cams, ds)
    log("render %.1fs" % (time.time() - t0))
    loss = float(np.sum((img[0, ..., :3] - tgt) ** 2))
    g_img = np.zeros_like(img)
    g_img[..., :3] = 2.0 * (img[..., :3] - tgt[None])
    cells = cells_of(adv, ds)
    g_cells = RM.raymarch_grid_transform_grad(cells, img, g_img, M, V, P, F, OUT_DHW, LINEAR, BORDER,
                                              B.BEER_LAMBERT)[0]
    log("render adjoint %.1fs" % (time.time() - t0))
    g_adv = ds * (g_cells[..., :3] @ LIGHT + g_cells[..., 3])[..., None]
    g_rho, g_vc, rho_keep, vel_keep, info = AM.oracle_adjoint(rho, vc, g_adv, 2, "MC_SMOOTH", BORDER)
    g_faces = A.centered_from_staggered_adjoint(g_vc)
    face_keep = [~m for m in AM.faces_of_cells(~vel_keep)]
    log("advect adjoint %.1fs, left out: %s" % (time.time() - t0, info))

    rays = Rays(cams)
    # the local march reproduces the full render
    for v in range(VIEWS):
        sel = rays.touching(v, (60, 60, 60), (68, 68, 68))
        got = rays.march(cells, v, sel)
        assert np.allclose(got, img[0, v][sel], rtol=0, atol=1e-13 * np.max(np.abs(img))), "local march differs"
    state = (rho, faces, vc, adv, cells, tgt, rays, ds)
    rng = np.random.default_rng(128)
    out = {"loss": loss}
    taken = set()
    idx, fd, adj, redrawn = [], [], [], 0
    for c in pick(rng, g_rho, 200, taken):
        if len(idx) == 64:
            break
        if not rho_keep[c]:
            continue
        d = fd_point(state, "rho", c)
        if d is None:
            redrawn += 1
            continue
        idx.append(c)
        fd.append(d)
        adj.append(g_rho[c][0])
    assert len(idx) == 64
    gmax = float(np.max(np.abs(g_rho)))
    err = np.max(np.abs(np.array(fd) - np.array(adj))) / gmax
    log("rho: 64 differences (%d redrawn), max |fd - adjoint| / max|g| = %.2e (%.1fs)" % (redrawn, err,
                                                                                           time.time() - t0))
    assert err < 1e-6, err
    out.update(rho_idx=np.array(idx, np.int64), rho_fd=np.array(fd), rho_adj=np.array(adj), rho_redrawn=redrawn)
    out["rho_gmax"], out["rho_gnorm"], out["rho_proj"] = checksums(np.where(rho_keep[..., None], g_rho, 0.0), 11)
    out["rho_gmax"] = float(np.max(np.abs(g_rho)))
    out["rho_out"] = np.argwhere(~rho_keep).astype(np.int64)

    fidx, ffd, fadj, redrawn = [], [], [], 0
    for k, count in ((0, 22), (1, 21), (2, 21)):
        taken = set()
        got = 0
        for c in pick(rng, g_faces[k], 200, taken):
            if got == count:
                break
            if not face_keep[k][c][0]:
                continue
            d = fd_point(state, k, c)
            if d is None:
                redrawn += 1
                continue
            fidx.append((k,) + tuple(c))
            ffd.append(d)
            fadj.append(g_faces[k][c][0])
            got += 1
        assert got == count
    fmax = np.array([np.max(np.abs(g)) for g in g_faces])
    err = max(abs(a - b) / fmax[i[0]] for i, a, b in zip(fidx, ffd, fadj))
    log("faces: 64 differences (%d redrawn), max |fd - adjoint| / max|g| = %.2e (%.1fs)" % (redrawn, err,
                                                                                             time.time() - t0))
    assert err < 1e-6, err
    out.update(face_idx=np.array(fidx, np.int64), face_fd=np.array(ffd), face_adj=np.array(fadj),
               face_redrawn=redrawn)
    sums = [checksums(np.where(m, g, 0.0), 12 + k) for k, (g, m) in enumerate(zip(g_faces, face_keep))]
    out["face_gmax"] = np.array([np.max(np.abs(g)) for g in g_faces])
    out["face_out"] = np.concatenate([np.concatenate([np.full((len(ix), 1), k), ix], 1)
                                      for k, ix in enumerate(np.argwhere(~m[..., 0]) for m in face_keep)]).astype(
        np.int64)
    out["face_gnorm"] = np.array([s[1] for s in sums])
    out["face_proj"] = np.array([s[2] for s in sums])

    out["adv_max"], out["adv_norm"], out["adv_proj"] = checksums(adv, 15)
    pix = sample_pixels(np.random.default_rng(16), 4096)
    out["img_pix"] = pix
    out["img_val"] = img[0][tuple(pix.T)]
    out["img_max"], out["img_norm"], out["img_proj"] = checksums(img[0], 17)
    np.savez_compressed(os.path.join(HERE, "fullsize_step128.npz"), **out)
    log("wrote fullsize_step128.npz (%.1fs)" % (time.time() - t0))


def sample_pixels(rng, count):
    return np.stack([rng.integers(0, VIEWS, count), rng.integers(0, IMG[0], count), rng.integers(0, IMG[1], count)],
                    -1).astype(np.int64)


def make_render():
    t0 = time.time()
    rho, _tgt, _faces, cams, ds = inputs()
    M, V, P, F = cams
    cells = cells_of(rho, ds)
    img = RM.raymarch_grid_transform(cells, M, V, P, F, OUT_DHW, LINEAR, BORDER, B.BEER_LAMBERT)[0]
    log("fused render %.1fs" % (time.time() - t0))
    plain = np.zeros_like(img)
    for v in range(VIEWS):
        fr = R.sample_grid_transform(cells, M, V[v:v + 1], P[v:v + 1], F[v:v + 1], OUT_DHW, LINEAR, BORDER)
        plain[v] = B.reduce_grid_blend(fr[:, 0], B.BEER_LAMBERT)[0]
        log("plain view %d %.1fs" % (v, time.time() - t0))
    pix = sample_pixels(np.random.default_rng(22), 4096)
    out = {"pix": pix, "fused_val": img[tuple(pix.T)], "plain_val": plain[tuple(pix.T)]}
    out["fused_max"], out["fused_norm"], out["fused_proj"] = checksums(img, 21)
    out["plain_max"], out["plain_norm"], out["plain_proj"] = checksums(plain, 21)
    np.savez_compressed(os.path.join(HERE, "fullsize_render128.npz"), **out)
    log("wrote fullsize_render128.npz (%.1fs)" % (time.time() - t0))


if __name__ == "__main__":
    which = sys.argv[1:] or ["step", "render"]
    if "render" in which:
        make_render()
    if "step" in which:
        make_step()
