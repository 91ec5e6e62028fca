"""Seeded synthetic inputs for the transport workloads (numpy, float32).

* ``blobs``: sum of K Gaussian blobs, centres uniform in the inner 60% of the
  grid, sigma U[4,10] voxels, amplitude U[0.2,1], clipped to [0,1].
* ``curl_noise_centered``: v = curl A of a Gaussian-filtered (sigma ~8 voxel)
  random vector potential, central differences, scaled to max |v| = vmax.
* ``curl_noise_staggered``: MAC face velocities x [D,H,W+1], y [D,H+1,W],
  z [D+1,H,W] from an edge-located vector potential; the discrete
  divergence of every cell is exactly 0; scaled to max |v_centered| = vmax.
* ``ring_setup``: the 5-camera ring of the render workloads: grid normalised
  to a unit cube centred at the origin, cameras on a horizontal ring of
  radius 3 grid extents, 72 degrees apart, horizontal FOV 40 degrees, looking
  at the centre, near / far = R -+ sqrt(3)/2.
* ``uneven_batch_layout``: the uneven-batch workload's per-sequence sizes.
"""
import numpy as np
from scipy.ndimage import gaussian_filter

from .camera import ring_cameras
from .transform import GridTransform


def blobs(shape_dhw, count, seed, sigma_range=(4.0, 10.0), amp_range=(0.2, 1.0)):
    rng = np.random.default_rng(seed)
    D, H, W = shape_dhw
    dims = np.array([W, H, D], dtype=np.float64)
    z = np.arange(D, dtype=np.float64)[:, None, None]
    y = np.arange(H, dtype=np.float64)[None, :, None]
    x = np.arange(W, dtype=np.float64)[None, None, :]
    out = np.zeros((D, H, W), dtype=np.float64)
    for _ in range(count):
        c = (0.2 + 0.6 * rng.uniform(0, 1, 3)) * dims
        s = rng.uniform(*sigma_range)
        a = rng.uniform(*amp_range)
        gx = np.exp(-(x - c[0]) ** 2 / (2 * s * s))
        gy = np.exp(-(y - c[1]) ** 2 / (2 * s * s))
        gz = np.exp(-(z - c[2]) ** 2 / (2 * s * s))
        out += a * gz * gy * gx
    return np.clip(out, 0.0, 1.0).astype(np.float32)[..., None]


def _smooth(rng, shape, sigma):
    return gaussian_filter(rng.standard_normal(shape), sigma, mode="wrap")


def curl_noise_centered(shape_dhw, seed, vmax=2.0, sigma=8.0):
    rng = np.random.default_rng(seed)
    ax, ay, az = (_smooth(rng, shape_dhw, sigma) for _ in range(3))
    # arrays are [z, y, x]
    dz = lambda f: np.gradient(f, axis=0)
    dy = lambda f: np.gradient(f, axis=1)
    dx = lambda f: np.gradient(f, axis=2)
    vx = dy(az) - dz(ay)
    vy = dz(ax) - dx(az)
    vz = dx(ay) - dy(ax)
    v = np.stack([vx, vy, vz], axis=-1)
    v *= vmax / np.max(np.linalg.norm(v, axis=-1))
    return v.astype(np.float32)


def curl_noise_staggered(shape_dhw, seed, vmax=2.0, sigma=8.0):
    rng = np.random.default_rng(seed)
    D, H, W = shape_dhw
    Ax = _smooth(rng, (D + 1, H + 1, W), sigma)   # x-edges
    Ay = _smooth(rng, (D + 1, H, W + 1), sigma)   # y-edges
    Az = _smooth(rng, (D, H + 1, W + 1), sigma)   # z-edges
    vx = (Az[:, 1:, :] - Az[:, :-1, :]) - (Ay[1:, :, :] - Ay[:-1, :, :])   # [D,H,W+1]
    vy = (Ax[1:, :, :] - Ax[:-1, :, :]) - (Az[:, :, 1:] - Az[:, :, :-1])   # [D,H+1,W]
    vz = (Ay[:, :, 1:] - Ay[:, :, :-1]) - (Ax[:, 1:, :] - Ax[:, :-1, :])   # [D+1,H,W]
    cx = 0.5 * (vx[:, :, 1:] + vx[:, :, :-1])
    cy = 0.5 * (vy[:, 1:] + vy[:, :-1])
    cz = 0.5 * (vz[1:] + vz[:-1])
    s = vmax / np.max(np.sqrt(cx * cx + cy * cy + cz * cz))
    return tuple((c * s).astype(np.float32)[..., None] for c in (vx, vy, vz))


def ring_setup(grid_dhw, image_hw=(144, 256), depth=128, views=5, fov=40.0, distance=3.0):
    """Grid transform + cameras of the render workloads."""
    D, H, W = grid_dhw
    ext = float(max(D, H, W))
    grid = GridTransform(list(grid_dhw), center=True, scale=[1.0 / ext] * 3)
    near = distance - np.sqrt(3.0) / 2.0
    far = distance + np.sqrt(3.0) / 2.0
    cams = ring_cameras(views, distance, fov, image_hw, depth, near, far)
    return grid, cams


UNEVEN_EDGES = (48, 64, 96, 128, 160)
UNEVEN_IMAGES = ((72, 128), (144, 256), (270, 480))


def uneven_batch_layout(seed=5, sequences=32):
    """Per sequence (cubic grid edge, views, (image H, W)): the edge drawn
    from 48, 64, 96, 128, 160, 1-8 ring views, the image from 128x72,
    256x144, 480x270 (in that order per sequence, from one seeded stream);
    each view marches edge-many depth slices."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(sequences):
        e = int(rng.choice(UNEVEN_EDGES))
        nv = int(rng.integers(1, 9))
        hw = UNEVEN_IMAGES[int(rng.integers(0, 3))]
        out.append((e, nv, hw))
    return out


# --- device-side generators (large batches) ----------------------------------------

def blobs_torch(shape_dhw, count, seed, device="cuda", sigma_range=(4.0, 10.0), amp_range=(0.2, 1.0)):
    """``blobs`` evaluated on the device (same parameter distribution, float32
    arithmetic), [D,H,W,1]."""
    import torch
    rng = np.random.default_rng(seed)
    D, H, W = shape_dhw
    dims = np.array([W, H, D], dtype=np.float64)
    z = torch.arange(D, device=device, dtype=torch.float32)[:, None, None]
    y = torch.arange(H, device=device, dtype=torch.float32)[None, :, None]
    x = torch.arange(W, device=device, dtype=torch.float32)[None, None, :]
    out = torch.zeros((D, H, W), device=device)
    for _ in range(count):
        c = (0.2 + 0.6 * rng.uniform(0, 1, 3)) * dims
        s = float(rng.uniform(*sigma_range))
        a = float(rng.uniform(*amp_range))
        k = -1.0 / (2 * s * s)
        out += a * torch.exp(k * (z - c[2]) ** 2) * torch.exp(k * (y - c[1]) ** 2) * torch.exp(k * (x - c[0]) ** 2)
    return out.clamp_(0.0, 1.0)[..., None]


def curl_noise_staggered_torch(shape_dhw, seed, vmax=2.0, wavelength=24.0, modes=6, device="cuda"):
    """MAC face velocities (x [D,H,W+1,1], y [D,H+1,W,1], z [D+1,H,W,1]) as the
    discrete curl of edge potentials that are sums of random plane waves of
    the given wavelength (cells): exactly divergence-free per cell, scaled to
    max |v_centered| = vmax."""
    import torch
    rng = np.random.default_rng(seed)
    D, H, W = shape_dhw

    def potential(shape):
        dz, dy, dx = shape
        z = torch.arange(dz, device=device, dtype=torch.float32)[:, None, None]
        y = torch.arange(dy, device=device, dtype=torch.float32)[None, :, None]
        x = torch.arange(dx, device=device, dtype=torch.float32)[None, None, :]
        out = torch.zeros(shape, device=device)
        for _ in range(modes):
            k = rng.normal(size=3)
            k *= 2 * np.pi / wavelength / np.linalg.norm(k)
            ph = float(rng.uniform(0, 2 * np.pi))
            out += float(rng.normal()) * torch.sin(float(k[0]) * x + float(k[1]) * y + float(k[2]) * z + ph)
        return out

    Ax = potential((D + 1, H + 1, W))
    Ay = potential((D + 1, H, W + 1))
    Az = potential((D, H + 1, W + 1))
    vx = (Az[:, 1:, :] - Az[:, :-1, :]) - (Ay[1:, :, :] - Ay[:-1, :, :])
    vy = (Ax[1:, :, :] - Ax[:-1, :, :]) - (Az[:, :, 1:] - Az[:, :, :-1])
    vz = (Ay[:, :, 1:] - Ay[:, :, :-1]) - (Ax[:, 1:, :] - Ax[:, :-1, :])
    cx = 0.5 * (vx[:, :, 1:] + vx[:, :, :-1])
    cy = 0.5 * (vy[:, 1:] + vy[:, :-1])
    cz = 0.5 * (vz[1:] + vz[:-1])
    s = vmax / float(torch.sqrt(cx * cx + cy * cy + cz * cz).max())
    return tuple((c * s).contiguous()[..., None] for c in (vx, vy, vz))

