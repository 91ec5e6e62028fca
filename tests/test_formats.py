"""On-disk formats of the grids (CPU): density npz ('arr_0' + optional hull /
inflow / inflow_mask / inflow_offset, data_structures.py:465-485,704-726)
and velocity npz ('vel_x' / 'vel_y' / 'vel_z' + 'centered_shape',
data_structures.py:909-925,1197-1198).  Files are written here with plain
numpy in the reference's layout and read back by the grids, and the grids'
files are read with plain numpy; nothing is unpickled."""
import os

import numpy as np
import pytest
import torch

from ngt.data_structures import DensityGrid, VelocityGrid


def test_density_npz_reference_layout(tmp_path):
    rng = np.random.default_rng(0)
    d = rng.uniform(0, 1, (1, 6, 7, 8, 1)).astype(np.float32)
    hull = (d > 0.5).astype(np.float32)
    inflow = rng.uniform(-2, 1, (1, 2, 3, 4, 1)).astype(np.float32)
    mask = (rng.uniform(0, 1, inflow.shape) > 0.3).astype(np.float32)
    path = os.path.join(tmp_path, "density.npz")
    np.savez_compressed(path, d, hull=hull, inflow=inflow, inflow_mask=mask, inflow_offset=np.asarray([1, 2, 3]))
    g = DensityGrid.from_file(path, device="cpu")
    assert g.shape == [6, 7, 8] and g.inflow_offset == [1, 2, 3]
    np.testing.assert_array_equal(g.d.numpy(), d)
    np.testing.assert_array_equal(g.hull.numpy(), hull)
    # with_inflow = max(d + padded masked inflow, 0)
    full = np.zeros_like(d)
    full[:, 1:3, 2:5, 3:7] = inflow * mask
    np.testing.assert_allclose(g.with_inflow().numpy(), np.maximum(d + full, 0.0), rtol=0, atol=1e-7)
    out = os.path.join(tmp_path, "saved.npz")
    g.save(out)
    with np.load(out, allow_pickle=False) as f:
        assert sorted(f.files) == ["arr_0", "hull", "inflow", "inflow_mask", "inflow_offset"]
        np.testing.assert_array_equal(f["arr_0"], d)
        np.testing.assert_array_equal(f["inflow"], inflow)
        assert f["inflow_offset"].tolist() == [1, 2, 3]


def test_density_npz_plain_and_failures(tmp_path):
    d = np.full((1, 3, 4, 5, 1), 0.25, dtype=np.float32)
    path = os.path.join(tmp_path, "plain.npz")
    np.savez_compressed(path, d)
    g = DensityGrid.from_file(path, device="cpu")
    assert g.hull is None and g.inflow_offset is None
    assert torch.equal(g.with_inflow(), g.d)
    assert DensityGrid.from_file(os.path.join(tmp_path, "missing.npz"), device="cpu") is None
    # an object array would need unpickling: refused, not executed
    obj = os.path.join(tmp_path, "object.npz")
    np.savez(obj, np.array([{"a": 1}], dtype=object))
    assert DensityGrid.from_file(obj, device="cpu") is None
    with pytest.raises(ValueError, match="does not fit"):
        DensityGrid([3, 4, 5], d=torch.zeros((1, 3, 4, 5, 1)), inflow=np.zeros((1, 2, 2, 2, 1)),
                    inflow_offset=[2, 0, 0])


def test_velocity_npz_reference_layout(tmp_path):
    rng = np.random.default_rng(1)
    cs = [5, 6, 7]
    comps = [rng.standard_normal((1,) + tuple(s) + (1,)).astype(np.float32) for s in VelocityGrid.component_shapes(cs)]
    path = os.path.join(tmp_path, "velocity.npz")
    np.savez_compressed(path, centered_shape=cs, vel_x=comps[0], vel_y=comps[1], vel_z=comps[2])
    v = VelocityGrid.from_file(path, device="cpu")
    assert v.centered_shape == cs
    for t, c in zip((v.x, v.y, v.z), comps):
        np.testing.assert_array_equal(t.numpy(), c)
    # legacy files without centered_shape: derived from vel_x (x has W + 1 cells)
    legacy = os.path.join(tmp_path, "legacy.npz")
    np.savez_compressed(legacy, vel_x=comps[0], vel_y=comps[1], vel_z=comps[2])
    assert VelocityGrid.from_file(legacy, device="cpu").centered_shape == cs
    out = os.path.join(tmp_path, "saved.npz")
    v.save(out)
    with np.load(out, allow_pickle=False) as f:
        assert sorted(f.files) == ["centered_shape", "vel_x", "vel_y", "vel_z"]
        assert f["centered_shape"].tolist() == cs
        np.testing.assert_array_equal(f["vel_z"], comps[2])
    assert VelocityGrid.from_file(os.path.join(tmp_path, "missing.npz"), device="cpu") is None
