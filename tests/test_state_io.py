"""State I/O in the reference's npz layout (data_structures.py:174-193,
332-353, 487-503, 680-681, 1123-1128, 1213-1215, 1241-1258, 1302-1307).

The files are written here exactly as the reference's save methods write
them (np.savez_compressed with the density as arr_0 and the optional hull /
inflow fields; centered_shape + vel_x/y/z for velocities; the legacy velocity
files without centered_shape) and read back through gft.grids' loaders,
and gft's own writers round-trip. CPU only.
"""
import os

import numpy as np
import torch

from gft.grids import DensityGrid, Sequence, State, VelocityGrid

RNG = np.random.Generator(np.random.PCG64(31))


def _vel_components(shape):
    D, H, W = shape
    return (RNG.standard_normal((1, D, H, W + 1, 1)).astype(np.float32),
            RNG.standard_normal((1, D, H + 1, W, 1)).astype(np.float32),
            RNG.standard_normal((1, D + 1, H, W, 1)).astype(np.float32))


def test_density_file_in_reference_layout(tmp_path):
    d = RNG.uniform(0, 1, (1, 4, 6, 5, 1)).astype(np.float32)
    hull = (d > 0.5).astype(np.float32)
    inflow = RNG.uniform(0, 1, (1, 2, 3, 2, 1)).astype(np.float32)
    mask = np.ones_like(inflow)
    path = str(tmp_path / "density.npz")
    # what DensityGrid.save writes
    np.savez_compressed(path, d, hull=hull, inflow=inflow, inflow_mask=mask, inflow_offset=np.asarray([1, 2, 3]))
    g = DensityGrid.from_file(path, device="cpu")
    assert g.shape == [4, 6, 5]
    assert np.array_equal(g.d.numpy(), d)
    assert np.array_equal(g.hull.numpy(), hull)
    assert g.inflow_offset == [1, 2, 3]
    inf = g.inflow.numpy()
    assert inf.shape == d.shape
    assert np.array_equal(inf[:, 1:3, 2:5, 3:5], inflow)
    assert inf.sum() == inflow.sum()


def test_density_round_trip_and_missing_file(tmp_path):
    d = torch.from_numpy(RNG.uniform(0, 1, (1, 3, 4, 5, 1)).astype(np.float32))
    g = DensityGrid([3, 4, 5], d=d, device="cpu")
    path = str(tmp_path / "d.npz")
    g.save(path)
    with np.load(path) as f:
        assert sorted(f.files) == ["arr_0"]
    g2 = DensityGrid.from_file(path, device="cpu")
    assert torch.equal(g2.d, d)
    assert DensityGrid.from_file(str(tmp_path / "missing.npz"), device="cpu") is None


def test_velocity_file_in_reference_layout_and_legacy(tmp_path):
    shape = (4, 6, 5)
    x, y, z = _vel_components(shape)
    path = str(tmp_path / "velocity.npz")
    # what VelocityGrid.save writes
    np.savez_compressed(path, centered_shape=np.asarray(shape), vel_x=x, vel_y=y, vel_z=z)
    v = VelocityGrid.from_file(path, device="cpu")
    assert v.centered_shape == list(shape)
    assert np.array_equal(v.x.numpy(), x) and np.array_equal(v.y.numpy(), y) and np.array_equal(v.z.numpy(), z)
    # legacy file: no centered_shape; shape from vel_x with one fewer cell in x
    legacy = str(tmp_path / "legacy.npz")
    np.savez_compressed(legacy, vel_x=x, vel_y=y, vel_z=z)
    v2 = VelocityGrid.from_file(legacy, device="cpu")
    assert v2.centered_shape == list(shape)
    assert np.array_equal(v2.z.numpy(), z)
    # round trip through the writer
    out = str(tmp_path / "rt.npz")
    v.save(out)
    with np.load(out) as f:
        assert sorted(f.files) == ["centered_shape", "vel_x", "vel_y", "vel_z"]
        assert f["centered_shape"].tolist() == list(shape)
    v3 = VelocityGrid.from_file(out, device="cpu")
    assert torch.equal(v3.y, v.y)


def test_sequence_directory_layout(tmp_path):
    shape = (3, 4, 5)
    frames = [7, 8, 10]
    states = []
    for f in frames:
        d = torch.from_numpy(RNG.uniform(0, 1, (1,) + shape + (1,)).astype(np.float32))
        x, y, z = (torch.from_numpy(c) for c in _vel_components(shape))
        states.append(State(DensityGrid(list(shape), d=d, device="cpu"),
                            VelocityGrid(list(shape), x=x, y=y, z=z, device="cpu"), f))
    Sequence(states).save(str(tmp_path))
    for f in frames:
        sub = tmp_path / "frame_{:06d}".format(f)
        assert sorted(os.listdir(sub)) == ["density.npz", "velocity.npz"]
    seq = Sequence.from_file(str(tmp_path), frames, device="cpu")
    assert [s.frame for s in seq] == frames
    assert seq[0].prev is None and seq[0].next is seq[1] and seq[2].prev is seq[1]
    for a, b in zip(seq, states):
        assert torch.equal(a.density.d, b.density.d)
        assert torch.equal(a.velocity.x, b.velocity.x)
    # suffixed files
    states[0].save(str(tmp_path), suffix="it3")
    assert os.path.exists(tmp_path / "density_it3.npz") and os.path.exists(tmp_path / "velocity_it3.npz")
    st = State.from_file(str(tmp_path), 7, device="cpu", density_filename="density_it3.npz",
                         velocity_filename="velocity_it3.npz")
    assert torch.equal(st.density.d, states[0].density.d)
