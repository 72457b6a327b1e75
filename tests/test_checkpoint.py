"""The model.pt reader on the host (no GPU): the reference's save_pt layout
round-trips, missing keys and non-scene files are refused, and a file whose
pickle would run code is refused by the weights-only loader."""

import os

import numpy as np
import pytest
import torch

from powerfoam_amd.checkpoint import KEYS, PowerfoamCheckpoint, save_pt


def _scene(n=50, T=4, D=8, seed=0):
    rng = np.random.default_rng(seed)
    return dict(points=rng.normal(size=(n, 3)), density=rng.normal(size=n), radii=rng.normal(size=n),
                quaternions=rng.normal(size=(n, 4)), texel_sites=rng.normal(size=(n, T, 2)),
                texel_sv_axis=rng.normal(size=(n, T, 3 * D)), texel_sv_rgb=rng.normal(size=(n, T, 3 * D)),
                texel_height=rng.normal(size=(n, T)), adjacency=np.arange(2 * n, dtype=np.int32) % n,
                adjacency_offsets=np.arange(0, 2 * n + 1, 2, dtype=np.int32))


def test_round_trip(tmp_path):
    s = _scene()
    path = os.path.join(tmp_path, "model.pt")
    save_pt(path, **s)
    raw = torch.load(path, weights_only=True)
    assert set(raw) == set(KEYS)
    ck = PowerfoamCheckpoint.load(path, device="cpu")
    assert ck.n_sites == 50 and ck.n_texels == 4 and ck.sv_dof == 8
    np.testing.assert_array_equal(ck.glue.points.numpy(), s["points"].astype(np.float32))
    np.testing.assert_array_equal(ck.glue.radii.numpy(), s["radii"].astype(np.float32))
    np.testing.assert_array_equal(ck.texel_sv_rgb.numpy(), s["texel_sv_rgb"].astype(np.float32))
    np.testing.assert_array_equal(ck.adjacency_offsets.numpy(), s["adjacency_offsets"])
    half = PowerfoamCheckpoint.load(path, device="cpu", attr_dtype="half")
    np.testing.assert_array_equal(half.glue.density.numpy(),
                                  s["density"].astype(np.float32).astype(np.float16).astype(np.float32))


def test_refuses_incomplete_or_foreign_files(tmp_path):
    s = _scene()
    del s["texel_sv_rgb"]
    p1 = os.path.join(tmp_path, "a.pt")
    torch.save({k: torch.as_tensor(v) for k, v in s.items()}, p1)
    with pytest.raises(ValueError, match="texel_sv_rgb"):
        PowerfoamCheckpoint.load(p1, device="cpu")
    p2 = os.path.join(tmp_path, "b.pt")
    torch.save([torch.zeros(3)], p2)
    with pytest.raises(ValueError):
        PowerfoamCheckpoint.load(p2, device="cpu")
    p3 = os.path.join(tmp_path, "c.pt")
    torch.save({"points": torch.zeros(3, 3), "obj": _Opaque()}, p3)
    with pytest.raises(Exception):
        PowerfoamCheckpoint.load(p3, device="cpu")


class _Opaque:
    pass
