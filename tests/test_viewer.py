"""Point-cloud export of frames (CPU): the PLY bytes, identity order and the
speed attribute."""
import numpy as np
import pytest
import torch

from pbf import scenes, viewer


def _shuffled_frame(n=1000, stride=4):
    sc = scenes.double_dam()
    pos, vel = sc.as_layout(stride)
    rng = np.random.default_rng(3)
    vel = vel.copy()
    vel[:, :3] = rng.normal(size=(len(vel), 3)).astype(np.float32)
    perm = rng.permutation(len(pos))[:n]  # a step leaves particles in cell order, not identity order
    return pos[perm], vel[perm], sc.iid[perm]


def test_ply_bytes_are_the_documented_layout(tmp_path):
    pos, vel, iid = _shuffled_frame()
    path = str(tmp_path / "f.ply")
    viewer.write_ply(path, pos, vel, iid, comments=["frame 7"])
    data = open(path, "rb").read()
    header = (b"ply\nformat binary_little_endian 1.0\ncomment frame 7\nelement vertex 1000\n"
              b"property float x\nproperty float y\nproperty float z\nproperty float speed\n"
              b"property uint iid\nend_header\n")
    assert data.startswith(header)
    body = data[len(header):]
    assert len(body) == 1000 * 20
    rec = np.frombuffer(body, dtype=np.uint8).reshape(1000, 20)
    xyz = rec[:, :12].copy().view("<f4").reshape(1000, 3)
    speed = rec[:, 12:16].copy().view("<f4").ravel()
    ids = rec[:, 16:].copy().view("<u4").ravel()
    o = np.argsort(iid)
    np.testing.assert_array_equal(ids, iid[o])  # identity order
    np.testing.assert_array_equal(xyz, pos[o, :3])
    np.testing.assert_allclose(speed, np.linalg.norm(vel[o, :3].astype(np.float64), axis=1),
                               rtol=1e-6)


def test_ply_round_trip_and_options(tmp_path):
    pos, vel, iid = _shuffled_frame(stride=3)
    path = str(tmp_path / "g.ply")
    viewer.write_ply(path, torch.from_numpy(pos), torch.from_numpy(vel), torch.from_numpy(iid),
                     nparticle=600, identity_order=False)
    v = viewer.read_ply(path)
    assert len(v) == 600
    np.testing.assert_array_equal(v["iid"], iid[:600])  # cell order kept when asked
    np.testing.assert_array_equal(np.stack([v["x"], v["y"], v["z"]], 1), pos[:600])


def test_ply_rejects_bad_input(tmp_path):
    pos, vel, iid = _shuffled_frame()
    with pytest.raises(ValueError):
        viewer.write_ply(str(tmp_path / "h.ply"), pos, vel[:, :3], iid)
    with pytest.raises(ValueError):
        viewer.write_ply(str(tmp_path / "h.ply"), pos, vel, iid[:10])
    bad = tmp_path / "bad.ply"
    bad.write_bytes(b"ply\nformat ascii 1.0\nend_header\n")
    with pytest.raises(ValueError):
        viewer.read_ply(str(bad))
