"""Scene generators and the snapshot format (CPU)."""
import numpy as np
import pytest

from pbf import scenes


def test_splitmix64_published_vectors():
    # Vigna's reference splitmix64.c outputs
    assert [int(v) for v in scenes.splitmix64(0, 0, 2)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4]
    assert [int(v) for v in scenes.splitmix64(1234567, 0, 5)] == [
        6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
        16408922859458223821]
    # counter form: any slice equals the same slice of the full stream
    full = scenes.splitmix64(99, 0, 100)
    np.testing.assert_array_equal(scenes.splitmix64(99, 37, 20), full[37:57])


def test_uniform_and_normal_moments():
    u = scenes.uniform01(5, 0, 200000)
    assert u.min() >= 0 and u.max() < 1
    assert abs(u.mean() - 0.5) < 3e-3
    z = scenes.normal(6, 0, 200000)
    assert abs(z.mean()) < 1e-2 and abs(z.std() - 1) < 1e-2


def test_lattice_layout_and_jitter():
    x = scenes.lattice((4, 3, 2), (1.0, 2.0, 3.0), 0.05, 0.005, seed=1)
    assert x.shape == (24, 3) and x.dtype == np.float32
    base = np.array([[1.0 + (i + 0.5) * 0.05, 2.0 + (j + 0.5) * 0.05, 3.0 + (k + 0.5) * 0.05]
                     for i in range(4) for j in range(3) for k in range(2)])  # x outer, z inner
    d = x - base
    assert np.abs(d).max() <= 0.005 + 1e-6
    assert np.abs(d).max() > 0.002


def test_workload_scenes():
    dd = scenes.double_dam()
    assert dd.n == 32768 and dd.ulim == pytest.approx((4.0, 2.0, 2.0))
    assert np.all(dd.pos >= 0) and np.all(dd.pos <= np.array(dd.ulim))
    assert np.array_equal(dd.iid, np.arange(32768))
    sd = scenes.single_dam(counts=(8, 4, 8), box_cells=(16, 8, 8))
    assert sd.n == 256
    big = scenes.single_dam.__defaults__
    assert big[2] == (128, 64, 128) and big[3] == (256, 128, 128)
    # deterministic
    np.testing.assert_array_equal(scenes.double_dam().pos, dd.pos)


def test_snapshot_round_trip(tmp_path):
    sc = scenes.block((5, 4, 3), (2.0, 2.0, 2.0), seed=3, vel_sigma=0.5)
    for layout in (3, 4):
        pos, vel = sc.as_layout(layout)
        f = tmp_path / f"s{layout}.pbfs"
        scenes.save_snapshot(str(f), pos, vel, sc.iid, sc.ulim, sc.llim)
        p2, v2, i2, ul, ll = scenes.load_snapshot(str(f))
        np.testing.assert_array_equal(p2, pos)
        np.testing.assert_array_equal(v2, vel)
        np.testing.assert_array_equal(i2, sc.iid)
        assert ul == pytest.approx(sc.ulim) and ll == pytest.approx(sc.llim)
    bad = tmp_path / "bad.pbfs"
    bad.write_bytes(b"XXXX" + bytes(40))
    with pytest.raises(ValueError):
        scenes.load_snapshot(str(bad))
