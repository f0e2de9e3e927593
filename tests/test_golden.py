"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the oracle): the oracle must reproduce them (CPU), and the GPU must equal
their FP32 outputs bit for bit (marked gpu)."""
import glob
import os

import numpy as np
import pytest

from conftest import ROOT, from_dev, to_dev
from pbf import params as P

FILES = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "*.npz")))


def _params(a):
    return P.FluidParams(niter=int(a[0]), pho0=float(a[1]), g=float(a[2]), h=float(a[3]),
                         dt=float(a[4]), lambda_eps=float(a[5]), delta_q=float(a[6]),
                         k_corr=float(a[7]), n_corr=float(a[8]), k_boundaryDensity=float(a[9]),
                         c_XSPH=float(a[10]))


def _load(f):
    z = np.load(f)  # plain arrays only (allow_pickle defaults to False)
    return {k: z[k] for k in z.files}


def test_fixtures_present():
    assert len(FILES) >= 5


@pytest.mark.parametrize("f", FILES, ids=[os.path.basename(f)[:-4] for f in FILES])
def test_oracle_reproduces_fixture(orc, f):
    z = _load(f)
    p = _params(z["params"])
    m = orc.step(p, z["ulim"], z["llim"], z["in_pos"], z["in_vel"], z["in_iid"], fp32=True)
    for k in ("pos", "vel", "npos", "nvel", "iid"):
        np.testing.assert_array_equal(m[k], z["f32_" + k], err_msg=k)
    np.testing.assert_array_equal(m["keys"], z["keys"])
    o = orc.step(p, z["ulim"], z["llim"], z["in_pos"], z["in_vel"], z["in_iid"])
    np.testing.assert_allclose(o["x_final"], z["f64_x"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(o["lambda"], z["f64_lambda"], rtol=1e-12, atol=1e-18)


@pytest.mark.gpu
@pytest.mark.parametrize("f", FILES, ids=[os.path.basename(f)[:-4] for f in FILES])
def test_gpu_matches_fixture(pbf_mod, device, f):
    import torch

    z = _load(f)
    p = _params(z["params"])
    n = z["in_pos"].shape[0]
    with pbf_mod.Simulator(p, tuple(z["ulim"]), tuple(z["llim"]), capacity=n) as sim:
        pos, vel = to_dev(z["in_pos"], device), to_dev(z["in_vel"], device)
        npos, nvel = torch.zeros_like(pos), torch.zeros_like(vel)
        iid = to_dev(z["in_iid"], device)
        sim.step(pos, npos, vel, nvel, iid, n)
        torch.cuda.synchronize()
        got = {"pos": pos, "vel": vel, "npos": npos, "nvel": nvel}
        for k, t in got.items():
            a = from_dev(t).reshape(n, 4)
            np.testing.assert_array_equal(a.view(np.uint32), z["f32_" + k].view(np.uint32),
                                          err_msg=k)
        np.testing.assert_array_equal(from_dev(iid, dtype=np.uint32), z["f32_iid"])
