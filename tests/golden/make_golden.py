"""Writes the golden fixtures of the PBF step (tests/golden/*.npz).

The reference ships no test vectors, so these are the project oracle's own
outputs on small seeded scenes (N <= 4096), frozen as regression fixtures:
inputs, the FP32 mirror's outputs (the GPU must equal them bit for bit) and
the FP64 oracle's diagnostics.  Regenerate only when the step's definition
changes:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "pbf-cuda_amd"))

from oracle import oracle as orc  # noqa: E402
from pbf import params as P  # noqa: E402
from pbf import scenes  # noqa: E402


def cases():
    dam = P.dam_params()
    ref = P.reference_defaults()
    blk = scenes.block((16, 8, 16), (1.6, 0.8, 1.6), seed=101, vel_sigma=0.3)
    yield "block2k_dam", dam, blk
    yield "block2k_reference", ref, blk
    wall = scenes.block((16, 16, 8), (1.0, 1.0, 1.0), seed=102, vel_sigma=1.0)
    yield "block2k_reference_wall", ref.replace(k_boundaryDensity=0.5), wall
    lat = scenes.lattice((12, 12, 12), (1.0, 1.0, 1.0), 0.05, 0.0, seed=1)
    yield "lattice1728_dam", dam.replace(g=0.0, niter=2), scenes.Scene(
        lat, np.zeros_like(lat), np.arange(lat.shape[0], dtype=np.uint32), (3.0, 3.0, 3.0),
        (0.0, 0.0, 0.0))
    two = np.array([[1.0, 1.0, 1.0], [1.06, 1.0, 1.0]], np.float32)
    yield "two_particles", P.FluidParams(niter=2, pho0=1000.0, g=0.0, h=0.1, dt=1 / 60,
                                         lambda_eps=600.0, delta_q=0.02, k_corr=0.1, n_corr=4.0,
                                         k_boundaryDensity=0.0, c_XSPH=0.5), scenes.Scene(
        two, np.zeros_like(two), np.arange(2, dtype=np.uint32), (2.0, 2.0, 2.0), (0.0, 0.0, 0.0))


def params_array(p):
    return np.array([p.niter, p.pho0, p.g, p.h, p.dt, p.lambda_eps, p.delta_q, p.k_corr,
                     p.n_corr, p.k_boundaryDensity, p.c_XSPH], np.float64)


def main():
    for name, p, sc in cases():
        pos, vel = sc.as_layout(4)
        m = orc.step(p, sc.ulim, sc.llim, pos, vel, sc.iid, fp32=True)
        o = orc.step(p, sc.ulim, sc.llim, pos, vel, sc.iid)
        np.savez_compressed(
            os.path.join(HERE, name + ".npz"), params=params_array(p),
            ulim=np.array(sc.ulim, np.float32), llim=np.array(sc.llim, np.float32),
            in_pos=pos, in_vel=vel, in_iid=sc.iid,
            f32_pos=m["pos"], f32_vel=m["vel"], f32_npos=m["npos"], f32_nvel=m["nvel"],
            f32_iid=m["iid"], keys=m["keys"], perm=m["perm"],
            f64_x=o["x_final"], f64_v=o["v_final"], f64_lambda=o["lambda"], f64_rho=o["rho"])
        print(f"{name}: n={sc.n}")


if __name__ == "__main__":
    main()
