"""Writes tests/golden/density_ensemble_dam.json: the frame-averaged mean and
max density error of 200 frames of the 32,768-particle double dam with the
dam constants, for an ensemble of FP64 oracle runs whose inputs differ by
at most 1 ulp per coordinate (member 0 unperturbed).

Those constants are chaotic over 200 frames, so the FP32 sequence (the GPU's,
bit for bit) can only be compared with the spread of such an ensemble.  The
ensemble costs minutes of CPU, so it is frozen here; the test re-runs member
0 to pin the fixture to the oracle.  Regenerate when the step's definition
changes:  python tests/golden/make_density_ensemble.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "pbf-cuda_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle import oracle as orc  # noqa: E402
from pbf import params as P  # noqa: E402
from pbf import scenes  # noqa: E402
from test_oracle import _density_run, _perturb_ulp  # noqa: E402

MEMBERS = 8
FRAMES = 200


def main():
    sc = scenes.double_dam()
    p = P.dam_params()
    pos, _ = sc.as_layout(4)
    out = []
    for s in range(MEMBERS):
        st = _density_run(orc, p, sc, FRAMES, False, _perturb_ulp(pos, s) if s else None)
        out.append([float(st[0]), float(st[1])])
        print(f"member {s}: mean {st[0]:.6f} max {st[1]:.6f}", flush=True)
    with open(os.path.join(HERE, "density_ensemble_dam.json"), "w") as f:
        json.dump({"scene": "double_dam_32k", "params": "dam", "frames": FRAMES,
                   "perturbation": "member s > 0: each position coordinate moved by -1, 0 or +1 "
                                   "ulp (numpy default_rng(s)); member 0 unperturbed",
                   "fp64_mean_max": out}, f, indent=1)


if __name__ == "__main__":
    main()
