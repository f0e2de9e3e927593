"""GPU test of the allocation-free step (pbf_step_noalloc, through the C ABI):
two calls in a row, the second on a smaller scene in another box, each
equal bit for bit to pbf_step on the same input (itself bit-exact against
the oracle's FP32 mirror in test_gpu_parity.py), and its limits rejected
with PBF_ERR_CAPACITY."""
import ctypes

import numpy as np
import pytest
import torch

from conftest import from_dev, to_dev
from pbf import _lib
from pbf import params as P
from pbf import scenes

pytestmark = pytest.mark.gpu


def _vec(v):
    return _lib.PbfVec3(float(v[0]), float(v[1]), float(v[2]))


def _arrays(sc, device):
    pos, vel = sc.as_layout(4)
    return [to_dev(pos, device), torch.zeros(sc.n * 4, device=device), to_dev(vel, device),
            torch.zeros(sc.n * 4, device=device), to_dev(sc.iid, device)]


def _out(arrs, n):
    torch.cuda.synchronize()
    return {k: from_dev(t) for k, t in zip(("pos", "npos", "vel", "nvel", "iid"), arrs)}


def _noalloc(lib, p, sc, arrs, n, stream=None):
    pc = p.to_c()
    return lib.pbf_step_noalloc(ctypes.byref(pc), _vec(sc.ulim), _vec(sc.llim),
                                *[ctypes.c_void_p(t.data_ptr()) for t in arrs], n,
                                ctypes.c_void_p(stream.cuda_stream) if stream else None)


def _step(pbf_mod, p, sc, arrs, n):
    with pbf_mod.Simulator(p, sc.ulim, sc.llim, capacity=n) as sim:
        sim.step(*arrs, n)
        torch.cuda.synchronize()


def test_noalloc_step_equals_pbf_step_on_two_calls(pbf_mod, device):
    lib = _lib.load()
    p = P.dam_params()
    big = scenes.double_dam()  # 32,768 particles: the limit
    small = scenes.block((16, 16, 16), (1.2, 1.0, 1.4), seed=21, vel_sigma=1.0)
    assert big.n == 32768 and small.n < big.n
    stream = torch.cuda.Stream()
    for k, sc in enumerate((big, small)):
        print(f"no-alloc call {k}: {sc.n} particles", flush=True)
        want = _arrays(sc, device)
        _step(pbf_mod, p, sc, want, sc.n)
        got = _arrays(sc, device)
        with torch.cuda.stream(stream):
            rc = _noalloc(lib, p, sc, got, sc.n, stream)
        assert rc == 0, lib.pbf_last_error(None)
        g, w = _out(got, sc.n), _out(want, sc.n)
        for key in ("pos", "npos", "vel", "nvel", "iid"):
            assert np.array_equal(g[key].view(np.uint32), w[key].view(np.uint32)), \
                f"call {k}: {key} differs from pbf_step"


def test_noalloc_step_rejects_sizes_beyond_its_scratch(pbf_mod, device):
    lib = _lib.load()
    p = P.dam_params()
    sc = scenes.block((8, 8, 8), (1.0, 1.0, 1.0), seed=3)
    arrs = _arrays(sc, device)
    assert _noalloc(lib, p, sc, arrs, 32769) == _lib.PBF_ERR_CAPACITY
    assert b"32768" in lib.pbf_last_error(None)
    wide = scenes.Scene(sc.pos, sc.vel, sc.iid, (10.0, 10.0, 10.0), (0.0, 0.0, 0.0))
    assert _noalloc(lib, p, wide, arrs, sc.n) == _lib.PBF_ERR_CAPACITY  # 1e6 cells
    assert _noalloc(lib, p, sc, arrs, 0) == 0
