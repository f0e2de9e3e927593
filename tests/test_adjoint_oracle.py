"""The oracle's rollout and adjoint (CPU).

The adjoint is pinned by FP64 central finite differences of the FP64
rollout, which need nothing from the adjoint's own derivation:

* reference constants (well conditioned): gradients with respect to initial
  velocities, positions and the rest density agree with central
  differences to 1e-6 (velocities) and 1e-5 (positions, rest density)
  relative over 1 and 10 frames;
* dam constants: about 40% of the first iteration's corrections exceed
  MAX_DP, so the rollout is piecewise smooth with kinks about 1e-6 apart;
  central differences converge to the adjoint once the step is below that
  (step 1e-8), over one frame.  Over more frames the exact derivative grows
  by ~1e3-1e4 per frame (3 frames: ~1e7) and no step resolves it, which is
  recorded here as a property of those constants.

The rollout itself is pinned to the step: its FP32 form equals chained
orc_step_f32 calls bit for bit, and its FP64 form agrees with orc_step.
test_planted_error_is_detected checks that the finite-difference
comparison is sharp enough to see a wrong gradient of small size.
"""
import numpy as np
import pytest

from pbf import params as P
from pbf import scenes


def _scene(seed=5):
    return scenes.block((8, 8, 8), (1.0, 1.0, 1.0), seed=seed, vel_sigma=0.5)


def _target(sc, seed=1):
    rng = np.random.default_rng(seed)
    return sc.pos.astype(np.float64) + rng.normal(0, 0.05, (sc.n, 3))


def test_rollout_fp32_equals_chained_steps(orc):
    p = P.dam_params()
    sc = _scene()
    r = orc.rollout(p, sc.ulim, sc.llim, sc.pos, sc.vel, sc.iid, 4, fp32=True, grad=False)
    pos, vel = sc.as_layout(4)
    iid = sc.iid.copy()
    for _ in range(4):
        m = orc.step(p, sc.ulim, sc.llim, pos, vel, iid, fp32=True)
        pos, vel, iid = m["npos"], m["nvel"], m["iid"]
    by_iid = np.zeros((sc.n, 3), np.float32)
    by_iid[iid] = pos[:, :3]
    np.testing.assert_array_equal(r["x_final"].astype(np.float32), by_iid)


def test_rollout_fp64_one_frame_matches_step(orc):
    p = P.reference_defaults()
    sc = _scene()
    r = orc.rollout(p, sc.ulim, sc.llim, sc.pos, sc.vel, sc.iid, 1, grad=False)
    pos, vel = sc.as_layout(4)
    o = orc.step(p, sc.ulim, sc.llim, pos, vel, sc.iid)
    by_iid = np.zeros((sc.n, 3))
    by_iid[o["iid"]] = o["x_final"]
    # orc_step rounds its predicted positions to FP32; the rollout keeps FP64
    assert np.abs(r["x_final"] - by_iid).max() < 1e-6 * p.h


@pytest.mark.parametrize("nframes", [1, 10])
def test_adjoint_matches_central_differences_reference_constants(orc, nframes):
    p = P.reference_defaults()
    sc = _scene()
    tgt = _target(sc)
    r = orc.rollout(p, sc.ulim, sc.llim, sc.pos, sc.vel, sc.iid, nframes, tgt)
    loss = lambda **kw: orc.rollout(p, sc.ulim, sc.llim, kw.get("x", sc.pos), kw.get("v", sc.vel),
                                    sc.iid, nframes, tgt, pho0=kw.get("rho0", 0.0),
                                    grad=False)["loss"]
    rng = np.random.default_rng(7)
    for _ in range(2):
        d = rng.normal(size=(sc.n, 3))
        eps = 1e-5
        fd = (loss(v=sc.vel + eps * d) - loss(v=sc.vel - eps * d)) / (2 * eps)
        an = float((r["g_vel0"] * d).sum())
        assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)
        eps = 1e-7  # positions: smaller step (1e-4 crosses clamp kinks)
        fd = (loss(x=sc.pos + eps * d) - loss(x=sc.pos - eps * d)) / (2 * eps)
        an = float((r["g_pos0"] * d).sum())
        assert abs(fd - an) <= 1e-5 * abs(an), (fd, an)
    rho0 = float(np.float32(p.pho0))
    e = 1e-4 * rho0
    fd = (loss(rho0=rho0 + e) - loss(rho0=rho0 - e)) / (2 * e)
    assert abs(fd - r["g_rho0"]) <= 1e-5 * abs(r["g_rho0"]), (fd, r["g_rho0"])


def test_adjoint_dam_constants_one_frame(orc):
    p = P.dam_params()
    sc = _scene()
    tgt = _target(sc)
    r = orc.rollout(p, sc.ulim, sc.llim, sc.pos, sc.vel, sc.iid, 1, tgt)
    rng = np.random.default_rng(3)
    d = rng.normal(size=(sc.n, 3))
    loss = lambda v: orc.rollout(p, sc.ulim, sc.llim, sc.pos, v, sc.iid, 1, tgt, grad=False)["loss"]
    an = float((r["g_vel0"] * d).sum())
    eps = 1e-8
    fd = (loss(sc.vel + eps * d) - loss(sc.vel - eps * d)) / (2 * eps)
    assert abs(fd - an) <= 1e-5 * abs(an), (fd, an)
    # the kinks: a step of 1e-5 crosses clamp boundaries and misses
    fd_big = (loss(sc.vel + 1e-5 * d) - loss(sc.vel - 1e-5 * d)) / 2e-5
    assert abs(fd_big - an) > 1e-2 * abs(an)


def test_dam_constants_gradient_growth(orc):
    """The dam constants' tangent dynamics: the exact derivative grows by
    orders of magnitude per frame (a property of the constants, not of the
    adjoint; the reference constants' gradients stay O(1))."""
    sc = _scene()
    tgt = _target(sc)
    g = {}
    for name, p in (("dam", P.dam_params()), ("ref", P.reference_defaults())):
        for nf in (1, 3):
            r = orc.rollout(p, sc.ulim, sc.llim, sc.pos, sc.vel, sc.iid, nf, tgt)
            g[name, nf] = np.abs(r["g_vel0"]).max()
    assert g["dam", 3] > 1e4 * g["dam", 1]
    assert g["ref", 3] < 10 * g["ref", 1]


def test_planted_error_is_detected(orc):
    """The comparison above is sharp: the adjoint of the loss with the
    target moved by 1e-3 (a stand-in for a wrong term of that size) already
    misses the unperturbed finite differences by far more than the bound."""
    p = P.reference_defaults()
    sc = _scene()
    tgt = _target(sc)
    wrong = orc.rollout(p, sc.ulim, sc.llim, sc.pos, sc.vel, sc.iid, 1, tgt + 1e-3)
    rng = np.random.default_rng(7)
    d = rng.normal(size=(sc.n, 3))
    loss = lambda v: orc.rollout(p, sc.ulim, sc.llim, sc.pos, v, sc.iid, 1, tgt, grad=False)["loss"]
    fd = (loss(sc.vel + 1e-4 * d) - loss(sc.vel - 1e-4 * d)) / 2e-4
    an = float((wrong["g_vel0"] * d).sum())
    assert abs(fd - an) > 1e-3 * abs(fd)


def test_adjoint_with_a_moving_wall(orc, pbf_mod):
    """Per-frame boxes (a wall pushing into the fluid, so box clamps are
    active): the adjoint still matches central differences."""
    p = P.reference_defaults()
    sc = _scene()
    tgt = _target(sc)
    sw = pbf_mod.Sweep(sc.ulim, sc.llim, a_ulim=(-0.75, 0.0, 0.0), w=0.5)
    boxes = [sw.box(f) for f in range(3)]
    assert boxes[2][0][0] < 0.4  # the wall has crossed the block's far face
    r = orc.rollout(p, sc.ulim, sc.llim, sc.pos, sc.vel, sc.iid, 3, tgt, boxes=boxes)
    rng = np.random.default_rng(5)
    d = rng.normal(size=(sc.n, 3))
    loss = lambda v: orc.rollout(p, sc.ulim, sc.llim, sc.pos, v, sc.iid, 3, tgt, grad=False,
                                 boxes=boxes)["loss"]
    fd = (loss(sc.vel + 1e-6 * d) - loss(sc.vel - 1e-6 * d)) / 2e-6
    an = float((r["g_vel0"] * d).sum())
    assert abs(fd - an) <= 1e-5 * abs(an), (fd, an)


@pytest.mark.slow
def test_c4_fp64_adjoint_against_central_differences_over_30_frames(orc):
    """The differentiable workload at its own size (16,384 particles, v0 ~
    N(0, 0.5^2), 30 frames, reference constants): the oracle's FP64 adjoint
    w.r.t. v0 (three random unit directions, JVP form) and pho0 agrees with
    FP64 central differences (step 1e-4) within 1e-3."""
    p = P.reference_defaults()
    sc = scenes.differentiable()
    tgt = scenes.target_cloud(sc)
    nf = 30
    o = orc.rollout(p, sc.ulim, sc.llim, sc.pos, sc.vel, sc.iid, nf, tgt)

    def loss(v=sc.vel, rho0=0.0):
        return orc.rollout(p, sc.ulim, sc.llim, sc.pos, v, sc.iid, nf, tgt, pho0=rho0,
                           grad=False)["loss"]

    rng = np.random.default_rng(3)
    for _ in range(3):
        d = rng.normal(size=sc.vel.shape)
        d /= np.linalg.norm(d)
        fd = (loss(sc.vel + 1e-4 * d) - loss(sc.vel - 1e-4 * d)) / 2e-4
        an = float((o["g_vel0"] * d).sum())
        assert abs(fd - an) <= 1e-3 * abs(fd), (fd, an)
    rho0 = float(np.float32(p.pho0))
    fd = (loss(rho0=rho0 + 1e-4) - loss(rho0=rho0 - 1e-4)) / 2e-4
    assert abs(fd - o["g_rho0"]) <= 1e-3 * abs(fd), (fd, o["g_rho0"])
