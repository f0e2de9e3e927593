"""GPU tests of the differentiable rollout (pbf_tape_* through the C ABI and
pbf.autograd):

* the tape's forward run equals plain steps bit for bit;
* the GPU adjoint (FP32 kernels) against the oracle's FP64 adjoint along the
  same FP32 trajectory (the oracle's rollout with fp32=True replays the GPU's
  forward exactly), so the comparison isolates the backward arithmetic;
* the GPU gradient against FP64 central finite differences of the FP64
  rollout (reference constants, where the rollout is smooth at the step);
* the autograd.Function wrapper gives the same gradients as the C ABI;
* both ways of collecting the pair terms the stale grid makes one-sided
  (scatter: float atomics; gather: each pass over a grid of its own
  positions) match the oracle, the gather mode repeats bit for bit.
The oracle's adjoint itself is pinned by finite differences in
tests/test_adjoint_oracle.py.
"""
import numpy as np
import pytest
import torch

from conftest import from_dev, to_dev
from pbf import params as P
from pbf import scenes

pytestmark = pytest.mark.gpu


def _scene(counts=(8, 8, 8), box=(1.0, 1.0, 1.0), seed=5):
    return scenes.block(counts, box, seed=seed, vel_sigma=0.5)


def _target(sc, seed=1):
    rng = np.random.default_rng(seed)
    return sc.pos.astype(np.float64) + rng.normal(0, 0.05, (sc.n, 3))


def gpu_grad(pbf_mod, sc, p, nframes, target, layout=4, store=False, sweep=None, adjoint=None,
             rescale=False, scale_out=None):
    """Loss gradients from the C ABI: (g_pos0, g_vel0 as N x 3 in input order,
    g_rho0, final positions by identity).  With rescale, the gradients as the
    tape wrote them (true x 2^-e) and e appended to scale_out."""
    n = sc.n
    sim = pbf_mod.Simulator(p, sc.ulim, sc.llim, capacity=n, layout=layout)
    tape = pbf_mod.Tape(sim, n, nframes, store=store, adjoint=adjoint, rescale=rescale)
    pos, vel = sc.as_layout(layout)
    dev = "cuda:0"
    tape.forward(to_dev(pos, dev), to_dev(vel, dev), to_dev(sc.iid, dev), sweep=sweep)
    xf = torch.zeros(n * layout, device=dev)
    iidf = torch.zeros(n, dtype=torch.int32, device=dev)
    tape.read(nframes, pos=xf, iid=iidf)
    torch.cuda.synchronize()
    X = from_dev(xf).reshape(n, layout)[:, :3].astype(np.float64)
    I = from_dev(iidf).astype(np.int64)
    g = np.zeros((n, layout), np.float32)
    g[:, :3] = 2.0 * (X - target[I]) / n
    gp0 = torch.zeros(n * layout, device=dev)
    gv0 = torch.zeros(n * layout, device=dev)
    grho = tape.backward(to_dev(g, dev), None, gp0, gv0)
    if scale_out is not None:
        scale_out.append(tape.gradient_log2_scale())
    x_by_iid = np.zeros((n, 3))
    x_by_iid[I] = X
    return (from_dev(gp0).reshape(n, layout)[:, :3].astype(np.float64),
            from_dev(gv0).reshape(n, layout)[:, :3].astype(np.float64), grho, x_by_iid)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_tape_forward_equals_steps(pbf_mod, device):
    from test_gpu_parity import Run

    p = P.dam_params()
    sc = _scene()
    run = Run(pbf_mod, sc, p, 4)
    sim = pbf_mod.Simulator(p, sc.ulim, sc.llim, capacity=sc.n)
    tape = pbf_mod.Tape(sim, sc.n, 5)
    pos, vel = sc.as_layout(4)
    tape.forward(to_dev(pos, device), to_dev(vel, device), to_dev(sc.iid, device))
    for f in range(1, 6):
        g = run.step()
        xp = torch.zeros(sc.n * 4, device=device)
        xv = torch.zeros(sc.n * 4, device=device)
        xi = torch.zeros(sc.n, dtype=torch.int32, device=device)
        tape.read(f, xp, xv, xi)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(from_dev(xp).reshape(-1, 4), g["npos"])
        np.testing.assert_array_equal(from_dev(xv).reshape(-1, 4), g["nvel"])
        np.testing.assert_array_equal(from_dev(xi).astype(np.uint32), g["iid"])


@pytest.mark.parametrize("adjoint", ["scatter", "gather"])
@pytest.mark.parametrize("pname,nframes,layout", [("reference", 3, 4), ("reference", 3, 3),
                                                  ("dam", 1, 4)])
def test_adjoint_matches_oracle_on_the_same_trajectory(pbf_mod, orc, device, pname, nframes,
                                                       layout, adjoint):
    p = P.dam_params() if pname == "dam" else P.reference_defaults()
    sc = _scene()
    tgt = _target(sc)
    print(f"adjoint {pname} frames {nframes} layout {layout} n {sc.n} {adjoint}", flush=True)
    gp, gv, gr, xf = gpu_grad(pbf_mod, sc, p, nframes, tgt, layout, adjoint=adjoint)
    o = orc.rollout(p, sc.ulim, sc.llim, sc.pos, sc.vel, sc.iid, nframes, tgt, fp32=True)
    np.testing.assert_array_equal(xf.astype(np.float32), o["x_final"].astype(np.float32))
    assert rel(gv, o["g_vel0"]) < 1e-4, rel(gv, o["g_vel0"])
    assert rel(gp, o["g_pos0"]) < 1e-4, rel(gp, o["g_pos0"])
    assert abs(gr - o["g_rho0"]) < 1e-3 * abs(o["g_rho0"]), (gr, o["g_rho0"])


@pytest.mark.parametrize("adjoint", ["scatter", "gather"])
def test_adjoint_with_scattered_pairs(pbf_mod, orc, device, adjoint):
    """Fast particles leave their stale cells, so candidate lists become
    asymmetric: the scatter mode adds the one-sided terms with atomics, the
    gather mode finds them on its query grid; both equal the oracle."""
    p = P.reference_defaults()
    sc = scenes.block((8, 8, 8), (1.0, 1.0, 1.0), seed=9, vel_sigma=6.0)
    tgt = _target(sc)
    print(f"one-sided pairs, {adjoint}", flush=True)
    gp, gv, gr, _ = gpu_grad(pbf_mod, sc, p, 2, tgt, adjoint=adjoint)
    o = orc.rollout(p, sc.ulim, sc.llim, sc.pos, sc.vel, sc.iid, 2, tgt, fp32=True)
    assert rel(gv, o["g_vel0"]) < 1e-4, rel(gv, o["g_vel0"])
    assert rel(gp, o["g_pos0"]) < 1e-4, rel(gp, o["g_pos0"])


def test_gpu_gradient_against_fp64_central_differences(pbf_mod, orc, device):
    """GPU gradient (FP32 trajectory and adjoint) against FP64 central
    differences (step 1e-4) of the FP64 rollout: within 1e-3 relative."""
    p = P.reference_defaults()
    sc = _scene()
    tgt = _target(sc)
    nf = 5
    gp, gv, gr, _ = gpu_grad(pbf_mod, sc, p, nf, tgt)
    rng = np.random.default_rng(11)
    loss = lambda v=sc.vel, rho0=0.0: orc.rollout(p, sc.ulim, sc.llim, sc.pos, v, sc.iid, nf, tgt,
                                                  pho0=rho0, grad=False)["loss"]
    for _ in range(3):
        d = rng.normal(size=(sc.n, 3))
        fd = (loss(sc.vel + 1e-4 * d) - loss(sc.vel - 1e-4 * d)) / 2e-4
        an = float((gv * d).sum())
        assert abs(fd - an) <= 1e-3 * abs(fd), (fd, an)
    rho0 = float(np.float32(p.pho0))
    fd = (loss(rho0=rho0 * (1 + 1e-4)) - loss(rho0=rho0 * (1 - 1e-4))) / (2e-4 * rho0)
    assert abs(fd - gr) <= 1e-3 * abs(fd), (fd, gr)


def test_autograd_function(pbf_mod, orc, device):
    from pbf.autograd import rollout

    p = P.reference_defaults()
    sc = _scene()
    tgt = _target(sc)
    n = sc.n
    sim = pbf_mod.Simulator(p, sc.ulim, sc.llim, capacity=n)
    tape = pbf_mod.Tape(sim, n, 3)
    pos, vel = sc.as_layout(4)
    pos0 = to_dev(pos, device).view(n, 4).requires_grad_(True)
    vel0 = to_dev(vel, device).view(n, 4).requires_grad_(True)
    rho0 = torch.tensor(float(np.float32(p.pho0)), dtype=torch.float64, device=device,
                        requires_grad=True)
    iid0 = to_dev(sc.iid, device)
    x = rollout(sim, tape, pos0, vel0, rho0, iid0)
    loss = ((x.double() - torch.from_numpy(tgt).to(device)) ** 2).sum() / n
    loss.backward()
    o = orc.rollout(p, sc.ulim, sc.llim, sc.pos, sc.vel, sc.iid, 3, tgt, fp32=True)
    assert abs(loss.item() - o["loss"]) < 1e-6 * o["loss"]
    gv = vel0.grad.cpu().numpy()[:, :3].astype(np.float64)
    gp = pos0.grad.cpu().numpy()[:, :3].astype(np.float64)
    assert rel(gv, o["g_vel0"]) < 1e-4 and rel(gp, o["g_pos0"]) < 1e-4
    assert abs(rho0.grad.item() - o["g_rho0"]) < 1e-3 * abs(o["g_rho0"])


def test_tape_argument_errors(pbf_mod, device):
    p = P.reference_defaults()
    sc = _scene()
    sim = pbf_mod.Simulator(p, sc.ulim, sc.llim, capacity=sc.n)
    with pytest.raises(pbf_mod.PBFError):
        pbf_mod.Tape(sim, sc.n + 1, 3)
    with pytest.raises(pbf_mod.PBFError):
        pbf_mod.Tape(sim, sc.n, 0)
    tape = pbf_mod.Tape(sim, sc.n, 2)
    g = torch.zeros(sc.n * 4, device=device)
    with pytest.raises(pbf_mod.PBFError) as e:
        tape.backward(None, None, g, g.clone())
    assert e.value.code == 5  # no forward run yet
    pos, vel = sc.as_layout(4)
    tape.forward(to_dev(pos, device), to_dev(vel, device), to_dev(sc.iid, device))
    sim.loadParams(p.replace(c_XSPH=0.1))
    with pytest.raises(pbf_mod.PBFError) as e:
        tape.backward(None, None, g, g.clone())
    assert e.value.code == 5  # parameters changed
    bsim = pbf_mod.BatchedSimulator(p, [sc.ulim], [sc.llim], [sc.n])
    with pytest.raises(pbf_mod.PBFError) as e:
        pbf_mod.Tape(bsim, sc.n, 2)
    assert e.value.code == 5


@pytest.mark.parametrize("adjoint", ["scatter", "gather"])
def test_adjoint_with_a_moving_wall_matches_oracle(pbf_mod, orc, device, adjoint):
    p = P.reference_defaults()
    sc = _scene()
    tgt = _target(sc)
    sw = pbf_mod.Sweep(sc.ulim, sc.llim, a_ulim=(-0.75, 0.0, 0.0), w=0.5)
    n = sc.n
    print(f"moving wall, {adjoint}", flush=True)
    sim = pbf_mod.Simulator(p, sc.ulim, sc.llim, capacity=n)
    tape = pbf_mod.Tape(sim, n, 3, adjoint=adjoint)
    pos, vel = sc.as_layout(4)
    tape.forward(to_dev(pos, device), to_dev(vel, device), to_dev(sc.iid, device), sweep=sw)
    assert sim.getLim()[0][0] == pytest.approx(sw.box(2)[0][0])
    xf = torch.zeros(n * 4, device=device)
    iidf = torch.zeros(n, dtype=torch.int32, device=device)
    tape.read(3, pos=xf, iid=iidf)
    torch.cuda.synchronize()
    X = from_dev(xf).reshape(n, 4)[:, :3].astype(np.float64)
    I = from_dev(iidf).astype(np.int64)
    g = np.zeros((n, 4), np.float32)
    g[:, :3] = 2.0 * (X - tgt[I]) / n
    gp0 = torch.zeros(n * 4, device=device)
    gv0 = torch.zeros(n * 4, device=device)
    tape.backward(to_dev(g, device), None, gp0, gv0)
    gv = from_dev(gv0).reshape(n, 4)[:, :3].astype(np.float64)
    o = orc.rollout(p, sc.ulim, sc.llim, sc.pos, sc.vel, sc.iid, 3, tgt, fp32=True,
                    boxes=[sw.box(f) for f in range(3)])
    x_by_iid = np.zeros((n, 3))
    x_by_iid[I] = X
    np.testing.assert_array_equal(x_by_iid.astype(np.float32), o["x_final"].astype(np.float32))
    assert rel(gv, o["g_vel0"]) < 1e-4, rel(gv, o["g_vel0"])


@pytest.mark.parametrize("pname,layout,moving", [("reference", 4, False), ("dam", 3, False),
                                                 ("reference", 4, True)])
def test_stored_tape_gives_the_same_gradients(pbf_mod, device, pname, layout, moving):
    """A tape that stores its iterations (PBF_TAPE_STORE) gives the gradients
    of one that recomputes them, including under a moving wall (per-frame
    grids of different sizes).  The forward states are the same bits; the
    gradients agree to the run-to-run spread of the adjoint's atomic scatter
    sums (pairs seen from one side only), whose order of addition varies."""
    p = P.dam_params() if pname == "dam" else P.reference_defaults()
    sc = scenes.block((8, 8, 8), (1.0, 1.0, 1.0), seed=9, vel_sigma=3.0)
    tgt = _target(sc)
    sw = (pbf_mod.Sweep(sc.ulim, sc.llim, a_ulim=(-0.75, 0.0, 0.0), w=0.5) if moving else None)
    print(f"stored tape {pname} layout {layout} moving {moving}", flush=True)
    a = gpu_grad(pbf_mod, sc, p, 4, tgt, layout, store=False, sweep=sw)
    b = gpu_grad(pbf_mod, sc, p, 4, tgt, layout, store=True, sweep=sw)
    np.testing.assert_array_equal(a[3], b[3])  # final positions
    assert rel(b[0], a[0]) < 1e-4, rel(b[0], a[0])
    assert rel(b[1], a[1]) < 1e-4, rel(b[1], a[1])
    assert abs(b[2] - a[2]) <= 1e-4 * abs(a[2]), (a[2], b[2])


def test_stored_tape_backward_twice_and_after_steps(pbf_mod, device):
    """A stored tape's backward pass can run again, also after unrelated
    steps of the solver, with the same result."""
    p = P.reference_defaults()
    sc = _scene()
    tgt = _target(sc)
    n = sc.n
    sim = pbf_mod.Simulator(p, sc.ulim, sc.llim, capacity=n)
    tape = pbf_mod.Tape(sim, n, 3, store=True)
    pos, vel = sc.as_layout(4)
    tape.forward(to_dev(pos, device), to_dev(vel, device), to_dev(sc.iid, device))
    g = torch.from_numpy(np.tile(np.float32([0.1, -0.2, 0.3, 0.0]), n)).to(device)
    outs = []
    for rep in range(2):
        gp0 = torch.zeros(n * 4, device=device)
        gv0 = torch.zeros(n * 4, device=device)
        gr = tape.backward(g, None, gp0, gv0)
        outs.append((from_dev(gp0), from_dev(gv0), gr))
        # unrelated steps in between
        x = [to_dev(pos, device), to_dev(vel, device)]
        y = [torch.zeros(n * 4, device=device), torch.zeros(n * 4, device=device)]
        ids = to_dev(sc.iid, device)
        sim.step(x[0], y[0], x[1], y[1], ids, n)
        torch.cuda.synchronize()
    assert rel(outs[1][0], outs[0][0]) < 1e-5
    assert rel(outs[1][1], outs[0][1]) < 1e-5
    assert abs(outs[1][2] - outs[0][2]) <= 1e-5 * abs(outs[0][2])


def test_tape_rejects_unknown_flags(pbf_mod, device):
    import ctypes

    from pbf import _lib

    p = P.reference_defaults()
    sc = _scene()
    sim = pbf_mod.Simulator(p, sc.ulim, sc.llim, capacity=sc.n)
    lib = _lib.load()
    t = ctypes.c_void_p()
    assert lib.pbf_tape_create_ex(sim._handle(), sc.n, 2, 16, ctypes.byref(t)) == 1
    assert b"flags" in lib.pbf_last_error(sim._handle())
    assert lib.pbf_tape_create_ex(sim._handle(), sc.n, 2, 2 | 4, ctypes.byref(t)) == 1
    assert b"exclude" in lib.pbf_last_error(sim._handle())
    with pytest.raises(ValueError):
        pbf_mod.Tape(sim, sc.n, 2, adjoint="atomic")


def test_gather_adjoint_needs_small_grids(pbf_mod, device):
    """A forced gather-mode tape on a grid wider than 1024 cells fails its
    backward pass with PBF_ERR_INVALID_ARGUMENT; the automatic mode falls
    back to scattering there and still gives finite gradients."""
    p = P.reference_defaults()
    sc = scenes.block((8, 8, 8), (1.0, 1.0, 1.0), seed=5, vel_sigma=0.5)
    ulim = (110.0, 1.0, 1.0)  # 1100 cells along x at h = 0.1
    n = sc.n
    out = {}
    for adjoint in ("gather", None):
        sim = pbf_mod.Simulator(p, ulim, sc.llim, capacity=n)
        tape = pbf_mod.Tape(sim, n, 1, adjoint=adjoint)
        pos, vel = sc.as_layout(4)
        tape.forward(to_dev(pos, device), to_dev(vel, device), to_dev(sc.iid, device))
        g = torch.ones(n * 4, device=device)
        gp0 = torch.zeros(n * 4, device=device)
        gv0 = torch.zeros(n * 4, device=device)
        if adjoint == "gather":
            with pytest.raises(RuntimeError, match="1024"):
                tape.backward(g, None, gp0, gv0)
        else:
            tape.backward(g, None, gp0, gv0)
            out["g"] = from_dev(gv0)
        tape.close()
    assert np.all(np.isfinite(out["g"])) and np.abs(out["g"]).sum() > 0


@pytest.mark.parametrize("store", [False, True])
def test_gather_adjoint_repeats_bit_for_bit_and_matches_scatter(pbf_mod, device, store):
    """The gather mode adds every term in a fixed order: two backward passes
    (and a second tape) give the same bits; the scatter mode, whose atomics
    add the one-sided terms in varying order, agrees within 1e-5."""
    p = P.reference_defaults()
    sc = scenes.block((10, 10, 10), (1.0, 1.0, 1.0), seed=9, vel_sigma=4.0)
    tgt = _target(sc)
    a = gpu_grad(pbf_mod, sc, p, 3, tgt, store=store, adjoint="gather")
    b = gpu_grad(pbf_mod, sc, p, 3, tgt, store=store, adjoint="gather")
    c = gpu_grad(pbf_mod, sc, p, 3, tgt, store=store, adjoint="scatter")
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    assert a[2] == b[2]
    np.testing.assert_array_equal(a[3], c[3])
    assert rel(a[0], c[0]) < 1e-5, rel(a[0], c[0])
    assert rel(a[1], c[1]) < 1e-5, rel(a[1], c[1])
    assert abs(a[2] - c[2]) <= 1e-5 * abs(c[2]), (a[2], c[2])


# ---- the differentiable workload at its own size --------------------------
# 16,384 particles (32x16x32 jittered lattice), v0 ~ N(0, 0.5^2), 30 frames,
# 4 iterations; loss = mean squared distance to a seeded target cloud by
# identity (scenes.differentiable / scenes.target_cloud).

def _c4():
    sc = scenes.differentiable()
    return sc, scenes.target_cloud(sc)


def _unit_dirs(shape, k, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(k):
        d = rng.normal(size=shape)
        out.append(d / np.linalg.norm(d))
    return out


@pytest.mark.parametrize("store,adjoint", [(False, None), (True, None), (False, "gather")])
def test_c4_adjoint_matches_oracle_over_30_frames(pbf_mod, orc, device, store, adjoint):
    """Reference constants, all 30 frames: the GPU's final state equals the
    oracle's FP32 trajectory bit for bit and its gradients (checkpointed tape
    with the iterations recomputed, and the stored tape; the automatic mode
    scatters at this size, and the gather mode) equal the oracle's adjoint
    along that trajectory within 1e-4."""
    p = P.reference_defaults()
    sc, tgt = _c4()
    print(f"c4 adjoint, reference constants, 30 frames, store={store} {adjoint}", flush=True)
    gp, gv, gr, xf = gpu_grad(pbf_mod, sc, p, 30, tgt, store=store, adjoint=adjoint)
    o = orc.rollout(p, sc.ulim, sc.llim, sc.pos, sc.vel, sc.iid, 30, tgt, fp32=True)
    np.testing.assert_array_equal(xf.astype(np.float32), o["x_final"].astype(np.float32))
    assert np.all(np.isfinite(gv)) and np.all(np.isfinite(gp)) and np.isfinite(gr)
    assert rel(gv, o["g_vel0"]) < 1e-4, rel(gv, o["g_vel0"])
    assert rel(gp, o["g_pos0"]) < 1e-4, rel(gp, o["g_pos0"])
    assert abs(gr - o["g_rho0"]) < 1e-4 * abs(o["g_rho0"]), (gr, o["g_rho0"])


def test_c4_gpu_gradient_against_fp64_central_differences(pbf_mod, orc, device):
    """Reference constants: the GPU gradient w.r.t. v0 (three random unit
    directions, JVP form) and pho0 against FP64 central differences (step
    1e-4) of the FP64 rollout, within 1e-3, over 15 frames.  The FP32 and FP64 trajectories drift apart with the horizon
    (gradient difference ~7e-6 at 5 frames, 6e-5 at 15, 2e-4 at 20, 5e-3 at
    30, measured on the oracle), so the full 30 frames are checked against
    the oracle's adjoint of the same trajectory above and the oracle's FP64
    adjoint against these differences in tests/test_adjoint_oracle.py."""
    p = P.reference_defaults()
    sc, tgt = _c4()
    nf = 15
    print(f"c4 gradient vs FP64 central differences, {nf} frames", flush=True)
    gp, gv, gr, _ = gpu_grad(pbf_mod, sc, p, nf, tgt)

    def loss(v=sc.vel, rho0=0.0):
        return orc.rollout(p, sc.ulim, sc.llim, sc.pos, v, sc.iid, nf, tgt, pho0=rho0,
                           grad=False)["loss"]

    for d in _unit_dirs(sc.vel.shape, 3, 3):
        fd = (loss(sc.vel + 1e-4 * d) - loss(sc.vel - 1e-4 * d)) / 2e-4
        an = float((gv * d).sum())
        assert abs(fd - an) <= 1e-3 * abs(fd), (fd, an)
    rho0 = float(np.float32(p.pho0))
    fd = (loss(rho0=rho0 + 1e-4) - loss(rho0=rho0 - 1e-4)) / 2e-4
    assert abs(fd - gr) <= 1e-3 * abs(fd), (fd, gr)


@pytest.mark.parametrize("nframes", [1, 2])
def test_c4_dam_constants_adjoint_matches_oracle(pbf_mod, orc, device, nframes):
    """Dam constants (pho0 6378, eps 600, s_corr k 0.1): the exact derivative
    grows by ~1e3-1e4 per frame, so only the first frames stay finite in
    FP32.  There the GPU adjoint equals the oracle's adjoint along the same
    FP32 trajectory within 1e-4."""
    p = P.dam_params()
    sc, tgt = _c4()
    print(f"c4 adjoint, dam constants, {nframes} frame(s)", flush=True)
    gp, gv, gr, xf = gpu_grad(pbf_mod, sc, p, nframes, tgt)
    o = orc.rollout(p, sc.ulim, sc.llim, sc.pos, sc.vel, sc.iid, nframes, tgt, fp32=True)
    np.testing.assert_array_equal(xf.astype(np.float32), o["x_final"].astype(np.float32))
    assert np.all(np.isfinite(gv)) and np.all(np.isfinite(gp))
    assert rel(gv, o["g_vel0"]) < 1e-4, rel(gv, o["g_vel0"])
    assert rel(gp, o["g_pos0"]) < 1e-4, rel(gp, o["g_pos0"])
    assert abs(gr - o["g_rho0"]) < 1e-4 * abs(o["g_rho0"]), (gr, o["g_rho0"])


@pytest.mark.parametrize("adjoint", ["gather", "scatter"])
def test_rescaled_backward_equals_unscaled(pbf_mod, device, adjoint):
    """PBF_TAPE_RESCALE multiplies the adjoint state by powers of two between
    frames: its gradients times 2^e equal the unscaled ones (scaling by a
    power of two is exact, and the adjoint is linear: in gather mode, whose
    sums do not depend on timing, they agree to rounding of values that reach
    the subnormal range in one of the two passes only; in scatter mode within
    the atomics' run-to-run spread), and dL/dpho0 is the same value."""
    p = P.reference_defaults()
    sc = scenes.block((8, 8, 8), (1.0, 1.0, 1.0), seed=9, vel_sigma=3.0)
    tgt = _target(sc)
    e = []
    a = gpu_grad(pbf_mod, sc, p, 4, tgt, adjoint=adjoint)
    b = gpu_grad(pbf_mod, sc, p, 4, tgt, adjoint=adjoint, rescale=True, scale_out=e)
    print(f"{adjoint}: log2 scale {e[0]}", flush=True)
    assert e[0] != 0
    gp, gv = np.ldexp(b[0], e[0]), np.ldexp(b[1], e[0])
    if adjoint == "gather":
        assert rel(gp, a[0]) < 1e-7 and rel(gv, a[1]) < 1e-7
        assert b[2] == pytest.approx(a[2], rel=1e-9)
    else:
        assert rel(gp, a[0]) < 1e-5 and rel(gv, a[1]) < 1e-5
        assert b[2] == pytest.approx(a[2], rel=1e-5)


def test_rescaled_c4_dam_constants_gradient_is_finite(pbf_mod, device):
    """The dam constants' 30-frame gradient of the 16K workload exceeds the
    FP32 range (~1e110 in FP64); with PBF_TAPE_RESCALE the tape's outputs
    stay finite, with the exponent carrying the magnitude, and over the
    first frame (where the unscaled gradient is finite) they equal the
    unscaled ones times 2^e."""
    p = P.dam_params()
    sc, tgt = _c4()
    e30 = []
    gp, gv, gr, _ = gpu_grad(pbf_mod, sc, p, 30, tgt, rescale=True, scale_out=e30)
    print(f"c4 dam constants, 30 frames: log2 scale {e30[0]}, |g_v| scaled "
          f"{np.linalg.norm(gv):.3g}, dL/dpho0 {gr:.3g}", flush=True)
    assert np.all(np.isfinite(gv)) and np.all(np.isfinite(gp)) and np.isfinite(gr)
    assert np.linalg.norm(gv) > 0 and e30[0] > 100
    e1 = []
    a = gpu_grad(pbf_mod, sc, p, 1, tgt)
    b = gpu_grad(pbf_mod, sc, p, 1, tgt, rescale=True, scale_out=e1)
    assert rel(np.ldexp(b[1], e1[0]), a[1]) < 1e-5
