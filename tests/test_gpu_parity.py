"""GPU parity of the HIP step, through the C ABI (pbf.Simulator over
libpbf.so), against the CPU oracle:

* bit for bit against the oracle's FP32 mirror (the same FP32 operation
  sequence) at every size, layout and frame count, which pins ordering,
  neighbour sets, summation order and every rounding;
* within 1e-4 h of the FP64 oracle where the step is well conditioned (the
  reference's default constants); the dam constants are ill-conditioned (see
  test_oracle.test_conditioning_of_the_dam_constants).
"""
import numpy as np
import pytest
import torch

from conftest import from_dev, to_dev
from pbf import params as P
from pbf import scenes

pytestmark = pytest.mark.gpu


class Run:
    """One scene on the GPU with tic-toc buffers (fluids/FluidSystem.cpp:112-117)."""

    def __init__(self, pbf_mod, sc, p, layout=4, capacity=None, max_cells=0, device="cuda:0"):
        self.L = layout
        self.n = sc.n
        pos, vel = sc.as_layout(layout)
        cap = capacity or max(sc.n, 1)
        self.sim = pbf_mod.Simulator(p, sc.ulim, sc.llim, capacity=cap, max_cells=max_cells,
                                     layout=layout)
        self.a = [to_dev(pos, device), to_dev(vel, device)]
        self.b = [torch.zeros(max(cap, 1) * layout, device=device),
                  torch.zeros(max(cap, 1) * layout, device=device)]
        self.iid = to_dev(sc.iid, device)

    def step(self, stream=None):
        self.sim.step(self.a[0], self.b[0], self.a[1], self.b[1], self.iid, self.n, stream)
        out = self.state()
        self.a, self.b = self.b, self.a
        return out

    def state(self):
        torch.cuda.synchronize()
        L, n = self.L, self.n
        return dict(pos=from_dev(self.a[0])[: n * L].reshape(n, L),
                    vel=from_dev(self.a[1])[: n * L].reshape(n, L),
                    npos=from_dev(self.b[0])[: n * L].reshape(n, L),
                    nvel=from_dev(self.b[1])[: n * L].reshape(n, L),
                    iid=from_dev(self.iid, dtype=np.uint32)[:n])


def assert_bitexact(g, m, what=""):
    L = g["npos"].shape[1]
    np.testing.assert_array_equal(g["iid"], m["iid"], err_msg=f"{what} iid")
    for k in ("pos", "vel", "npos", "nvel"):
        ga, ma = g[k][:, :3], m[k][:, :3]
        if not np.array_equal(ga.view(np.uint32), ma.view(np.uint32)):
            bad = np.where(np.any(ga.view(np.uint32) != ma.view(np.uint32), axis=1))[0]
            raise AssertionError(f"{what} {k}: {len(bad)} particles differ, first {bad[:5]}: "
                                 f"gpu {ga[bad[0]]} mirror {ma[bad[0]]}")
    if L == 4:
        assert np.all(g["npos"][:, 3] == 0) and np.all(g["nvel"][:, 3] == 0)


def mirror(orc, p, sc, pos, vel, iid):
    return orc.step(p, sc.ulim, sc.llim, pos, vel, iid, fp32=True)


@pytest.mark.parametrize("layout", [4, 3])
@pytest.mark.parametrize("pname", ["dam", "reference"])
def test_double_dam_step_bitexact(pbf_mod, orc, device, layout, pname):
    p = P.dam_params() if pname == "dam" else P.reference_defaults()
    sc = scenes.double_dam()
    run = Run(pbf_mod, sc, p, layout)
    g = run.step()
    pos, vel = sc.as_layout(layout)
    assert_bitexact(g, mirror(orc, p, sc, pos, vel, sc.iid), f"{pname}/L{layout}")


def test_double_dam_within_fp64_tolerance_when_well_conditioned(pbf_mod, orc, device):
    p = P.reference_defaults()
    sc = scenes.double_dam()
    g = Run(pbf_mod, sc, p).step()
    pos, vel = sc.as_layout(4)
    o = orc.step(p, sc.ulim, sc.llim, pos, vel, sc.iid)
    np.testing.assert_array_equal(g["iid"], o["iid"])
    np.testing.assert_array_equal(g["pos"], o["pos"])  # sorted start positions: exact
    tol = 1e-4 * p.h
    assert np.abs(g["npos"][:, :3] - o["x_final"]).max() <= tol
    assert np.abs(g["nvel"][:, :3] - o["v_final"]).max() <= tol / p.dt
    assert np.abs(g["vel"][:, :3] - o["v_pre"]).max() <= tol / p.dt


@pytest.mark.parametrize("n", [1, 2, 63, 64, 255, 256, 257, 2047, 2048, 2049, 4097, 6000])
def test_ragged_sizes_bitexact(pbf_mod, orc, device, n):
    full = scenes.block((20, 20, 16), (2.0, 2.0, 2.0), seed=21, vel_sigma=0.4)
    sc = scenes.Scene(full.pos[:n], full.vel[:n], full.iid[:n], full.ulim, full.llim)
    p = P.dam_params()
    g = Run(pbf_mod, sc, p).step()
    pos, vel = sc.as_layout(4)
    assert_bitexact(g, mirror(orc, p, sc, pos, vel, sc.iid), f"n={n}")


def test_zero_particles_is_a_noop(pbf_mod, device):
    sc = scenes.block((2, 2, 2), (1.0, 1.0, 1.0), seed=1)
    run = Run(pbf_mod, sc, P.dam_params())
    before = [t.clone() for t in run.a + run.b]
    run.sim.step(run.a[0], run.b[0], run.a[1], run.b[1], run.iid, 0)
    torch.cuda.synchronize()
    for x, y in zip(before, run.a + run.b):
        assert torch.equal(x, y)


def test_outside_box_and_piled_particles_bitexact(pbf_mod, orc, device):
    """Particles outside the box map to border cells; a pile of identical
    positions puts every key in one cell (all-equal keys in the sort)."""
    rng = np.random.default_rng(8)
    out = rng.uniform(-0.6, 2.6, size=(3000, 3)).astype(np.float32)
    pile = np.tile(np.array([[0.55, 0.55, 0.05]], np.float32), (1500, 1))
    pos = np.concatenate([out, pile])
    vel = rng.normal(0, 2.0, size=pos.shape).astype(np.float32)
    sc = scenes.Scene(pos, vel, np.arange(pos.shape[0], dtype=np.uint32), (2.0, 2.0, 2.0),
                      (0.0, 0.0, 0.0))
    for p in (P.dam_params(), P.reference_defaults().replace(k_boundaryDensity=0.5)):
        g = Run(pbf_mod, sc, p).step()
        P4, V4 = sc.as_layout(4)
        assert_bitexact(g, mirror(orc, p, sc, P4, V4, sc.iid), "outside/pile")


def test_tiny_box_single_cell(pbf_mod, orc, device):
    sc = scenes.block((4, 4, 4), (0.1, 0.1, 0.1), seed=2, lo=(0.0, 0.0, 0.0))
    sc.pos[:] = np.clip(sc.pos, 0.0, 0.099)
    p = P.reference_defaults()
    g = Run(pbf_mod, sc, p).step()
    pos, vel = sc.as_layout(4)
    assert_bitexact(g, mirror(orc, p, sc, pos, vel, sc.iid), "one cell")


def test_multiframe_bitexact_200_frames(pbf_mod, orc, device):
    """200 frames of the double dam with the dam constants (chaotic): the GPU
    stays bit-identical to the FP32 mirror on every frame."""
    p = P.dam_params()
    sc = scenes.double_dam()
    run = Run(pbf_mod, sc, p)
    pos, vel = sc.as_layout(4)
    iid = sc.iid
    for f in range(200):
        g = run.step()
        m = mirror(orc, p, sc, pos, vel, iid)
        assert_bitexact(g, m, f"frame {f}")
        pos, vel, iid = m["npos"], m["nvel"], m["iid"]


def test_moving_box_bitexact(pbf_mod, orc, device):
    """setLim between frames (fluids/FluidSystem.cpp:104-110): the wall at
    x = ulim moves sinusoidally; grid dims change with it."""
    p = P.dam_params()
    sc = scenes.block((32, 16, 32), (3.0, 1.6, 2.0), seed=31)
    run = Run(pbf_mod, sc, p, max_cells=4 * 40 * 16 * 20)
    pos, vel = sc.as_layout(4)
    iid = sc.iid
    for f in range(12):
        ux = float(np.float32(3.0 + 0.8 * np.sin(0.5 * f)))
        ul = (ux, 1.6, 2.0)
        run.sim.setLim(ul, sc.llim)
        g = run.step()
        m = orc.step(p, ul, sc.llim, pos, vel, iid, fp32=True)
        assert_bitexact(g, m, f"frame {f}")
        pos, vel, iid = m["npos"], m["nvel"], m["iid"]


def test_single_dam_1m_bitexact(pbf_mod, orc, device):
    p = P.dam_params()
    sc = scenes.single_dam()
    g = Run(pbf_mod, sc, p).step()
    pos, vel = sc.as_layout(4)
    assert_bitexact(g, mirror(orc, p, sc, pos, vel, sc.iid), "single dam 1M")


def test_single_dam_1m_fp64_tolerance_reference_constants(pbf_mod, orc, device):
    p = P.reference_defaults()
    sc = scenes.single_dam()
    g = Run(pbf_mod, sc, p).step()
    pos, vel = sc.as_layout(4)
    o = orc.step(p, sc.ulim, sc.llim, pos, vel, sc.iid)
    np.testing.assert_array_equal(g["iid"], o["iid"])
    assert np.abs(g["npos"][:, :3] - o["x_final"]).max() <= 1e-4 * p.h


def test_density_error_matches_oracle(pbf_mod, orc, device):
    p = P.dam_params()
    sc = scenes.double_dam()
    run = Run(pbf_mod, sc, p)
    for _ in range(3):
        g = run.step()
    mean_g, max_g = run.sim.density_error(run.a[0], sc.n)
    mean_o, max_o, _ = orc.density_error_f32(p, sc.ulim, sc.llim, g["npos"])
    assert mean_g == pytest.approx(mean_o, rel=1e-12)
    assert max_g == pytest.approx(max_o, rel=1e-15)
    d64 = orc.density_error(p, sc.ulim, sc.llim, g["npos"])
    assert mean_g == pytest.approx(d64[0], rel=1e-4)


def test_explicit_stream_and_determinism(pbf_mod, device):
    p = P.dam_params()
    sc = scenes.double_dam()
    s = torch.cuda.Stream()
    r1 = Run(pbf_mod, sc, p)
    r2 = Run(pbf_mod, sc, p)
    with torch.cuda.stream(s):
        for _ in range(3):
            g1 = r1.step(stream=s)
    for _ in range(3):
        g2 = r2.step()
    for k in g1:
        np.testing.assert_array_equal(g1[k], g2[k])


def test_capacity_and_argument_errors(pbf_mod, device):
    p = P.dam_params()
    sc = scenes.block((4, 4, 4), (1.0, 1.0, 1.0), seed=1)
    run = Run(pbf_mod, sc, p, capacity=64, max_cells=2000)
    raw = [t.data_ptr() for t in (run.a[0], run.b[0], run.a[1], run.b[1], run.iid)]
    with pytest.raises(pbf_mod.PBFError) as e:  # raw pointers reach the C ABI's own check
        run.sim.step(*raw, 65)
    assert e.value.code == 2
    with pytest.raises(ValueError):  # the Python wrapper checks tensor sizes first
        run.sim.step(run.a[0], run.b[0], run.a[1], run.b[1], run.iid, 65)
    with pytest.raises(pbf_mod.PBFError) as e:
        run.sim.setLim((3.0, 3.0, 3.0), (0, 0, 0))  # 27000 cells > 2000
    assert e.value.code == 2
    assert run.sim.getLim()[0] == pytest.approx((1.0, 1.0, 1.0))  # unchanged
    base = run.a[0].data_ptr()
    with pytest.raises(pbf_mod.PBFError) as e:  # misaligned float4 array
        run.sim.step(base + 4, run.b[0], run.a[1], run.b[1], run.iid, 8)
    assert e.value.code == 1
    with pytest.raises(pbf_mod.PBFError) as e:  # aliased arrays
        run.sim.step(run.a[0], run.a[0], run.a[1], run.b[1], run.iid, 8)
    assert e.value.code == 1
    with pytest.raises(pbf_mod.PBFError) as e:
        run.sim.loadParams(p.replace(h=0.05))  # 20^3 cells > max_cells
    assert e.value.code in (1, 2)
    run.sim.loadParams(p.replace(niter=2))
    assert run.sim.saveParams().niter == 2
    run.step()  # still usable after the rejected calls


def test_phase_timing(pbf_mod, device):
    p = P.dam_params()
    sc = scenes.double_dam()
    run = Run(pbf_mod, sc, p)
    run.sim.set_timing(True)
    run.step()
    run.step()
    t = run.sim.get_timing()
    assert t["lambda"][1] == 2 * p.niter and t["delta"][1] == 2 * p.niter
    assert t["step"][1] == 2 and t["step"][0] > 0
    assert run.sim.get_timing()["step"][1] == 0  # reset


def test_run_frames_sweeping_wall_bitexact(pbf_mod, orc, device):
    """The native multi-frame driver (tic-toc + sweeping wall,
    fluids/FluidSystem.cpp:99-120) equals stepping the FP32 mirror frame by
    frame with the same boxes."""
    p = P.dam_params()
    sc = scenes.block((32, 16, 32), (3.0, 1.6, 2.0), seed=41)
    sw = pbf_mod.Sweep(sc.ulim, sc.llim, a_ulim=(0.6, 0.0, 0.0), w=2 * np.pi / 20, start_frame=2)
    run = Run(pbf_mod, sc, p, max_cells=4 * 40 * 16 * 20)
    which = run.sim.run_frames(run.a[0], run.a[1], run.b[0], run.b[1], run.iid, sc.n, 0, 15, sw)
    torch.cuda.synchronize()
    assert which == 1  # 15 frames: odd count ends in the b buffers
    out = run.b if which else run.a
    pos, vel = sc.as_layout(4)
    iid = sc.iid
    for f in range(15):
        ul, ll = sw.box(f)
        m = orc.step(p, ul, ll, pos, vel, iid, fp32=True)
        pos, vel, iid = m["npos"], m["nvel"], m["iid"]
    assert run.sim.getLim() == sw.box(14)
    g_pos = from_dev(out[0]).reshape(-1, 4)
    g_vel = from_dev(out[1]).reshape(-1, 4)
    np.testing.assert_array_equal(from_dev(run.iid, dtype=np.uint32), iid)
    np.testing.assert_array_equal(g_pos.view(np.uint32), pos.view(np.uint32))
    np.testing.assert_array_equal(g_vel.view(np.uint32), vel.view(np.uint32))


def test_sweep_dam_4m_piled_frame_bitexact(pbf_mod, orc, device):
    """The 4,194,304-particle sweeping-wall workload (6 iterations) at full
    size: 60 frames on the GPU drive ~10,000 particles into one corner cell;
    frame 60, stepped by the GPU from that state, equals the FP32 mirror's
    step from the same state bit for bit."""
    p = P.dam_params(niter=6)
    sc = scenes.sweep_dam()
    sw = pbf_mod.Sweep(*scenes.sweep_of_sweep_dam())
    print(f"sweep dam 4M: {sc.n} particles, 60 frames then one compared", flush=True)
    run = Run(pbf_mod, sc, p)
    which = run.sim.run_frames(run.a[0], run.a[1], run.b[0], run.b[1], run.iid, sc.n, 0, 60, sw)
    cur, nxt = (run.b, run.a) if which else (run.a, run.b)
    torch.cuda.synchronize()
    pos = from_dev(cur[0]).reshape(-1, 4)
    vel = from_dev(cur[1]).reshape(-1, 4)
    iid = from_dev(run.iid, dtype=np.uint32)
    which = run.sim.run_frames(cur[0], cur[1], nxt[0], nxt[1], run.iid, sc.n, 60, 1, sw)
    assert which == 1
    torch.cuda.synchronize()
    ul, ll = sw.box(60)
    m = orc.step(p, ul, ll, pos, vel, iid, fp32=True)
    assert int(np.bincount(m["keys"]).max()) > 5000  # the corner pile is there
    np.testing.assert_array_equal(from_dev(run.iid, dtype=np.uint32), m["iid"])
    np.testing.assert_array_equal(from_dev(nxt[0]).reshape(-1, 4).view(np.uint32),
                                  m["npos"].view(np.uint32))
    np.testing.assert_array_equal(from_dev(nxt[1]).reshape(-1, 4).view(np.uint32),
                                  m["nvel"].view(np.uint32))


def test_hit_masks_around_their_candidate_limit(pbf_mod, orc, device):
    """The lambda pass keeps hit masks for the delta-p pass of targets with at
    most 2048 candidates (32 blocks of 64).  A compressed lattice (spacing
    0.45h, ~60 to ~340 candidates) with two piles of identical positions in
    its corners (1,800 and 2,300 particles) has targets on both sides of that
    limit and with every block count up to it: bit-exact against the FP32
    mirror for three frames, with the adaptive heavy threshold, with every
    target above 64 candidates in the chunk kernels, and with no target in
    the chunk kernels (threshold 2^30: the light lanes store and load mask
    blocks 4..31 themselves)."""
    p = P.dam_params()
    base = scenes.block((24, 24, 24), (2.0, 2.0, 2.0), seed=57, vel_sigma=0.3)
    piles = [np.tile(np.array([[0.03, 0.03, 0.03]], np.float32), (1800, 1)),
             np.tile(np.array([[1.97, 1.97, 0.03]], np.float32), (2300, 1))]
    pos3 = np.concatenate([(base.pos * np.float32(0.9)).astype(np.float32)] + piles)
    vel3 = np.concatenate([base.vel] + [np.zeros_like(q) for q in piles])
    sc = scenes.Scene(pos3, vel3, np.arange(pos3.shape[0], dtype=np.uint32), base.ulim,
                      base.llim)
    pos, vel = sc.as_layout(4)
    m0 = mirror(orc, p, sc, pos, vel, sc.iid)
    # candidate counts of the first pass (27 cells around each sorted key)
    D = (20, 20, 20)
    cnt = np.bincount(m0["keys"].astype(np.int64), minlength=D[0] * D[1] * D[2]).reshape(D)
    pad = np.pad(cnt, 1)
    nb = sum(pad[dx:dx + D[0], dy:dy + D[1], dz:dz + D[2]]
             for dx in range(3) for dy in range(3) for dz in range(3))
    m = nb.reshape(-1)[m0["keys"].astype(np.int64)]
    assert (m <= 256).sum() > 1000 and ((m > 256) & (m <= 2048)).sum() > 1000, (m.min(), m.max())
    assert (m > 2048).sum() > 1000, m.max()
    for thr in (0, 64, 1 << 30):
        run = Run(pbf_mod, sc, p)
        run.sim.set_heavy_threshold(thr)
        P4, V4, I = pos, vel, sc.iid
        for f in range(3):
            g = run.step()
            mm = mirror(orc, p, sc, P4, V4, I)
            assert_bitexact(g, mm, f"threshold {thr} frame {f}")
            P4, V4, I = mm["npos"], mm["nvel"], mm["iid"]


def test_cpp_host_program_equals_run_frames(pbf_mod, device, tmp_path):
    """pbf-cuda_amd/examples/stepsim, FluidSystem::stepSimulate written as a
    C++ host program over include/pbf_simulator.hpp (setLim + step per frame,
    tic-toc buffers), run as a child process, equals the library's
    run_frames over the same 15 frames of a sweeping wall bit for bit."""
    import os
    import subprocess

    from conftest import ROOT

    exe = os.path.join(ROOT, "pbf-cuda_amd", "examples", "stepsim")
    if not os.path.exists(exe):
        pytest.fail("pbf-cuda_amd/examples/stepsim is not built (make -C pbf-cuda_amd)")
    p = P.dam_params()
    sc = scenes.block((32, 16, 32), (3.0, 1.6, 2.0), seed=41)
    sw = pbf_mod.Sweep(sc.ulim, sc.llim, a_ulim=(0.6, 0.0, 0.0), w=2 * np.pi / 20, start_frame=2)
    nf = 15
    pos, vel = sc.as_layout(4)
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    src.write_bytes(np.array([sc.n, nf], np.int32).tobytes() + bytes(p.to_c()) + bytes(sw.c)
                    + pos.tobytes() + vel.tobytes() + sc.iid.astype(np.uint32).tobytes())
    print(f"stepsim: {sc.n} particles, {nf} frames", flush=True)
    r = subprocess.run([exe, str(src), str(dst)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    raw = np.frombuffer(dst.read_bytes(), np.uint8)
    nb = sc.n * 16
    c_pos = raw[:nb].view(np.float32).reshape(-1, 4)
    c_vel = raw[nb:2 * nb].view(np.float32).reshape(-1, 4)
    c_iid = raw[2 * nb:].view(np.uint32)
    run = Run(pbf_mod, sc, p)
    which = run.sim.run_frames(run.a[0], run.a[1], run.b[0], run.b[1], run.iid, sc.n, 0, nf, sw)
    torch.cuda.synchronize()
    out = run.b if which else run.a
    np.testing.assert_array_equal(c_iid, from_dev(run.iid, dtype=np.uint32))
    np.testing.assert_array_equal(c_pos.view(np.uint32),
                                  from_dev(out[0]).reshape(-1, 4).view(np.uint32))
    np.testing.assert_array_equal(c_vel.view(np.uint32),
                                  from_dev(out[1]).reshape(-1, 4).view(np.uint32))


@pytest.mark.parametrize("box", [(30.0, 20.0, 15.0), (6.4, 6.4, 3.2)])
def test_large_and_small_grids_bitexact(pbf_mod, orc, device, box):
    """A sparse 300x200x150-cell box (9M cells: the cell-start scan runs its
    three-kernel form) and a 64x64x32 one (two-kernel form), two frames each,
    bit-exact against the FP32 mirror."""
    p = P.dam_params()
    sc = scenes.block((16, 16, 16), box, seed=21, vel_sigma=1.0)
    print(f"grid box {box} n {sc.n}", flush=True)
    run = Run(pbf_mod, sc, p, 4)
    pos, vel = sc.as_layout(4)
    iid = sc.iid
    for f in range(2):
        g = run.step()
        m = mirror(orc, p, sc, pos, vel, iid)
        assert_bitexact(g, m, f"box {box} frame {f}")
        pos, vel, iid = m["npos"], m["nvel"], m["iid"]
