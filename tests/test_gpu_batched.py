"""GPU tests of the additions around the step, through the C ABI:

* batched scenes (pbf_create_batched / pbf_step_batched): every scene of a
  batch equals, bit for bit, the same scene stepped alone (itself bit-exact
  against the oracle's FP32 mirror), over several frames and both layouts,
  and one scene of the batch is checked against the mirror directly;
* graph mode: recorded-and-replayed steps equal eager steps bit for bit,
  replay after replay, with eager work and parameter changes in between,
  and the reference's per-frame loadParams + sweeping setLim keep replaying
  two graphs;
* launch shapes and the autotuner: every shape gives the same bits; the
  tuner leaves its inputs untouched and installs one of the candidates.
"""
import numpy as np
import pytest
import torch

from conftest import from_dev, to_dev
from pbf import params as P
from pbf import scenes
from test_gpu_parity import Run, assert_bitexact, mirror

pytestmark = pytest.mark.gpu


class BatchRun:
    """A batch of scenes on the GPU with tic-toc buffers."""

    def __init__(self, pbf_mod, scs, p, layout=4, device="cuda:0"):
        self.L = layout
        pos, vel, iid, counts, ul, ll, offs = scenes.concat(scs, layout)
        self.counts, self.offs = counts, offs
        self.n = int(sum(counts))
        self.sim = pbf_mod.BatchedSimulator(p, ul, ll, counts, layout=layout)
        self.a = [to_dev(pos, device), to_dev(vel, device)]
        self.b = [torch.zeros(self.n * layout, device=device),
                  torch.zeros(self.n * layout, device=device)]
        self.iid = to_dev(iid, device)

    def step(self):
        self.sim.step_batched(self.a[0], self.b[0], self.a[1], self.b[1], self.iid)
        out = self.state()
        self.a, self.b = self.b, self.a
        return out

    def state(self):
        torch.cuda.synchronize()
        L, n = self.L, self.n
        return dict(pos=from_dev(self.a[0]).reshape(n, L), vel=from_dev(self.a[1]).reshape(n, L),
                    npos=from_dev(self.b[0]).reshape(n, L), nvel=from_dev(self.b[1]).reshape(n, L),
                    iid=from_dev(self.iid, dtype=np.uint32))

    def scene(self, st, k):
        o, c = int(self.offs[k]), self.counts[k]
        return {key: v[o:o + c] for key, v in st.items()}


def _small_batch(seed=0x5EED00B1):
    scs = scenes.batch_scenes(nscenes=5, seed=seed, nmin=300, nmax=6000)
    # one scene with particles piled outside its box and an empty scene
    pile = scenes.block((8, 8, 8), (1.0, 1.0, 1.0), seed=7, vel_sigma=2.0)
    pile.pos[:100] += np.float32(1.5)
    empty = scenes.Scene(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32),
                         np.zeros(0, np.uint32), (0.5, 0.5, 0.5), (0.0, 0.0, 0.0))
    return scs[:2] + [pile, empty] + scs[2:]


@pytest.mark.parametrize("layout", [4, 3])
def test_batched_equals_each_scene_alone(pbf_mod, device, layout):
    p = P.dam_params()
    scs = _small_batch()
    print(f"batch of {len(scs)} scenes, counts {[s.n for s in scs]}, layout {layout}", flush=True)
    br = BatchRun(pbf_mod, scs, p, layout)
    assert br.sim.scenes() == [s.n for s in scs]
    singles = [Run(pbf_mod, s, p, layout) if s.n else None for s in scs]
    for frame in range(4):
        gb = br.step()
        for k, r in enumerate(singles):
            if r is None:
                continue
            assert_bitexact(br.scene(gb, k), r.step(), f"frame {frame} scene {k}")


def test_batched_scene_matches_mirror(pbf_mod, orc, device):
    p = P.dam_params()
    scs = _small_batch()
    br = BatchRun(pbf_mod, scs, p, 4)
    gb = br.step()
    for k in (0, 2):
        sc = scs[k]
        pos, vel = sc.as_layout(4)
        assert_bitexact(br.scene(gb, k), mirror(orc, p, sc, pos, vel, sc.iid), f"scene {k}")


def test_batched_boxes_are_fixed_and_n_checked(pbf_mod, device):
    p = P.dam_params()
    br = BatchRun(pbf_mod, _small_batch(), p, 4)
    with pytest.raises(pbf_mod.PBFError) as e:
        br.sim.setLim((2, 2, 2), (0, 0, 0))
    assert e.value.code == 5
    with pytest.raises(pbf_mod.PBFError) as e:
        br.sim.step(br.a[0], br.b[0], br.a[1], br.b[1], br.iid, br.n - 1)
    assert e.value.code == 1
    # a parameter change reaches every scene (h smaller: more cells, within 4x)
    br.sim.loadParams(p.replace(h=0.09, delta_q=0.018))
    br.step()


def test_full_batch_scenes_match_single(pbf_mod, device):
    """The 64-scene batch (4.03M particles): three scenes of the batch equal
    the same scenes stepped alone, after two frames."""
    p = P.dam_params()
    scs = scenes.batch_scenes()
    print(f"full batch: {sum(s.n for s in scs)} particles", flush=True)
    br = BatchRun(pbf_mod, scs, p, 4)
    pick = [0, int(np.argmax([s.n for s in scs])), 63]
    singles = {k: Run(pbf_mod, scs[k], p, 4) for k in pick}
    for frame in range(2):
        gb = br.step()
        for k, r in singles.items():
            assert_bitexact(br.scene(gb, k), r.step(), f"frame {frame} scene {k}")


def test_graph_replays_equal_eager(pbf_mod, device):
    p = P.dam_params()
    sc = scenes.double_dam()
    eager = Run(pbf_mod, sc, p, 4)
    graph = Run(pbf_mod, sc, p, 4)
    graph.sim.set_graph(True)
    stream = torch.cuda.Stream()
    for frame in range(8):
        if frame == 5:  # a parameter change: the recorded graphs are updated
            for r in (eager, graph):
                r.sim.loadParams(p.replace(c_XSPH=0.3))
        with torch.cuda.stream(stream):
            ge = graph.step(stream)
        ee = eager.step()
        assert_bitexact(ge, ee, f"graph frame {frame}")
        # eager work on the same solver between replays
        graph.sim.density_error(graph.a[0], graph.n, stream)


def test_graph_survives_per_frame_params_and_sweeping_box(pbf_mod, device):
    """The reference's frame loop (fluids/FluidSystem.cpp:99-120): loadParams
    every frame, a sweeping setLim, tic-toc buffers.  Graph mode keeps two
    graphs (updated in place when the box moves, a parameter really changes
    or the launch shapes change) and every frame equals eager stepping bit
    for bit, with eager work on the same solver between replays."""
    p = P.dam_params()
    sc = scenes.double_dam()
    sw = pbf_mod.Sweep(sc.ulim, sc.llim, (0.5, 0.0, 0.0), (0.0, 0.0, 0.0), 2 * np.pi / 12, 0)
    eager = Run(pbf_mod, sc, p, 4)
    graph = Run(pbf_mod, sc, p, 4)
    graph.sim.set_graph(True)
    stream = torch.cuda.Stream()
    frames = 14
    for frame in range(frames):
        pf = p if frame < 9 else p.replace(c_XSPH=0.3)
        ul, ll = sw.box(frame)
        for r in (eager, graph):
            r.sim.loadParams(pf)
            r.sim.setLim(ul, ll)
        if frame == 11:
            graph.sim.set_shapes(128, 1024)
        with torch.cuda.stream(stream):
            ge = graph.step(stream)
        ee = eager.step()
        assert_bitexact(ge, ee, f"graph frame {frame}")
        graph.sim.density_error(graph.a[0], graph.n, stream)
    info = graph.sim.graph_info()
    print("graph info", info)
    assert info["live"] == 2 and info["recorded"] == 2, info
    assert info["replayed"] == frames, info
    # every frame after the first two moved the box: the graph of its buffer
    # pair was recorded again and updated (or, if refused, re-instantiated)
    assert info["updated"] + info["instantiated"] - info["recorded"] == frames - 2, info


def test_graph_kept_when_params_reloaded_unchanged(pbf_mod, device):
    """loadParams with the values already set changes nothing: the recorded
    graphs are replayed as they are."""
    p = P.dam_params()
    r = Run(pbf_mod, scenes.double_dam(), p, 4)
    r.sim.set_graph(True)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        for _ in range(6):
            r.sim.loadParams(p)
            r.sim.setLim(*r.sim.getLim())
            r.step(stream)
    info = r.sim.graph_info()
    assert info == {"recorded": 2, "instantiated": 2, "updated": 0, "replayed": 6, "live": 2}, info


def test_graph_steps_record_step_timing(pbf_mod, device):
    sc = scenes.double_dam()
    r = Run(pbf_mod, sc, P.dam_params(), 4)
    r.sim.set_graph(True)
    r.sim.set_timing(True)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        for _ in range(3):
            r.step(stream)
    t = r.sim.get_timing()
    assert t["step"][1] == 3 and t["step"][0] > 0
    assert t["lambda"][1] == 0


@pytest.mark.parametrize("shape", [(64, 512), (128, 1024), (256, 4096), (64, 64)])
def test_launch_shapes_do_not_change_results(pbf_mod, device, shape):
    p = P.dam_params()
    sc = scenes.double_dam()
    base = Run(pbf_mod, sc, p, 4)
    other = Run(pbf_mod, sc, p, 4)
    other.sim.set_shapes(*shape)
    assert other.sim.get_shapes() == shape
    for frame in range(12):  # frames with corner piles exercise the chunk kernels
        assert_bitexact(other.step(), base.step(), f"shape {shape} frame {frame}")


def test_shapes_validation(pbf_mod, device):
    r = Run(pbf_mod, scenes.double_dam(), P.dam_params(), 4)
    for bad in ((32, 2048), (512, 2048), (256, 10), (256, 1 << 20)):
        with pytest.raises(pbf_mod.PBFError) as e:
            r.sim.set_shapes(*bad)
        assert e.value.code == 1
    assert r.sim.get_shapes() == (256, 2048)


def test_autotune_leaves_inputs_and_installs_a_candidate(pbf_mod, device):
    p = P.dam_params()
    sc = scenes.double_dam()
    r = Run(pbf_mod, sc, p, 4)
    for _ in range(6):
        r.step()
    before = r.state()
    (lb, cb), ms = r.sim.autotune(r.a[0], r.a[1], r.iid, r.n, reps=2)
    assert lb in (64, 128, 256) and cb in (512, 1024, 2048, 4096) and ms > 0
    assert r.sim.get_shapes() == (lb, cb)
    after = r.state()
    for k in ("pos", "vel", "iid"):
        np.testing.assert_array_equal(before[k], after[k])
    ref = Run(pbf_mod, sc, p, 4)
    for _ in range(6):
        ref.step()
    assert_bitexact(r.step(), ref.step(), "after autotune")


@pytest.mark.parametrize("batched", [False, True])
def test_heavy_threshold_changes_no_bit(pbf_mod, device, batched):
    """Which targets are summed by one lane and which are spread over the
    chunk kernel is a scheduling choice only: fixed thresholds of 64, 128,
    256 and 2^30 (no target spread unless over the list's capacity) and the
    adaptive default give the same bits, frame after frame, in a scene with
    piles (crowded corners)."""
    p = P.dam_params()
    sc = scenes.block((24, 24, 24), (0.8, 0.8, 0.8), seed=33, vel_sigma=4.0)
    print(f"heavy threshold batched={batched} n={sc.n}", flush=True)
    results = {}
    for thr in (0, 64, 128, 256, 1 << 30):
        if batched:
            run = BatchRun(pbf_mod, [sc, sc], p)
        else:
            run = Run(pbf_mod, sc, p, 4)
        run.sim.set_heavy_threshold(thr)
        assert run.sim.get_heavy_threshold() == thr
        frames = [run.step() for _ in range(4)]
        results[thr] = frames
    base = results[0]
    for thr, frames in results.items():
        for f, (g, b) in enumerate(zip(frames, base)):
            for k in ("npos", "nvel", "iid"):
                np.testing.assert_array_equal(g[k], b[k], err_msg=f"threshold {thr} frame {f} {k}")
    with pytest.raises(pbf_mod.PBFError):
        Run(pbf_mod, sc, p, 4).sim.set_heavy_threshold(10)
