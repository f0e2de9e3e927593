#!/usr/bin/env python3
"""Benchmark of the PBF forward step on one MI355X.

Workload: the 1,048,576-particle single dam (128x64x128 jittered lattice,
spacing 0.5h, in a 256h x 128h x 128h box of h-sized cells; h=0.1,
pho0=6378, eps=600, s_corr k=0.1 n=4 dq=0.2h, g=9.8, dt=1/60, 4 solver
iterations, FP32), advanced frame after frame with ping-pong buffers as the
reference's FluidSystem does.  W frames warm up, K frames are timed.

value = N * niter / (time per step): particle-solver-iterations per second,
inputs resident in HBM, bracketed by torch.cuda.synchronize(), timed with
HIP events on the launch stream.  The launch shapes come from the library's
autotuner run on the warmed-up state; the timed steps replay a recorded HIP
graph (the eager figure is reported beside it, from a second simulation
over the same frames).  Also measured: the step rate of later frame
windows, the 64-scene batched rate, the backward/forward ratio of the
differentiable rollout (16,384 particles, 30 frames), the 4,194,304-particle
sweeping-wall workload with its backward over the last 10 frames, and the
step time of one GPU's share of an 8-way x-slab split (slab plus one-cell
halo, stepped as its own scene).  One JSON line on stdout.

`roofline` describes the lambda pass (light + chunk + combine kernels of one
lambda launch), the pass with the largest share of the step: algorithmic
bytes 20 B per particle (read x 12 B, write lambda and rho 8 B), divided by
its mean launch duration measured with HIP events on the launch stream over
the timed frames of the eager run.  `traffic` and the VALU figures come from
the rocprofv3 summaries committed under profiles/ (counter runs of this same
command); `python bench.py --summarize DIR` turns a rocprofv3 output folder
of this command into the per-kernel figures of the timed frames.
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pbf-cuda_amd"))

METRIC = "particle-solver-iters/sec at 1M particles; batched rate; backward/forward ratio"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per particle of one lambda launch: read x (12 B), write
# lambda and rho (8 B).
LAMBDA_BYTES_PER_PARTICLE = 20.0
# Algorithmic bytes per particle of the predict kernel (advect + cell key):
# read x, v (32 B), write the predicted x (16 B), key and index (8 B).
PREDICT_BYTES_PER_PARTICLE = 56.0
# Non-packed FP32 VALU issue peak: 256 CUs x 64 lanes x 2.4 GHz, i.e. one
# wave64 instruction per CU per clock (the build keeps packed FP32 off).
VALU_WAVE_INSTS_PER_S = 256 * 2.4e9
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "lambda_pass_single_dam_1m_traffic.md")
SQ_FILE = os.path.join(ROOT, "profiles", "lambda_pass_single_dam_1m_sq_counters.md")
STEP_FILE = os.path.join(ROOT, "profiles", "step_kernels_single_dam_1m.md")


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-steps", type=int, default=3,
                    help="timed CPU oracle steps per configuration (after one warm-up step)")
    ap.add_argument("--no-graph", action="store_true", help="time eager steps only")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batched rate")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the backward/forward ratio")
    ap.add_argument("--no-slab", action="store_true",
                    help="skip the one-slab share of an 8-way x-slab split")
    ap.add_argument("--no-sweep", action="store_true",
                    help="skip the 4,194,304-particle sweeping-wall workload (120 frames, 6 "
                         "iterations, backward over the last 10 frames)")
    ap.add_argument("--windows", type=int, default=5,
                    help="later windows of `steps` frames timed after the headline window")
    ap.add_argument("--copy-only", action="store_true",
                    help="only run a 256 MiB device-to-device copy 3 times (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against known bytes)")
    ap.add_argument("--summarize", metavar="DIR",
                    help="no GPU work: summarise a rocprofv3 output folder of this command "
                         "(kernel trace or counter collection CSV) over the timed frames")
    return ap.parse_args()


def _profile_fields(path, keys):
    """Numbers `key: value` from a committed rocprofv3 summary, or None."""
    try:
        text = open(path).read()
    except OSError:
        return None
    got = {}
    for key in keys:
        m = re.search(re.escape(key) + r":\s*([0-9.eE+-]+)", text)
        if not m:
            return None
        got[key] = float(m.group(1))
    return got


def measured_traffic():
    """HBM bytes per lambda launch (FETCH_SIZE + WRITE_SIZE, scaled by the
    copy-kernel factors) from the committed rocprofv3 summary, or None."""
    got = _profile_fields(TRAFFIC_FILE, ("traffic_bytes_per_launch",))
    return None if got is None else got["traffic_bytes_per_launch"]


def measured_valu():
    """VALU issue of the lambda pass from the committed rocprofv3 SQ summary,
    against the non-packed peak (one wave64 instruction per CU per clock), or
    None."""
    got = _profile_fields(SQ_FILE, ("valu_wave_insts_per_launch", "traced_launch_us",
                                    "lane_utilisation"))
    if got is None:
        return None
    issue_us = got["valu_wave_insts_per_launch"] / VALU_WAVE_INSTS_PER_S * 1e6
    return {"kernel": "lambda pass (rocprofv3 SQ_INSTS_VALU of k_light + k_chunks + k_combine "
                      "over the timed frames)",
            "wave_insts_per_launch": got["valu_wave_insts_per_launch"],
            "issue_us": round(issue_us, 2), "launch_us": got["traced_launch_us"],
            "frac": issue_us / got["traced_launch_us"],
            "lane_utilisation": got["lane_utilisation"],
            "peak": "256 CUs x 64 FP32 lanes x 2.4 GHz (non-packed)",
            "source": os.path.relpath(SQ_FILE, ROOT)}


def traced_predict_us():
    """Mean duration of the predict kernel in the committed kernel-trace
    summary of this benchmark's single dam (eager run), or None."""
    try:
        text = open(STEP_FILE).read()
    except OSError:
        return None
    m = re.search(r"\| `k_predict_key<4, OneScene>` \| [0-9.]+ \| ([0-9.]+) \|", text)
    return float(m.group(1)) if m else None


def step_bytes(n, niter, cells):
    """Algorithmic HBM bytes of one whole step: one compulsory read or write of
    each logical array per phase, neighbour re-reads counted as cache hits;
    per particle predict 48, cell key 16, sort (key, index) 16, payload
    permute 80, cell ranges 4, per iteration 48 (lambda: x 12 -> lambda, rho
    8; delta-p: x 12, lambda 4 -> x 12), velocity 36, XSPH 40; plus 8 per
    cell for the cell table."""
    return n * (240 + 48 * niter) + 8 * cells


def cpu_baseline(cases, steps):
    """FP64 oracle (the checker, test infrastructure) timed on the host cores:
    one warm-up step then `steps` timed steps of each (name, params, scene,
    pos, vel, iid, sample) case, single-threaded and on all OpenMP threads.
    value = the 1M single dam on all threads."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as orc

    cores = orc.max_threads()
    figures = []
    for name, p, sc, pos0, vel0, iid0, sample in cases:
        for threads in (1, cores):
            pos, vel, iid = pos0, vel0, iid0
            r = orc.step(p, sc.ulim, sc.llim, pos, vel, iid, threads=threads)  # warm-up
            pos, vel, iid = r["npos"], r["nvel"], r["iid"]
            t0 = time.perf_counter()
            for _ in range(steps):
                r = orc.step(p, sc.ulim, sc.llim, pos, vel, iid, threads=threads)
                pos, vel, iid = r["npos"], r["nvel"], r["iid"]
            dt = (time.perf_counter() - t0) / steps
            figures.append({"workload": name, "threads": threads, "steps": steps,
                            "s_per_step": round(dt, 4), "rate": sc.n * p.niter / dt,
                            "sample": sample})
            print(f"cpu: {name} {threads} thread(s) {dt:.3f} s/step", file=sys.stderr,
                  flush=True)
    head = [f for f in figures if f["workload"] == "single_dam_1m" and f["threads"] == cores][0]
    return {"value": head["rate"], "unit": "particle-solver-iters/s", "cores": cores,
            "kind": "port",
            "sample": f"FP64 oracle, {steps} timed steps after 1 warm-up step, "
                      f"{cores} OpenMP threads, {head['sample']}",
            "figures": figures}


class SingleRun:
    """The single dam on the GPU with ping-pong buffers, from its initial state."""

    def __init__(self, pbf, p, sc, dev):
        import numpy as np
        import torch

        self.n = sc.n
        pos, vel = sc.as_layout(4)
        self.bufs = [[torch.from_numpy(pos.ravel()).to(dev), torch.from_numpy(vel.ravel()).to(dev)],
                     [torch.zeros(self.n * 4, device=dev), torch.zeros(self.n * 4, device=dev)]]
        self.iid = torch.from_numpy(sc.iid.astype(np.int32)).to(dev)
        self.sim = pbf.Simulator(p, sc.ulim, sc.llim, capacity=self.n)
        self.f = 0

    def step(self, stream):
        a, b = self.bufs[self.f % 2], self.bufs[(self.f + 1) % 2]
        self.sim.step(a[0], b[0], a[1], b[1], self.iid, self.n, stream)
        self.f += 1

    def current(self):
        return self.bufs[self.f % 2]


class BatchRun:
    """The 64-scene batch on the GPU with ping-pong buffers."""

    def __init__(self, pbf, p, scs, dev):
        import numpy as np
        import torch

        from pbf import scenes

        pos, vel, iid, counts, ul, ll, _ = scenes.concat(scs, 4)
        self.n = int(sum(counts))
        self.bufs = [[torch.from_numpy(pos.ravel()).to(dev), torch.from_numpy(vel.ravel()).to(dev)],
                     [torch.zeros(self.n * 4, device=dev), torch.zeros(self.n * 4, device=dev)]]
        self.iid = torch.from_numpy(iid.astype(np.int32)).to(dev)
        self.sim = pbf.BatchedSimulator(p, ul, ll, counts)
        self.f = 0

    def step(self, stream):
        a, b = self.bufs[self.f % 2], self.bufs[(self.f + 1) % 2]
        self.sim.step_batched(a[0], b[0], a[1], b[1], self.iid, stream)
        self.f += 1

    def current(self):
        return self.bufs[self.f % 2]


def timed(run, steps, stream):
    """ms per step over `steps` steps, bracketed by synchronisation, HIP
    events on the launch stream."""
    import torch

    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(steps):
        run.step(stream)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


GRAD_NOTE = ("tapes with PBF_TAPE_RESCALE: the adjoint is renormalised by a power of two before "
             "each frame, the outputs are the true gradients x 2^-log2_scale; with the dam "
             "constants the exact derivative of the rollout grows by ~1e3-1e4 per frame (the FP64 "
             "oracle's 30-frame gradient is ~1e110, beyond the FP32 range), which the exponent "
             "carries; with the reference constants the 30-frame gradient is checked in "
             "tests/test_gpu_adjoint.py against the oracle's adjoint of the same trajectory (1e-4) "
             "and, over 15 frames, against FP64 central differences (1e-3)")


SWEEP_GRAD_NOTE = ("by frame 110 the moving wall has piled thousands of particles into a box "
                   "corner at distances just above KERNAL_EPS, where the spiky gradient's "
                   "derivative scales as 1/r: the derivative of the frame map grows by ~1e5 per "
                   "frame with the reference constants and ~1e7-1e8 with the dam constants, so "
                   "through all 10 frames it exceeds the FP32 range; the tapes use "
                   "PBF_TAPE_RESCALE (a power-of-two renormalisation of the adjoint before each "
                   "frame), so their outputs are finite: the true gradients x 2^-log2_scale; "
                   "tests/test_gpu_sweep_adjoint.py checks the unscaled gradients through the "
                   "last 5 (reference) and 2 (dam) frames, both tapes, the last frame against the "
                   "oracle's adjoint at full size, and the rescaled pass through all 10 frames")


def adjoint_ratio(pbf, p, dev, stream, reps=3, store=True):
    """Backward / forward time of the 30-frame differentiable rollout of the
    16,384-particle scene (4 iterations, constants `p`): forward = the tape's
    30 frames, backward = dL/d(pos0, vel0, pho0) through the 30 frames.
    store=False: the forward run keeps only per-frame checkpoints and the
    backward pass recomputes each frame's solver iterations (the mode the
    workload specifies); store=True: the forward run keeps every frame's
    iterations, sort and grid on the device and the backward pass reads
    them."""
    import numpy as np
    import torch

    from pbf import scenes

    sc = scenes.differentiable()
    n, nf = sc.n, 30
    tgt = torch.from_numpy(scenes.target_cloud(sc)).to(dev)
    sim = pbf.Simulator(p, sc.ulim, sc.llim, capacity=n)
    tape = pbf.Tape(sim, n, nf, store=store, rescale=True)
    pos, vel = sc.as_layout(4)
    pos0 = torch.from_numpy(pos.ravel()).to(dev)
    vel0 = torch.from_numpy(vel.ravel()).to(dev)
    iid0 = torch.from_numpy(sc.iid.astype(np.int32)).to(dev)
    xf = torch.empty(n * 4, device=dev)
    iidf = torch.empty(n, dtype=torch.int32, device=dev)
    gpos = torch.zeros(n, 4, device=dev)
    gp0 = torch.empty(n * 4, device=dev)
    gv0 = torch.empty(n * 4, device=dev)
    loss = [0.0]

    def forward():
        tape.forward(pos0, vel0, iid0, stream)

    def backward():
        tape.read(nf, pos=xf, iid=iidf, stream=stream)
        x = xf.view(n, 4)[:, :3].double()
        dx = x - tgt[iidf.long()]
        loss[0] = dx
        gpos[:, :3] = (2.0 * dx / n).float()
        return tape.backward(gpos.view(-1), None, gp0, gv0, stream)

    def plain():
        a = [pos0.clone(), vel0.clone()]
        b = [torch.empty_like(pos0), torch.empty_like(vel0)]
        ids = iid0.clone()
        for _ in range(nf):
            sim.step(a[0], b[0], a[1], b[1], ids, n, stream)
            a, b = b, a

    def time_it(fn):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            e0.record(stream)
            fn()
            e1.record(stream)
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        return sorted(ts)[reps // 2]

    with torch.cuda.stream(stream):
        forward()
        grho = backward()
        t_f = time_it(forward)
        t_b = time_it(backward)
        t_p = time_it(plain)
    gv = gv0.view(n, 4)[:, :3]
    finite = bool(torch.isfinite(gv).all().item()) and bool(np.isfinite(grho))
    return {"value": t_b / t_f, "unit": "backward time / forward time",
            "tape": "stored iterations" if store else "checkpoints, iterations recomputed",
            "forward_ms": t_f, "backward_ms": t_b, "plain_forward_ms": t_p,
            "backward_over_plain_forward": t_b / t_p,
            "loss": float((loss[0] ** 2).sum().item() / n),
            "grad_vel0_norm_scaled": float(gv.double().norm().item()),
            "log2_scale": tape.gradient_log2_scale(), "grad_pho0": float(grho),
            "gradients_finite": finite}


def c4_workload(pbf, dev, stream):
    """The differentiable workload under both constant sets: the reference
    defaults (finite, checked gradients; the headline) and the dam constants
    (timing only, see GRAD_NOTE); each with the recomputing and the storing
    tape."""
    out = {}
    for name, p in (("reference_constants", pbf.reference_defaults()),
                    ("dam_constants", pbf.dam_params(niter=4))):
        rec = adjoint_ratio(pbf, p, dev, stream, store=False)
        sto = adjoint_ratio(pbf, p, dev, stream, store=True)
        out[name] = {"recompute": rec, "stored": sto}
        print(f"adjoint {name}: forward {rec['forward_ms']:.2f} ms, backward "
              f"{rec['backward_ms']:.2f} ms, ratio {rec['value']:.3f} (stored tape "
              f"{sto['value']:.3f}), finite {rec['gradients_finite']}", file=sys.stderr,
              flush=True)
    head = out["reference_constants"]["recompute"]
    return {"value": head["value"], "unit": "backward time / forward time",
            "tape": head["tape"], "constants": "reference defaults (FluidSystem.cpp:15-25)",
            "adjoint_mode": "scatter: one-sided pair terms added with float atomics (the "
                            "library's choice below 2^18 particles)",
            "forward_ms": head["forward_ms"], "backward_ms": head["backward_ms"],
            "gradients_finite": head["gradients_finite"],
            "workload": "differentiable_16k: 16,384 particles (32x16x32 jittered lattice, "
                        "v0 ~ N(0, 0.5^2)), box 40h x 20h x 40h, 30 frames, 4 iterations; loss = "
                        "mean squared distance to a seeded target cloud; gradients w.r.t. pos0, "
                        "vel0, pho0",
            "gradients_note": GRAD_NOTE, "by_constants": out}


def slab_share(pbf, p, pos, vel, iid, ulim, llim, shapes, dev, stream, parts=8, warmup=3,
               steps=20):
    """One GPU's share of an x-slab decomposition, stepped at its real shape:
    the particles are split into `parts` slabs of whole h-cells in x with
    balanced particle counts, and the slab with the most particles plus a
    one-cell halo on each side is stepped as a scene of its own (box = the
    slab and halo cells).  The halo exchange, migration and the slab walls'
    physics are not modelled; this times the kernels on one share."""
    import numpy as np
    import torch

    h = p.h
    n = pos.shape[0]
    dims_x = int(np.ceil((ulim[0] - llim[0]) / h))
    cx = np.clip(((pos[:, 0] - llim[0]) / h).astype(np.int64), 0, dims_x - 1)
    per_cell = np.bincount(cx, minlength=dims_x)
    cum = np.cumsum(per_cell)
    cuts = [0] + [int(np.searchsorted(cum, k * n / parts)) + 1 for k in range(1, parts)] + [dims_x]
    best = None
    for a, b in zip(cuts[:-1], cuts[1:]):
        own = int(per_cell[a:b].sum())
        if best is None or own > best[2]:
            best = (a, b, own)
    a, b, own = best
    lo, hi = max(a - 1, 0), min(b + 1, dims_x)
    sel = (cx >= lo) & (cx < hi)
    m = int(sel.sum())
    sl_ll = (llim[0] + lo * h, llim[1], llim[2])
    sl_ul = (llim[0] + hi * h, ulim[1], ulim[2])
    P = np.ascontiguousarray(pos[sel], dtype=np.float32)
    V = np.ascontiguousarray(vel[sel], dtype=np.float32)
    I = np.arange(m, dtype=np.int32)
    sim = pbf.Simulator(p, sl_ul, sl_ll, capacity=m)
    sim.set_shapes(*shapes)
    sim.set_graph(True)
    bufs = [[torch.from_numpy(P.ravel()).to(dev), torch.from_numpy(V.ravel()).to(dev)],
            [torch.empty(m * 4, device=dev), torch.empty(m * 4, device=dev)]]
    ids = torch.from_numpy(I).to(dev)

    class _Run:
        f = 0

        def step(self, st):
            x, y = bufs[self.f % 2], bufs[(self.f + 1) % 2]
            sim.step(x[0], y[0], x[1], y[1], ids, m, st)
            self.f += 1

    run = _Run()
    with torch.cuda.stream(stream):
        for _ in range(warmup):
            run.step(stream)
        ms = timed(run, steps, stream)
    return {"parts": parts, "x_cells": [a, b], "slab_particles": own, "halo_particles": m - own,
            "ms_per_step": ms, "rate": own * p.niter / (ms * 1e-3), "niter": p.niter,
            "mode": "hip graph replay", "unit": "particle-solver-iters/s (slab particles)"}


def sweep_workload(pbf, dev, stream, frames=120, back=10):
    """The sweeping-wall workload: 4,194,304 particles, the +x wall moving as
    160h + 20h sin(2 pi f / 120), dam constants with 6 iterations.  Frames
    0 .. frames-back-1 run plainly (timed), the last `back` frames run on a
    tape (timed), then dL/d(state at frame frames-back) is computed through
    them (timed), L = mean squared distance to a seeded cloud; once with a
    tape that stores the iterations and once with one that recomputes them."""
    import numpy as np
    import torch

    from pbf import scenes

    p = pbf.dam_params(niter=6)
    sc = scenes.sweep_dam()
    sw = pbf.Sweep(*scenes.sweep_of_sweep_dam())
    n = sc.n
    tgt = torch.from_numpy(scenes.target_cloud(sc)).to(dev)
    sim = pbf.Simulator(p, sc.ulim, sc.llim, capacity=n)
    pos, vel = sc.as_layout(4)
    a = [torch.from_numpy(pos.ravel()).to(dev), torch.from_numpy(vel.ravel()).to(dev)]
    b = [torch.empty_like(a[0]), torch.empty_like(a[1])]
    iid = torch.from_numpy(sc.iid.astype(np.int32)).to(dev)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    plain = frames - back

    def span(fn):
        torch.cuda.synchronize()
        e0.record(stream)
        r = fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1), r

    out = {"particles": n, "frames": frames, "niter": p.niter}
    with torch.cuda.stream(stream):
        ms, which = span(lambda: sim.run_frames(a[0], a[1], b[0], b[1], iid, n, 0, plain, sw,
                                                stream))
        out["plain_ms_per_frame"] = ms / plain
        out["plain_rate"] = n * p.niter / (ms / plain * 1e-3)
        cur = b if which else a
        for store in (True, False):
            free0 = torch.cuda.mem_get_info()[0]
            tape = pbf.Tape(sim, n, back, store=store, rescale=True)
            tape_bytes = free0 - torch.cuda.mem_get_info()[0]
            ms_f, _ = span(lambda: tape.forward(cur[0], cur[1], iid, stream, first_frame=plain,
                                                sweep=sw))
            xf = torch.empty(n * 4, device=dev)
            iidf = torch.empty(n, dtype=torch.int32, device=dev)
            tape.read(back, pos=xf, iid=iidf, stream=stream)
            g = torch.zeros(n, 4, device=dev)
            g[:, :3] = (2.0 * (xf.view(n, 4)[:, :3].double() - tgt[iidf.long()]) / n).float()
            gp0 = torch.empty(n * 4, device=dev)
            gv0 = torch.empty(n * 4, device=dev)
            ms_b, grho = span(lambda: tape.backward(g.view(-1), None, gp0, gv0, stream))
            gv = gv0.view(n, 4)[:, :3].double()
            e = tape.gradient_log2_scale()
            key = "stored" if store else "recompute"
            gnorm = float(gv.norm().item())
            out[key] = {"tape_forward_ms": ms_f, "backward_ms": ms_b,
                        "backward_forward_ratio": ms_b / ms_f, "tape_device_bytes": int(tape_bytes),
                        "gradients_finite": bool(torch.isfinite(gv).all().item())
                        and bool(np.isfinite(grho)),
                        "grad_vel_norm_scaled": gnorm, "log2_scale": e,
                        "grad_vel_norm_log10": (float(np.log10(gnorm)) + e * float(np.log10(2.0))
                                                if gnorm > 0 else None),
                        "grad_pho0": float(grho)}
            gv_true = gv * 2.0 ** e
            if store:
                gv_stored = gv_true.clone()
            else:
                out["recompute"]["rel_diff_to_stored_tape"] = float(
                    ((gv_true - gv_stored).norm() / gv_stored.norm()).item())
            tape.close()
            del tape, xf, iidf, g, gp0, gv0
    out["backward_forward_ratio"] = out["recompute"]["backward_forward_ratio"]
    out["tape"] = "checkpoints, iterations recomputed"
    out["adjoint_mode"] = ("gather: each backward pass over a grid of its own positions, no "
                           "atomics (the library's choice from 2^18 particles)")
    out["gradients_note"] = SWEEP_GRAD_NOTE
    out["workload"] = ("sweep_dam_4m: 256x64x256 jittered lattice in a 160h x 64h x 160h box, "
                       "+x wall 160h + 20h sin(2 pi f/120), dam constants, 6 iterations; "
                       f"frames 0..{plain - 1} plain, {plain}..{frames - 1} on a tape, "
                       f"backward through the last {back}")
    return out


def _short_kernel(name):
    n = name.replace("pbf::(anonymous namespace)::", "").replace("pbf::", "")
    if n.startswith("void "):
        n = n[5:]
    return n.split("(")[0]


def summarize(folder, warmup, steps, windows, n=1048576):
    """Per-kernel figures of the timed frames from a rocprofv3 output folder
    of this command (no GPU work).  Finds the single-dam steps by their
    predict kernel (k_predict_key<4, OneScene> over n particles): the eager
    run's `warmup` steps, the autotuner's steps, its `steps` timed steps, then
    the graph run's `warmup` + steps * (1 + windows) steps.  Prints, for the
    eager and the graph run's headline window, each kernel's launches per
    step and mean duration (kernel trace) or mean counter values per launch
    (counter collection), and the lambda pass's totals per launch."""
    import collections
    import csv
    import glob

    trace = glob.glob(os.path.join(folder, "**", "*kernel_trace.csv"), recursive=True)
    ctrs = glob.glob(os.path.join(folder, "**", "*counter_collection.csv"), recursive=True)
    if not trace and not ctrs:
        sys.exit(f"--summarize: no kernel_trace.csv or counter_collection.csv under {folder}")
    disp = {}  # dispatch id -> [name, grid, start, end, {counter: value}]
    for path in trace + ctrs:
        for r in csv.DictReader(open(path)):
            d = int(r["Dispatch_Id"])
            if d not in disp:
                grid = int(r.get("Grid_Size_X") or r.get("Grid_Size") or 0)
                disp[d] = [r["Kernel_Name"], grid, int(r["Start_Timestamp"]),
                           int(r["End_Timestamp"]), {}]
            if "Counter_Name" in r:
                c = disp[d][4]
                c[r["Counter_Name"]] = c.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    order = sorted(disp)
    pred = [k for k in order if "k_predict_key<4, pbf::OneScene>" in disp[k][0]
            and disp[k][1] >= n and disp[k][1] < n + 256]
    tuned = len(pred) - 2 * warmup - steps * (2 + windows)
    if tuned < 0:
        sys.exit(f"--summarize: {len(pred)} single-dam steps, fewer than the command runs")
    spans = {"eager": (warmup + tuned, warmup + tuned + steps),
             "graph": (2 * warmup + tuned + steps, 2 * warmup + tuned + 2 * steps)}
    lines = [f"# rocprofv3 summary of `bench.py --steps {steps} --warmup {warmup}` "
             f"({len(pred)} single-dam steps, {tuned} of them by the autotuner)",
             f"# frames {warmup}..{warmup + steps - 1} of the 1,048,576-particle single dam"]
    for run, (a, b) in spans.items():
        lo = pred[a]
        # the window ends where the next step starts (the last step of the
        # command: as many dispatches as the step before it)
        hi = pred[b] if b < len(pred) else pred[b - 1] + (pred[b - 1] - pred[b - 2])
        sel = [k for k in order if lo <= k < hi]
        per = collections.OrderedDict()
        for k in sel:
            nm = _short_kernel(disp[k][0])
            per.setdefault(nm, []).append(disp[k])
        lines.append(f"\n## {run} run, {steps} steps (dispatches {lo}..{hi - 1})")
        tot_ns = sum(d[3] - d[2] for k in sel for d in [disp[k]])
        lines.append(f"kernel time per step: {tot_ns / steps / 1e3:.1f} us")
        names = sorted(per, key=lambda x: -sum(d[3] - d[2] for d in per[x]))
        for nm in names:
            ds = per[nm]
            mean_us = sum(d[3] - d[2] for d in ds) / len(ds) / 1e3
            row = f"- {nm}: {len(ds) / steps:g} per step, mean {mean_us:.2f} us"
            cnames = sorted({c for d in ds for c in d[4]})
            if cnames:
                row += "; per launch " + ", ".join(
                    f"{c} {sum(d[4].get(c, 0.0) for d in ds) / len(ds):.6g}" for c in cnames)
            lines.append(row)
        lam = [nm for nm in per if "<LambdaPass, OneScene>" in nm]
        if lam:
            launches = len(per[[x for x in lam if x.startswith("k_light")][0]])
            us = sum(d[3] - d[2] for x in lam for d in per[x]) / launches / 1e3
            lines.append(f"lambda pass (k_light + k_chunks + k_combine) per launch: "
                         f"{us:.2f} us over {launches} launches")
            cn = sorted({c for x in lam for d in per[x] for c in d[4]})
            for c in cn:
                v = sum(d[4].get(c, 0.0) for x in lam for d in per[x]) / launches
                lines.append(f"lambda pass {c} per launch: {v:.6g}")
    print("\n".join(lines))


def main():
    args = parse_args()
    if args.summarize:
        summarize(args.summarize, args.warmup, args.steps, args.windows)
        return
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus}: the benchmark measures one GPU only "
              "(the step has no multi-GPU path; the single scene would shard as x-slabs, whose "
              "one-GPU share the slab_share figure times)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import pbf
    from pbf import scenes

    dev = torch.device("cuda:0")
    if args.copy_only:
        a = torch.ones(64 << 20, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel() * 4, "copies": 3}), flush=True)
        return
    p = pbf.dam_params(niter=4)
    sc = scenes.single_dam()
    n = sc.n
    stream = torch.cuda.Stream()  # a capturable (non-null) stream for graph mode

    # eager run: warm up, tune the launch shapes on the warmed-up state, then
    # time with per-phase events
    eager = SingleRun(pbf, p, sc, dev)
    with torch.cuda.stream(stream):
        for _ in range(args.warmup):
            eager.step(stream)
        shapes = eager.sim.get_shapes()
        if not args.no_tune:
            cur = eager.current()
            shapes, _ = eager.sim.autotune(cur[0], cur[1], eager.iid, n, reps=5, stream=stream)
        eager.sim.set_timing(True)
        eager.sim.get_timing()  # reset
        ms_eager = timed(eager, args.steps, stream)
        phases = eager.sim.get_timing()
        eager.sim.set_timing(False)
    print(f"eager: {ms_eager:.4f} ms/step, shapes {shapes}", file=sys.stderr, flush=True)

    # graph run: same frames from the same initial state, one graph launch per step
    ms_graph = None
    windows = None
    if not args.no_graph:
        graph = SingleRun(pbf, p, sc, dev)
        graph.sim.set_shapes(*shapes)
        graph.sim.set_graph(True)
        with torch.cuda.stream(stream):
            for _ in range(args.warmup):
                graph.step(stream)
            ms_graph = timed(graph, args.steps, stream)
            # later windows of the same run: the rate depends on how far the
            # dam has spread (dense early frames are the slowest)
            windows = [{"frames": [args.warmup, args.warmup + args.steps - 1],
                        "ms_per_step": ms_graph,
                        "rate": n * p.niter / (ms_graph * 1e-3)}]
            for k in range(args.windows):
                f0 = args.warmup + (k + 1) * args.steps
                wms = timed(graph, args.steps, stream)
                windows.append({"frames": [f0, f0 + args.steps - 1], "ms_per_step": wms,
                                "rate": n * p.niter / (wms * 1e-3)})
        print(f"graph: {ms_graph:.4f} ms/step; later windows "
              + ", ".join(f"{w['frames'][0]}-{w['frames'][1]}: {w['ms_per_step']:.3f}"
                          for w in windows[1:]), file=sys.stderr, flush=True)
        del graph
    ms = ms_graph if ms_graph is not None else ms_eager
    value = n * p.niter / (ms * 1e-3)

    phase_ms = {k: round(t / c, 4) for k, (t, c) in phases.items() if c}
    l_ms = phases["lambda"][0] / phases["lambda"][1]
    achieved = LAMBDA_BYTES_PER_PARTICLE * n / (l_ms * 1e-3) / 1e9
    traffic = measured_traffic()
    roofline = {"bound": "hbm", "kernel": "lambda pass (k_light + k_chunks + k_combine)",
                "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT) if traffic else None,
                "launch_us": round(l_ms * 1e3, 2),
                "launch_timing": "HIP events on the launch stream around each lambda launch, "
                                 f"eager run, frames {args.warmup}..{args.warmup + args.steps - 1}",
                "algorithmic_bytes": LAMBDA_BYTES_PER_PARTICLE * n,
                "algorithmic_bytes_per_particle": LAMBDA_BYTES_PER_PARTICLE}

    p_ms = phases["predict"][0] / phases["predict"][1]
    p_gbs = PREDICT_BYTES_PER_PARTICLE * n / (p_ms * 1e-3) / 1e9
    predict_roofline = {"bound": "hbm", "kernel": "k_predict_key (advect + cell key)",
                        "achieved": round(p_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": p_gbs / HBM_PEAK_GBS, "launch_us": round(p_ms * 1e3, 2),
                        "algorithmic_bytes_per_particle": PREDICT_BYTES_PER_PARTICLE,
                        "launch_timing": "HIP events around each predict launch, eager run "
                                         "(the event pair adds several us to a ~10 us kernel)"}
    tp = traced_predict_us()
    if tp:
        t_gbs = PREDICT_BYTES_PER_PARTICLE * n / (tp * 1e-6) / 1e9
        predict_roofline.update({"traced_us": tp, "traced_achieved": round(t_gbs, 1),
                                 "traced_frac": t_gbs / HBM_PEAK_GBS,
                                 "traced_source": os.path.relpath(STEP_FILE, ROOT)})

    # batched: 64 scenes of uneven size in one launch sequence, graph mode
    batched = None
    if not args.no_batch:
        scs = scenes.batch_scenes()
        br = BatchRun(pbf, p, scs, dev)
        with torch.cuda.stream(stream):
            for _ in range(args.warmup):
                br.step(stream)
            bshapes = br.sim.get_shapes()
            if not args.no_tune:
                cur = br.current()
                bshapes, _ = br.sim.autotune(cur[0], cur[1], br.iid, br.n, reps=5, stream=stream)
            if not args.no_graph:
                br.sim.set_graph(True)
            ms_b = timed(br, args.steps, stream)
        rate = br.n * p.niter / (ms_b * 1e-3)
        del br
        # the same 64 scenes as 64 single-scene solvers stepped one after the
        # other on the stream (same frames, same shapes, graph mode): what
        # batching saves
        runs = [SingleRun(pbf, p, sc_k, dev) for sc_k in scs]
        for r in runs:
            r.sim.set_shapes(*bshapes)
            r.sim.set_graph(not args.no_graph)

        class _All:
            def step(self, st):
                for r in runs:
                    r.step(st)

        with torch.cuda.stream(stream):
            for _ in range(args.warmup):
                _All().step(stream)
            ms_seq = timed(_All(), args.steps, stream)
        del runs
        batched = {"value": rate, "unit": "particle-solver-iters/s", "ms_per_step": ms_b,
                   "scenes": len(scs), "particles": int(sum(x.n for x in scs)),
                   "frac_of_single": rate / value,
                   "one_solver_per_scene_ms_per_step": ms_seq,
                   "speedup_over_one_solver_per_scene": ms_seq / ms_b,
                   "shapes": list(bshapes), "graph": not args.no_graph,
                   "workload": "64 box scenes, counts log-uniform in [4096, 262144] (seeded), "
                               "fill 0.2-0.5, one jittered lattice block each; same constants"}
        print(f"batched: {ms_b:.4f} ms/step, {rate:.4g} /s; one solver per scene "
              f"{ms_seq:.3f} ms/step", file=sys.stderr, flush=True)

    # one GPU's share of an 8-way x-slab split of the single dam (its state
    # after the timed frames) and of the 4M sweep dam (its initial state)
    slab = None
    if not args.no_slab:
        cur = eager.current()
        P4 = cur[0].cpu().numpy().reshape(n, 4)
        V4 = cur[1].cpu().numpy().reshape(n, 4)
        slab = {"single_dam_1m": slab_share(pbf, p, P4, V4, None, sc.ulim, sc.llim, shapes, dev,
                                            stream)}
        sd = scenes.sweep_dam()
        p6 = pbf.dam_params(niter=6)
        sp, sv = sd.as_layout(4)
        slab["sweep_dam_4m"] = slab_share(pbf, p6, sp, sv, None, sd.ulim, sd.llim, shapes, dev,
                                          stream)
        print(f"slab: 1M share {slab['single_dam_1m']['ms_per_step']:.3f} ms/step, "
              f"4M share {slab['sweep_dam_4m']['ms_per_step']:.3f} ms/step", file=sys.stderr,
              flush=True)

    adjoint = None
    if not args.no_adjoint:
        adjoint = c4_workload(pbf, dev, stream)

    sweep = None
    if not args.no_sweep:
        sweep = sweep_workload(pbf, dev, stream)
        st_, rc_ = sweep["stored"], sweep["recompute"]
        print(f"sweep: {sweep['plain_ms_per_frame']:.2f} ms/frame; stored tape backward "
              f"{st_['backward_ms']:.1f} ms / forward {st_['tape_forward_ms']:.1f} ms; recomputing "
              f"tape {rc_['backward_ms']:.1f} / {rc_['tape_forward_ms']:.1f} ms", file=sys.stderr,
              flush=True)

    cpu = None
    if not args.no_cpu:
        cur = eager.current()
        P4 = cur[0].cpu().numpy().reshape(n, 4)
        V4 = cur[1].cpu().numpy().reshape(n, 4)
        I = eager.iid.cpu().numpy().astype(np.uint32)
        c1 = scenes.double_dam()
        p1, v1 = c1.as_layout(4)
        done = args.warmup + args.steps
        cases = [("single_dam_1m", p, sc, P4, V4, I,
                  f"1,048,576-particle single dam from the GPU's state after frame {done - 1}"),
                 ("double_dam_32k", p, c1, p1, v1, c1.iid,
                  "32,768-particle double dam from its initial state")]
        cpu = cpu_baseline(cases, args.cpu_steps)

    sb = step_bytes(n, p.niter, 256 * 128 * 128)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-solver-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"single_dam_1m: {n} particles, 128x64x128 jittered lattice (seeded "
                        "SplitMix64), 256x128x128 cells, h=0.1, pho0=6378, eps=600, k_corr=0.1, "
                        "n_corr=4, dq=0.2h, g=9.8, dt=1/60, c_XSPH=0.5, 4 iterations; frames "
                        f"{args.warmup}..{args.warmup + args.steps - 1} of the running simulation",
            "particles": n, "niter": p.niter, "parallelism": "single GPU",
            "mode": "hip graph replay" if ms_graph is not None else "eager",
            "shapes": list(shapes)},
        "eager_ms_per_step": ms_eager,
        "graph_ms_per_step": ms_graph,
        "roofline": roofline,
        "step_roofline": {"bound": "hbm", "algorithmic_bytes": sb,
                          "achieved": round(sb / (ms * 1e-3) / 1e9, 3), "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": sb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "valu": measured_valu(),
        "predict_roofline": predict_roofline,
        "cpu_baseline": cpu,
        "phase_ms": phase_ms,
        "batched_rate": batched,
        "backward_forward_ratio": adjoint,
        "sweep_workload": sweep,
        "frame_windows": windows,
        "slab_share": slab,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
