#!/usr/bin/env python3
"""Benchmark of the PBF forward step on one MI355X.

Workload: the 1,048,576-particle single dam (128x64x128 jittered lattice,
spacing 0.5h, in a 256h x 128h x 128h box of h-sized cells; h=0.1,
pho0=6378, eps=600, s_corr k=0.1 n=4 dq=0.2h, g=9.8, dt=1/60, 4 solver
iterations, FP32), advanced frame after frame with ping-pong buffers as the
reference's FluidSystem does.  W frames warm up, K frames are timed.

value = N * niter / (time per step): particle-solver-iterations per second,
inputs resident in HBM, torch's current stream, bracketed by
torch.cuda.synchronize().  One JSON line on stdout.
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
# Algorithmic bytes per particle of one delta-p launch: read x (12 B) and
# lambda (4 B), write x (12 B).
DELTA_BYTES_PER_PARTICLE = 28.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_delta_single_dam_1m_traffic.md")


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-steps", type=int, default=2)
    return ap.parse_args()


def measured_traffic():
    """HBM bytes per k_delta launch from the committed rocprofv3 summary, or None."""
    try:
        text = open(TRAFFIC_FILE).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", text)
    return float(m.group(1)) if m else None


def cpu_baseline(p, sc, pos_np, vel_np, iid_np, steps):
    """FP64 oracle (the checker, test infrastructure) timed on the host cores
    on the same frame the GPU last stepped."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as orc

    cores = orc.max_threads()
    orc.step(p, sc.ulim, sc.llim, pos_np[:4096], vel_np[:4096], iid_np[:4096])  # load, warm
    t0 = time.perf_counter()
    done = 0
    pos, vel, iid = pos_np, vel_np, iid_np
    for _ in range(steps):
        r = orc.step(p, sc.ulim, sc.llim, pos, vel, iid, threads=cores)
        pos, vel, iid = r["npos"], r["nvel"], r["iid"]
        done += 1
        if time.perf_counter() - t0 > 30.0:
            break
    dt = (time.perf_counter() - t0) / done
    return {"value": sc.n * p.niter / dt, "unit": "particle-solver-iters/s", "cores": cores,
            "kind": "port",
            "sample": f"{done} FP64 oracle step(s) of the {sc.n}-particle single dam from the "
                      f"GPU's state after the timed frames, {cores} OpenMP threads; "
                      f"{dt:.2f} s per step"}


def main():
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus}: this build measures one GPU only "
              "(the step has no multi-GPU path; see DESIGN.md on sharding)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import pbf
    from pbf import scenes

    dev = torch.device("cuda:0")
    p = pbf.dam_params(niter=4)
    sc = scenes.single_dam()
    n = sc.n
    pos, vel = sc.as_layout(4)
    bufs = [[torch.from_numpy(pos.ravel()).to(dev), torch.from_numpy(vel.ravel()).to(dev)],
            [torch.zeros(n * 4, device=dev), torch.zeros(n * 4, device=dev)]]
    iid = torch.from_numpy(sc.iid.astype(np.int32)).to(dev)
    sim = pbf.Simulator(p, sc.ulim, sc.llim, capacity=n)
    stream = torch.cuda.current_stream()

    def step(f):
        a, b = bufs[f % 2], bufs[(f + 1) % 2]
        sim.step(a[0], b[0], a[1], b[1], iid, n, stream)

    for f in range(args.warmup):
        step(f)
    torch.cuda.synchronize()
    sim.set_timing(True)
    sim.get_timing()  # reset
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for f in range(args.warmup, args.warmup + args.steps):
        step(f)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    phases = sim.get_timing()
    sim.set_timing(False)
    value = n * p.niter / (ms * 1e-3)

    phase_ms = {k: round(t / c, 4) for k, (t, c) in phases.items() if c}
    d_ms = phases["delta"][0] / phases["delta"][1]
    achieved = DELTA_BYTES_PER_PARTICLE * n / (d_ms * 1e-3) / 1e9
    traffic = measured_traffic()
    roofline = {"bound": "hbm", "kernel": "k_delta", "achieved": round(achieved, 3),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                "traffic": traffic, "launch_us": round(d_ms * 1e3, 2),
                "algorithmic_bytes": DELTA_BYTES_PER_PARTICLE * n}

    cpu = None
    if not args.no_cpu:
        last = (args.warmup + args.steps) % 2
        P4 = bufs[last][0].cpu().numpy().reshape(n, 4)
        V4 = bufs[last][1].cpu().numpy().reshape(n, 4)
        I = iid.cpu().numpy().astype(np.uint32)
        cpu = cpu_baseline(p, sc, P4, V4, I, args.cpu_steps)

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
            "particles": n, "niter": p.niter, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "phase_ms": phase_ms,
        "batched_rate": None,
        "backward_forward_ratio": None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
