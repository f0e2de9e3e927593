"""CPU checks of bench.py's contract that need no GPU: the figures it reads
from the committed rocprofv3 summaries parse and are consistent with each
other, the kernel-trace summarizer reads a trace of the shape rocprofv3
writes, and more than one GPU is refused with a message."""
import csv
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_committed_profiles_parse():
    traffic = bench.measured_traffic()
    assert traffic is not None
    algorithmic = bench.LAMBDA_BYTES_PER_PARTICLE * 1048576
    assert traffic > algorithmic  # counted HBM bytes include the design's extra traffic
    valu = bench.measured_valu()
    assert valu is not None
    assert 0.0 < valu["frac"] < 1.0 and 0.0 < valu["lane_utilisation"] <= 1.0
    # issue time = instructions / (256 CUs x 2.4 GHz), one wave64 instruction per CU per clock
    assert valu["issue_us"] == pytest.approx(valu["wave_insts_per_launch"] / (256 * 2.4e9) * 1e6,
                                             rel=1e-3)
    tp = bench.traced_predict_us()
    assert tp is not None and 1.0 < tp < 1000.0


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def _trace(path, steps_per_run):
    """A synthetic kernel trace: per step a predict kernel over 1M particles,
    four 100-ns lambda light kernels and one 50-ns permute, back to back."""
    rows, t, d = [], 0, 0
    name_p = "void pbf::(anonymous namespace)::k_predict_key<4, pbf::OneScene>(float const*)"
    name_l = ("void pbf::(anonymous namespace)::k_light<pbf::(anonymous namespace)::LambdaPass, "
              "pbf::OneScene>(int)")
    name_c = "void pbf::(anonymous namespace)::k_permute<4>(unsigned int const*)"
    for _ in range(steps_per_run):
        for name, grid, dur in [(name_p, 1048576, 10)] + [(name_l, 1048576, 100)] * 4 + [
                (name_c, 1048576, 50)]:
            d += 1
            rows.append({"Dispatch_Id": d, "Kernel_Name": name, "Grid_Size_X": grid,
                         "Start_Timestamp": t, "End_Timestamp": t + dur})
            t += dur + 5
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=list(rows[0]))
        w.writeheader()
        w.writerows(rows)


def test_summarize_reads_a_kernel_trace(tmp_path, capsys):
    warmup, steps, windows, tuned = 2, 3, 1, 4
    total = 2 * warmup + tuned + steps * (2 + windows)
    _trace(str(tmp_path / "kt_kernel_trace.csv"), total)
    bench.summarize(str(tmp_path), warmup, steps, windows)
    out = capsys.readouterr().out
    assert f"{total} single-dam steps, {tuned} of them by the autotuner" in out
    assert "k_light<LambdaPass, OneScene>: 4 per step, mean 0.10 us" in out
    assert "lambda pass (k_light + k_chunks + k_combine) per launch: 0.10 us over" in out
