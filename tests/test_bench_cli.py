"""bench.py host logic on the CPU: the one-GPU refusal, the per-apply pair
and FLOP counts of the headline workload, and the trace-window summary."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_headline_pair_and_flop_counts():
    # 30,720 segments + 16,002 blobs: every (target, source) pair of the four
    # blocks, and 92 / 56 / 56 / 40 algorithmic flops per ss / sb / bs / bb pair
    pairs, flops = bench.grand_pairs_flops(30720, 16002)
    assert pairs == 46722 ** 2 == 2182945284
    assert flops == 92 * 30720 ** 2 + 2 * 56 * 30720 * 16002 + 40 * 16002 ** 2
    assert flops == 152121774240


def test_summarize_trace_keeps_the_window(tmp_path, capsys):
    log = tmp_path / "run.log"
    log.write_text("# step x: profiled window opens at 1000 ns (CLOCK_MONOTONIC)\n"
                   "# step x: profiled window closes at 11000 ns (CLOCK_MONOTONIC)\n")
    csv = tmp_path / "trace.csv"
    rows = [("void gmi::(anonymous namespace)::k_a(int)", 500, 1500),    # before: dropped
            ("void gmi::(anonymous namespace)::k_a(int)", 2000, 4000),
            ("gmi::(anonymous namespace)::k_b(double*)", 5000, 6000),
            ("void gmi::(anonymous namespace)::k_a(int)", 7000, 9000),
            ("gmi::(anonymous namespace)::k_b(double*)", 10500, 12000)]  # after: dropped
    csv.write_text("Kernel_Name,Start_Timestamp,End_Timestamp\n" +
                   "".join(f'"{n}",{s},{e}\n' for n, s, e in rows))
    bench.summarize_trace(str(csv), str(log), 2)
    out = capsys.readouterr().out
    # 10 us window over 2 steps; 5 us inside kernels
    assert "window 0.01 ms over 2 steps: 0.005 ms/step wall, 0.003 ms/step in kernels (50% busy)" in out
    assert "| `k_a` | 1.00 | 2.0 | 2.00 | 80.0% |" in out
    assert "| `k_b` | 0.50 | 0.5 | 1.00 | 20.0% |" in out
