"""CPU checks of bench.py's command-line contract (no GPU work)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_help_lists_the_workloads_and_legs():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    for flag in ("--steps", "--warmup", "--workload", "--fp16", "--no-learned", "--tau"):
        assert flag in r.stdout
    assert "--batch-shape" in r.stdout


def test_profiled_traffic_is_read_from_the_committed_summary():
    """The roofline's traffic figure comes from the rocprof summary of the
    same workload: it parses, and it is at least the bytes of sweep state
    the algorithm itself moves per launch ((16 V + 20) B per pixel)."""
    sys.path.insert(0, ROOT)
    import bench
    traffic, src = bench.profiled_traffic("c2_realistic")
    assert src == os.path.join("profiles", "sweep_kernel_c2_realistic.md")
    state = (16 * 10 + 20) * 2048 * 1536
    assert state < traffic < 20 * state
    assert bench.profiled_traffic("no_such_workload") == (None, None)
