"""The measurement script's contract pieces that run without a GPU: a request
for more than one GPU is refused with a message, and the CPU baseline (the
oracle timed on a band of rows) reports the fields the bench line carries."""
import os
import subprocess
import sys
import types

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_cpu_baseline_fields_and_repeats():
    sys.path.insert(0, ROOT)
    import bench
    from prender_amd import generators as G
    d = G.config1(width=32, height=24, samples=1)
    args = types.SimpleNamespace(cpu_rows=4, seed=3, cpu_min_s=0.05)
    r = bench.cpu_baseline(d, args, 1e-3, None)
    assert r["kind"] == "port" and r["unit"] == "samples/s" and r["cores"] >= 1
    assert r["value"] > 0 and r["mrays_per_s"] > 0 and r["gpu_over_cpu"] > 0
    # a tiny band is rendered again with further seeds until the minimum time
    reps = int(r["sample"].split(" render(s)")[0].rsplit(" ", 1)[-1])
    assert reps >= 1
    assert "rows 10..13" in r["sample"]
