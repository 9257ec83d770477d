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
    args = types.SimpleNamespace(cpu_rows=4, seed=3, cpu_min_s=0.05, rmse_seeds=8)
    band = bench.cpu_band(d, args)
    assert band == (10, 14)
    r = bench.cpu_baseline(d, args, 1e-3, band)
    assert r["kind"] == "port" and r["unit"] == "samples/s" and r["cores"] >= 1
    assert r["value"] > 0 and r["mrays_per_s"] > 0 and r["gpu_over_cpu"] > 0
    # a tiny band is rendered again with further seeds until the minimum time
    reps = int(r["sample"].split(" render(s)")[0].rsplit(" ", 1)[-1])
    assert reps >= 8 and len(r["_band_images"]) == 8   # at least the 8 RMSE seeds
    assert "rows 10..13" in r["sample"]


def test_fp64_band_rmse_matches_equal_spp_renders():
    # two oracle renders at the workload's spp stand in for the GPU frames:
    # the GPU and CPU figures are the same estimator's RMSE against the
    # 16x-spp FP64 band, so they agree up to noise
    sys.path.insert(0, ROOT)
    import bench
    from oracle.oracle import OracleScene
    from prender_amd import generators as G
    d = G.config1(width=32, height=24, samples=2)
    args = types.SimpleNamespace(cpu_rows=4, seed=3, cpu_min_s=0.01, rmse_seeds=3)
    band = bench.cpu_band(d, args)
    cpu = bench.cpu_baseline(d, args, 1e-3, band)
    o = OracleScene(d)
    frames = [o.render(seed=100 + i, rows=band)[0] for i in range(3)]
    res = bench.fp64_band_rmse(d, args, band, frames, cpu)
    assert "_band_images" not in cpu
    assert len(res["gpu_band_by_seed"]) == 3 and res["gpu_band_median"] > 0
    assert res["cpu_band_median"] > 0
    assert 0.5 < res["cpu_over_gpu_rmse"] < 2.0


def test_cpu_batch_and_adjoint_baselines():
    sys.path.insert(0, ROOT)
    import bench
    from prender_amd import generators as G
    descs = G.config4(n_scenes=4, tri_range=(500, 2000), resolutions=(16,), spps=(8,))
    r = bench.cpu_batch_baseline(descs, 1e6, every=2, rows=4)
    assert r["kind"] == "port" and r["value"] > 0 and "2 of the 4 scenes" in r["sample"]
    d = G.config5(width=16, height=16, samples=1)
    a = bench.cpu_adjoint_baseline(d, 1e-3)
    assert a["unit"] == "grad evals/s" and a["value"] > 0 and a["forward_s"] > 0 and a["adjoint_s"] > 0
