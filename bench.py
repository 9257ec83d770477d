"""Benchmark of the path-tracing hot path on one MI355X.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--cpu-rows R] [--no-cpu]

A step is one full render (Render::Run, render.h:1125-1455) of the workload:
the quad Cornell box with the level-6 displaced icosphere (81,944 triangles),
a 0.25x0.25 area light, 512x512, 64 spp (SAMPLING 16 x SUPERSAMPLING 2),
DEPTH 5 / DEPTH_LIMIT 8, next-event estimation.  Inputs (scene, BVH) are
resident in HBM before the timed region; the image stays on the device.

Prints one JSON line: samples/s over the timed steps, Mrays/s, the roofline
of the dominant traversal kernel (algorithmic bytes per launch / its average
duration from HIP events on the library's stream) and the CPU baseline (the
FP64 oracle on a band of rows of the same workload, all host threads).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def progress(msg):
    """A phase line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)
METRIC = "Mrays/s + samples/s at matched RMSE vs host CPU; adjoint grad evals/s, 1 GPU"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-rows", type=int, default=32,
                    help="rows of the workload rendered by the CPU baseline and the FP64 RMSE reference")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-min-s", type=float, default=2.0, help="least time the CPU baseline is timed over")
    ap.add_argument("--no-rmse", action="store_true")
    ap.add_argument("--rmse-seeds", type=int, default=8, help="untimed GPU frames whose median RMSE is reported")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--workload", default="cfg2", choices=["cfg1", "cfg2", "cfg3", "cfg4", "cfg5", "mlt"],
                    help="cfg2 (default) is the metric's configuration; the others are extra measurements "
                         "(cfg4: 64-scene batch vs per-scene renders, cfg5: inverse rendering)")
    ap.add_argument("--no-tune", action="store_true", help="cfg4: skip the batch tuner (library defaults)")
    ap.add_argument("--batch-opts", default=None,
                    help='cfg4: fixed batch options as JSON (e.g. \'{"persistent": 2, "stack_lds": 8}\'); no tuner')
    ap.add_argument("--no-trials", action="store_true",
                    help="cfg4: time only the batched renders (no persistent-engine trials, no per-scene renders)")
    ap.add_argument("--adam-steps", type=int, default=200, help="cfg5: optimisation steps")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run a 1 GiB device copy first (reference dispatch for PMC byte counters)")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box here has one MI355X (asked for --gpus {args.gpus}); "
              "splitting camera samples across GPUs is not built", file=sys.stderr)
        sys.exit(2)
    import torch
    from prender_amd import generators as G
    from prender_amd import prt

    dev = torch.device("cuda:0")
    info = prt.device_info()
    if args.calibrate_copy:
        # 1 GiB read + 1 GiB write of known size: the FETCH_SIZE / WRITE_SIZE
        # counters of this dispatch scale those of the traversal kernels
        a = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    if args.workload == "cfg4":
        print(json.dumps(bench_batch(args, torch, G, prt, info)))
        return
    if args.workload == "cfg5":
        print(json.dumps(bench_inverse(args, torch, G, prt, info)))
        return
    if args.workload == "mlt":
        print(json.dumps(bench_mlt(args, torch, G, prt, info)))
        return
    desc, wl_text = workload(G, args.workload)
    progress(f"{args.workload}: scene build")
    W, H, spp = desc.camera.width, desc.camera.height, desc.spp
    t0 = time.time()
    scene = prt.Scene(desc)
    build_s = time.time() - t0
    sinfo = scene.info()
    rgb = torch.empty((H, W, 3), dtype=torch.float32, device=dev)
    var = torch.empty((H, W, 3), dtype=torch.float32, device=dev)
    for i in range(args.warmup):
        scene.render_device(rgb.data_ptr(), var.data_ptr(), seed=1000 + i, timing=True)
    torch.cuda.synchronize()
    progress(f"timed renders: {args.steps}")
    stats = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        stats.append(scene.render_device(rgb.data_ptr(), var.data_ptr(), seed=args.seed + i, timing=True))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    samples = W * H * spp * args.steps
    sum_ = lambda k: sum(s[k] for s in stats)
    rays = sum_("primary_rays") + sum_("extension_rays") + sum_("shadow_rays")
    ms_ext, ms_sh, ms_shade = sum_("ms_extend"), sum_("ms_shadow"), sum_("ms_shade")
    launches = sum_("launches_extend")
    # engine 0's choice: scenes without light sampling render in one
    # persistent launch (k_persist traces every ray of the frame), scenes
    # with glass through the streaming tile queue (k_stream_extend /
    # _shadow), the others depth by depth (k_extend / k_shadow)
    persistent = not (desc.sampling.next_event_estimation and sinfo["lights"] > 0)
    streamed = not persistent and any(m.reflection == "reflection" for m in desc.materials)
    if persistent:
        kx = ks = "k_persist"
        kernels = {kx: (sum_("bytes_extend") + sum_("bytes_shadow"), ms_ext, launches)}
    else:
        kx, ks = ("k_stream_extend", "k_stream_shadow") if streamed else ("k_extend", "k_shadow")
        kernels = {
            kx: (sum_("bytes_extend"), ms_ext, launches),
            ks: (sum_("bytes_shadow"), ms_sh, launches),
        }
    # the traversal kernel moving the most algorithmic bytes
    dom = max(kernels, key=lambda k: kernels[k][0])
    b, ms, nl = kernels[dom]
    achieved = b / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": samples / elapsed,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": wl_text,
                   "width": W, "height": H, "spp": spp, "triangles": sinfo["tris"], "bvh_nodes": sinfo["nodes"],
                   "parallelism": "1 GPU"},
        "mrays_per_s": rays / elapsed / 1e6,
        "rays": {"primary": sum_("primary_rays"), "extension": sum_("extension_rays"),
                 "shadow": sum_("shadow_rays"), "pickup_hits_no_ray": sum_("pickup_hits")},
        "kernel_ms_per_step": {"extend": ms_ext / args.steps, "shade": ms_shade / args.steps,
                               "nee": sum_("ms_nee") / args.steps, "shadow": ms_sh / args.steps,
                               "other": sum_("ms_other") / args.steps},
        "mrays_per_s_by_class": {
            "primary": (sum_("primary_rays") / sum_("ms_extend_primary") / 1e3
                        if sum_("ms_extend_primary") > 0 else None),
            "secondary": sum_("extension_rays") / max(1e-9, ms_ext - sum_("ms_extend_primary")) / 1e3,
            "shadow": sum_("shadow_rays") / max(1e-9, ms_sh) / 1e3,
            "note": "rays / time of their traversal kernels"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_launch": b / max(1, nl), "avg_launch_ms": ms / max(1, nl),
                     "note": "algorithmic bytes = 128 B per BVH node + 48 B per triangle + 48 B per ray; "
                             "node and triangle reuse in L2 / MALL lets this exceed the HBM figure"},
        "traversal_union": {
            "kernels": kx if persistent else kx + " + " + ks,
            "engine": "persistent" if persistent else ("streaming tile queue" if streamed else "depth-synchronous"),
            "ms_per_step": sum_("ms_traversal") / args.steps,
            "achieved": (sum_("bytes_extend") + sum_("bytes_shadow")) / max(1e-12, sum_("ms_traversal") * 1e-3) / 1e9,
            "unit": "GB/s", "peak": HBM_PEAK_GBS,
            "note": ("the shadow kernel runs on a second stream (depth-synchronous engine: beside the next "
                     "depth's shade and NEE; tile queue: beside the next iteration's closest-hit kernel), so "
                     "its launch time includes the time it shares the chip; this is both traversal kernels' "
                     "algorithmic bytes over the time during which either runs")},
        "device": info["arch"],
        "scene_build_s": build_s,
    }
    tu = out["traversal_union"]
    tu["frac"] = tu["achieved"] / HBM_PEAK_GBS
    if persistent:
        out["roofline"]["note"] += ("; k_persist also shades and samples lights, so its launch time is the "
                                    "whole frame's and its bound is latency rather than HBM")
        out["traversal_union"]["note"] = "one launch traces, shades and resolves every path of the frame"
        out["mrays_per_s_by_class"].update(
            secondary=None, shadow=None,
            note="one launch traces every class of ray; the camera-ray rate is measured on its own below")
    if streamed or persistent:
        # the streaming and persistent engines trace camera rays in the same
        # launches as bounce rays; their rate is measured on their own
        out["mrays_per_s_by_class"]["primary"] = primary_trace_rate(torch, G, desc, scene)
        out["mrays_per_s_by_class"]["primary_note"] = (
            "the frame's W*H*SS^2 camera rays traced on their own through prt_trace_closest_device "
            "(device buffers, HIP-event time of the trace kernel, median of 5)")
    traffic = measured_traffic(dom, args.workload)
    if traffic is not None:
        out["roofline"]["traffic"] = traffic["traffic_bytes_per_launch"]
        out["roofline"]["traffic_source"] = ("profiles/%s_%s_hbm_traffic.md: rocprofv3 FETCH_SIZE x %.3f + "
                                             "WRITE_SIZE x %.3f per dispatch (scales from a 1 GiB copy / fill)"
                                             % (dom, args.workload, traffic["fetch_scale"], traffic["write_scale"]))
    band = cpu_band(desc, args)
    gpu_frames = None
    if not args.no_rmse:
        # untimed frames with fixed seeds (so the figure does not depend on
        # --steps): single-frame RMSE is dominated by rare bright paths, so
        # the median over the frames is reported
        progress(f"RMSE frames: {args.rmse_seeds}")
        gpu_frames = []
        for i in range(args.rmse_seeds):
            scene.render_device(rgb.data_ptr(), None, seed=RMSE_SEED0 + i)
            torch.cuda.synchronize()
            gpu_frames.append(rgb.cpu().numpy())
        ref = torch.empty_like(rgb)
        scene.render_device(ref.data_ptr(), None, seed=999_999, samples=desc.sampling.samples * 16)
        torch.cuda.synchronize()
        ref_np = ref.cpu().numpy()
        whole = [float(np.sqrt(np.mean((f - ref_np) ** 2))) for f in gpu_frames]
        out["rmse_vs_16x_spp_gpu_image"] = {"median": float(np.median(whole)), "by_seed": whole,
                                            "reference": f"GPU render, {desc.spp * 16} spp, whole image"}
    if not args.no_cpu:
        progress(f"CPU baseline: rows {band[0]}..{band[1] - 1}")
        out["cpu_baseline"] = cpu_baseline(desc, args, elapsed / args.steps, band)
    if not args.no_rmse:
        progress("FP64 reference band at 16x spp")
        out["rmse_vs_fp64"] = fp64_band_rmse(desc, args, band, gpu_frames, out.get("cpu_baseline"))
        out["rmse_median"] = out["rmse_vs_fp64"]["gpu_band_median"]
    if "cpu_baseline" in out:
        out["cpu_baseline"].pop("_band_images", None)
    print(json.dumps(out))


def primary_trace_rate(torch, G, desc, scene, reps=5):
    """Mrays/s of the frame's unique camera rays through the closest-hit
    entry point (rays and hits resident on the device)."""
    from prender_amd.scene import HIT_DTYPE
    cam = G.camera_rays(desc)
    n = len(cam)
    rays = torch.from_numpy(cam.view(np.uint8).reshape(-1)).to("cuda")
    hits = torch.empty(n * HIT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    scene.trace_device(rays.data_ptr(), n, hits.data_ptr())
    ms = sorted(scene.trace_device(rays.data_ptr(), n, hits.data_ptr())["ms_total"] for _ in range(reps))
    return n / ms[reps // 2] / 1e3


RMSE_SEED0 = 1          # GPU frames for the RMSE statistics: seeds 1..rmse_seeds
FP64_REF_SEED = 424_242  # the FP64 reference band


def cpu_band(desc, args):
    """Rows [r0, r1) of the reference's y numbering that the CPU baseline
    times and the FP64 reference covers (the band across the image centre)."""
    H = desc.camera.height
    rows = min(args.cpu_rows, H)
    r0 = H // 2 - rows // 2
    return r0, r0 + rows


def fp64_band_rmse(desc, args, band, gpu_frames, cpu):
    """RMSE of the GPU frames and of the CPU baseline's band renders against
    an FP64 oracle render of the same band at 16x the spp: both sides render
    the same estimator at the same spp, so the two figures match up to noise
    and the throughput ratio is the speed-up at matched RMSE."""
    from oracle.oracle import OracleScene
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    o = OracleScene(desc)
    H = desc.camera.height
    r0, r1 = band
    t0 = time.perf_counter()
    ref, _, _ = o.render(seed=FP64_REF_SEED, samples=desc.sampling.samples * 16, rows=(r0, r1), threads=threads)
    ref_s = time.perf_counter() - t0
    rows = slice(H - r1, H - r0)  # image rows are stored top-down (render.h:1234)
    gpu = [float(np.sqrt(np.mean((f[rows] - ref[rows]) ** 2))) for f in gpu_frames]
    res = {"reference": f"FP64 oracle, rows {r0}..{r1 - 1}, {desc.spp * 16} spp, seed {FP64_REF_SEED} "
                        f"({ref_s:.1f} s on {threads} threads)",
           "gpu_band_by_seed": gpu, "gpu_band_median": float(np.median(gpu)),
           "gpu_seeds": f"{RMSE_SEED0}..{RMSE_SEED0 + len(gpu) - 1}"}
    if cpu is not None and cpu.get("_band_images"):
        c = [float(np.sqrt(np.mean((img[rows] - ref[rows]) ** 2))) for img in cpu["_band_images"]]
        res["cpu_band_by_seed"] = c
        res["cpu_band_median"] = float(np.median(c))
        res["cpu_over_gpu_rmse"] = res["cpu_band_median"] / max(1e-30, res["gpu_band_median"])
    if cpu is not None:
        cpu.pop("_band_images", None)
    return res


def measured_traffic(kernel, workload_name):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary of the same workload (PMC counters need their own profiler run),
    or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload_name}_hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d if d.get("kernel") == kernel and d.get("workload") == workload_name else None


def workload(G, name):
    if name == "cfg1":
        return G.config1(), ("cfg1: sphere Cornell box (6 r=1e5 walls, mirror, glass, r=5 light), 320x240, "
                             "16 spp (4x2x2), DEPTH 5 / DEPTH_LIMIT 8, no NEE")
    if name == "cfg3":
        return G.config3(), ("cfg3: ~2.1M triangles in 400 instances (displaced icospheres L3-L6 and boxes) on a "
                             "100x100 ground, 70/15/15 diffuse/mirror/glass, 1 large + 8 small area lights, "
                             "1024x768, 64 spp (16x2x2), DEPTH 5 / DEPTH_LIMIT 10, NEE")
    return G.config2(), ("cfg2: quad Cornell box (24 tris) + L6 displaced icosphere (81,920 tris), "
                         "0.25^2 area light, 512x512, 64 spp (16x2x2), DEPTH 5 / DEPTH_LIMIT 8, NEE")


def bench_mlt(args, torch, G, prt, info):
    """Primary-sample-space Metropolis renders (prt_render_mlt_device) of the
    cfg2 scene: SAMPLING 16 chains of the reference's kind, mutation 4,
    pre_sample 0.4, as 2^20 chains side by side.  Reports paths traced per
    second (seed paths + proposals) and the RMSE of the Metropolis image and
    of a path-traced image with as many paths, both against a 1024 spp
    path-traced reference."""
    desc = G.config2()
    desc.sampling.luminance_cutoff = 0.0
    W, H = desc.camera.width, desc.camera.height
    scene = prt.Scene(desc)
    dev = torch.device("cuda:0")
    img = torch.empty((H, W, 3), dtype=torch.float32, device=dev)
    kw = dict(mutation=4.0, pre_sample=0.4, chains=1 << 20)
    for i in range(args.warmup):
        scene.render_mlt_device(img.data_ptr(), seed=500 + i, **kw)
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        stats.append(scene.render_mlt_device(img.data_ptr(), seed=args.seed + i, **kw))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    paths = sum(s["camera_samples"] for s in stats)
    props = sum(s["mlt_proposals"] for s in stats)
    out = {
        "metric": "MLT paths/s (seed paths + proposals), 1 GPU", "value": paths / elapsed, "unit": "paths/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": "mlt: cfg2 scene (512x512, 81,944 triangles, NEE), PSSMLT with SAMPLING 16 x "
                               "W*H*4 steps over 2^20 chains, pre_sample 0.4", "chains": 1 << 20,
                   "paths_per_render": paths // args.steps, "parallelism": "1 GPU"},
        "proposals_per_s": props / elapsed,
        "accept_rate": sum(s["mlt_accepts"] for s in stats) / max(1, props),
        "mrays_per_s": sum(s["primary_rays"] + s["extension_rays"] + s["shadow_rays"] for s in stats)
        / elapsed / 1e6,
        "device": info["arch"],
    }
    if not args.no_rmse:
        ref = torch.empty_like(img)
        scene.render_device(ref.data_ptr(), None, seed=999_999, samples=256)
        pt = torch.empty_like(img)
        spp_eq = max(1, round(paths / args.steps / (W * H) / 4))  # SUPERSAMPLING 2: 4 rays per sample
        scene.render_device(pt.data_ptr(), None, seed=77, samples=spp_eq)
        torch.cuda.synchronize()
        def blocks(a, k=8):
            return a.reshape(H // k, k, W // k, k, 3).mean(dim=(1, 3))
        out["rmse_vs_1024spp"] = {
            "mlt": float(torch.sqrt(torch.mean((img - ref) ** 2)).item()),
            "path_tracing_equal_paths": float(torch.sqrt(torch.mean((pt - ref) ** 2)).item()),
            "mlt_8x8_blocks": float(torch.sqrt(torch.mean((blocks(img) - blocks(ref)) ** 2)).item()),
            "path_tracing_8x8_blocks": float(torch.sqrt(torch.mean((blocks(pt) - blocks(ref)) ** 2)).item()),
            "path_tracing_spp": spp_eq * 4,
            "note": "the Metropolis image blurs pixel edges with the reference's tent filter, which the "
                    "path-traced reference does not"}
    return out


def bench_batch(args, torch, G, prt, info):
    """64 seeded scenes (sphere-box / quad-box generators, meshes of 5k-500k
    triangles, 128^2-512^2, 8-64 spp) rendered by one batched call per step,
    against the same scenes rendered one prt_render_device call at a time.
    Both produce bit-identical images (same seed), so they are compared at
    equal RMSE by construction; the check is part of the output."""
    descs = G.config4()
    t0 = time.time()
    scenes = [prt.Scene(d) for d in descs]
    build_s = time.time() - t0
    dev = torch.device("cuda:0")
    outs = [torch.empty((s.height, s.width, 3), dtype=torch.float32, device=dev) for s in scenes]
    ptrs = [o.data_ptr() for o in outs]
    samples = sum(s.width * s.height * s.desc.spp for s in scenes)
    tris = sum(s.info()["tris"] for s in scenes)
    t0 = time.time()
    if args.batch_opts:
        opts = prt.batch_opts(**json.loads(args.batch_opts))
        tune_ns = None
    elif args.no_tune:
        opts = prt.batch_opts()
        tune_ns = None
    else:
        progress("batch tuner")
        opts, tune_ns = prt.tune_batch(scenes, ptrs)
    tune_s = time.time() - t0
    for i in range(args.warmup):
        opts.seed = 1000 + i
        prt.render_batch_device(scenes, ptrs, opts=opts)
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        opts.seed = args.seed + i
        stats.append(prt.render_batch_device(scenes, ptrs, opts=opts))
    torch.cuda.synchronize()
    batched_s = (time.perf_counter() - t0) / args.steps
    last = [o.clone() for o in outs]
    # per-kernel device times: one more batch with plain launches and events
    topts = prt.PrtBatchOpts.from_buffer_copy(opts)
    topts.use_graph, topts.timing, topts.seed = -1, 1, args.seed + args.steps - 1
    tst = prt.render_batch_device(scenes, ptrs, opts=topts)
    # the traversal kernels' algorithmic bytes and device time: (bytes, ms,
    # launches); the split form's halves are timed on their own so that the
    # persistent kernel and the tile queue's closest-hit kernel each report
    # their own figure
    roof = {}
    if opts.persistent == 2:
        nee = [i for i, s in enumerate(scenes) if s.desc.sampling.next_event_estimation and s.info()["lights"] > 0]
        rest = [i for i in range(len(scenes)) if i not in nee]
        if rest:
            o1 = prt.PrtBatchOpts.from_buffer_copy(topts)
            o1.persistent, o1.persist_waves = 1, opts.persist_waves or 2
            s1 = prt.render_batch_device([scenes[i] for i in rest], [ptrs[i] for i in rest], opts=o1)
            roof["k_persist"] = (s1["bytes_extend"] + s1["bytes_shadow"], s1["ms_extend"], 1)
        if nee:
            o0 = prt.PrtBatchOpts.from_buffer_copy(topts)
            o0.persistent = 0
            s0 = prt.render_batch_device([scenes[i] for i in nee], [ptrs[i] for i in nee], opts=o0)
            roof["k_stream_extend"] = (s0["bytes_extend"], s0["ms_extend"], s0["launches_extend"])
    elif opts.persistent == 1:
        roof["k_persist"] = (tst["bytes_extend"] + tst["bytes_shadow"], tst["ms_extend"], 1)
    else:
        roof["k_stream_extend"] = (tst["bytes_extend"], tst["ms_extend"], tst["launches_extend"])

    def roofline_of(k):
        b, ms, nl = roof[k]
        a = b / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        r = {"bound": "hbm", "kernel": k, "achieved": a, "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": a / HBM_PEAK_GBS, "traffic": None, "bytes_per_launch": b / max(1, nl),
             "avg_launch_ms": ms / max(1, nl)}
        tr = measured_traffic(k, "cfg4")
        cur = opts.as_dict()
        if tr is not None and all(cur.get(key) == v for key, v in (tr.get("batch_opts") or {}).items()):
            r["traffic"] = tr["traffic_bytes_per_launch"]
            r["traffic_source"] = ("profiles/%s_cfg4_hbm_traffic.md: rocprofv3 FETCH_SIZE x %.3f + WRITE_SIZE x %.3f "
                                   "per dispatch, batch options %s" % (k, tr["fetch_scale"], tr["write_scale"],
                                                                      tr.get("batch_opts")))
        return r
    dom = max(roof, key=lambda k: roof[k][1])  # the kernel with the most device time
    pers = seq_s = max_diff = seq_diff = None
    if not args.no_trials:
        # the persistent engine: one launch over every sample of the batch; one
        # trial per register budget (2, 3 or 4 waves per SIMD), the fastest timed
        progress("persistent engine")
        pers = {}
        trials = {}
        for wv in (2, 3, 4):
            po = prt.batch_opts(seed=args.seed + args.steps - 1, persistent=True, trav_blocks_per_cu=wv)
            trials[wv] = prt.render_batch_device(scenes, ptrs, opts=po)["ms_total"]
        best_w = min(trials, key=trials.get)
        torch.cuda.synchronize()
        pers_diff = [i for i, (a, b) in enumerate(zip(last, outs)) if not torch.equal(a, b)]
        pers_same = not pers_diff
        t0 = time.perf_counter()
        for i in range(args.steps):
            po = prt.batch_opts(seed=args.seed + i, persistent=True, trav_blocks_per_cu=best_w)
            pst = prt.render_batch_device(scenes, ptrs, opts=po)
        torch.cuda.synchronize()
        pers_s = (time.perf_counter() - t0) / args.steps
        pers = {"samples_per_s": samples / pers_s, "ms_per_step": 1e3 * pers_s,
                "waves_per_simd": best_w, "trial_ms_by_waves": trials,
                "images_identical_to_batched": pers_same, "differing_scenes": pers_diff[:16],
                "over_batched": batched_s / pers_s, "over_per_scene": None,
                "kernel_ms": pst["ms_extend"], "resolve_ms": pst["ms_other"]}
        progress("per-scene renders of the same scenes")
        # sequential per-scene renders of the same work
        for s, o in zip(scenes, outs):  # warm-up: workspaces
            s.render_device(o.data_ptr(), seed=1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(args.steps):
            for s, o in zip(scenes, outs):
                s.render_device(o.data_ptr(), seed=args.seed + i)
        torch.cuda.synchronize()
        seq_s = (time.perf_counter() - t0) / args.steps
        max_diff = max(float((a - b).abs().max()) for a, b in zip(last, outs))
        seq_diff = [i for i, (a, b) in enumerate(zip(last, outs)) if not torch.equal(a, b)]
        pers["over_per_scene"] = seq_s / pers_s
    sum_ = lambda k: sum(s[k] for s in stats)
    rays = sum_("primary_rays") + sum_("extension_rays") + sum_("shadow_rays")
    ext_ms = tst["ms_extend"]
    out = {
        "metric": METRIC, "value": samples / batched_s, "unit": "samples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * batched_s, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (64 seeded procedural scenes)",
        "config": {"workload": "cfg4: 64 scenes (sphere-box / quad-box + displaced lat-long mesh of 5k-500k tris, "
                               "log-uniform), 128^2..512^2, 8..64 spp (SUPERSAMPLING 2), DEPTH 5 / DEPTH_LIMIT 8",
                   "scenes": len(scenes), "camera_samples": samples, "triangles": tris, "parallelism": "1 GPU"},
        "mrays_per_s": rays / (batched_s * args.steps) / 1e6,
        "per_scene_launches": {"samples_per_s": samples / seq_s, "ms_per_step": 1e3 * seq_s} if seq_s else None,
        "batched_over_per_scene": seq_s / batched_s if seq_s else None,
        "persistent": pers,
        "images_identical": max_diff == 0.0 if seq_s else None, "max_abs_diff": max_diff,
        "differing_scenes": ([(i, bool(descs[i].sampling.next_event_estimation)) for i in seq_diff[:16]]
                             if seq_s else None),
        "batch_options": opts.as_dict(), "tuner_ns_per_sample": tune_ns, "tune_s": tune_s,
        "iterations_per_step": stats[-1]["launches_extend"],
        "rays_per_step": {"primary": stats[-1]["primary_rays"], "extension": stats[-1]["extension_rays"],
                          "shadow": stats[-1]["shadow_rays"]},
        "mixed_scene_passes_per_step": stats[-1]["mixed_scene_passes"],
        "kernel_ms_per_step_plain_launches": {"extend": ext_ms, "shade": tst["ms_shade"], "nee": tst["ms_nee"],
                                              "shadow": tst["ms_shadow"], "other": tst["ms_other"],
                                              "total": tst["ms_total"]},
        "roofline": roofline_of(dom),
        "roofline_other_kernels": [roofline_of(k) for k in roof if k != dom],
        "device": info["arch"], "scene_build_s": build_s,
    }
    if not args.no_cpu:
        progress("CPU baseline: oracle on a sample of the scenes")
        out["cpu_baseline"] = cpu_batch_baseline(descs, samples / batched_s)
    return out


def cpu_batch_baseline(descs, gpu_samples_per_s, every=8, rows=8):
    """The FP64 oracle on a bounded sample of the batch: every `every`-th
    scene, a band of `rows` rows through its middle at the scene's spp (the
    reference renders one scene per process; the sample's samples/s is the
    host rate for this mix of scenes)."""
    from oracle.oracle import OracleScene
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    n = rays = 0
    dt = 0.0
    picked = descs[::every]
    for d in picked:
        o = OracleScene(d)
        H, W = d.camera.height, d.camera.width
        r0 = H // 2 - rows // 2
        t0 = time.perf_counter()
        _, _, st = o.render(seed=1, rows=(r0, r0 + rows), threads=threads)
        dt += time.perf_counter() - t0
        n += W * rows * d.spp
        rays += st["primary"] + st["extension"] + st["shadow"] + st["pickup"]
        o.close()
    v = n / dt
    return {"value": v, "unit": "samples/s", "cores": threads, "kind": "port",
            "sample": f"{len(picked)} of the {len(descs)} scenes (every {every}th), {rows} rows through the middle "
                      f"of each at its spp ({n} camera samples, {dt:.1f} s; BVH builds excluded)",
            "mrays_per_s": rays / dt / 1e6, "gpu_over_cpu": gpu_samples_per_s / v}


def bench_inverse(args, torch, G, prt, info):
    """Inverse rendering: recover 81,920 per-face RGB albedos (+ the light's
    emission) of the quad box + displaced icosphere from a 1024-spp target
    rendered with seeded U[0.1,0.9] albedos; forward and path-replay adjoint at
    32 spp, L2 image loss, Adam."""
    from prender_amd.autograd import render
    W = H = 256
    d_t = G.config5(width=W, height=H, target=True)
    sc_t = prt.Scene(d_t)
    target = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    sc_t.render_device(target.data_ptr(), None, seed=7, samples=256)  # 256 x 2^2 = 1024 spp
    d = G.config5(width=W, height=H)
    sc = prt.Scene(d)
    nf = d.entities[1].indices.shape[0]
    alb_true = torch.from_numpy(d_t.entities[1].face_albedo).cuda()
    alb = torch.full((nf, 3), 0.5, device="cuda", requires_grad=True)
    em = torch.tensor(40.0, requires_grad=True)
    opt = torch.optim.Adam([{"params": [alb], "lr": 0.02}, {"params": [em], "lr": 0.5}])
    # forward vs adjoint cost on the same scene
    img = torch.empty_like(target)
    gimg = torch.ones_like(target)
    galb = torch.empty((nf, 3), dtype=torch.float32, device="cuda")
    sc.render_device(img.data_ptr(), None, seed=1)
    sc.adjoint_device(gimg.data_ptr(), 1, galb.data_ptr(), 2, seed=2)
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for i in range(reps):
        sc.render_device(img.data_ptr(), None, seed=10 + i)
    torch.cuda.synchronize()
    fwd_s = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for i in range(reps):
        sc.adjoint_device(gimg.data_ptr(), 1, galb.data_ptr(), 2, seed=20 + i)
    torch.cuda.synchronize()
    adj_s = (time.perf_counter() - t0) / reps
    # observability of each face: |d(sum of the image)/d(albedo)| at 64 spp;
    # most of the 81,920 faces cover less than a pixel of the 256x256 image,
    # so only the pixel averages of neighbouring faces are constrained
    sc.set_face_albedo_device(1, alb.detach().contiguous().data_ptr())
    obs = torch.empty((nf, 3), dtype=torch.float32, device="cuda")
    sc.adjoint_device(gimg.data_ptr(), 1, obs.data_ptr(), 2, seed=5, samples=64)
    obs = obs.abs().sum(dim=1)
    top = {"top1pct": obs >= torch.quantile(obs, 0.99), "top10pct": obs >= torch.quantile(obs, 0.90)}
    losses = []
    alb_err0 = float((alb.detach() - alb_true).abs().mean())
    t0 = time.perf_counter()
    progress(f"Adam steps: {args.adam_steps}")
    for it in range(args.adam_steps):
        opt.zero_grad()
        out = render(alb, em, sc, 1, 2, seed=2 * it + 100, adjoint_seed=2 * it + 101)
        loss = ((out - target) ** 2).mean()
        loss.backward()
        opt.step()
        with torch.no_grad():
            alb.clamp_(0.0, 1.0)
        losses.append(float(loss.detach()))
    torch.cuda.synchronize()
    opt_s = time.perf_counter() - t0
    recovery = albedo_recovery(G, d, sc, alb.detach(), alb_true)
    # faces seen by the camera move towards the truth; report the error over
    # faces whose albedo changed
    moved = (alb.detach() - 0.5).abs().sum(dim=1) > 1e-3
    alb_err = float((alb.detach() - alb_true).abs().mean())
    alb_err_moved0 = float((0.5 - alb_true[moved]).abs().mean()) if bool(moved.any()) else None
    alb_err_moved = float((alb.detach()[moved] - alb_true[moved]).abs().mean()) if bool(moved.any()) else None
    out = {
        "metric": METRIC, "value": args.adam_steps / opt_s, "unit": "grad evals/s", "n_gpus": 1,
        "steps": args.adam_steps, "warmup": 1, "ms_per_step": 1e3 * opt_s / args.adam_steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded procedural scene and albedos)",
        "config": {"workload": "cfg5: inverse rendering of 81,920 per-face RGB albedos + light emission, quad box + "
                               "L6 displaced icosphere, 256x256, forward 32 spp, adjoint 32 spp, target 1024 spp, "
                               "L2 loss, Adam", "parallelism": "1 GPU"},
        "forward_ms": 1e3 * fwd_s, "adjoint_ms": 1e3 * adj_s, "adjoint_over_forward": adj_s / fwd_s,
        "loss_first": losses[0], "loss_last": losses[-1], "loss_last10_mean": float(np.mean(losses[-10:])),
        "emission_final": float(em.detach()), "emission_true": 60.0,
        "albedo_mean_abs_err": {"start": alb_err0, "end": alb_err, "faces_updated": int(moved.sum()),
                                "updated_start": alb_err_moved0, "updated_end": alb_err_moved,
                                **{f"{k}_start": float((0.5 - alb_true[m]).abs().mean()) for k, m in top.items()},
                                **{f"{k}_end": float((alb.detach()[m] - alb_true[m]).abs().mean())
                                   for k, m in top.items()},
                                "note": "faces ranked by |d image sum / d albedo|; most faces cover less than a "
                                        "pixel, so only pixel averages of neighbouring faces are constrained"},
        "albedo_recovery": recovery,
        "device": info["arch"],
    }
    if not args.no_cpu:
        progress("CPU baseline: FP64 forward + adjoint")
        out["cpu_baseline"] = cpu_adjoint_baseline(d, fwd_s + adj_s)
    return out


def albedo_recovery(G, desc, scene, alb, alb_true, mesh=1):
    """What the camera can identify: the albedo seen through each pixel
    (the mean over its SS^2 camera rays of the albedo of the mesh face each
    ray hits first), and the albedo of faces hit by at least SS^2 camera rays
    (a footprint of one pixel or more); mean absolute error of the initial
    guess (0.5) and of the recovered albedo against the truth."""
    cam = G.camera_rays(desc)
    hits, _ = scene.trace(cam)
    ss2 = desc.sampling.supersamples ** 2
    ent = hits["entity"].reshape(-1, ss2)
    face = hits["prim"].reshape(-1, ss2)
    a = alb.cpu().numpy().astype(np.float64)
    t = alb_true.cpu().numpy().astype(np.float64)
    on = (ent == mesh).all(axis=1)                    # every ray of the pixel meets the mesh first
    pf = face[on]
    seen_t, seen_a = t[pf].mean(axis=1), a[pf].mean(axis=1)
    counts = np.bincount(face[ent == mesh], minlength=a.shape[0])
    big = counts >= ss2
    return {"pixels_on_mesh": int(on.sum()),
            "pixel_albedo_mae": {"start": float(np.abs(0.5 - seen_t).mean()),
                                 "end": float(np.abs(seen_a - seen_t).mean())},
            "faces_with_pixel_footprint": int(big.sum()),
            "footprint_faces_mae": {"start": float(np.abs(0.5 - t[big]).mean()) if big.any() else None,
                                    "end": float(np.abs(a[big] - t[big]).mean()) if big.any() else None},
            "note": "camera rays of the fixed sub-pixel grid traced through prt_trace_closest; a pixel counts "
                    "when all its rays hit the mesh first"}


def cpu_adjoint_baseline(desc, gpu_eval_s):
    """One gradient evaluation of the inverse-rendering workload on the host:
    the FP64 oracle's forward render and its adjoint (the CPU restatement of
    the path-replay adjoint) of the full image at the workload's spp."""
    from oracle.oracle import OracleScene
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    o = OracleScene(desc)
    H, W = desc.camera.height, desc.camera.width
    t0 = time.perf_counter()
    o.render(seed=1, threads=threads)
    t1 = time.perf_counter()
    o.adjoint(seed=2, dimage=np.ones((H, W, 3)), entity=1, light=2, threads=threads)
    t2 = time.perf_counter()
    return {"value": 1.0 / (t2 - t0), "unit": "grad evals/s", "cores": threads, "kind": "port",
            "sample": f"one forward ({t1 - t0:.1f} s) + one adjoint ({t2 - t1:.1f} s) of the full "
                      f"{W}x{H}x{desc.spp}spp image",
            "forward_s": t1 - t0, "adjoint_s": t2 - t1, "gpu_over_cpu": (t2 - t0) / gpu_eval_s}


def cpu_baseline(desc, args, gpu_step_s, band):
    """The FP64 oracle (a restatement of the reference's Render::Run) on a
    band of rows of the same workload, timed on this host's cores."""
    from oracle.oracle import OracleScene
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    o = OracleScene(desc)
    H, W = desc.camera.height, desc.camera.width
    r0, r1 = band
    # the band is rendered with the seeds of the GPU's RMSE frames (so the
    # two RMSE medians are over the same random numbers), and again with
    # further seeds until at least --cpu-min-s seconds are timed
    min_reps = 1 if getattr(args, "no_rmse", False) else max(1, args.rmse_seeds)
    dt, reps, rays, images = 0.0, 0, 0, []
    while True:
        t0 = time.perf_counter()
        img, var, st = o.render(seed=RMSE_SEED0 + reps, rows=(r0, r1), threads=threads)
        dt += time.perf_counter() - t0
        rays += st["primary"] + st["extension"] + st["shadow"] + st["pickup"]
        if len(images) < args.rmse_seeds:
            images.append(img)
        reps += 1
        if (dt >= args.cpu_min_s and reps >= min_reps) or reps >= 64:
            break
    n = W * (r1 - r0) * desc.spp * reps
    res = {"value": n / dt, "unit": "samples/s", "cores": threads, "kind": "port",
           "sample": f"rows {r0}..{r1 - 1} of the {W}x{H}x{desc.spp}spp workload, {reps} render(s), seeds "
                     f"{RMSE_SEED0}..{RMSE_SEED0 + reps - 1} ({n} camera samples, {dt:.1f} s)",
           "mrays_per_s": rays / dt / 1e6, "_band_images": images}
    res["gpu_over_cpu"] = (desc.spp * W * H / gpu_step_s) / res["value"]
    return res


if __name__ == "__main__":
    main()
