"""Batched multi-scene rendering through the shared tile queue: every scene's
image (and per-pixel variance) must be bit-identical to a per-scene render
with the same seed, whatever the tile size, in-flight tile count, LDS stack
depth, traversal grid or launch mode (the per-sample sums are exact integers
and a sample's random numbers depend only on its pixel and index) -- also
in the persistent engine, where one launch follows every sample's whole path."""
import numpy as np
import pytest

from prender_amd import generators as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch_env():
    import torch
    from prender_amd import prt
    assert prt.device_info()["arch"].startswith("gfx950")
    descs = G.config4(n_scenes=6, seed=7, tri_range=(2_000, 20_000), resolutions=(32, 48, 64), spps=(8, 16, 32))
    scenes = [prt.Scene(d) for d in descs]
    ref = []
    for sc in scenes:
        rgb = torch.empty((sc.height, sc.width, 3), dtype=torch.float32, device="cuda")
        var = torch.empty_like(rgb)
        sc.render_device(rgb.data_ptr(), var.data_ptr(), seed=21)
        ref.append((rgb, var))
    torch.cuda.synchronize()
    return torch, prt, scenes, ref


def _run_batch(torch, prt, scenes, **kw):
    outs = [torch.full((sc.height, sc.width, 3), -1.0, dtype=torch.float32, device="cuda") for sc in scenes]
    vars_ = [torch.full_like(o, -1.0) for o in outs]
    st = prt.render_batch_device(scenes, [o.data_ptr() for o in outs], [v.data_ptr() for v in vars_], seed=21, **kw)
    torch.cuda.synchronize()
    return outs, vars_, st


@pytest.mark.parametrize("kw", [
    {},
    {"use_graph": True},
    {"tile_samples": 256, "inflight_tiles": 7},
    {"tile_samples": 1024, "inflight_tiles": 3, "stack_lds": 8, "trav_blocks_per_cu": 4},
    {"tile_samples": 16384, "stack_lds": 32, "trav_blocks_per_cu": 16},
    {"persistent": True},
    {"persistent": True, "trav_blocks_per_cu": 3, "stack_lds": 8},
    {"persistent": True, "trav_blocks_per_cu": 4},
    {"persistent": True, "persist_waves": 3},
    {"persistent": 2},
    {"persistent": 2, "tile_samples": 1024, "stack_lds": 8, "use_graph": False},
    {"persistent": 2, "persist_waves": 4},
])
def test_batch_matches_per_scene_renders(batch_env, kw):
    torch, prt, scenes, ref = batch_env
    print("batch options", kw, flush=True)
    outs, vars_, st = _run_batch(torch, prt, scenes, **kw)
    total = sum(sc.width * sc.height * sc.desc.spp for sc in scenes)
    assert st["camera_samples"] == total
    for i, (o, v) in enumerate(zip(outs, vars_)):
        r, rv = ref[i]
        assert torch.equal(o, r), f"scene {i}: max |diff| {float((o - r).abs().max())}"
        assert torch.equal(v, rv), f"scene {i}: variance differs"
    print("stats", st)


def test_batch_subsample_renders_only_its_tiles(batch_env):
    torch, prt, scenes, ref = batch_env
    outs, _, st = _run_batch(torch, prt, scenes, tile_samples=512, tile_stride=3)
    total = sum(sc.width * sc.height * sc.desc.spp for sc in scenes)
    assert 0 < st["camera_samples"] < total
    # rendered pixels equal the reference, the others are untouched
    for i, o in enumerate(outs):
        done = (o != -1.0).all(dim=2)
        assert bool(done.any())
        assert torch.equal(o[done], ref[i][0][done])


def test_tuner_returns_valid_options(batch_env):
    torch, prt, scenes, ref = batch_env
    outs = [torch.empty((sc.height, sc.width, 3), dtype=torch.float32, device="cuda") for sc in scenes]
    best, ns = prt.tune_batch(scenes, [o.data_ptr() for o in outs], tile_stride=2)
    d = best.as_dict()
    print("tuned", d, "ns/sample", ns)
    assert ns > 0
    assert d["stack_lds"] in (8, 12, 16, 32) and d["trav_blocks_per_cu"] in (4, 8, 16)
    # inflight_tiles 0: the automatic in-flight size (the tuner keeps it)
    assert d["tile_samples"] in (1024, 4096, 16384) and d["inflight_tiles"] == 0
    assert d["persistent"] in (0, 1, 2) and d["persist_waves"] in (0, 2, 3, 4)
    assert d["use_graph"] in (0, 1)
    # the tuned options still render identically
    o2, v2, _ = _run_batch(torch, prt, scenes, **{k: d[k] for k in ("tile_samples", "inflight_tiles", "stack_lds",
                                                                      "trav_blocks_per_cu", "persistent",
                                                                      "persist_waves")}, use_graph=bool(d["use_graph"]))
    for i, o in enumerate(o2):
        assert torch.equal(o, ref[i][0])
