"""The batched workload at full size: 64 seeded scenes from the sphere-box and
quad-box generators, one displaced mesh each with 5k-500k triangles
(log-uniform), 128^2-512^2 pixels, 8-64 spp, all rendered in one
prt_render_batch_device call.

* Every scene's image and variance must equal its own prt_render_device
  render with the same seed bit for bit (a sample's random numbers depend
  only on its pixel and index, and the per-sample sums are exact integers),
  for the replayed tile-queue engine, for the persistent one-launch
  engine and for the split form (each engine on its half), and the batch must report exactly the sum of W*H*spp camera
  samples.
* Four of the scenes (two per generator, the largest mesh of each and a
  small one) are rendered at reduced size on common random numbers against
  the FP64 oracle, with the criterion of the other parity tests.
"""
import copy

import numpy as np
import pytest

from oracle.oracle import OracleScene
from parity_util import assert_crn
from prender_amd import generators as G

pytestmark = pytest.mark.gpu

SEED = 31


@pytest.fixture(scope="module")
def cfg4_descs():
    descs = G.config4()
    print("cfg4:", len(descs), "scenes,", sum(d.n_triangles for d in descs), "triangles,",
          sum(d.camera.width * d.camera.height * d.spp for d in descs), "camera samples", flush=True)
    return descs


def test_cfg4_full_batch_equals_per_scene_renders(cfg4_descs):
    import torch
    from prender_amd import prt
    assert prt.device_info()["arch"].startswith("gfx950")
    scenes = [prt.Scene(d) for d in cfg4_descs]
    outs = [torch.full((sc.height, sc.width, 3), -1.0, dtype=torch.float32, device="cuda") for sc in scenes]
    vars_ = [torch.full_like(o, -1.0) for o in outs]
    print("batched render of the 64 scenes", flush=True)
    st = prt.render_batch_device(scenes, [o.data_ptr() for o in outs], [v.data_ptr() for v in vars_], seed=SEED)
    torch.cuda.synchronize()
    total = sum(sc.width * sc.height * sc.desc.spp for sc in scenes)
    assert st["camera_samples"] == total
    assert st["truncated_paths"] == 0
    print("batch stats", {k: st[k] for k in ("camera_samples", "extension_rays", "shadow_rays", "ms_total")},
          flush=True)
    # the persistent engine: one launch over every sample of the 64 scenes
    po = [torch.full_like(o, -1.0) for o in outs]
    pv = [torch.full_like(o, -1.0) for o in outs]
    print("persistent render of the 64 scenes", flush=True)
    sp = prt.render_batch_device(scenes, [o.data_ptr() for o in po], [v.data_ptr() for v in pv], seed=SEED,
                                 persistent=True)
    torch.cuda.synchronize()
    assert sp["camera_samples"] == total and sp["launches_extend"] == 1
    for i in range(len(scenes)):
        assert torch.equal(po[i], outs[i]), f"persistent scene {i}: max |diff| {float((po[i] - outs[i]).abs().max())}"
        assert torch.equal(pv[i], vars_[i]), f"persistent scene {i}: variance differs"
    print("persistent stats", {k: sp[k] for k in ("camera_samples", "extension_rays", "shadow_rays", "ms_total")},
          flush=True)
    # split form: the no-NEE scenes through the persistent engine, the NEE
    # scenes through the tile queue
    print("split render of the 64 scenes", flush=True)
    po = [torch.full_like(o, -1.0) for o in outs]
    pv = [torch.full_like(o, -1.0) for o in outs]
    ss = prt.render_batch_device(scenes, [o.data_ptr() for o in po], [v.data_ptr() for v in pv], seed=SEED,
                                 persistent=2)
    torch.cuda.synchronize()
    assert ss["camera_samples"] == total
    bad = [i for i in range(len(scenes)) if not (torch.equal(po[i], outs[i]) and torch.equal(pv[i], vars_[i]))]
    assert not bad, f"split form differs on scenes {bad}"
    del po, pv
    # per-scene renders one scene at a time; each handle (and the render
    # workspace it allocated) is released after its comparison
    for i, sc in enumerate(scenes):
        rgb = torch.empty_like(outs[i])
        var = torch.empty_like(outs[i])
        s1 = sc.render_device(rgb.data_ptr(), var.data_ptr(), seed=SEED)
        torch.cuda.synchronize()
        assert s1["camera_samples"] == sc.width * sc.height * sc.desc.spp
        assert torch.equal(outs[i], rgb), f"scene {i}: max |diff| {float((outs[i] - rgb).abs().max())}"
        assert torch.equal(vars_[i], var), f"scene {i}: variance differs"
        assert float(rgb.mean()) > 0
        if i:
            sc.close()
            outs[i] = vars_[i] = None
    print("all 64 batch images equal their per-scene renders", flush=True)


def _pick(descs):
    """Two scenes of each generator: the largest mesh and the smallest."""
    even = [i for i in range(len(descs)) if i % 2 == 0]   # quad box + area light, NEE
    odd = [i for i in range(len(descs)) if i % 2 == 1]    # sphere box with mirror and glass
    out = []
    for grp in (even, odd):
        by_size = sorted(grp, key=lambda i: descs[i].n_triangles)
        out += [by_size[-1], by_size[0]]
    return out


def test_cfg4_scenes_common_random_numbers(cfg4_descs):
    from prender_amd import prt
    for i in _pick(cfg4_descs):
        d = copy.deepcopy(cfg4_descs[i])
        d.camera.width = d.camera.height = 32
        d.sampling.samples = 2
        print(f"cfg4 scene {i}: {d.n_triangles} triangles, 32x32x8spp seed {SEED}", flush=True)
        rg, vg, st = prt.Scene(d).render(seed=SEED)
        ro, vo, _ = OracleScene(d).render(seed=SEED)
        assert_crn(rg, vg, ro, vo, f"cfg4 scene {i}")


def test_cfg4_repeated_tile_queue_renders(cfg4_descs):
    """Eight renders of the batch (split form, 1024-sample tiles: a tile of an
    8- or 16-spp scene holds 128 or 64 pixels, so two waves of the resolve
    kernel fold one tile), each NEE scene's image against its own per-scene
    render of the same seed.  A resolve kernel that freed a tile's slot while
    another wave of its block had yet to read the slot record left half such
    a tile unwritten and its sums in the slot's next tile, in about one render
    in twenty."""
    import torch
    from prender_amd import prt
    scenes = [prt.Scene(d) for d in cfg4_descs]
    nee = [i for i, s in enumerate(scenes) if s.desc.sampling.next_event_estimation]
    outs = [torch.empty((s.height, s.width, 3), dtype=torch.float32, device="cuda") for s in scenes]
    ref = {i: torch.empty_like(outs[i]) for i in nee}
    opts = dict(tile_samples=1024, inflight_tiles=4096, stack_lds=8, trav_blocks_per_cu=8, persistent=2)
    for k in range(8):
        seed = 200 + k
        prt.render_batch_device(scenes, [o.data_ptr() for o in outs], seed=seed, **opts)
        torch.cuda.synchronize()
        for i in nee:
            scenes[i].render_device(ref[i].data_ptr(), seed=seed)
        torch.cuda.synchronize()
        bad = [i for i in nee if not torch.equal(outs[i], ref[i])]
        print(f"render {k} seed {seed}: scenes differing {bad}", flush=True)
        assert not bad
