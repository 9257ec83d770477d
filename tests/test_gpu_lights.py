"""NEE with several lights (LightList::randomLight, light.h:132-155) through
the three ways the NEE kernels read light records: one light in wave-uniform
registers, up to 16 lights in per-wave LDS tables, and more lights straight
from global memory.

* common random numbers against the oracle for 3 and 20 sphere lights;
* the depth-synchronous and streaming engines give identical images;
* a batch of scenes with 1, 3, 5 and 20 lights equals per-scene renders bit
  for bit (the batch's NEE kernel reads each chunk's scene's lights, with
  scenes changing from chunk to chunk).
"""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from prender_amd import generators as G
from prender_amd.scene import Material, Sphere

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def prt():
    from prender_amd import prt as P
    info = P.device_info()
    print("device", info, flush=True)
    assert info["arch"].startswith("gfx950")
    return P


def lit_scene(n_extra, width=32, height=32, samples=2, level=3, seed=2):
    """The cfg2 box and mesh at reduced size, its area light plus n_extra small
    sphere lights of different powers under the ceiling."""
    d = G.config2(width=width, height=height, samples=samples, level=level, seed=seed)
    rng = np.random.default_rng(100 + n_extra)
    for k in range(n_extra):
        d.materials.append(Material(color=(0.0, 0.0, 0.0), emission=(float(rng.uniform(5.0, 40.0)),) * 3))
        pos = (float(rng.uniform(-0.8, 0.8)), float(rng.uniform(1.6, 1.9)), float(rng.uniform(-0.8, 2.0)))
        d.entities.append(Sphere(position=pos, radius=float(rng.uniform(0.02, 0.06)), material=len(d.materials) - 1))
    return d


def crn_close(g, o, frac=0.98):
    close = np.abs(g - o) <= 1e-4 * np.abs(o) + 1e-6
    assert close.mean() >= frac, close.mean()
    assert abs(float(g.mean()) - float(o.mean())) <= 1e-2 * max(float(o.mean()), 1e-6)


@pytest.mark.parametrize("n_extra", [2, 19])
def test_many_lights_match_oracle(prt, n_extra):
    d = lit_scene(n_extra)
    print(f"{n_extra + 1} lights 32x32x8spp CRN seed 6", flush=True)
    rg, _, st = prt.Scene(d).render(seed=6)
    ro, _, so = OracleScene(d).render(seed=6)
    assert st["shadow_rays"] > 0
    crn_close(rg, ro)


@pytest.mark.parametrize("n_extra", [0, 2, 19])
def test_many_lights_engines_identical(prt, n_extra):
    d = lit_scene(n_extra, width=40, height=24)
    s = prt.Scene(d)
    print(f"{n_extra + 1} lights engines 1 / 2", flush=True)
    a, va, sa = s.render(seed=9, engine=1)
    b, vb, sb = s.render(seed=9, engine=2)
    assert np.array_equal(a, b) and np.array_equal(va, vb)
    assert sa["shadow_rays"] == sb["shadow_rays"] > 0


@pytest.mark.parametrize("extras", [(0, 2, 4), (0, 2, 19)])
def test_batch_with_mixed_light_counts(prt, extras):
    import torch
    descs = [lit_scene(n, width=24 + 8 * i, height=24, seed=2 + i) for i, n in enumerate(extras)]
    scenes = [prt.Scene(d) for d in descs]
    print("batch of scenes with", [n + 1 for n in extras], "lights", flush=True)
    ref = []
    for sc in scenes:
        rgb = torch.empty((sc.height, sc.width, 3), dtype=torch.float32, device="cuda")
        sc.render_device(rgb.data_ptr(), None, seed=21, engine=1)
        ref.append(rgb)
    outs = [torch.full((sc.height, sc.width, 3), -1.0, dtype=torch.float32, device="cuda") for sc in scenes]
    prt.render_batch_device(scenes, [o.data_ptr() for o in outs], None, seed=21, tile_samples=256)
    torch.cuda.synchronize()
    for i, (o, r) in enumerate(zip(outs, ref)):
        assert torch.equal(o, r), f"scene {i}: max |diff| {float((o - r).abs().max())}"
