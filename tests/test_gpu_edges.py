"""Edge cases and size-independent properties of the GPU render path, through
the C ABI.

* at the full cfg2 size (512x512, 64 spp): the image does not depend on how
  the frame is cut into wavefront batches (fixed-point sums and
  counter-based random numbers make it a pure function of the seed), and the
  same seed gives the same image bit for bit;
* degenerate and ragged sizes: a 1x1 image and a 3x5 image on common random
  numbers with the oracle;
* an empty ray batch, NEE on a scene without emitters, the vertex cap
  (truncation counted as the oracle counts it) and bad descriptions.
"""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from prender_amd import generators as G
from prender_amd.scene import RAY_DTYPE

pytestmark = pytest.mark.gpu


def crn_close(g, o, frac=0.98):
    """Common random numbers: nearly every pixel equal to float rounding (a
    sample whose double-precision path differs in its last bits can flip a
    Russian-roulette or BSDF choice), and the image means agree."""
    close = np.abs(g - o) <= 1e-4 * np.abs(o) + 1e-6
    assert close.mean() >= frac, close.mean()
    assert abs(float(g.mean()) - float(o.mean())) <= 1e-2 * max(float(o.mean()), 1e-6)



@pytest.fixture(scope="module")
def prt():
    from prender_amd import prt as P
    info = P.device_info()
    print("device", info, flush=True)
    assert info["arch"].startswith("gfx950")
    return P


def test_full_size_image_is_independent_of_batching(prt):
    import torch
    d = G.config2()  # 512x512, 64 spp, 81,944 triangles, NEE
    s = prt.Scene(d)
    H, W = d.camera.height, d.camera.width
    a = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    c = torch.empty_like(a)
    print("cfg2 512x512x64spp: default batches, 3001-pixel batches, repeat", flush=True)
    sa = s.render_device(a.data_ptr(), None, seed=21)
    sb = s.render_device(b.data_ptr(), None, seed=21, batch_pixels=3001)  # ragged last batch
    s.render_device(c.data_ptr(), None, seed=21)
    torch.cuda.synchronize()
    assert torch.equal(a, b) and torch.equal(a, c)
    assert sa["camera_samples"] == sb["camera_samples"] == W * H * 64
    assert sa["shadow_rays"] == sb["shadow_rays"] and sa["extension_rays"] == sb["extension_rays"]
    assert float(a.mean()) > 0.0


@pytest.mark.parametrize("W,H", [(1, 1), (3, 5)])
def test_tiny_and_ragged_images(prt, W, H):
    d = G.config2(width=W, height=H, samples=4, level=3)
    print(f"cfg2 {W}x{H} CRN", flush=True)
    rg, _, st = prt.Scene(d).render(seed=5)
    ro, _, _ = OracleScene(d).render(seed=5)
    assert rg.shape == (H, W, 3) and st["camera_samples"] == W * H * 16
    crn_close(rg, ro, frac=0.9)


def test_empty_ray_batch(prt):
    s = prt.Scene(G.config1(width=8, height=8, samples=1))
    hits, st = s.trace(np.zeros(0, dtype=RAY_DTYPE))
    assert hits.shape == (0,) and st["primary_rays"] == 0


def test_nee_without_emitters(prt):
    # NEE on, no light entity: no shadow rays, the image is what the oracle
    # gives (here: black walls under a black background)
    d = G.config1(width=16, height=12, samples=2, nee=1)
    d.entities = d.entities[:-1]  # drop the light sphere
    rg, _, st = prt.Scene(d).render(seed=3)
    ro, _, _ = OracleScene(d).render(seed=3)
    assert st["shadow_rays"] == 0
    np.testing.assert_allclose(rg, ro, atol=1e-6)


def test_vertex_cap_truncates_like_the_oracle(prt):
    # a cap of 2 vertices cuts every path after its second hit; both sides
    # count the cut paths the same way on common random numbers
    d = G.config1(width=16, height=12, samples=2)
    rg, _, st = prt.Scene(d).render(seed=8, max_depth=2)
    o = OracleScene(d)
    o.set_max_vertices(2)
    ro, _, so = o.render(seed=8)
    assert st["truncated_paths"] > 0
    assert abs(st["truncated_paths"] - so["truncated"]) <= 0.01 * so["truncated"]
    crn_close(rg, ro)


def test_bad_descriptions_fail_loudly(prt):
    d = G.config1(width=8, height=8, samples=1)
    d.entities[0].material = 99  # no such material
    with pytest.raises(prt.PrtError):
        prt.Scene(d)
    d = G.config1(width=8, height=8, samples=1)
    d.camera.width = 0
    with pytest.raises(prt.PrtError):
        prt.Scene(d)


@pytest.mark.parametrize("name", ["cfg1", "cfg2", "cfg3"])
def test_engines_give_identical_images(prt, name):
    # depth-synchronous batches (engine 1), the streaming tile queue (engine
    # 2; the default for scenes with glass and NEE) and one persistent launch
    # (engine 3; the default for scenes without NEE) are three schedules of
    # the same per-sample computation
    d = {"cfg1": lambda: G.config1(width=64, height=48, samples=4),
         "cfg2": lambda: G.config2(width=48, height=48, samples=4, level=4),
         "cfg3": lambda: G.config3(width=64, height=48, samples=2)}[name]()
    s = prt.Scene(d)
    print(f"{name} engines 1 / 2 / 3 / auto", flush=True)
    a, va, sa = s.render(seed=13, engine=1)
    b, vb, sb = s.render(seed=13, engine=2)
    p, vp, sp = s.render(seed=13, engine=3)
    c, vc, _ = s.render(seed=13)
    assert np.array_equal(a, b) and np.array_equal(a, c) and np.array_equal(a, p)
    assert np.array_equal(va, vb) and np.array_equal(va, vc) and np.array_equal(va, vp)
    assert sa["camera_samples"] == sb["camera_samples"] == sp["camera_samples"]
    assert sa["shadow_rays"] == sb["shadow_rays"] == sp["shadow_rays"]
    assert sp["launches_extend"] == 1
    for engine in (2, 3):
        with pytest.raises(prt.PrtError):
            s.render(seed=13, engine=engine, samples=1)


@pytest.mark.parametrize("make", [lambda: G.config2(width=64, height=48, samples=2),
                                  lambda: G.config1(width=64, height=48, samples=2)], ids=["cfg2", "cfg1_glass"])
def test_progress_dumps_between_batches(prt, make, tmp_path):
    # Render::Run's periodic OutputBmp (render.h:1390-1401) as prt_render_dump:
    # with a 0-second interval a dump follows every batch but the last; each
    # holds the finished pixels (those of the batches done, in the order the
    # batches cover the reference's rows y = 0, 1, ...) and zeros elsewhere,
    # and the final image equals a render without dumps bit for bit
    d = make()
    s = prt.Scene(d)
    W, H = d.camera.width, d.camera.height
    bp = 512
    n_batches = -(-W * H // bp)
    seen = []
    rgb, var, st = s.render_dump(0.0, callback=lambda img, i, f: seen.append((img, i, f)),
                                 directory=str(tmp_path), seed=8, batch_pixels=bp)
    ref, ref_var, _ = s.render(seed=8)
    assert np.array_equal(rgb, ref) and np.array_equal(var, ref_var)
    assert [i for _, i, _ in seen] == list(range(n_batches - 1))
    flat_ref = ref[::-1].reshape(-1, 3)          # pixel y*W + x of the reference's numbering
    for img, i, f in seen:
        done = (i + 1) * bp
        assert f == pytest.approx(done / (W * H))
        flat = img[::-1].reshape(-1, 3)
        assert np.array_equal(flat[:done], flat_ref[:done])
        assert not flat[done:].any()
        name = tmp_path / f"output_W{i:06d}.bmp"
        again = tmp_path / f"again_{i}.bmp"
        prt.write_bmp(str(again), img)
        assert name.read_bytes() == again.read_bytes()
    # a negative interval never dumps
    seen.clear()
    rgb2, _, _ = s.render_dump(-1.0, callback=lambda img, i, f: seen.append(i), seed=8, batch_pixels=bp)
    assert not seen and np.array_equal(rgb2, ref)
