"""Path-replay adjoint on the GPU against the oracle's CPU adjoint (common
random numbers) and against central finite differences of GPU renders
(deterministic: sample radiance is summed in fixed point)."""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from prender_amd import generators as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_prt():
    import torch
    from prender_amd import prt
    assert prt.device_info()["arch"].startswith("gfx950")
    return torch, prt


def _gpu_loss(torch, sc, dimg_t, seed, samples=None):
    # double-precision image: a single face moves a pixel by far less than a
    # float32 ulp of its value
    img = torch.empty((sc.height, sc.width, 3), dtype=torch.float64, device="cuda")
    sc.render_device(img.data_ptr(), None, seed=seed, samples=samples, image_f64=True)
    torch.cuda.synchronize()
    return float((img * dimg_t.double()).sum().item())


def test_gpu_adjoint_matches_cpu_adjoint(torch_prt):
    torch, prt = torch_prt
    d = G.config5(width=32, height=32, samples=2)
    dimg = np.random.default_rng(1).uniform(-1, 1, (32, 32, 3)).astype(np.float32)
    sc = prt.Scene(d)
    dimg_t = torch.from_numpy(dimg).cuda()
    nf = d.entities[1].indices.shape[0]
    g_t = torch.zeros((nf, 3), dtype=torch.float32, device="cuda")
    print("gpu adjoint cfg5 32x32x8spp", flush=True)
    ge, st = sc.adjoint_device(dimg_t.data_ptr(), 1, g_t.data_ptr(), 2, seed=5)
    g_gpu = g_t.cpu().numpy().astype(np.float64)
    g_cpu, ge_cpu = OracleScene(d).adjoint(seed=5, dimage=dimg, entity=1, light=2)
    # same random numbers: the same paths, so the gradients agree up to float rounding
    err = np.abs(g_gpu - g_cpu).sum() / np.abs(g_cpu).sum()
    print("relative L1 error albedo", err, "emission", ge, ge_cpu, "stats", st)
    assert err < 1e-3
    np.testing.assert_allclose(ge, ge_cpu, rtol=1e-3)


@pytest.mark.parametrize("russian_roulette", [False, True])
def test_gpu_adjoint_matches_finite_differences(torch_prt, russian_roulette):
    torch, prt = torch_prt
    d = G.config5(width=64, height=64, samples=8)
    if not russian_roulette:
        # DEPTH beyond the vertex cap: no Russian roulette, so no decision can
        # flip between the +eps and -eps renders
        d.sampling.depth = d.sampling.depth_limit = 250
    sc = prt.Scene(d)
    dimg = np.random.default_rng(2).uniform(-1, 1, (64, 64, 3)).astype(np.float32)
    dimg_t = torch.from_numpy(dimg).cuda()
    nf = d.entities[1].indices.shape[0]
    g_t = torch.zeros((nf, 3), dtype=torch.float32, device="cuda")
    seed = 11
    ge, _ = sc.adjoint_device(dimg_t.data_ptr(), 1, g_t.data_ptr(), 2, seed=seed)
    g = g_t.cpu().numpy().astype(np.float64)
    # 256 random parameters among those a difference quotient can resolve:
    # deep bounces give almost every face a gradient of ~1e-20 of the
    # largest, far below what a change of 1e-3 moves the image by
    rng = np.random.default_rng(3)
    fc = np.argwhere(np.abs(g) >= 1e-4 * np.abs(g).max())
    pick = fc[rng.choice(len(fc), 256, replace=False)]
    faces, chans = pick[:, 0], pick[:, 1]
    alb0 = torch.from_numpy(d.entities[1].face_albedo.copy()).cuda()
    eps = 1e-3
    rel = []
    if russian_roulette:
        # the adjoint holds RR probabilities at their primal values; the
        # difference quotient does the same, so no RR decision flips
        sc.hold_rr_albedo(True)
    for f, c in zip(faces, chans):
        a = alb0.clone(); a[f, c] += eps
        sc.set_face_albedo_device(1, a.data_ptr()); gp = _gpu_loss(torch, sc, dimg_t, seed)
        a = alb0.clone(); a[f, c] -= eps
        sc.set_face_albedo_device(1, a.data_ptr()); gm = _gpu_loss(torch, sc, dimg_t, seed)
        # the float32 step actually taken
        step = float(np.float32(alb0[f, c].item() + eps)) - float(np.float32(alb0[f, c].item() - eps))
        fd = (gp - gm) / step
        rel.append(abs(g[f, c] - fd) / max(abs(fd), 1e-12))
    sc.set_face_albedo_device(1, alb0.data_ptr())
    sc.hold_rr_albedo(False)
    rel = np.array(rel)
    print("FD check (RR %s): median rel err" % russian_roulette, np.median(rel), "p90", np.quantile(rel, 0.9),
          "max", rel.max(), "within 2%:", np.mean(rel < 0.02))
    assert np.mean(rel < 0.02) >= 0.99
    e0 = d.materials[2].emission[0]
    sc.set_emission(2, (e0 + eps,) * 3); gp = _gpu_loss(torch, sc, dimg_t, seed)
    sc.set_emission(2, (e0 - eps,) * 3); gm = _gpu_loss(torch, sc, dimg_t, seed)
    sc.set_emission(2, (e0,) * 3)
    fd_e = (gp - gm) / (2 * eps)
    print("emission: adjoint", ge.sum(), "fd", fd_e)
    assert abs(ge.sum() - fd_e) < 0.02 * abs(fd_e)


def test_gpu_adjoint_finite_differences_uniform_parameters(torch_prt):
    # 256 parameters drawn uniformly from all 3 * 81,920 face albedos (no
    # selection by gradient).  Faces reached only by deep bounces have
    # gradients many orders below the largest; a 1e-3 step moves the
    # fixed-point image of such a face by less than its rounding, so the
    # difference quotient cannot resolve them.  Resolvable parameters
    # (|g| >= 1e-4 max |g|) must agree within 2 %; for the others adjoint and
    # difference quotient must both be below that floor.
    torch, prt = torch_prt
    d = G.config5(width=64, height=64, samples=8)
    d.sampling.depth = d.sampling.depth_limit = 250      # no Russian roulette to flip
    sc = prt.Scene(d)
    dimg = np.random.default_rng(4).uniform(-1, 1, (64, 64, 3)).astype(np.float32)
    dimg_t = torch.from_numpy(dimg).cuda()
    nf = d.entities[1].indices.shape[0]
    g_t = torch.zeros((nf, 3), dtype=torch.float32, device="cuda")
    seed = 13
    sc.adjoint_device(dimg_t.data_ptr(), 1, g_t.data_ptr(), 2, seed=seed)
    g = g_t.cpu().numpy().astype(np.float64)
    rng = np.random.default_rng(6)
    flat = rng.choice(nf * 3, 256, replace=False)
    faces, chans = flat // 3, flat % 3
    alb0 = torch.from_numpy(d.entities[1].face_albedo.copy()).cuda()
    eps = 1e-3
    floor = 1e-4 * np.abs(g).max()
    res_rel, low = [], []
    for f, c in zip(faces, chans):
        a = alb0.clone(); a[f, c] += eps
        sc.set_face_albedo_device(1, a.data_ptr()); gp = _gpu_loss(torch, sc, dimg_t, seed)
        a = alb0.clone(); a[f, c] -= eps
        sc.set_face_albedo_device(1, a.data_ptr()); gm = _gpu_loss(torch, sc, dimg_t, seed)
        step = float(np.float32(alb0[f, c].item() + eps)) - float(np.float32(alb0[f, c].item() - eps))
        fd = (gp - gm) / step
        if abs(g[f, c]) >= floor:
            res_rel.append(abs(g[f, c] - fd) / max(abs(fd), 1e-12))
        else:
            low.append((g[f, c], fd))
    sc.set_face_albedo_device(1, alb0.data_ptr())
    res_rel = np.array(res_rel)
    low = np.array(low).reshape(-1, 2)
    print(f"uniform FD check: {len(res_rel)} of 256 resolvable, within 2%: "
          f"{np.mean(res_rel < 0.02) if len(res_rel) else float('nan'):.3f}, median rel err "
          f"{np.median(res_rel) if len(res_rel) else float('nan'):.2e}; {len(low)} below the floor {floor:.2e}: "
          f"max |adjoint| {np.abs(low[:, 0]).max() if len(low) else 0:.2e}, "
          f"max |fd| {np.abs(low[:, 1]).max() if len(low) else 0:.2e}", flush=True)
    assert len(res_rel) >= 16
    assert np.mean(res_rel < 0.02) >= 0.99
    if len(low):
        assert np.abs(low[:, 1]).max() < 2 * floor


def test_autograd_function_and_adam_decrease_loss(torch_prt):
    torch, prt = torch_prt
    from prender_amd.autograd import render
    d_t = G.config5(width=64, height=64, samples=8, target=True)
    sc_t = prt.Scene(d_t)
    target = torch.empty((64, 64, 3), dtype=torch.float32, device="cuda")
    sc_t.render_device(target.data_ptr(), None, seed=100, samples=64)
    d = G.config5(width=64, height=64, samples=4)
    sc = prt.Scene(d)
    alb = torch.full((d.entities[1].indices.shape[0], 3), 0.5, device="cuda", requires_grad=True)
    em = torch.tensor(60.0, requires_grad=True)
    opt = torch.optim.Adam([alb], lr=0.02)
    losses = []
    for it in range(30):
        opt.zero_grad()
        img = render(alb, em, sc, 1, 2, seed=2 * it, adjoint_seed=2 * it + 1)
        loss = ((img - target) ** 2).mean()
        loss.backward()
        opt.step()
        with torch.no_grad():
            alb.clamp_(0.0, 1.0)
        losses.append(float(loss))
    print("losses", losses[0], losses[-1])
    assert alb.grad is not None and torch.isfinite(alb.grad).all()
    assert np.mean(losses[-5:]) < 0.8 * np.mean(losses[:5])
