"""Comparison criteria shared by the GPU parity tests (see test_gpu_parity.py
for what each one asserts and why)."""
import math

import numpy as np


def assert_hits_match(hg, ho):
    assert (hg["entity"] == ho["entity"]).all(), np.flatnonzero(hg["entity"] != ho["entity"])[:10]
    hit = ho["entity"] >= 0
    assert (hg["t"][hit] == ho["t"][hit]).all()
    diff_prim = hg["prim"] != ho["prim"]
    # a different triangle only on an exact tie of t
    assert (hg["t"][diff_prim] == ho["t"][diff_prim]).all()
    same = ~diff_prim & hit & (ho["prim"] >= 0)
    assert (hg["u"][same] == ho["u"][same]).all() and (hg["v"][same] == ho["v"][same]).all()
    return int(diff_prim.sum())


def assert_crn(rg, vg, ro, vo, label, rel=1e-5):
    """Common random numbers: >= 99.9 % of channels within float32 rounding,
    the rest within 6 standard errors of the pixel mean."""
    tol = rel * np.abs(ro) + 1e-9
    err = np.abs(rg - ro)
    close = err <= tol
    frac = float(close.mean())
    se = np.sqrt(vg + vo)
    far = ~close
    worst = float((err[far] / (se[far] + 1e-12)).max()) if far.any() else 0.0
    print(f"{label} CRN: {frac * 100:.3f}% of {close.size} channels within {rel:g} relative; "
          f"{int(far.sum())} others, worst {worst:.2f} standard errors; max |diff| {float(err.max()):.3g}; "
          f"means gpu {float(rg.mean()):.6f} cpu {float(ro.mean()):.6f}", flush=True)
    assert frac >= 0.999, frac
    assert worst <= 6.0, worst


def assert_independent(a, va, b, vb, label, tile=8):
    """Independent seeds: per-channel 3-sigma exceedance, clustering, tile
    means and the image mean."""
    s = np.sqrt(va + vb) + 1e-6 * np.abs(b) + 1e-12
    ex = np.abs(a - b) > 3 * s
    H, W, _ = a.shape
    max_tile, max_tz = 0, 0.0
    for i in range(0, H - tile + 1, tile):
        for j in range(0, W - tile + 1, tile):
            sl = (slice(i, i + tile), slice(j, j + tile))
            max_tile = max(max_tile, int(ex[sl].sum()))
            n = tile * tile
            se_t = np.sqrt(va[sl].mean((0, 1)) / n + vb[sl].mean((0, 1)) / n) + 1e-12
            max_tz = max(max_tz, float((np.abs(a[sl].mean((0, 1)) - b[sl].mean((0, 1))) / se_t).max()))
    gz = abs(float(a.mean() - b.mean())) / math.sqrt(float(va.mean() + vb.mean()) / a[..., 0].size)
    frac = float(ex.mean())
    print(f"{label} independent seeds: {frac * 100:.3f}% beyond 3 sigma, at most {max_tile} per {tile}x{tile} tile, "
          f"max tile |z| {max_tz:.2f}, image |z| {gz:.2f}", flush=True)
    assert frac <= 0.006, frac
    assert max_tile <= 6, max_tile
    assert max_tz < 4.5, max_tz
    assert gz < 4.0, gz
