"""prt_trace_closest_device_noalloc against prt_trace_closest_device.

The allocation-free entry point runs the same traversal with its counters in
the library's per-GPU scratch array; its hits must equal the allocating entry
point's byte for byte and its counters must equal that call's counters, on
two calls in a row (the second smaller, so a counter the fill kernel failed
to clear would show), through the C ABI.
"""
import numpy as np
import pytest

from prender_amd import generators as G
from prender_amd.scene import HIT_DTYPE

pytestmark = pytest.mark.gpu


def _device(torch, a: np.ndarray):
    t = torch.empty(a.nbytes, dtype=torch.uint8, device="cuda")
    t.copy_(torch.from_numpy(a.view(np.uint8).reshape(-1)))
    return t


def _hits(torch, buf, n):
    return buf.cpu().numpy().view(HIT_DTYPE)[:n].copy()


def test_noalloc_trace_matches_allocating_trace():
    import torch
    from prender_amd import prt
    d = G.config2(width=64, height=64, samples=1, level=5)
    sc = prt.Scene(d)
    rays = G.hit_parity_rays(d, n=1 << 16)
    for n in (1 << 16, 3001):                       # the second call is smaller
        print(f"noalloc trace: cfg2 L5 scene, {n} rays (seed 0x5EED)", flush=True)
        r = _device(torch, np.ascontiguousarray(rays[:n]))
        h_ref = torch.zeros(n * HIT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
        h_new = torch.full_like(h_ref, 0xA5)
        st_ref = sc.trace_device(r.data_ptr(), n, h_ref.data_ptr())
        st_new = sc.trace_device_noalloc(r.data_ptr(), n, h_new.data_ptr())
        torch.cuda.synchronize()
        assert torch.equal(h_ref, h_new)             # every byte of every hit
        hits = _hits(torch, h_new, n)
        assert (hits["entity"] >= 0).mean() > 0.9
        for k in ("node_visits", "tri_tests", "extension_rays", "bytes_algorithmic"):
            assert st_new[k] == st_ref[k], (k, st_new[k], st_ref[k])
        assert st_new["node_visits"] > 0
    # stats are optional, and an empty batch is a valid call
    r = _device(torch, np.ascontiguousarray(rays[:512]))
    h = torch.zeros(512 * HIT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    assert sc.trace_device_noalloc(r.data_ptr(), 512, h.data_ptr(), stats=False) is None
    st0 = sc.trace_device_noalloc(r.data_ptr(), 0, h.data_ptr())
    assert st0["node_visits"] == 0 and st0["extension_rays"] == 0
