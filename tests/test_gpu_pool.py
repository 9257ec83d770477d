"""Scene handles on one device share one pool of render memory.

Eight handles of a 256x256x64-spp NEE scene (4.2 M camera samples; its
queues, shadow queues and accumulators take about 3 GB) rendered one after
another must add at most about one such workspace of device memory (none if
an earlier, larger call already grew the pool), where one workspace per
handle would add eight; and a handle's image must not depend on what the
other handles rendered in between (same seed, same bits)."""
import copy

import pytest

from prender_amd import generators as G

pytestmark = pytest.mark.gpu


def test_handles_share_the_device_pool():
    import torch
    from prender_amd import prt
    d = copy.deepcopy(G.config2())
    d.camera.width = d.camera.height = 256
    scenes = [prt.Scene(d) for _ in range(8)]
    outs = [torch.empty((256, 256, 3), dtype=torch.float32, device="cuda") for _ in scenes]
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for i, (sc, o) in enumerate(zip(scenes, outs)):
        sc.render_device(o.data_ptr(), seed=5 + i)
    torch.cuda.synchronize()
    first = outs[0].clone()
    free1, _ = torch.cuda.mem_get_info()
    grown = (free0 - free1) / 1e9
    print("eight handles took %.2f GB of device memory" % grown, flush=True)
    assert grown < 6.0, "render memory grows with the number of handles"
    # the first handle again, after the others used the pool: same bits
    scenes[0].render_device(outs[0].data_ptr(), seed=5)
    torch.cuda.synchronize()
    assert torch.equal(outs[0], first)
    assert float(first.mean()) > 0
    for sc in scenes:
        sc.close()
