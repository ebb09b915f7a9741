"""rt_hip_group_spots_noalloc: the spots entry point that allocates nothing.

Its partial moments live in a static device array instead of a buffer the
library allocates; the kernels and their summation order are those of
rt_hip_group_spots, so the bar (written here) is equality bit for bit with
rt_hip_group_spots on the same rays, on two calls in a row (the second on a
smaller input, so stale partials of the first call would show), plus the
size limit: an input whose partial moments exceed the array is rejected
with RT_ERR_BAD_VALUE before any device work.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def random_rays(n, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    rays = torch.zeros((7, n), dtype=torch.float64)
    rays[0] = torch.randn(n, generator=g, dtype=torch.float64) * 3.0 + 1.5
    rays[1] = torch.randn(n, generator=g, dtype=torch.float64) * 0.7 - 2.0
    flag = (torch.rand(n, generator=g) < 0.8).to(torch.int8)
    return rays.cuda(), flag.cuda()


def host_spots(rays, flag, G):
    x, y, f = rays[0].cpu().numpy(), rays[1].cpu().numpy(), flag.cpu().numpy()
    out = np.zeros((G, 6))
    for g in range(G):
        m = f[g::G] == 1
        xs, ys = x[g::G][m], y[g::G][m]
        if xs.size:
            out[g] = [xs.size, xs.mean(), ys.mean(), xs.var(), ys.var(), np.sqrt(xs.var() + ys.var())]
    return out


@pytest.mark.parametrize("calls", [((1_000_000, 64, 11), (30_000, 13, 12))])
def test_noalloc_spots_equal_allocating_entry_point(hip, calls):
    for n, G, seed in calls:
        print(f"[noalloc spots] n={n} groups={G} seed={seed}", flush=True)
        rays, flag = random_rays(n, seed)
        ref = hip.group_spots(rays, flag, G)
        # poison the output so an unwritten group would show
        out = torch.full((G, hip.SPOT_STRIDE), float("nan"), dtype=torch.float64, device="cuda")
        hip.group_spots(rays, flag, G, out=out, noalloc=True)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), f"n={n} G={G}: spots differ from rt_hip_group_spots"
        h = host_spots(rays, flag, G)
        got = out.cpu().numpy()
        assert np.array_equal(got[:, 0], h[:, 0])
        assert np.allclose(got[:, 1:], h[:, 1:], rtol=1e-12, atol=1e-12)


def test_noalloc_spots_reject_oversize(hip):
    print("[noalloc spots] oversize rejection", flush=True)
    n, G = 3_000_000, 64
    rays, flag = random_rays(n, 5)
    out = torch.zeros((G, hip.SPOT_STRIDE), dtype=torch.float64, device="cuda")
    with pytest.raises(hip.EngineError, match="no-allocation"):
        hip.group_spots(rays, flag, G, out=out, noalloc=True)
    # the rejected call left the output alone and the next call still works
    torch.cuda.synchronize()
    assert torch.count_nonzero(out).item() == 0
    small, sf = random_rays(4096, 6)
    ref = hip.group_spots(small, sf, 8)
    got = hip.group_spots(small, sf, 8, noalloc=True)
    torch.cuda.synchronize()
    assert torch.equal(got, ref)
