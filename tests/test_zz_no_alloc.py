"""GPU: the allocation-free direct sum (fss_velocity_direct_noalloc, through
the C ABI) equals fss_velocity_direct bit for bit, on two calls in a row
whose second is smaller (the shared scratch keeps the first call's records
beyond the second's n), in both precisions; sizes beyond the scratch are
refused before any launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(fss, n, seed):
    import torch
    x, y, z = fss.grid_random(n, seed)
    a = np.full(n, 4 * np.pi / n)
    v = fss.gauss_vortices(x, y, z, a, 4, 2.0, 0.05, 0.3, seed=seed)
    return [torch.from_numpy(t).cuda() for t in (x, y, z, v, a)]


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_direct_equals_direct(fsslib, precision):
    import torch
    ctx = fsslib.Context(precision=precision)
    for n, seed in ((40000, 3), (9000, 4)):
        print(f"noalloc direct precision={precision} n={n}", flush=True)
        d = _case(fsslib, n, seed)
        got = [t.cpu().numpy() for t in ctx.velocity_direct_noalloc(*d)]
        torch.cuda.synchronize()
        ref = [t.cpu().numpy() for t in ctx.velocity_direct(*d)]
        for g, r in zip(got, ref):
            assert np.isfinite(g).all()
            np.testing.assert_array_equal(g, r)


def test_noalloc_rejects_sizes_beyond_scratch(fsslib):
    import torch
    limit = {0: (32 << 20) // 32, 1: (32 << 20) // 16}
    for precision, nmax in limit.items():
        ctx = fsslib.Context(precision=precision)
        n = nmax + 1
        # allocated but never read: the call must refuse before launching
        d = [torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(5)]
        with pytest.raises(fsslib.FssError) as e:
            ctx.velocity_direct_noalloc(*d)
        assert e.value.status == 1 and str(nmax) in str(e.value)
