"""The allocation-free entry point (bcrt_run_device_noalloc, residue planes in
the library's fixed device scratch array) against bcrt_run_device, bit for
bit, through the C ABI: two calls in a row, the second with a smaller input,
then a third at N=65,536 and a batch above the scratch size refused."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def batch(pkg, N, W, q, B, seed):
    dg = pkg.datagen
    a = np.stack([dg.operands(N, q, seed + i, edge_fraction=0.1) for i in range(B)])
    b = np.stack([dg.operands(N, q, seed + 100 + i, edges_first=(i == 0)) for i in range(B)])
    L = a.shape[-1]
    to = lambda x: torch.from_numpy(x.view(np.int32).copy()).cuda().view(B, N, L)
    return to(a), to(b)


def test_noalloc_matches_run_device(pkg):
    dg = pkg.datagen
    N, W = 32768, 868
    q = dg.modulus(W, 4242)
    e = pkg.Engine(N, W, q, max_batch=4)
    # first call: four products (the whole scratch array at this geometry)
    a, b = batch(pkg, N, W, q, 4, 10)
    want = torch.zeros_like(a)
    e.run_device(a, b, want, 4)
    got = torch.full_like(a, -1)
    e.run_device_noalloc(a, b, got, 4)
    torch.cuda.synchronize()
    assert torch.equal(got, want)
    # second call, smaller: two other products; the scratch holds the first call's planes
    a2, b2 = batch(pkg, N, W, q, 2, 50)
    want2 = torch.zeros_like(a2)
    e.run_device(a2, b2, want2, 2)
    got2 = torch.full_like(a2, -1)
    e.run_device_noalloc(a2, b2, got2, 2)
    torch.cuda.synchronize()
    assert torch.equal(got2, want2)
    # another geometry on the same scratch: one product at N=65,536 (planes 59.8 MB)
    N3, W3 = 65536, 1747
    q3 = dg.modulus(W3, 4343)
    e3 = pkg.Engine(N3, W3, q3, max_batch=1)
    a3, b3 = batch(pkg, N3, W3, q3, 1, 90)
    want3 = torch.zeros_like(a3)
    e3.run_device(a3, b3, want3, 1)
    got3 = torch.full_like(a3, -1)
    e3.run_device_noalloc(a3, b3, got3, 1, stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    assert torch.equal(got3, want3)
    # two products at N=65,536 would need 119.5 MB of planes: refused before any launch
    with pytest.raises(ValueError, match="scratch"):
        e3.run_device_noalloc(a3, b3, got3, 2)
    with pytest.raises(ValueError):
        e3.run_device_noalloc(a3, b3, got3, 0)
