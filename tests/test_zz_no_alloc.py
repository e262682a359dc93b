"""fmm3dpp_direct_device_noalloc (the direct lattice sum with its partial sums in the library's static
device array) against fmm3dpp_direct_device (the same kernels with a fresh allocation), through the
C-ABI: two calls in a row, the second with fewer bodies and targets, each bit-identical to the
allocating entry point, and targets beyond the array's limit refused with an argument error."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(torch, n, ntar, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    pts = rng.random((n, 3)) - 0.5
    q = rng.uniform(-1, 1, n)
    q -= q.mean()
    tar = np.sort(rng.choice(n, ntar, replace=False)).astype(np.int32)
    dev = torch.device("cuda", 0)
    cols = [torch.tensor(pts[:, k].copy(), dtype=torch.float64, device=dev) for k in range(3)]
    return cols, torch.tensor(q, device=dev), torch.tensor(tar, device=dev)


def _call(fn, torch, cols, dq, dt, n, ntar, images):
    dev = torch.device("cuda", 0)
    phi = torch.full((ntar,), float("nan"), dtype=torch.float64, device=dev)
    grad = torch.full((ntar, 3), float("nan"), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    vp = ctypes.c_void_p
    rc = fn(vp(0), n, vp(cols[0].data_ptr()), vp(cols[1].data_ptr()), vp(cols[2].data_ptr()), vp(dq.data_ptr()),
            ntar, vp(dt.data_ptr()), 1.0, images, vp(phi.data_ptr()), vp(grad.data_ptr()))
    return rc, phi.cpu().numpy(), grad.cpu().numpy()


def test_noalloc_direct_equals_allocating_entry_point():
    import torch
    import pyfmm3dpp
    L = pyfmm3dpp.lib()
    for n, ntar, images, seed in ((20000, 3000, 1, 1), (5000, 700, 0, 2)):
        print(f"case n={n} ntar={ntar} images={images}", flush=True)
        cols, dq, dt = _inputs(torch, n, ntar, seed)
        rc0, p0, g0 = _call(L.fmm3dpp_direct_device, torch, cols, dq, dt, n, ntar, images)
        rc1, p1, g1 = _call(L.fmm3dpp_direct_device_noalloc, torch, cols, dq, dt, n, ntar, images)
        assert rc0 == pyfmm3dpp.OK and rc1 == pyfmm3dpp.OK, (rc0, rc1)
        assert np.all(np.isfinite(p1)) and np.all(np.isfinite(g1))
        assert np.array_equal(p0, p1) and np.array_equal(g0, g1)


def test_noalloc_direct_rejects_too_many_targets():
    import torch
    import pyfmm3dpp
    L = pyfmm3dpp.lib()
    n = 600000
    dev = torch.device("cuda", 0)
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    t = torch.arange(n, dtype=torch.int32, device=dev)
    out = torch.zeros(3 * n, dtype=torch.float64, device=dev)
    vp = ctypes.c_void_p
    rc = L.fmm3dpp_direct_device_noalloc(vp(0), n, vp(x.data_ptr()), vp(x.data_ptr()), vp(x.data_ptr()),
                                         vp(x.data_ptr()), n, vp(t.data_ptr()), 1.0, 0, vp(out.data_ptr()),
                                         vp(out.data_ptr()))
    assert rc == pyfmm3dpp.ERR_ARG
