"""fmm3dpp_solve_device (tree + lists + evaluation in one call, the near-field list build on a second
stream beside the far field) against set_particles_device followed by evaluate_device: the same
kernels on the same data, so the results are identical (the bound below is 0)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(n, seed, clustered):
    rng = np.random.Generator(np.random.PCG64(seed))
    if clustered:
        x = np.concatenate([rng.normal(0.1, 0.05, (n // 2, 3)), rng.random((n - n // 2, 3)) - 0.5])
        x = (x + 0.5) % 1.0 - 0.5
    else:
        x = rng.random((n, 3)) - 0.5
    q = rng.uniform(-1, 1, n)
    q -= q.mean()
    return x, q


@pytest.mark.parametrize("precision", [0, 1])
def test_solve_device_equals_two_calls(precision):
    import torch
    import pyfmm3dpp
    dev = torch.device("cuda", 0)
    plan = pyfmm3dpp.LaplaceFMM(6, 48, 2, 1.0, dipole_mode=pyfmm3dpp.DIPOLE_PHYSICAL, precision=precision, device=0)
    ref = pyfmm3dpp.LaplaceFMM(6, 48, 2, 1.0, dipole_mode=pyfmm3dpp.DIPOLE_PHYSICAL, precision=precision, device=0)
    for n, seed, clustered in ((40000, 1, False), (25000, 2, True), (40000, 3, False)):   # reuse, shrink, regrow
        print(f"case n={n} clustered={clustered} precision={precision}", flush=True)
        x, q = _case(n, seed, clustered)
        d = [torch.from_numpy(np.ascontiguousarray(x[:, k])).to(dev) for k in range(3)]
        dq = torch.from_numpy(q).to(dev)
        phi, grad = torch.empty(n, dtype=torch.float64, device=dev), torch.empty((n, 3), dtype=torch.float64, device=dev)
        phi2, grad2 = torch.empty_like(phi), torch.empty_like(grad)
        torch.cuda.synchronize()
        plan.solve_device(n, d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), dq.data_ptr(), phi.data_ptr(),
                          grad.data_ptr())
        ref.set_particles_device(n, d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr())
        ref.evaluate_device(dq.data_ptr(), phi2.data_ptr(), grad2.data_ptr())
        assert plan.timing_ms("evaluate") > 0 and ref.timing_ms("evaluate") > 0   # waits for both streams
        torch.cuda.synchronize()
        dp = float((phi - phi2).abs().max())
        dg = float((grad - grad2).abs().max())
        print(f"  max |diff| phi {dp:.3e} grad {dg:.3e}", flush=True)
        assert dp == 0.0 and dg == 0.0
        assert plan.counts() == ref.counts()


@pytest.mark.parametrize("precision", [0, 1])
def test_near_field_stream_order(precision):
    """The P2P runs on a second stream of the plan, forked from and joined back to the plan's
    stream.  Two evaluations queued back to back (no host sync between them) on a caller's stream
    equal the same evaluations run one at a time on a plan's own stream, bit for bit."""
    import torch
    import pyfmm3dpp
    dev = torch.device("cuda", 0)
    n = 30000
    x, q1 = _case(n, 11, True)
    q2 = np.random.Generator(np.random.PCG64(12)).uniform(-1, 1, n)
    d = [torch.from_numpy(np.ascontiguousarray(x[:, k])).to(dev) for k in range(3)]
    dq = [torch.from_numpy(q).to(dev) for q in (q1, q2)]
    out = [(torch.empty(n, dtype=torch.float64, device=dev), torch.empty((n, 3), dtype=torch.float64, device=dev))
           for _ in range(4)]
    mk = lambda: pyfmm3dpp.LaplaceFMM(5, 40, 2, 1.0, precision=precision, device=0)
    a, b = mk(), mk()
    stream = torch.cuda.Stream(device=dev)
    b.set_stream(stream.cuda_stream)
    torch.cuda.synchronize()
    for p in (a, b):
        p.set_particles_device(n, d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr())
    for k in range(2):   # one at a time on the plan's own stream
        a.evaluate_device(dq[k].data_ptr(), out[k][0].data_ptr(), out[k][1].data_ptr())
        torch.cuda.synchronize()
    for k in range(2):   # back to back on the caller's stream
        b.evaluate_device(dq[k].data_ptr(), out[2 + k][0].data_ptr(), out[2 + k][1].data_ptr())
    stream.synchronize()
    for k in range(2):
        dp = float((out[k][0] - out[2 + k][0]).abs().max())
        dg = float((out[k][1] - out[2 + k][1]).abs().max())
        print(f"evaluation {k}: max |diff| phi {dp:.3e} grad {dg:.3e}", flush=True)
        assert dp == 0.0 and dg == 0.0
    assert float(out[0][0].abs().max()) > 0 and not torch.equal(out[0][0], out[1][0])


@pytest.mark.parametrize("precision", [0, 1])
def test_near_field_stream_switch(precision):
    """fmm3dpp_set_near_field_stream(plan, 0) runs the near field on the plan's own stream: the same
    kernels in another order, so evaluate_device and solve_device give the same bits."""
    import torch
    import pyfmm3dpp
    dev = torch.device("cuda", 0)
    n = 30000
    x, q = _case(n, 21, False)
    d = [torch.from_numpy(np.ascontiguousarray(x[:, k])).to(dev) for k in range(3)]
    dq = torch.from_numpy(q).to(dev)
    res = []
    for side in (1, 0):
        plan = pyfmm3dpp.LaplaceFMM(6, 40, 2, 1.0, precision=precision, device=0)
        plan.set_near_field_stream(side)
        out = []
        for fused in (False, True):
            phi = torch.empty(n, dtype=torch.float64, device=dev)
            grad = torch.empty((n, 3), dtype=torch.float64, device=dev)
            torch.cuda.synchronize()
            if fused:
                plan.solve_device(n, d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), dq.data_ptr(), phi.data_ptr(),
                                  grad.data_ptr())
            else:
                plan.set_particles_device(n, d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr())
                plan.evaluate_device(dq.data_ptr(), phi.data_ptr(), grad.data_ptr())
            torch.cuda.synchronize()
            out.append((phi, grad))
        res.append(out)
        plan.close()
    for k in range(2):
        for a, b in zip(res[0][k], res[1][k]):
            print(f"max |diff| {float((a - b).abs().max()):.3e}", flush=True)
            assert torch.equal(a, b)
        assert torch.equal(res[0][0][0], res[0][1][0])
