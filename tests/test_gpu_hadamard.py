"""The 8x8 M2L Hadamard copies each chunk's distinct source spectra into LDS before the matrix ops
(the default) or reads every source vector from HBM / L2 (fmm3dpp_set_hadamard_staging(plan, 0)).
Both forms multiply the same values in the same order, so potentials and gradients agree bit for bit.
Trees: uniform and clustered (adaptive, several chunks straddling levels), both precisions, with
and without image levels; the float64 cases also run the Hadamard over the image targets."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [  # n, P, ncrit, images, clustered, precision
    (60000, 6, 32, 2, False, 0),
    (60000, 6, 32, 2, False, 1),
    (50000, 8, 24, 3, True, 1),
    (50000, 4, 16, 0, True, 0),
    (30000, 8, 64, 1, False, 1),
]


def _case(n, seed, clustered):
    rng = np.random.Generator(np.random.PCG64(seed))
    if clustered:
        x = np.concatenate([rng.normal(-0.2, 0.04, (n // 2, 3)), rng.random((n - n // 2, 3)) - 0.5])
        x = (x + 0.5) % 1.0 - 0.5
    else:
        x = rng.random((n, 3)) - 0.5
    q = rng.uniform(-1, 1, n)
    q -= q.mean()
    return x, q


@pytest.mark.parametrize("n,P,ncrit,images,clustered,precision", CASES)
def test_staged_equals_unstaged(n, P, ncrit, images, clustered, precision):
    import torch
    import pyfmm3dpp
    print(f"case n={n} P={P} ncrit={ncrit} images={images} clustered={clustered} precision={precision}", flush=True)
    dev = torch.device("cuda", 0)
    x, q = _case(n, 100 + n % 97 + P, clustered)
    d = [torch.from_numpy(np.ascontiguousarray(x[:, k])).to(dev) for k in range(3)]
    dq = torch.from_numpy(q).to(dev)
    out = []
    for staging in (1, 0):
        plan = pyfmm3dpp.LaplaceFMM(P, ncrit, images, 1.0, dipole_mode=pyfmm3dpp.DIPOLE_PHYSICAL,
                                    precision=precision, device=0)
        plan.set_hadamard_staging(staging)
        phi = torch.empty(n, dtype=torch.float64, device=dev)
        grad = torch.empty((n, 3), dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        plan.solve_device(n, d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), dq.data_ptr(), phi.data_ptr(),
                          grad.data_ptr())
        torch.cuda.synchronize()
        out.append((phi, grad))
        counts = plan.counts()
    dp = float((out[0][0] - out[1][0]).abs().max())
    dg = float((out[0][1] - out[1][1]).abs().max())
    print(f"  parents {counts['n_tree_cells']} max |diff| phi {dp:.3e} grad {dg:.3e}", flush=True)
    assert float(out[0][0].abs().max()) > 0
    assert dp == 0.0 and dg == 0.0
