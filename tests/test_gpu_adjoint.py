"""Adjoint of the FMM (fmm3dpp_evaluate_adjoint) and its torch.autograd binding.

Checked against float64 central finite differences of the FMM's own energy E = 1/2 sum q phi
(h = 1e-6 for positions and charges), against the bilinear identity w.(A q) = (A^T w).q, and
dE/dx = q grad(phi) against the CPU restatement's gradient (operators shared, 1e-12).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fmm():
    import pyfmm3dpp
    pyfmm3dpp.lib()
    return pyfmm3dpp


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _system(n, seed, gaussian=False):
    rng = np.random.Generator(np.random.PCG64(seed))
    x = rng.random((n, 3)) - 0.5
    if gaussian:
        q = rng.standard_normal(n).astype(np.float32).astype(np.float64)
    else:
        q = rng.uniform(-1, 1, n)
    q -= q.mean()
    return x, q


def _energy(plan, x, q):
    phi, _ = plan.solve(x[:, 0], x[:, 1], x[:, 2], q)
    return 0.5 * float(np.dot(q, phi))


def _fd_check(fmm, plan, x, q, nsample, seed, h=1e-6):
    """Central differences of E on nsample random bodies (all 3 coordinates and the charge)."""
    import torch
    from pyfmm3dpp.autograd import energy
    dev = torch.device("cuda", 0)
    xt = torch.tensor(x, device=dev, requires_grad=True)
    qt = torch.tensor(q, device=dev, requires_grad=True)
    E = energy(plan, xt, qt)
    E.backward()
    gx, gq = xt.grad.cpu().numpy(), qt.grad.cpu().numpy()
    rng = np.random.Generator(np.random.PCG64(seed))
    idx = rng.choice(len(q), nsample, replace=False)
    fdx, adx, fdq = [], [], []
    for i in idx:
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, d] += h
            xm[i, d] -= h
            fdx.append((_energy(plan, xp, q) - _energy(plan, xm, q)) / (2 * h))
            adx.append(gx[i, d])
        qp, qm = q.copy(), q.copy()
        qp[i] += h
        qm[i] -= h
        fdq.append((_energy(plan, x, qp) - _energy(plan, x, qm)) / (2 * h))
    ex, eq = _rel(np.array(adx), np.array(fdx)), _rel(gq[idx], np.array(fdq))
    print(f"  FD: dE/dx rel {ex:.2e}, dE/dq rel {eq:.2e}", flush=True)
    return ex, eq


@pytest.mark.parametrize("images,mode", [(0, 0), (2, 2)])
def test_energy_gradient_vs_finite_differences(fmm, images, mode):
    x, q = _system(4096, 17)
    plan = fmm.LaplaceFMM(6, 64, images, 1.0, dipole_mode=mode, device=0)
    plan.precompute()
    ex, eq = _fd_check(fmm, plan, x, q, 16, 5)
    assert ex <= 1e-4 and eq <= 1e-4


def test_energy_gradient_differentiated_workload(fmm):
    """N=262144 uniform bodies, zero-sum float32 Gaussian charges (sigma 1), P=8, images=2,
    infinite-lattice gauge; 64 sampled bodies, all coordinates and charges."""
    x, q = _system(262144, 2024, gaussian=True)
    plan = fmm.LaplaceFMM(8, 64, 2, 1.0, dipole_mode=fmm.DIPOLE_PHYSICAL, device=0)
    plan.precompute()
    ex, eq = _fd_check(fmm, plan, x, q, 64, 6)
    assert ex <= 1e-4 and eq <= 1e-4


def test_adjoint_bilinear_identity(fmm):
    """(A^T w).q = w.(A q): the adjoint's dq for L = w.phi, against the forward with charges q.  The
    FMM operator is symmetric up to its approximation error, so the two agree to that error relative
    to |w| |A q| (the dot products themselves cancel heavily)."""
    x, q = _system(20000, 3)
    rng = np.random.Generator(np.random.PCG64(4))
    w = rng.uniform(-1, 1, len(q))
    plan = fmm.LaplaceFMM(8, 64, 3, 1.0, dipole_mode=fmm.DIPOLE_PHYSICAL, device=0)
    plan.precompute()
    phi, grad = plan.solve(x[:, 0], x[:, 1], x[:, 2], q)
    dq, dx = plan.evaluate_adjoint(q, w, grad)
    lhs, rhs = float(np.dot(dq, q)), float(np.dot(w, phi))
    scale = float(np.linalg.norm(w) * np.linalg.norm(phi))
    print(f"  w.(Aq) {rhs:.12e}  (A^T w).q {lhs:.12e}  difference / (|w| |Aq|) {abs(lhs - rhs) / scale:.2e}",
          flush=True)
    assert abs(lhs - rhs) <= 1e-8 * scale


def test_energy_force_matches_reference_gradient(fmm, oracle_mod):
    """dE/dx = q grad(phi), with grad(phi) from the CPU restatement of the reference (shared
    operators, no dipole term)."""
    import pyfmm3dpp
    x, q = _system(8192, 8)
    plan = fmm.LaplaceFMM(6, 64, 2, 1.0, dipole_mode=fmm.DIPOLE_OFF, device=0)
    plan.precompute()
    phi, grad = plan.solve(x[:, 0], x[:, 1], x[:, 2], q)
    dq, dx = plan.evaluate_adjoint(q, 0.5 * q, grad)
    dE_dq, dE_dx = 0.5 * phi + dq, dx        # E = 1/2 q.phi(x, q): q enters explicitly and through phi
    orc = oracle_mod.OracleFMM(6, 64, 2, 1.0, dipole=0, threads=16)
    orc.precompute()
    for name in pyfmm3dpp.OPERATOR_NAMES:
        orc.set_operator(name, plan.get_operator(name))
    S = plan.counts()["S"]
    orc.set_operator("L2L_IMG", plan.get_operator("L2L_IMG"))
    orc.set_operator("M2M_IMG0", plan.get_operator("M2M_IMG_SUM"))
    for k in range(1, 27):
        orc.set_operator(f"M2M_IMG{k}", np.zeros(S * S))
    g = np.zeros_like(plan.get_operator("GK27")).reshape(26, -1, 28, 2)
    g[:, :, 0, :] = plan.get_operator("GK27_SUM").reshape(26, -1, 2)
    orc.set_operator("GK27", g)
    orc.set_particles(x[:, 0], x[:, 1], x[:, 2])
    p_ref, g_ref = orc.evaluate(q)
    e1, e2 = _rel(dE_dx, q[:, None] * g_ref), _rel(dE_dq, p_ref)
    print(f"  dE/dx vs q grad_ref {e1:.2e}, dE/dq vs phi_ref {e2:.2e}", flush=True)
    assert e1 <= 1e-11 and e2 <= 1e-11


def test_reference_dipole_mode_has_no_adjoint(fmm):
    x, q = _system(2048, 1)
    plan = fmm.LaplaceFMM(4, 64, 1, 1.0, dipole_correction=1, device=0)
    plan.precompute()
    phi, grad = plan.solve(x[:, 0], x[:, 1], x[:, 2], q)
    with pytest.raises(fmm.FMMError):
        plan.evaluate_adjoint(q, q, grad)
