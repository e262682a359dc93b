"""Parity of the HIP float64 path (libfmm3dpp.so through its C-ABI) with the CPU restatement.

Operators are shared for the 1e-12 comparisons: the oracle receives the plan's own operator set
(fmm3dpp_get_operator -> orc_set_operator), because the check-to-equivalent pseudo-inverse has
condition number ~3.6e12 at P=8 and two correct SVDs give operators that differ at the 1e-5 level.
With shared operators every remaining difference is summation order, so potentials and gradients
must agree to 1e-12 relative L2 (the bar the reference's FP64 path is held to).  The end-to-end
runs with independently computed operators are checked against the reference's logged energies.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def fmm():
    import pyfmm3dpp
    pyfmm3dpp.lib()
    return pyfmm3dpp


def _share_operators(plan, orc, images):
    """Give the oracle the plan's operators in the structure the device applies them: factored
    P2M/L2P pseudo-inverses as the reference, and the 27 image M2M matrices and 27 lattice kernels
    pre-summed (sums taken in extended precision; differences to term-wise sums stay <= 1e-13)."""
    import pyfmm3dpp
    for name in pyfmm3dpp.OPERATOR_NAMES:
        orc.set_operator(name, plan.get_operator(name))
    if images > 0:
        S = plan.counts()["S"]
        orc.set_operator("L2L_IMG", plan.get_operator("L2L_IMG"))
        orc.set_operator("M2M_IMG0", plan.get_operator("M2M_IMG_SUM"))
        for k in range(1, 27):
            orc.set_operator(f"M2M_IMG{k}", np.zeros(S * S))
    if images >= 2:
        g = np.zeros_like(plan.get_operator("GK27")).reshape(26, -1, 28, 2)
        g[:, :, 0, :] = plan.get_operator("GK27_SUM").reshape(26, -1, 2)
        orc.set_operator("GK27", g)


def _rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _uniform(n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    x = rng.random((n, 3)) - 0.5
    q = rng.uniform(-1, 1, n)
    q -= q.mean()
    return x[:, 0].copy(), x[:, 1].copy(), x[:, 2].copy(), q


def _case(fmm, oracle_mod, x, y, z, q, P, ncrit, images, cycle, dipole, div4pi=0, tol=1e-12):
    print(f"case n={len(x)} P={P} ncrit={ncrit} images={images} dipole={dipole}", flush=True)
    plan = fmm.LaplaceFMM(P, ncrit, images, cycle, dipole_correction=dipole, divide_4pi=div4pi, device=0)
    plan.precompute()
    plan.set_particles(x, y, z)
    phi, grad = plan.evaluate(q)
    orc = oracle_mod.OracleFMM(P, ncrit, images, cycle, dipole=dipole, divide_4pi=div4pi, threads=16)
    orc.precompute()
    _share_operators(plan, orc, images)
    orc.set_particles(x, y, z)
    p_ref, g_ref = orc.evaluate(q)
    ep, eg = _rel_l2(phi, p_ref), _rel_l2(grad, g_ref)
    print(f"  rel L2 phi {ep:.2e} grad {eg:.2e}", flush=True)
    assert ep <= tol and eg <= tol, (ep, eg)
    return plan, phi, grad


# The pseudo-inverses amplify rounding along their small singular directions, and their condition
# number grows with P (3.6e12 at P=8): two float64 evaluations that round in different orders agree
# to 1e-12 up to P=8 and to 2-4e-12 at P=10 (measured 1.6e-12 potential, 3.7e-12 gradient).
P_TOL = {10: 1e-11}


@pytest.mark.parametrize("P,images", [(2, 1), (4, 0), (4, 1), (6, 2), (8, 3), (10, 1)])
def test_fp64_matches_oracle_small(fmm, oracle_mod, P, images):
    x, y, z, q = _uniform(4096, 1000 + P + images)
    _case(fmm, oracle_mod, x, y, z, q, P, 64, images, 1.0, dipole=1, tol=P_TOL.get(P, 1e-12))


def test_fp64_adaptive_clustered_tree(fmm, oracle_mod):
    rng = np.random.Generator(np.random.PCG64(7))
    n = 20000
    pts = rng.normal(0.0, 0.06, (n, 3))
    pts = np.mod(pts + 0.5, 1.0) - 0.5
    q = np.where(rng.permutation(n) < n // 2, 1.0, -1.0)
    _case(fmm, oracle_mod, pts[:, 0].copy(), pts[:, 1].copy(), pts[:, 2].copy(), q, 6, 32, 2, 1.0, dipole=1)


def test_fp64_config1_workload(fmm, oracle_mod):
    """N=65,536 uniform, zero-sum charges, P=6, ncrit=64, 2 image levels (the CPU-runnable case)."""
    x, y, z, q = _uniform(65536, 2024)
    _case(fmm, oracle_mod, x, y, z, q, 6, 64, 2, 1.0, dipole=1)


def test_fp64_nonperiodic_divide4pi(fmm, oracle_mod):
    x, y, z, q = _uniform(10000, 9)
    _case(fmm, oracle_mod, x, y, z, q, 6, 64, 0, 1.0, dipole=0, div4pi=1)


def test_fp64_reference_log_energies(fmm, oracle_mod):
    """Independent operators end to end: the reference's logged energies (N=24000 drand48 seed 5)."""
    fix = json.load(open(os.path.join(HERE, "golden", "log_energies.json")))
    cyc = 2 * np.pi
    x, y, z, q = oracle_mod.drand48_bodies(24000, cyc / 2, seed=5)
    plans = {}
    for row in fix["periodic"]:
        if row["images"] not in (0, 1, 2, 3, 5):
            continue
        P, i = row["P"], row["images"]
        plan = fmm.LaplaceFMM(P, 128, i, cyc, dipole_correction=1, device=0)
        phi, grad = plan.solve(x, y, z, q)
        e = float(np.sum(q * phi))
        rel = abs(e - row["energy_fmm"]) / abs(row["energy_fmm"])
        print(f"P={P} images={i} energy {e:.12e} log {row['energy_fmm']:.12e} rel {rel:.1e}", flush=True)
        assert rel <= 2e-10


@pytest.mark.parametrize("clustered", [False, True])
def test_fp64_stage_vectors_match_oracle(fmm, oracle_mod, clustered):
    """Per-stage parity with shared operators: the equivalent densities after the upward pass
    (P2M, M2M, image M2M) and the check potentials after the downward pass (M2L, L2L) of every
    cell, GPU against the restated reference, cell for cell (the trees are identical).  The check
    potentials are smooth fields and agree to rounding; the equivalent densities carry the
    pseudo-inverse's amplification of rounding along its small singular directions, so they agree
    to a looser relative bound (their fields, compared above, agree to 1e-12): 1e-8 here."""
    if clustered:
        rng = np.random.Generator(np.random.PCG64(31))
        pts = np.mod(rng.normal(0.0, 0.07, (12000, 3)) + 0.5, 1.0) - 0.5
        x, y, z = (pts[:, k].copy() for k in range(3))
        q = rng.uniform(-1, 1, 12000)
        q -= q.mean()
    else:
        x, y, z, q = _uniform(8192, 77)
    P, ncrit, images = 6, 48, 2
    print(f"stage vectors clustered={clustered}", flush=True)
    plan = fmm.LaplaceFMM(P, ncrit, images, 1.0, dipole_correction=1, device=0)
    plan.precompute()
    plan.set_particles(x, y, z)
    plan.evaluate(q)
    orc = oracle_mod.OracleFMM(P, ncrit, images, 1.0, dipole=1, threads=16)
    orc.precompute()
    _share_operators(plan, orc, images)
    orc.set_particles(x, y, z)
    orc.evaluate(q)
    qg, pg = plan.stage_vectors()
    qo, po = orc.stage_vectors()
    assert qg.shape == qo.shape and pg.shape == po.shape
    eq, ep = _rel_l2(qg, qo), _rel_l2(pg, po)
    print(f"  q_equiv rel L2 {eq:.2e}, p_check rel L2 {ep:.2e}", flush=True)
    assert ep <= 1e-12 and eq <= 1e-8, (eq, ep)   # measured 1.7e-13 / 1.5e-9 (uniform), 1.7e-14 / 1.9e-9
