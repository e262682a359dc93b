"""Per-stage entry points (the operator API of the reference's LaplaceKernel) through the C-ABI.

* The stage sequence over every cell (charges, p2m, m2m, m2l, l2l, l2p, p2p, finalize) equals
  evaluate bit for bit: float64 and mixed precision, the reference's dipole term and the
  infinite-lattice gauge, a uniform and an adaptive tree.
* Cell lists: a stage run over the cells split into several lists equals the stage over all cells,
  bit for bit (every output column's arithmetic is the same).
* Each stage against the same stage of the restated reference, started from the same state (the
  reference's state before the stage is loaded into the plan), shared operators: check potentials
  and body fields to 1e-12; equivalent densities (pseudo-inverse outputs, which amplify rounding
  along their small singular directions) to 1e-8, as the whole-evaluation stage-vector test.
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
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _points(kind, n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    if kind == "uniform":
        p = rng.random((n, 3)) - 0.5
    else:
        p = np.mod(rng.normal(0.0, 0.07, (n, 3)) + 0.5, 1.0) - 0.5
    q = rng.uniform(-1, 1, n)
    q -= q.mean()
    return p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), q


def _stages(plan, q, lists=None):
    lists = lists or {}
    plan.stage_charges(q)
    for s in ("p2m", "m2m", "m2l", "l2l", "l2p", "p2p"):
        if s in lists:
            for part in lists[s]:
                plan.stage(s, part)
        else:
            plan.stage(s)
    return plan.stage_finalize()


@pytest.mark.parametrize("kind,prec,mode", [("uniform", 0, 1), ("uniform", 1, 2), ("clustered", 0, 2),
                                            ("clustered", 1, 1)])
def test_stage_sequence_equals_evaluate(fmm, kind, prec, mode):
    x, y, z, q = _points(kind, 12000, 5)
    print(f"stages {kind} precision {prec} dipole mode {mode}", flush=True)
    plan = fmm.LaplaceFMM(6, 48, 2, 1.0, dipole_mode=mode, precision=prec, device=0)
    plan.precompute()
    plan.set_particles(x, y, z)
    p0, g0 = plan.evaluate(q)
    p1, g1 = _stages(plan, q)
    assert np.array_equal(p0, p1) and np.array_equal(g0, g1)


def test_stage_cell_lists_equal_whole_stages(fmm):
    """p2m / l2p / p2p over the leaves in three lists, m2m over the parents one level per list
    (deepest first, image cells last) and l2l over all parents in one list (run level by level from
    the outermost image cell): the same bits as the stages over every cell."""
    x, y, z, q = _points("clustered", 12000, 6)
    plan = fmm.LaplaceFMM(6, 32, 3, 1.0, dipole_mode=fmm.DIPOLE_PHYSICAL, device=0)
    plan.precompute()
    plan.set_particles(x, y, z)
    p0, g0 = plan.evaluate(q)
    ints, _ = plan.cells()
    depth, cnum = ints[:, 2], ints[:, 4]
    idx = np.arange(len(ints), dtype=np.int32)
    leaves = idx[(cnum == 0) & (depth >= 0)]
    parents = idx[cnum > 0]
    split = np.array_split(leaves, 3)
    by_level = [parents[depth[parents] == d] for d in sorted(set(depth[parents].tolist()), reverse=True)]
    lists = {"p2m": split, "m2m": by_level, "l2l": [parents], "l2p": split, "p2p": split}
    # the listed p2m does not clear p_check (the whole-stage form does): start from a cleared state
    plan.set_cell_vectors(1, np.zeros((len(ints), plan.counts()["S"])))
    p1, g1 = _stages(plan, q, lists)
    print(f"  lists vs evaluate: max |dphi| {np.max(np.abs(p1 - p0)):.2e}", flush=True)
    assert np.array_equal(p0, p1) and np.array_equal(g0, g1)


def _share_operators(plan, orc, images):
    import pyfmm3dpp
    for name in pyfmm3dpp.OPERATOR_NAMES:
        orc.set_operator(name, plan.get_operator(name))
    S = plan.counts()["S"]
    orc.set_operator("L2L_IMG", plan.get_operator("L2L_IMG"))
    orc.set_operator("M2M_IMG0", plan.get_operator("M2M_IMG_SUM"))
    for k in range(1, 27):
        orc.set_operator(f"M2M_IMG{k}", np.zeros(S * S))
    if images >= 2:
        g = np.zeros_like(plan.get_operator("GK27")).reshape(26, -1, 28, 2)
        g[:, :, 0, :] = plan.get_operator("GK27_SUM").reshape(26, -1, 2)
        orc.set_operator("GK27", g)


@pytest.mark.parametrize("kind", ["uniform", "clustered"])
def test_each_stage_matches_restated_reference(fmm, oracle_mod, kind):
    P, ncrit, images = 6, 48, 2
    x, y, z, q = _points(kind, 10000, 8)
    print(f"per-stage parity {kind}", flush=True)
    plan = fmm.LaplaceFMM(P, ncrit, images, 1.0, dipole_correction=1, device=0)
    plan.precompute()
    plan.set_particles(x, y, z)
    orc = oracle_mod.OracleFMM(P, ncrit, images, 1.0, dipole=1, threads=16)
    orc.precompute()
    _share_operators(plan, orc, images)
    orc.set_particles(x, y, z)
    plan.stage_charges(q)
    orc.stage("charges", q)
    errs = {}
    # P2M from the charges alone
    plan.stage("p2m")
    orc.stage("p2m")
    errs["p2m q_equiv"] = (_rel(plan.cell_vectors(0), orc.stage_vectors()[0]), 1e-8)
    # every later stage starts from the restated reference's state before it
    for s in ("m2m", "m2l", "l2l"):
        qo, po = orc.stage_vectors()
        plan.set_cell_vectors(0, qo)
        plan.set_cell_vectors(1, po)
        plan.stage(s)
        orc.stage(s)
        qo, po = orc.stage_vectors()
        if s == "m2m":
            errs["m2m q_equiv"] = (_rel(plan.cell_vectors(0), qo), 1e-8)
        else:
            errs[f"{s} p_check"] = (_rel(plan.cell_vectors(1), po), 1e-12)
    _, po = orc.stage_vectors()
    plan.set_cell_vectors(1, po)
    plan.stage("l2p")
    orc.stage("l2p")
    pg, gg = plan.body_fields()
    pr, gr = orc.body_fields()
    errs["l2p phi"], errs["l2p grad"] = (_rel(pg, pr), 1e-12), (_rel(gg, gr), 1e-12)
    plan.stage("p2p", accumulate=0)
    orc.stage("p2p_only")
    pg, gg = plan.body_fields()
    pr, gr = orc.body_fields()
    errs["p2p phi"], errs["p2p grad"] = (_rel(pg, pr), 1e-12), (_rel(gg, gr), 1e-12)
    print("  " + ", ".join(f"{k} {v[0]:.2e}" for k, v in errs.items()), flush=True)
    for k, (e, tol) in errs.items():
        assert e <= tol, (k, e)
    # the reference's finalize from identical body fields
    plan.set_body_fields(pr, gr)
    p1, g1 = plan.stage_finalize()
    p2, g2 = orc.stage_finalize()
    assert _rel(p1, p2) <= 1e-14 and _rel(g1, g2) <= 1e-14
