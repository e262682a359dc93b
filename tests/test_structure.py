"""Tree and interaction lists of the product vs the restated reference (Tree::build, Traverser).

Bit-exact comparison: every cell (BFS index, octant, depth, child range, body range, radius,
centre), the body order, the P2P map rows (target, source, periodic offset) and the M2L parent map
rows (target, source, offset, image flag).  Structure-only plans build on the host (CPU tests);
device plans build on the GPU (tests marked gpu).
"""
import numpy as np
import pytest

CASES = {
    "uniform_free": dict(kind="uniform", n=4096, ncrit=64, images=0, cycle=1.0),
    "uniform_i1": dict(kind="uniform", n=4096, ncrit=64, images=1, cycle=1.0),
    "uniform_i3": dict(kind="uniform", n=20000, ncrit=64, images=3, cycle=1.0),
    "clustered_i2": dict(kind="clustered", n=20000, ncrit=32, images=2, cycle=1.0),
    "drand48_2pi": dict(kind="drand48", n=24000, ncrit=128, images=3, cycle=2 * np.pi),
    "root_leaf": dict(kind="uniform", n=10, ncrit=64, images=1, cycle=1.0),
    "one_split": dict(kind="uniform", n=65, ncrit=64, images=2, cycle=1.0),
    "ncrit1": dict(kind="uniform", n=300, ncrit=1, images=1, cycle=1.0),
}


def _points(oracle_mod, kind, n, cycle):
    if kind == "drand48":
        x, y, z, _ = oracle_mod.drand48_bodies(n, cycle / 2, seed=5)
        return x, y, z
    rng = np.random.Generator(np.random.PCG64(n))
    if kind == "uniform":
        p = rng.random((n, 3)) - 0.5
    else:
        p = np.mod(rng.normal(0.0, 0.05, (n, 3)) + 0.5, 1.0) - 0.5
        p[: n // 5] = rng.random((n // 5, 3)) - 0.5
    return p[:, 0] * cycle, p[:, 1] * cycle, p[:, 2] * cycle


def _rows(t, s, o, extra=None):
    r = list(zip(t.tolist(), s.tolist(), map(tuple, np.round(o, 9).tolist())))
    if extra is not None:
        r = [a + (b,) for a, b in zip(r, extra.tolist())]
    return sorted(r)


def _compare(oracle_mod, device, case):
    import pyfmm3dpp
    c = CASES[case]
    x, y, z = _points(oracle_mod, c["kind"], c["n"], c["cycle"])
    orc = oracle_mod.OracleFMM(4, c["ncrit"], c["images"], c["cycle"])
    orc.set_particles(x, y, z)
    plan = pyfmm3dpp.LaplaceFMM(4, c["ncrit"], c["images"], c["cycle"], device=device)
    plan.set_particles(x, y, z)
    oi, od = orc.cells()
    pi, pd = plan.cells()
    assert oi.shape == pi.shape
    assert np.array_equal(oi, pi)
    assert np.array_equal(od, pd)          # radii and centres bit for bit
    assert np.array_equal(orc.body_order(), plan.body_order())
    assert _rows(*orc.p2p_list()) == _rows(*plan.p2p_list())
    t1, s1, o1, i1 = orc.m2l_list()
    t2, s2, o2, i2 = plan.m2l_list()
    assert _rows(t1, s1, o1, i1) == _rows(t2, s2, o2, i2)
    assert plan.counts()["n_lost"] == orc.lost_pairs()


@pytest.mark.parametrize("case", sorted(CASES))
def test_host_structure_matches_reference(oracle_mod, case):
    import pyfmm3dpp
    _compare(oracle_mod, pyfmm3dpp.DEVICE_NONE, case)


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_structure_matches_reference(oracle_mod, case):
    print(f"gpu structure case {case}", flush=True)
    _compare(oracle_mod, 0, case)


@pytest.mark.gpu
def test_gpu_structure_reused_plan(oracle_mod):
    """One device plan, several point sets in a row.  After the first build the keys are sorted on
    the previous tree's depth + 2 levels only; a deeper tree (clustered after uniform) restarts the
    build with every level sorted.  Every build must equal the restated reference."""
    import pyfmm3dpp
    ncrit, images, cycle = 32, 2, 1.0
    plan = pyfmm3dpp.LaplaceFMM(4, ncrit, images, cycle, device=0)
    depths = []
    for kind, n in (("uniform", 4096), ("clustered", 20000), ("uniform", 20000), ("clustered", 20000),
                    ("uniform", 300)):
        print(f"reused plan: {kind} n={n}", flush=True)
        x, y, z = _points(oracle_mod, kind, n, cycle)
        orc = oracle_mod.OracleFMM(4, ncrit, images, cycle)
        orc.set_particles(x, y, z)
        plan.set_particles(x, y, z)
        oi, od = orc.cells()
        pi, pd = plan.cells()
        assert np.array_equal(oi, pi) and np.array_equal(od, pd)
        assert np.array_equal(orc.body_order(), plan.body_order())
        assert _rows(*orc.p2p_list()) == _rows(*plan.p2p_list())
        t1, s1, o1, i1 = orc.m2l_list()
        t2, s2, o2, i2 = plan.m2l_list()
        assert _rows(t1, s1, o1, i1) == _rows(t2, s2, o2, i2)
        depths.append(int(pi[: plan.counts()["n_tree_cells"], 2].max()))
    print("tree depths", depths, flush=True)
    assert depths[1] > depths[0] + 2   # the clustered tree is deeper than the levels sorted for it


def _sorted_rows(cols):
    a = np.stack([np.asarray(c, dtype=np.float64) for c in cols], axis=1)
    return a[np.lexsort(a.T[::-1])]


def test_host_structure_deep_plummer_tree(oracle_mod):
    """262,144 bodies of a Plummer sphere (scale radius 0.05, wrapped into the unit cube), ncrit 32,
    3 image levels: an adaptive tree nine or more levels deep, host build vs the restated reference,
    bit for bit, with no lost leaf pairs."""
    import sys
    import os
    import pyfmm3dpp
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    x, y, z, _ = bench.plummer_workload(1 << 18, 19)
    orc = oracle_mod.OracleFMM(4, 32, 3, 1.0)
    orc.set_particles(x, y, z)
    plan = pyfmm3dpp.LaplaceFMM(4, 32, 3, 1.0, device=pyfmm3dpp.DEVICE_NONE)
    plan.set_particles(x, y, z)
    oi, od = orc.cells()
    pi, pd = plan.cells()
    ntree = plan.counts()["n_tree_cells"]
    assert int(pi[:ntree, 2].max()) >= 9
    assert np.array_equal(oi, pi) and np.array_equal(od, pd)
    assert np.array_equal(orc.body_order(), plan.body_order())
    for lo, lp in ((orc.p2p_list(), plan.p2p_list()), (orc.m2l_list(), plan.m2l_list())):
        ro = _sorted_rows([lo[0], lo[1]] + [lo[2][:, k] for k in range(3)] + list(lo[3:]))
        rp = _sorted_rows([lp[0], lp[1]] + [lp[2][:, k] for k in range(3)] + list(lp[3:]))
        assert np.array_equal(ro, rp)
    assert plan.counts()["n_lost"] == orc.lost_pairs() == 0
