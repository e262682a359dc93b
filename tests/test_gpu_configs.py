"""The benchmark configurations at their full sizes, through the C-ABI, with the bench's own inputs.

* N=1,048,576 uniform, P=8, ncrit=64, 3 image levels (the headline workload): float64 against the
  restated reference with shared operators (1e-12), mixed precision against float64, and the raw
  potential and gradient against the float64 Ewald sum on 4096 sampled targets, at 3 image levels
  (the configuration's own lattice truncation) and at 5 and 6 (below 1e-5).
* N=4,194,304 Plummer sphere (scale radius 0.05, wrapped), +-1 charges, P=8, adaptive tree with
  ncrit=128, 3 image levels: tree and lists bit-exact against the restated reference with no lost
  leaf pairs, float64 against the restated reference (1e-12), mixed precision against float64, the
  adaptive FMM against the direct sum over the same 27-box lattice (1 image level) on sampled
  targets, and the infinite-lattice gauge against Ewald.
* 48 systems, N log-uniform in [4096, 262144], P 6 and 8, 2 image levels, mixed precision: the
  batch replayed from its HIP graph equals one plan per system.
"""
import os
import tempfile

import numpy as np
import pytest

import bench

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fmm():
    import pyfmm3dpp
    pyfmm3dpp.lib()
    return pyfmm3dpp


@pytest.fixture(scope="module")
def opdir():
    with tempfile.TemporaryDirectory(prefix="fmm3dpp_ops_") as d:
        yield d


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _plan(fmm, opdir, P, ncrit, images, **kw):
    """A device plan whose operators come from one precompute per (P, images) (saved and reloaded:
    bit-identical)."""
    plan = fmm.LaplaceFMM(P, ncrit, images, 1.0, device=0, **kw)
    path = os.path.join(opdir, f"ops_P{P}_i{images}.bin")
    if os.path.exists(path):
        plan.load_operators(path)
    else:
        plan.precompute()
        plan.save_operators(path)
    return plan


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


def _oracle_parity(fmm, oracle_mod, opdir, x, y, z, q, P, ncrit, images, tol=1e-12):
    """float64 plan (reference dipole term) against the restated reference with the plan's operators."""
    plan = _plan(fmm, opdir, P, ncrit, images, dipole_correction=1)
    plan.set_particles(x, y, z)
    phi, grad = plan.evaluate(q)
    orc = oracle_mod.OracleFMM(P, ncrit, images, 1.0, dipole=1, threads=16)
    orc.precompute()
    _share_operators(plan, orc, images)
    orc.set_particles(x, y, z)
    pr, gr = orc.evaluate(q)
    ep, eg = _rel(phi, pr), _rel(grad, gr)
    print(f"  float64 vs restated reference: phi {ep:.2e} grad {eg:.2e}", flush=True)
    assert ep <= tol and eg <= tol
    return plan, orc


def _sorted_rows(cols):
    """Rows of equal-length columns (ints, or offsets rounded to 1e-9) in lexicographic order."""
    a = np.stack([np.asarray(c, dtype=np.int64) for c in cols], axis=1)
    return a[np.lexsort(a.T[::-1])]


def _offsets(o):
    return [np.round(o[:, k] * 1e9).astype(np.int64) for k in range(3)]


def _physical(fmm, opdir, x, y, z, q, P, ncrit, images, precision):
    plan = _plan(fmm, opdir, P, ncrit, images, dipole_mode=fmm.DIPOLE_PHYSICAL, precision=precision)
    phi, grad = plan.solve(x, y, z, q)
    plan.close()
    return phi, grad


def test_headline_uniform_1m_full_size(fmm, oracle_mod, opdir):
    n, P, ncrit = 1 << 20, 8, 64
    x, y, z, q = bench.workload(n, 42)
    print(f"uniform N={n} P={P} ncrit={ncrit} images=3", flush=True)
    plan, _ = _oracle_parity(fmm, oracle_mod, opdir, x, y, z, q, P, ncrit, 3)
    plan.close()
    tar = np.sort(np.random.Generator(np.random.PCG64(43)).choice(n, 4096, replace=False)).astype(np.int32)
    pe, ge = fmm.ewald(x, y, z, q, 1.0, 23, alpha=12.0, cutoff=0.5, targets=tar)
    err = {}
    for images in (3, 5, 6):
        p64, g64 = _physical(fmm, opdir, x, y, z, q, P, ncrit, images, fmm.PREC_FP64)
        pmx, gmx = _physical(fmm, opdir, x, y, z, q, P, ncrit, images, fmm.PREC_MIXED)
        err[images] = {"fp64": (_rel(p64[tar], pe), _rel(g64[tar], ge)),
                       "mixed": (_rel(pmx[tar], pe), _rel(gmx[tar], ge)),
                       "mixed_vs_fp64": (_rel(pmx, p64), _rel(gmx, g64))}
        print(f"  images {images}: " + ", ".join(f"{k} phi {v[0]:.2e} grad {v[1]:.2e}" for k, v in err[images].items()),
              flush=True)
        assert err[images]["mixed_vs_fp64"][0] <= 2e-6 and err[images]["mixed_vs_fp64"][1] <= 1e-5
    for path in ("fp64", "mixed"):
        # 3 image levels: the lattice truncation of the configuration itself (2.4e-4 / 9e-6 measured)
        assert err[3][path][0] <= 4e-4 and err[3][path][1] <= 2e-5
        # 5 and 6 levels: the raw potential and the gradient within the 1e-5 target
        assert err[5][path][0] <= 1e-5 and err[5][path][1] <= 1e-5
        assert err[6][path][0] <= 1e-5 and err[6][path][1] <= 1e-5
    assert err[6]["fp64"][0] < err[5]["fp64"][0] < err[3]["fp64"][0]


def test_clustered_plummer_4m_full_size(fmm, oracle_mod, opdir):
    n, P, ncrit, images = 1 << 22, 8, 128, 3
    x, y, z, q = bench.plummer_workload(n, 45)
    print(f"Plummer N={n} P={P} ncrit={ncrit} images={images}", flush=True)
    plan, orc = _oracle_parity(fmm, oracle_mod, opdir, x, y, z, q, P, ncrit, images)
    # tree and lists bit-exact, no dropped leaf pairs
    oi, od = orc.cells()
    pi, pd = plan.cells()
    assert np.array_equal(oi, pi) and np.array_equal(od, pd)
    assert np.array_equal(orc.body_order(), plan.body_order())
    t1, s1, o1 = orc.p2p_list()
    t2, s2, o2 = plan.p2p_list()
    assert np.array_equal(_sorted_rows([t1, s1] + _offsets(o1)), _sorted_rows([t2, s2] + _offsets(o2)))
    t1, s1, o1, i1 = orc.m2l_list()
    t2, s2, o2, i2 = plan.m2l_list()
    assert np.array_equal(_sorted_rows([t1, s1, i1] + _offsets(o1)), _sorted_rows([t2, s2, i2] + _offsets(o2)))
    c = plan.counts()
    ntree = c["n_tree_cells"]
    print(f"  cells {ntree}, max depth {int(pi[:ntree, 2].max())}, lost leaf pairs {c['n_lost']} "
          f"(restated reference {orc.lost_pairs()})", flush=True)
    assert c["n_lost"] == 0 and orc.lost_pairs() == 0
    assert int(pi[:ntree, 2].max()) >= 8   # the adaptive tree is much deeper than a uniform one
    plan.close()
    del orc
    # mixed precision against float64, infinite-lattice gauge against Ewald
    tar = np.sort(np.random.Generator(np.random.PCG64(44)).choice(n, 4096, replace=False)).astype(np.int32)
    pe, ge = fmm.ewald(x, y, z, q, 1.0, 23, alpha=12.0, cutoff=0.5, targets=tar)
    p64, g64 = _physical(fmm, opdir, x, y, z, q, P, ncrit, images, fmm.PREC_FP64)
    pmx, gmx = _physical(fmm, opdir, x, y, z, q, P, ncrit, images, fmm.PREC_MIXED)
    e = {"mixed_vs_fp64": (_rel(pmx, p64), _rel(gmx, g64)), "fp64_vs_ewald": (_rel(p64[tar], pe), _rel(g64[tar], ge)),
         "mixed_vs_ewald": (_rel(pmx[tar], pe), _rel(gmx[tar], ge))}
    print("  " + ", ".join(f"{k} phi {v[0]:.2e} grad {v[1]:.2e}" for k, v in e.items()), flush=True)
    assert e["mixed_vs_fp64"][0] <= 2e-6 and e["mixed_vs_fp64"][1] <= 1e-5
    # CPU estimate on 262,144 bodies of the same distribution: 2.4e-6 / 2.8e-7
    assert e["fp64_vs_ewald"][0] <= 1e-5 and e["fp64_vs_ewald"][1] <= 2e-6
    assert e["mixed_vs_ewald"][0] <= 1e-5 and e["mixed_vs_ewald"][1] <= 1e-5
    # the adaptive FMM against the direct sum over the same lattice of 27 boxes (1 image level);
    # no dipole term on either side.  CPU estimate on 262,144 bodies: 1.6e-7 / 2.8e-7
    p1 = _plan(fmm, opdir, P, ncrit, 1, dipole_mode=fmm.DIPOLE_OFF)
    ph1, gr1 = p1.solve(x, y, z, q)
    p1.close()
    tar1 = tar[::4].copy()
    pd_, gd_ = fmm.direct(x, y, z, q, 1, 1.0, targets=tar1)
    ed = (_rel(ph1[tar1], pd_), _rel(gr1[tar1], gd_))
    print(f"  float64, 1 image level, vs direct on {len(tar1)} targets: phi {ed[0]:.2e} grad {ed[1]:.2e}", flush=True)
    assert ed[0] <= 1e-6 and ed[1] <= 2e-6


def test_batch_48_systems_full_size(fmm, opdir):
    """The bench's batch: 48 systems, N log-uniform in [4096, 262144] (3.1e6 bodies), P 6 and 8,
    mixed precision, 2 image levels, infinite-lattice gauge.  Replays of the captured graph (new
    charges each time, eager work and synchronisations between them) against one plan per system:
    the same arithmetic in the same order, so the differences are expected to be 0 (bound 1e-13)."""
    import torch
    from pyfmm3dpp.batch import Batch
    counts, Ps, pts, q = bench.batch_workload(48, 4096, 262144, 49)
    n = int(counts.sum())
    off = np.concatenate([[0], np.cumsum(counts)])
    print(f"batch: 48 systems, {n} bodies, P {sorted(set(int(p) for p in Ps))}", flush=True)
    systems = [(int(P), 64) for P in Ps]
    b = Batch(systems, 2, 1.0, dipole_mode=fmm.DIPOLE_PHYSICAL, precision=fmm.PREC_MIXED)
    b.set_particles(counts, pts[:, 0], pts[:, 1], pts[:, 2])
    b.set_graph(True)
    dev = torch.device("cuda", 0)
    dq = torch.empty(n, dtype=torch.float64, device=dev)
    phi = torch.empty(n, dtype=torch.float64, device=dev)
    grad = torch.empty((n, 3), dtype=torch.float64, device=dev)
    rng = np.random.Generator(np.random.PCG64(50))
    charges = [q]
    for _ in range(2):   # new zero-sum charges per system for each further replay
        qq = rng.uniform(-1, 1, n)
        for s in range(len(counts)):
            qq[off[s]:off[s + 1]] -= qq[off[s]:off[s + 1]].mean()
        charges.append(qq)
    plans = []   # one plan per system, particles set once
    for s in range(len(counts)):
        plan = _plan(fmm, opdir, int(Ps[s]), 64, 2, dipole_mode=fmm.DIPOLE_PHYSICAL, precision=fmm.PREC_MIXED)
        plan.set_near_field_stream(0)   # one stream per plan (48 plans alive)
        sl = slice(int(off[s]), int(off[s + 1]))
        plan.set_particles(pts[sl, 0], pts[sl, 1], pts[sl, 2])
        plans.append((plan, sl))
    for rep, qq in enumerate(charges):
        dq.copy_(torch.from_numpy(qq))
        junk = torch.randn(1 << 20, device=dev).sum()   # eager work between replays
        torch.cuda.synchronize()
        b.evaluate(dq, phi, grad)
        torch.cuda.synchronize()
        pb, gb = phi.cpu().numpy(), grad.cpu().numpy()
        worst = 0.0
        for plan, sl in plans:
            ps, gs = plan.evaluate(qq[sl])
            worst = max(worst, _rel(pb[sl], ps), _rel(gb[sl], gs))
        print(f"  replay {rep}: worst per-system rel L2 vs one plan per system {worst:.2e} ({float(junk):.1f})",
              flush=True)
        assert worst <= 1e-13
    for plan, _ in plans:
        plan.close()
