"""Pins the oracle's restated third-party numerics and list construction against mathematics.

* DFT: the oracle's separable r2c/c2r equals numpy.fft.rfftn/irfftn (FFTW conventions).
* SVD: one-sided Jacobi reproduces LAPACK gesvd's singular values (scipy, the routine the
  reference calls at 3dpp/src/mathfunc.cpp:95-98) and reconstructs A.
* Lists: the same-depth grid lookup used for the M2L parent map equals the reference's brute
  force O(C^2 * 27) scan (3dpp/src/traverser.cpp:234-282), also on an adaptive clustered tree.
* FMM vs the finite image lattice it approximates: images=i covers (3^i)^3 copies, exactly the
  lattice of LaplaceKernel::direct (3dpp/src/kernel.cpp:11-77); a dropped term, a wrong sign or
  a wrong image scale fails these bounds by orders of magnitude.
"""
import numpy as np
import pytest
import scipy.linalg as sl


@pytest.mark.parametrize("N", [7, 11, 15])
def test_dft_r2c_c2r_match_numpy(oracle_mod, N):
    rng = np.random.default_rng(N)
    g = rng.standard_normal((N, N, N))
    G = oracle_mod.test_r2c(g)
    ref = np.fft.rfftn(g)  # axes (k, j, i), last axis halved: same as FFTW r2c on [k][j][i]
    assert np.max(np.abs(G - ref)) <= 1e-12 * np.max(np.abs(ref))
    back = oracle_mod.test_c2r(ref)
    assert np.max(np.abs(back / N ** 3 - g)) <= 1e-13 * np.max(np.abs(g))


def _surf(p, r):
    pts = [(-1 + i * 2.0 / (p - 1), -1 + j * 2.0 / (p - 1), -1 + k * 2.0 / (p - 1))
           for i in range(p) for j in range(p) for k in range(p)
           if i in (0, p - 1) or j in (0, p - 1) or k in (0, p - 1)]
    return np.array(pts) * r


@pytest.mark.parametrize("P", [4, 6, 8])
def test_svd_matches_lapack_gesvd(oracle_mod, P):
    chk, eqv = _surf(P, np.pi * 2.95), _surf(P, np.pi * 1.05)
    d = chk[:, None, :] - eqv[None, :, :]
    A = 1 / np.sqrt((d ** 2).sum(-1))
    U, s, V = oracle_mod.test_svd(A)
    s_ref = sl.svd(A.T, compute_uv=False, lapack_driver="gesvd")
    # singular values agree to n * eps * s_max (absolute accuracy of backward-stable SVDs)
    assert np.max(np.abs(s - s_ref)) <= 2 * len(s) * np.finfo(float).eps * s_ref[0]
    assert np.all(np.diff(s) <= 0)
    assert np.max(np.abs((U * s) @ V.T - A)) <= 1e-12 * np.abs(A).max()
    assert np.max(np.abs(U.T @ U - np.eye(len(s)))) <= 1e-12
    assert np.max(np.abs(V.T @ V - np.eye(len(s)))) <= 1e-12


def _lists_equal(fa, fb):
    ta, sa, oa, ia = fa.m2l_list()
    tb, sb, ob, ib = fb.m2l_list()
    ka = sorted(zip(ta.tolist(), sa.tolist(), map(tuple, np.round(oa, 9).tolist()), ia.tolist()))
    kb = sorted(zip(tb.tolist(), sb.tolist(), map(tuple, np.round(ob, 9).tolist()), ib.tolist()))
    return ka == kb, len(ka)


@pytest.mark.parametrize("images", [0, 1, 3])
@pytest.mark.parametrize("kind", ["uniform", "clustered"])
def test_parent_map_grid_equals_bruteforce(oracle_mod, images, kind):
    rng = np.random.default_rng(7)
    n = 6000
    if kind == "uniform":
        pts = rng.random((n, 3)) - 0.5
    else:
        pts = np.clip(rng.normal(0.1, 0.08, (n, 3)), -0.499, 0.499)
        pts[: n // 4] = rng.random((n // 4, 3)) - 0.5
    fa = oracle_mod.OracleFMM(4, 32, images, 1.0)
    fb = oracle_mod.OracleFMM(4, 32, images, 1.0, m2l_bruteforce=True)
    for f in (fa, fb):
        f.set_particles(pts[:, 0], pts[:, 1], pts[:, 2])
    same, cnt = _lists_equal(fa, fb)
    assert same and cnt > 0


@pytest.mark.parametrize("P,images,tol_f", [(6, 0, 2e-5), (6, 1, 2e-5), (8, 2, 2e-6)])
def test_fmm_approximates_finite_lattice(oracle_mod, P, images, tol_f):
    rng = np.random.default_rng(11)
    n = 3000
    x, y, z = (rng.random(n) - 0.5 for _ in range(3))
    q = rng.uniform(-1, 1, n)
    q -= q.mean()
    f = oracle_mod.OracleFMM(P, 32, images, 1.0, dipole=0)
    f.set_particles(x, y, z)
    p, g = f.evaluate(q)
    tar = np.arange(0, n, 7, dtype=np.int32)
    pd, gd = oracle_mod.direct(x, y, z, q, images, 1.0, targets=tar)
    l2p = np.linalg.norm(p[tar] - pd) / np.linalg.norm(pd)
    l2f = np.linalg.norm(g[tar] - gd) / np.linalg.norm(gd)
    assert l2f < tol_f and l2p < tol_f, (l2p, l2f)


@pytest.mark.parametrize("dipole", [1, 2])
def test_oracle_stages_equal_evaluate(oracle_mod, dipole):
    """The oracle's per-stage form (what the library's per-stage entry points are compared with) runs the
    same code as its evaluate: the stage sequence reproduces evaluate bit for bit."""
    rng = np.random.Generator(np.random.PCG64(3))
    n = 3000
    x = rng.random((n, 3)) - 0.5
    q = rng.uniform(-1, 1, n)
    q -= q.mean()
    f = oracle_mod.OracleFMM(6, 32, 2, 1.0, dipole=dipole, threads=4)
    f.precompute()
    f.set_particles(x[:, 0], x[:, 1], x[:, 2])
    p0, g0 = f.evaluate(q)
    f.stage("charges", q)
    for s in ("p2m", "m2m", "m2l", "l2l", "l2p", "p2p"):
        f.stage(s)
    p1, g1 = f.stage_finalize()
    assert np.array_equal(p0, p1) and np.array_equal(g0, g1)
