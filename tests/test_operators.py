"""Pins the product's operator precompute (long-double SVD and DFTs) by mathematics.

* e2c * (VSinv * (UT * b)) = b                       (3dpp/test/test_matmul.cpp:98-114 pattern)
* P2M then M2M reproduce the far field of the sources; M2M_IMG likewise for the 27 image copies
* Gk tables: FFT convolution of a child's equivalent charges equals the direct surface-to-surface
  sum for every non-adjacent child pair of a pattern (3dpp/test/test_fft.cpp:75-146 pattern)
* operators agree with the CPU restatement's float64-SVD operators on physical inputs
"""
import numpy as np
import pytest


def _surf(P, r, c=(0.0, 0.0, 0.0)):
    pts = [(-1.0 + i * 2.0 / (P - 1), -1.0 + j * 2.0 / (P - 1), -1.0 + k * 2.0 / (P - 1))
           for i in range(P) for j in range(P) for k in range(P)
           if i in (0, P - 1) or j in (0, P - 1) or k in (0, P - 1)]
    return np.array(pts) * r + np.array(c)


def _K(t, s):
    d = t[:, None, :] - s[None, :, :]
    r = np.sqrt((d ** 2).sum(-1))
    return np.where(r > 0, 1.0 / np.where(r > 0, r, 1.0), 0.0)


@pytest.fixture(scope="module", params=[4, 6, 8])
def ops(request):
    import pyfmm3dpp
    P = request.param
    plan = pyfmm3dpp.LaplaceFMM(P, 64, 2, 1.0, device=pyfmm3dpp.DEVICE_NONE)
    S = 6 * (P - 1) ** 2 + 2
    get = lambda n: plan.get_operator(n)
    return dict(P=P, S=S, plan=plan, UT=get("UT").reshape(S, S), VS=get("VSinv").reshape(S, S),
                UTl=get("UT_l2p").reshape(S, S), VSl=get("VSinv_l2p").reshape(S, S),
                M2M=[get(f"M2M{o}").reshape(S, S) for o in range(8)], GK8=get("GK8"),
                M2MI=[get(f"M2M_IMG{o}").reshape(S, S) for o in range(27)])


def test_pinv_factors_reproduce_check_potential(ops):
    P, S = ops["P"], ops["S"]
    r0 = 0.5
    A = _K(_surf(P, 2.95 * r0), _surf(P, 1.05 * r0))
    rng = np.random.default_rng(1)
    src = rng.uniform(-r0, r0, (40, 3))
    b = _K(_surf(P, 2.95 * r0), src) @ rng.uniform(-1, 1, 40)
    q = ops["VS"] @ (ops["UT"] @ b)
    assert np.linalg.norm(A @ q - b) <= 1e-11 * np.linalg.norm(b)
    # downward: check surface 1.05 r, equivalent surface 2.95 r (kernel.cpp:1461-1467)
    Ad = _K(_surf(P, 1.05 * r0), _surf(P, 2.95 * r0))
    far = rng.uniform(-1, 1, (40, 3)) * 4 * r0 + 6 * r0
    bd = _K(_surf(P, 1.05 * r0), far) @ rng.uniform(-1, 1, 40)
    qd = ops["VSl"] @ (ops["UTl"] @ bd)
    assert np.linalg.norm(Ad @ qd - bd) <= 1e-11 * np.linalg.norm(bd)


def _far_points(rng, n, rmin, rmax):
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    return v * rng.uniform(rmin, rmax, n)[:, None]


def test_m2m_translates_far_field(ops):
    P, S = ops["P"], ops["S"]
    tol = {4: 5e-3, 6: 2e-4, 8: 1e-5}[P]
    r0 = 0.5
    rng = np.random.default_rng(2)
    far = _far_points(rng, 200, 8 * r0, 12 * r0)
    for o in (0, 5):
        cc = np.array([(r0 / 2) if o & 4 else -(r0 / 2), (r0 / 2) if o & 2 else -(r0 / 2), (r0 / 2) if o & 1 else -(r0 / 2)])
        src = cc + rng.uniform(-r0 / 2, r0 / 2, (30, 3))
        qs = rng.uniform(-1, 1, 30)
        # child P2M at depth 1: q = 0.5 * VSinv (UT p_check)   (kernel.cpp:332-371)
        pc = _K(_surf(P, 2.95 * r0 / 2, cc), src) @ qs
        qc = 0.5 * (ops["VS"] @ (ops["UT"] @ pc))
        qp = ops["M2M"][o] @ qc
        exact = _K(far, src) @ qs
        approx = _K(far, _surf(P, 1.05 * r0)) @ qp
        assert np.linalg.norm(approx - exact) <= tol * np.linalg.norm(exact)


def test_m2m_image_translates_27_copies(ops):
    P, S = ops["P"], ops["S"]
    tol = {4: 5e-3, 6: 2e-4, 8: 1e-5}[P]
    r0 = 0.5                      # parent (image cell) radius; the child is the 3x smaller centre copy
    rc = r0 / 3
    rng = np.random.default_rng(3)
    src = rng.uniform(-rc, rc, (30, 3))
    qs = rng.uniform(-1, 1, 30)
    pc = _K(_surf(P, 2.95 * rc), src) @ qs
    # unit-scale operators at child radius rc: P2M scale (r0/rc) relative to the root operator radius
    qc = (rc / r0) * (ops["VS"] @ (ops["UT"] @ pc))
    qp = sum(M @ qc for M in ops["M2MI"])
    far = _far_points(rng, 200, 8 * r0, 12 * r0)
    shifts = np.array([(i, j, k) for k in (-1, 0, 1) for j in (-1, 0, 1) for i in (-1, 0, 1)]) * 2 * rc
    exact = sum(_K(far, src + s) @ qs for s in shifts)
    approx = _K(far, _surf(P, 1.05 * r0)) @ qp
    assert np.linalg.norm(approx - exact) <= tol * np.linalg.norm(exact)


def test_gk8_equals_direct_surface_sum(ops):
    """For pattern rel=(1,0,0) and every non-adjacent child pair: irfft(Gk * rfft(Q)) gathered at the
    shifted check lattice equals the direct sum from the source child's equivalent surface."""
    P, S = ops["P"], ops["S"]
    N, Nh = 2 * P - 1, P
    F = N * N * Nh
    r0 = 0.5                                  # M2L operators are tabulated at root radius, child radius r0
    lat = np.array([(i, j, k) for i in range(P) for j in range(P) for k in range(P)
                    if i in (0, P - 1) or j in (0, P - 1) or k in (0, P - 1)])
    pars = [(i, j, k) for k in (-1, 0, 1) for j in (-1, 0, 1) for i in (-1, 0, 1) if (i, j, k) != (0, 0, 0)]
    pi = pars.index((1, 0, 0))
    G = ops["GK8"].reshape(26, F, 64, 2)
    rng = np.random.default_rng(4)
    q = rng.uniform(-1, 1, S)
    grid = np.zeros((N, N, N))
    grid[lat[:, 2], lat[:, 1], lat[:, 0]] = q
    Qk = np.fft.rfftn(grid)
    checked = 0
    for isrc in range(8):
        for itar in range(8):
            d = np.array([2 * pars[pi][a] + ((isrc >> (2 - a)) & 1) - ((itar >> (2 - a)) & 1) for a in range(3)])
            gk = G[pi, :, isrc * 8 + itar, 0] + 1j * G[pi, :, isrc * 8 + itar, 1]
            if np.all(np.abs(d) <= 1):
                assert np.all(gk == 0)
                continue
            p = np.fft.irfftn(gk.reshape(N, N, Nh) * Qk, s=(N, N, N), axes=(0, 1, 2))
            fft_val = p[lat[:, 2] + P - 1, lat[:, 1] + P - 1, lat[:, 0] + P - 1]
            # source child equivalent surface at offset d (child diameters of 2 r0), target at origin
            direct = _K(_surf(P, 1.05 * r0), _surf(P, 1.05 * r0, d * 2 * r0)) @ q
            assert np.max(np.abs(fft_val - direct)) <= 1e-12 * np.max(np.abs(direct))
            checked += 1
    assert checked > 20


def test_operators_agree_with_restatement_on_physical_input(ops, oracle_mod):
    P, S = ops["P"], ops["S"]
    orc = oracle_mod.OracleFMM(P, 64, 2, 1.0)
    orc.precompute()
    g8 = orc.get_operator("GK8")
    assert np.max(np.abs(g8 - ops["GK8"])) <= 1e-13 * np.max(np.abs(g8))
    r0 = 0.5
    rng = np.random.default_rng(5)
    cc = np.array([-r0 / 2] * 3)
    src = cc + rng.uniform(-r0 / 2, r0 / 2, (30, 3))
    qs = rng.uniform(-1, 1, 30)
    pc = _K(_surf(P, 2.95 * r0 / 2, cc), src) @ qs
    qc = 0.5 * (ops["VS"] @ (ops["UT"] @ pc))
    far = _far_points(rng, 100, 8 * r0, 12 * r0)
    a = _K(far, _surf(P, 1.05 * r0)) @ (ops["M2M"][0] @ qc)
    b = _K(far, _surf(P, 1.05 * r0)) @ (orc.get_operator("M2M0").reshape(S, S) @ qc)
    assert np.linalg.norm(a - b) <= 1e-9 * np.linalg.norm(b)


def test_operator_cache_round_trip(tmp_path):
    """fmm3dpp_save_operators / fmm3dpp_load_operators: every operator comes back bit-identical;
    a file for another configuration or a damaged file is refused."""
    import pyfmm3dpp
    a = pyfmm3dpp.LaplaceFMM(4, 16, 2, 1.0, device=pyfmm3dpp.DEVICE_NONE)
    a.precompute()
    path = str(tmp_path / "ops_p4_i2.bin")
    a.save_operators(path)
    b = pyfmm3dpp.LaplaceFMM(4, 16, 2, 1.0, device=pyfmm3dpp.DEVICE_NONE)
    b.load_operators(path)
    names = pyfmm3dpp.OPERATOR_NAMES + pyfmm3dpp.IMAGE_OPERATOR_NAMES + ["GK27", "GK27_SUM", "M2M_IMG_SUM"]
    for name in names:
        assert np.array_equal(a.get_operator(name), b.get_operator(name)), name
    wrong = pyfmm3dpp.LaplaceFMM(4, 16, 3, 1.0, device=pyfmm3dpp.DEVICE_NONE)
    with pytest.raises(pyfmm3dpp.FMMError):
        wrong.load_operators(path)
    data = bytearray(open(path, "rb").read())
    data[len(data) // 2] ^= 0x55
    bad = str(tmp_path / "bad.bin")
    open(bad, "wb").write(bytes(data))
    c = pyfmm3dpp.LaplaceFMM(4, 16, 2, 1.0, device=pyfmm3dpp.DEVICE_NONE)
    with pytest.raises(pyfmm3dpp.FMMError):
        c.load_operators(bad)
    with pytest.raises(pyfmm3dpp.FMMError):
        c.load_operators(str(tmp_path / "missing.bin"))


def test_operator_cache_directory(tmp_path, monkeypatch):
    """FMM3DPP_OPERATOR_CACHE: the first precompute of a (P, images, cycle) writes its file to the
    directory, a later plan loads it instead of computing (bit-identical operators); a damaged file
    is recomputed and replaced; a directory that cannot be written only loses the file."""
    import glob
    import pyfmm3dpp
    names = pyfmm3dpp.OPERATOR_NAMES + pyfmm3dpp.IMAGE_OPERATOR_NAMES + ["GK27", "GK27_SUM", "M2M_IMG_SUM"]
    ref = pyfmm3dpp.LaplaceFMM(6, 16, 2, 1.0, device=pyfmm3dpp.DEVICE_NONE)
    ref.precompute()                                   # no cache directory: computed
    monkeypatch.setenv("FMM3DPP_OPERATOR_CACHE", str(tmp_path))

    def plan():
        p = pyfmm3dpp.LaplaceFMM(6, 16, 2, 1.0, device=pyfmm3dpp.DEVICE_NONE)
        p.precompute()
        for name in names:
            assert np.array_equal(p.get_operator(name), ref.get_operator(name)), name
        return p.timing_ms("precompute")

    t_first = plan()
    files = glob.glob(str(tmp_path / "fmm3dpp_ops_*.bin"))
    assert len(files) == 1 and not glob.glob(str(tmp_path / "*.tmp*"))
    t_loaded = plan()
    print(f"precompute {t_first:.0f} ms, from the cache directory {t_loaded:.0f} ms")
    assert t_loaded < 0.5 * t_first
    data = bytearray(open(files[0], "rb").read())
    data[len(data) // 3] ^= 0x10
    open(files[0], "wb").write(bytes(data))
    plan()                                             # damaged: recomputed, file replaced
    plan()                                             # and loaded again
    other = pyfmm3dpp.LaplaceFMM(6, 16, 2, 2.0, device=pyfmm3dpp.DEVICE_NONE)
    other.precompute()                                 # another cycle: a file of its own
    assert len(glob.glob(str(tmp_path / "fmm3dpp_ops_*.bin"))) == 2
    monkeypatch.setenv("FMM3DPP_OPERATOR_CACHE", str(tmp_path / "missing" / "dir"))
    plan()
