"""The mixed path applies the P2M and L2P pseudo-inverses at a truncated rank: the leading rows of UT
and columns of VSinv whose singular values exceed 1e-7 of the largest, rounded up to a multiple of 16
(DeviceOperators::truncated_rank in exafmm-rd_amd/csrc/device_ops.h, which reads the singular values
as the reciprocal column norms of VSinv = V diag(1/s)).  These checks run on the host plan's own
operators (long double SVD of the equivalent -> check matrix, src/kernel.cpp:1642-1651 and 1699) and
compare with the direct sum: the rank at P=8, and that the truncated upward (P2M) and downward (L2P)
applies reproduce the full-rank fields to 1e-7 (the mixed path's bound is 1e-5)."""
import numpy as np
import pytest


def _surface(P, r):   # the surface point order of the operators (surface.h:28-48)
    pts = []
    for i in range(P):
        for j in range(P):
            for k in range(P):
                if i in (0, P - 1) or j in (0, P - 1) or k in (0, P - 1):
                    pts.append([(-1 + i * 2 / (P - 1)) * r, (-1 + j * 2 / (P - 1)) * r, (-1 + k * 2 / (P - 1)) * r])
    return np.array(pts)


def _K(t, s):
    return 1.0 / np.linalg.norm(t[:, None, :] - s[None, :, :], axis=2)


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("P,rank", [(8, 192), (10, 208)])
def test_mixed_pinv_truncated_rank_keeps_the_field(P, rank):
    import pyfmm3dpp
    S = 6 * (P - 1) ** 2 + 2
    plan = pyfmm3dpp.LaplaceFMM(P, 64, 0, 2.0, device=pyfmm3dpp.DEVICE_NONE)   # cell half size r0 = 1
    ops = {n: plan.get_operator(n).reshape(S, S) for n in ("UT", "VSinv", "UT_l2p", "VSinv_l2p")}
    norms = np.linalg.norm(ops["VSinv"], axis=0)
    kept = int((norms > 0).sum())   # singular values below 4e-16 of the largest have zero columns
    assert np.all(norms[kept:] == 0)
    sv = 1.0 / norms[:kept]
    assert np.all(np.diff(sv) <= 1e-12 * sv.max())   # descending (degenerate triples to rounding)
    r = min(S, -(-int((sv > 1e-7 * sv.max()).sum()) // 16) * 16)
    assert r == rank
    # the same singular values behind the L2P factors (VSinv_l2p = U diag(1/s))
    assert np.allclose(1.0 / np.linalg.norm(ops["VSinv_l2p"], axis=0)[:kept], sv, rtol=1e-9)
    chk, eqv = _surface(P, 2.95), _surface(P, 1.05)
    rng = np.random.default_rng(7)
    q = rng.standard_normal(64)
    # P2M: sources in the cell, field at well-separated targets
    src = rng.uniform(-1, 1, (64, 3))
    tar = rng.uniform(-1, 1, (300, 3)) + np.array([6.0, 0, 0]) * rng.choice([-1, 1], (300, 1))
    c = _K(chk, src) @ q
    full = _K(tar, eqv) @ (ops["VSinv"] @ (ops["UT"] @ c))
    trunc = _K(tar, eqv) @ (ops["VSinv"][:, :r] @ (ops["UT"][:r] @ c))
    exact = _K(tar, src) @ q
    assert _rel(full, exact) < 1e-7 and _rel(trunc, full) < 1e-7
    # L2P: sources two cells out, field inside the cell (check surface 1.05, equivalent 2.95)
    src2 = rng.uniform(-1, 1, (64, 3))
    src2 += np.sign(src2) * 3
    tar2 = rng.uniform(-1, 1, (300, 3))
    c2 = _K(eqv, src2) @ q
    full2 = _K(tar2, chk) @ (ops["VSinv_l2p"] @ (ops["UT_l2p"] @ c2))
    trunc2 = _K(tar2, chk) @ (ops["VSinv_l2p"][:, :r] @ (ops["UT_l2p"][:r] @ c2))
    exact2 = _K(tar2, src2) @ q
    assert _rel(full2, exact2) < 1e-7 and _rel(trunc2, full2) < 1e-7
