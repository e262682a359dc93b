"""The host-side parts of bench.py, on plans built without a device: the stage-rate accounting
(algorithmic work counted from the plan's own tree and lists), the Hadamard's algorithmic flops,
the synthetic workloads, the committed traffic figure the bench line reads, and the one-GPU guard.
The timed legs themselves need a GPU."""
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def bench():
    sys.path.insert(0, ROOT)
    import bench as b
    return b


@pytest.fixture(scope="module")
def fmm():
    import pyfmm3dpp
    pyfmm3dpp.lib()
    return pyfmm3dpp


def _host_plan(fmm, prec, P=8, n=20000, seed=3):
    rng = np.random.Generator(np.random.PCG64(seed))
    x = rng.random((n, 3)) - 0.5
    plan = fmm.LaplaceFMM(P, 64, 3, 1.0, precision=prec, device=fmm.DEVICE_NONE)
    plan.precompute()
    plan.set_particles(x[:, 0].copy(), x[:, 1].copy(), x[:, 2].copy())
    return plan


def test_stage_rates_count_the_arithmetic_each_path_runs(bench, fmm):
    """Float64: full-rank pseudo-inverse applies against the float64 matrix-core peak; mixed at
    P=8: the truncated rank (192 of 296) against the float32 peak; mixed at P=10: truncated rank
    (208 of 488), float64 applies."""
    stages = {k: 0.1 for k in bench.STAGES}
    rates, nleaf = {}, {}
    for prec, P in ((fmm.PREC_FP64, 8), (fmm.PREC_MIXED, 8), (fmm.PREC_MIXED, 10)):
        plan = _host_plan(fmm, prec, P)
        rates[(prec, P)] = bench.stage_rooflines(plan, stages, prec == fmm.PREC_MIXED)
        c = plan.counts()
        ints, _ = plan.cells()
        nleaf[P] = int(np.sum(ints[: c["n_tree_cells"], 4] == 0))
        plan.close()
    r64, rmx, m10 = rates[(fmm.PREC_FP64, 8)], rates[(fmm.PREC_MIXED, 8)], rates[(fmm.PREC_MIXED, 10)]
    assert r64["P2M"]["gemm_gflop"] == round(4.0 * 296 * 296 * nleaf[8] / 1e9, 2)
    assert rmx["P2M"]["gemm_gflop"] == round(4.0 * 192 * 296 * nleaf[8] / 1e9, 2)
    assert m10["P2M"]["gemm_gflop"] == round(4.0 * 208 * 488 * nleaf[10] / 1e9, 2)
    assert r64["P2M"]["peak_tflops"] == bench.FP64_PEAK_TFLOPS and rmx["P2M"]["peak_tflops"] == bench.FP32_PEAK_TFLOPS
    assert rmx["M2M"]["peak_tflops"] == bench.FP32_PEAK_TFLOPS and m10["P2M"]["peak_tflops"] == bench.FP64_PEAK_TFLOPS
    assert rmx["M2M"]["gemm_gflop"] == r64["M2M"]["gemm_gflop"]
    for r in rates.values():
        for k in ("P2M", "L2P", "M2M", "L2L", "M2L_hadamard"):
            assert 0 < r[k]["frac_mfma"] < 1
        for k in ("M2L_up", "M2L_down", "P2P"):
            assert 0 < r[k]["frac_hbm"] < 1


def test_hadamard_flops_from_the_parent_pairs(bench, fmm):
    plan = _host_plan(fmm, fmm.PREC_MIXED)
    fl, pairs = bench.hadamard_flops(plan)
    tar, src, off, img = plan.m2l_list()
    assert pairs == int(np.sum(img == 0)) and pairs > 0
    assert fl == 8.0 * 64.0 * plan.counts()["n_freq"] * pairs
    plan.close()


def test_workloads(bench):
    x, y, z, q = bench.plummer_workload(100000, 7)
    for a in (x, y, z):
        assert a.min() >= -0.5 and a.max() < 0.5
    assert np.sum(q == 1.0) == np.sum(q == -1.0) == 50000
    r = np.sqrt(x * x + y * y + z * z)
    assert np.median(r) < 0.1   # concentrated: the Plummer sphere's half-mass radius is 0.065
    counts, Ps, pts, qq = bench.batch_workload(48, 4096, 262144, 49)
    assert len(counts) == 48 and counts.min() >= 4096 and counts.max() <= 262144
    assert set(Ps.tolist()) <= {6, 8} and len(pts) == len(qq) == counts.sum()
    off = np.concatenate([[0], np.cumsum(counts)])
    for k in range(48):
        assert abs(qq[off[k]:off[k + 1]].sum()) < 1e-9


def test_committed_traffic_figure(bench):
    t = bench.profiled_traffic("mixed")
    assert t is not None and 1.0e9 < t < 2.0e9


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""
