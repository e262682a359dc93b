#!/usr/bin/env python3
"""Benchmark of the periodic Laplace KIFMM evaluation path on one MI355X.

Workload (the configuration the headline metric is quoted on): N = 1,048,576 bodies uniform in the
unit cube from a seeded PCG64, zero-sum uniform [-1, 1] charges, P = 8, ncrit = 64, periodic with
3 image levels.  The unit cube maps to the reference frame by x - 0.5 with cycle = 1.

One step = fmm3dpp_solve_device: the adaptive octree and P2P/M2L lists built on the GPU + the evaluation
(P2M, M2M, M2L, L2L, L2P, P2P, dipole correction) with positions and charges already in HBM; the near-field
list build runs on a second stream beside the far-field stages.  The one-time
operator precompute depends only on (P, images) and is excluded (reported separately).  Both
precision paths are timed the same way on the same inputs; `value` is the mixed path's throughput
(`--path fp64` reports the float64 path instead), and the line carries the mixed-vs-float64
difference of potential and gradient.

The `batch` object times many uneven systems (48 by default, N log-uniform in [4096, 262144], P 6 or
8, images 2, mixed precision): one batch (every FMM stage one launch per P group, replayed from a
HIP graph) against one plan per system evaluated one after another, particles set once, `steps`
evaluations each.  Prints one JSON line.
"""
import argparse
import json
import os
import shutil
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "exafmm-rd_amd"))

FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 matrix
HBM_PEAK_GBS = 8000.0


def log(msg):
    """Progress on stderr (the JSON result line is the only stdout output)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def workload(n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    pts = rng.random((n, 3)) - 0.5
    q = rng.uniform(-1.0, 1.0, n)
    q -= q.mean()
    q[-1] -= q.sum()          # exact zero net charge
    return pts[:, 0].copy(), pts[:, 1].copy(), pts[:, 2].copy(), q


def hadamard_flops(plan):
    """Algorithmic flops of one 8x8 Hadamard launch: 8 flops per complex MAC, 64 MACs per (pair, f)."""
    import pyfmm3dpp
    tar, src, off, img = plan.m2l_list()
    pairs = int(np.sum(img == 0))
    c = plan.counts()
    return 8.0 * 64.0 * c["n_freq"] * pairs, pairs


def cpu_info():
    """Model, sockets and cores of this host from lscpu (empty fields if lscpu is missing)."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return info
    for line in out.splitlines():
        k, _, v = line.partition(":")
        k, v = k.strip(), v.strip()
        if k == "Model name":
            info["model"] = v
        elif k == "Socket(s)":
            info["sockets"] = int(v)
        elif k == "Core(s) per socket":
            info["cores_per_socket"] = int(v)
        elif k == "Thread(s) per core":
            info["threads_per_core"] = int(v)
    return info


def cpu_baseline(x, y, z, q, P, ncrit, images, threads, repeats=3):
    """The CPU restatement (oracle/, C++17 + OpenMP, built -O3 -march=x86-64-v3) timed on this host's
    cores on the same inputs: one full step (tree + lists + every stage, precompute excluded), median
    of `repeats`, on `threads` threads and again on 8 (the reference's logged runs use 8).  Its
    fidelity is checked in the same run: the energy of the reference's N=24000 drand48 fixture
    (P=8, 3 image levels, ncrit 128, box 2 pi) against the logged 2.758374398027e+01."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    def timed(nthreads):
        f = oracle.OracleFMM(P, ncrit, images, 1.0, dipole=1, threads=nthreads)
        f.precompute()
        ts = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            f.set_particles(x, y, z)
            f.evaluate(q)
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)), ts

    dt, ts = timed(threads)
    dt8, ts8 = timed(8)
    # the reference's logged run: N=24000 drand48 bodies, P=8, 3 image levels, ncrit 128, box 2 pi,
    # 8 threads of an EPYC 7453; timed here the same way (tree + lists + stages) on 8 threads
    fx, fy, fz, fq = oracle.drand48_bodies(24000, np.pi, seed=5, zero_netcharge=1)
    fix = oracle.OracleFMM(8, 128, 3, 2 * np.pi, dipole=1, threads=8)
    fix.precompute()
    fts = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        fix.set_particles(fx, fy, fz)
        fp, _ = fix.evaluate(fq)
        fts.append(time.perf_counter() - t0)
    energy = float(np.sum(fq * fp))
    logged = 2.758374398027e+01   # 3dpp/exp/slurm-153086.out:2899
    # build_and_traverse 0.02466 s + FMM_kernels 0.03467 s (slurm-153086.out:2869,2888)
    logged_s = 0.02466 + 0.03467
    return {"value": round(len(x) / dt, 1), "unit": "particles/s", "cores": threads, "kind": "port",
            "sample": f"one full step (tree + lists + all stages, precompute excluded) of the same N={len(x)} "
                      f"workload, median of {repeats}",
            "seconds": [round(t, 3) for t in ts],
            "threads8": {"value": round(len(x) / dt8, 1), "seconds": [round(t, 3) for t in ts8]},
            "host": cpu_info(),
            "fixture_energy": {"value": energy, "logged": logged, "rel_diff": float("%.2e" % (abs(energy - logged) / logged))},
            "fixture_time_8_threads": {"seconds": round(float(np.median(fts)), 4), "reference_logged_seconds": logged_s,
                                       "note": "tree + lists + stages of the N=24000 P=8 images=3 fixture; "
                                               "the reference's figure is from its own log on another CPU"}}


def plummer_workload(n, seed):
    """Plummer sphere of scale radius 0.05 centred at (0.5, 0.5, 0.5), wrapped into the periodic unit
    cube, charges +-1 in equal numbers by a seeded shuffle; returned in the reference frame (x - 0.5)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    u = np.clip(rng.random(n), 1e-12, 1.0 - 1e-12)
    r = 0.05 / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    cz = rng.uniform(-1.0, 1.0, n)
    ph = rng.uniform(0.0, 2 * np.pi, n)
    st = np.sqrt(1.0 - cz * cz)
    pts = np.stack([r * st * np.cos(ph), r * st * np.sin(ph), r * cz], axis=1) + 0.5
    pts = np.mod(pts, 1.0) - 0.5
    q = np.ones(n)
    q[: n // 2] = -1.0
    rng.shuffle(q)
    return pts[:, 0].copy(), pts[:, 1].copy(), pts[:, 2].copy(), q


def run_clustered(pyfmm3dpp, torch, args):
    """N = 4,194,304 Plummer bodies, P=8, adaptive tree with ncrit=128, 3 image levels, mixed
    precision: step time, stage breakdown and stage rates, and the difference to the float64 path."""
    n = args.clustered_n
    x, y, z, q = plummer_workload(n, args.seed + 3)
    dev = torch.device("cuda", 0)
    d_x, d_y, d_z, d_q = (torch.from_numpy(a).to(dev) for a in (x, y, z, q))
    sub = argparse.Namespace(**vars(args))
    sub.n, sub.P, sub.ncrit, sub.images = n, 8, 128, 3
    rmx = run_path(pyfmm3dpp, torch, sub, pyfmm3dpp.PREC_MIXED, d_x, d_y, d_z, d_q, steps=3, warmup=1)
    c = rmx["plan"].counts()
    ints, _ = rmx["plan"].cells()
    rmx = finish(rmx, True)
    r64 = finish(run_path(pyfmm3dpp, torch, sub, pyfmm3dpp.PREC_FP64, d_x, d_y, d_z, d_q, steps=1, warmup=1), False)
    leafn = ints[: c["n_tree_cells"], 6][ints[: c["n_tree_cells"], 4] == 0]
    # the adaptive tree's accuracy: the float64 FMM over one image level (27 boxes, no dipole term)
    # against the direct sum over the same boxes on 1024 sampled targets
    tar = np.sort(np.random.Generator(np.random.PCG64(args.seed + 5)).choice(n, 1024, replace=False)).astype(np.int32)
    p1 = pyfmm3dpp.LaplaceFMM(8, 128, 1, 1.0, dipole_mode=pyfmm3dpp.DIPOLE_OFF, device=0)
    p1.precompute()
    ph1, gr1 = p1.solve(x, y, z, q)
    n_lost_images1 = int(p1.counts()["n_lost"])
    p1.close()
    pd, gd = pyfmm3dpp.direct(x, y, z, q, 1, 1.0, targets=tar)
    vs_direct = {"phi": float("%.3e" % (np.linalg.norm(ph1[tar] - pd) / np.linalg.norm(pd))),
                 "grad": float("%.3e" % (np.linalg.norm(gr1[tar] - gd) / np.linalg.norm(gd))),
                 "targets": 1024, "images": 1, "n_lost": n_lost_images1}
    return {"workload": f"Plummer a=0.05 wrapped, N={n}, +-1 charges, P=8, ncrit=128, images=3, mixed",
            "particles_per_s": round(n / (rmx["ms_per_step"] / 1e3), 1), "ms_per_step": round(rmx["ms_per_step"], 3),
            "fp64_ms_per_step": round(r64["ms_per_step"], 3),
            "tree": {"cells": int(c["n_tree_cells"]), "leaves": int(len(leafn)),
                     "max_depth": int(ints[: c["n_tree_cells"], 2].max()),
                     "leaf_bodies_min_mean_max": [int(leafn.min()), round(float(leafn.mean()), 1), int(leafn.max())],
                     "n_lost": int(c["n_lost"])},
            "fp64_vs_direct": vs_direct,
            "breakdown_ms": rmx["stages"],
            "stage_rates": rmx["rates"],
            "phi_rel_l2_mixed_vs_fp64": float("%.3e" % rel_l2(torch, rmx["phi"], r64["phi"])),
            "grad_rel_l2_mixed_vs_fp64": float("%.3e" % rel_l2(torch, rmx["grad"], r64["grad"]))}


def batch_workload(nsys, nmin, nmax, seed):
    """nsys systems, N log-uniform in [nmin, nmax], P alternating 6 / 8, zero-sum charges."""
    rng = np.random.Generator(np.random.PCG64(seed))
    counts = np.exp(rng.uniform(np.log(nmin), np.log(nmax), nsys)).astype(np.int64)
    Ps = rng.choice([6, 8], nsys)
    xs, qs = [], []
    for n in counts:
        pts = rng.random((int(n), 3)) - 0.5
        q = rng.uniform(-1.0, 1.0, int(n))
        q -= q.mean()
        xs.append(pts)
        qs.append(q)
    return counts, Ps, np.concatenate(xs), np.concatenate(qs)


def run_batch(pyfmm3dpp, torch, args):
    """Many uneven systems: one Batch (merged launches, HIP graph) against one plan per system run
    one after another; both time `steps` evaluations with the particles set once."""
    from pyfmm3dpp.batch import Batch
    counts, Ps, pts, q = batch_workload(args.batch_systems, 4096, 262144, args.seed + 7)
    n = int(counts.sum())
    dev = torch.device("cuda", 0)
    d_x, d_y, d_z = (torch.from_numpy(pts[:, k].copy()).to(dev) for k in range(3))
    d_q = torch.from_numpy(q).to(dev)
    off = np.concatenate([[0], np.cumsum(counts)])
    prec = pyfmm3dpp.PREC_MIXED
    # sequential: one plan per system
    plans, cache = [], {}
    tmp = tempfile.mkdtemp(prefix="fmm3dpp_ops_")
    for k in range(len(counts)):
        P = int(Ps[k])
        p = pyfmm3dpp.LaplaceFMM(P, args.ncrit, 2, 1.0, dipole_mode=pyfmm3dpp.DIPOLE_PHYSICAL,
                                 precision=prec, device=0)
        if P in cache:               # operators depend on (P, images, cycle) only
            p.load_operators(cache[P])
        else:
            p.precompute()
            cache[P] = os.path.join(tmp, f"ops_P{P}.bin")
            p.save_operators(cache[P])
        p.set_near_field_stream(0)   # one stream per plan until the two-stream variant below
        s = slice(int(off[k]), int(off[k + 1]))
        p.set_particles_device(int(counts[k]), d_x[s].data_ptr(), d_y[s].data_ptr(), d_z[s].data_ptr())
        plans.append((p, s))
    phi_s = torch.empty(n, dtype=torch.float64, device=dev)
    grad_s = torch.empty((n, 3), dtype=torch.float64, device=dev)

    def seq_step():
        for p, s in plans:
            p.evaluate_device(d_q[s].data_ptr(), phi_s[s].data_ptr(), grad_s[s].data_ptr())

    def near_two_streams():   # each plan's second stream is created at its next evaluation
        for p, _ in plans:
            p.set_near_field_stream(1)

    b = Batch([(int(P), args.ncrit) for P in Ps], 2, 1.0, dipole_mode=pyfmm3dpp.DIPOLE_PHYSICAL, precision=prec)
    t0 = time.perf_counter()
    b.set_particles(counts, d_x, d_y, d_z)
    set_s = time.perf_counter() - t0
    b.set_graph(True)
    phi_b = torch.empty(n, dtype=torch.float64, device=dev)
    grad_b = torch.empty((n, 3), dtype=torch.float64, device=dev)

    def batch_step():
        b.evaluate(d_q, phi_b, grad_b)

    out = {}
    for name, step, setup in (("sequential", seq_step, None), ("sequential_two_streams", seq_step, near_two_streams),
                              ("batch", batch_step, None)):
        if setup:
            setup()
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        out[name] = (time.perf_counter() - t0) / args.steps
    seq_best = min(out["sequential"], out["sequential_two_streams"])
    diff = max(rel_l2(torch, phi_b, phi_s), rel_l2(torch, grad_b, grad_s))
    shutil.rmtree(tmp, ignore_errors=True)
    return {"systems": int(len(counts)), "bodies": n, "P": sorted(set(int(p) for p in Ps)),
            "sequential_particles_per_s": round(n / seq_best, 1),
            "batch_particles_per_s": round(n / out["batch"], 1),
            "speedup": round(seq_best / out["batch"], 3),
            "batch_ms_per_step": round(out["batch"] * 1e3, 3),
            "sequential_ms_per_step": round(seq_best * 1e3, 3),
            "sequential_ms_per_step_near_field_one_stream": round(out["sequential"] * 1e3, 3),
            "sequential_ms_per_step_near_field_two_streams": round(out["sequential_two_streams"] * 1e3, 3),
            "batch_set_particles_s": round(set_s, 3),
            "rel_l2_batch_vs_sequential": float("%.3e" % diff)}


def stage_rooflines(plan, stages, mixed):
    """Achieved rates of each stage from its algorithmic work (counted from the plan's own tree and
    lists) and its HIP-event time: flop rates against the matrix-core peak of the arithmetic the stage
    runs in for the operator applies (GEMM flops over the whole stage time, so P2M also pays its check
    potentials and L2P its field evaluation) and the Hadamard, bytes for the transforms and the near
    field."""
    ints, dbl = plan.cells()
    c = plan.counts()
    ntree = c["n_tree_cells"]
    cnum, bnum, depth = ints[:ntree, 4], ints[:ntree, 6], ints[:ntree, 2]
    leaf = cnum == 0
    nleaf, nparent = int(leaf.sum()), int((~leaf).sum())
    S, F = c["S"], c["n_freq"]
    tar, src, _ = plan.p2p_list()
    bn = ints[:, 6].astype(np.int64)
    pairs = float(np.sum(bn[tar] * bn[src]))
    flops_op = 2.0 * S * S
    # the mixed path applies the P2M / L2P pseudo-inverses at the truncated rank (singular values
    # above 1e-7 of the largest, rounded up to 16) and, for P = 6..8, all tree applies in float32
    rank = S
    f32_applies = mixed and 6 <= plan.P <= 8
    if mixed:
        vs = plan.get_operator("VSinv").reshape(S, S)
        n2 = (vs * vs).sum(0)
        sv = np.where(n2 > 0, 1.0 / np.sqrt(np.where(n2 > 0, n2, 1.0)), 0.0)
        rank = min(S, (int(np.nonzero(sv > 1e-7 * sv.max())[0].max()) + 1 + 15) // 16 * 16)
    pinv = 2 * (2.0 * rank * S)
    gemm = {"P2M": pinv * nleaf, "L2P": pinv * nleaf, "M2M": flops_op * (ntree - 1), "L2L": flops_op * (ntree - 1)}
    gpeak = FP32_PEAK_TFLOPS if f32_applies else FP64_PEAK_TFLOPS
    out = {}
    for k, fl in gemm.items():
        t = stages[k] * 1e-3
        out[k] = {"gemm_gflop": round(fl / 1e9, 2), "tflops": round(fl / t / 1e12, 2), "peak_tflops": gpeak,
                  "frac_mfma": round(fl / t / 1e12 / gpeak, 3)}
    had, npairs = hadamard_flops(plan)
    t = stages["M2L_hadamard"] * 1e-3
    peak = FP32_PEAK_TFLOPS if mixed else FP64_PEAK_TFLOPS
    out["M2L_hadamard"] = {"gflop": round(had / 1e9, 2), "tflops": round(had / t / 1e12, 2),
                           "frac_mfma": round(had / t / 1e12 / peak, 3)}
    csz = 8 if mixed else 16   # bytes per complex value of the spectra
    spec = float(nparent) * F * 8 * csz
    for k in ("M2L_up", "M2L_down"):
        t = stages[k] * 1e-3
        out[k] = {"spectrum_gb": round(spec / 1e9, 3), "gbs": round(spec / t / 1e9, 1),
                  "frac_hbm": round(spec / t / 1e9 / HBM_PEAK_GBS, 3)}
    n = int(bnum[0])
    w = 4 if mixed else 8   # bytes per word: targets 4 words in and 4 out, sources 4 words per row
    p2p_bytes = n * 8.0 * w + float(np.sum(bn[src])) * 4 * w
    t = stages["P2P"] * 1e-3
    out["P2P"] = {"pairs": pairs, "gpairs_per_s": round(pairs / t / 1e9, 1),
                  "gbs_algorithmic": round(p2p_bytes / t / 1e9, 1),
                  "frac_hbm": round(p2p_bytes / t / 1e9 / HBM_PEAK_GBS, 3)}
    return out


def profiled_traffic(path):
    """HBM bytes per launch of the dominant kernel from the committed counter summary (FETCH_SIZE +
    WRITE_SIZE, each divided by the same counter's reading on a copy kernel of known size), or None."""
    f = os.path.join(ROOT, "profiles", f"hadamard8_{path}_N1M_P8_hbm_traffic.md")
    if not os.path.exists(f):
        return None
    import re
    m = re.search(r"^traffic_bytes_per_launch: (\d+)", open(f).read(), re.M)
    return int(m.group(1)) if m else None


STAGES = ["P2M", "M2M", "M2L_up", "M2L_hadamard", "M2L_down", "L2L", "L2P", "P2P", "finalize", "evaluate",
          "near_start", "far_end"]


def run_path(pyfmm3dpp, torch, args, precision, d_x, d_y, d_z, d_q, images=None, steps=None, warmup=None):
    """Warm up, then time exactly `steps` steps (solve_device: tree, lists, all stages) of one
    precision path; per-stage times come from HIP events on the plan's stream."""
    n = args.n
    images = args.images if images is None else images
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    plan = pyfmm3dpp.LaplaceFMM(args.P, args.ncrit, images, 1.0, dipole_mode=pyfmm3dpp.DIPOLE_PHYSICAL,
                                precision=precision, device=0)
    t0 = time.perf_counter()
    plan.precompute()
    precompute_s = time.perf_counter() - t0
    dev = d_x.device
    d_phi = torch.empty(n, dtype=torch.float64, device=dev)
    d_grad = torch.empty((n, 3), dtype=torch.float64, device=dev)

    def step():   # tree + lists + all stages, the list build overlapping the far field
        plan.solve_device(n, d_x.data_ptr(), d_y.data_ptr(), d_z.data_ptr(), d_q.data_ptr(), d_phi.data_ptr(),
                          d_grad.data_ptr())

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    stages = {k: 0.0 for k in STAGES}
    set_ms = 0.0
    per_step = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        step()
        for k in STAGES:
            stages[k] += plan.timing_ms(k)   # HIP events on the plan's stream (synchronises)
        set_ms += plan.timing_ms("tree")   # host wall of the tree build ahead of the far field
        per_step.append((time.perf_counter() - ts) * 1e3)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = dt * 1e3 / max(steps, 1)
    stages = {k: round(v / max(steps, 1), 4) for k, v in stages.items()}
    stages["tree_build"] = round(set_ms / max(steps, 1), 3)
    return {"plan": plan, "ms_per_step": ms, "median_ms_per_step": float(np.median(per_step)) if per_step else ms,
            "stages": stages, "precompute_s": precompute_s, "phi": d_phi, "grad": d_grad}


def finish(r, mixed):
    """Stage rates and Hadamard work from the plan of a finished run, then release the plan."""
    r["rates"] = stage_rooflines(r["plan"], r["stages"], mixed)
    r["had"] = hadamard_flops(r["plan"])
    r["plan"].close()
    r["plan"] = None
    return r


def rel_l2(torch, a, b):
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--P", type=int, default=8)
    ap.add_argument("--ncrit", type=int, default=64)
    ap.add_argument("--images", type=int, default=3)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--path", choices=["mixed", "fp64"], default="mixed",
                    help="path whose throughput is the reported value (the other is timed beside it)")
    ap.add_argument("--single-path", action="store_true",
                    help="time only the --path path, no accuracy legs (for kernel profiles of one path)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--batch-systems", type=int, default=48,
                    help="systems of the batch leg (N log-uniform 4096..262144, P 6/8, images 2, mixed)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--clustered-n", type=int, default=1 << 22,
                    help="bodies of the clustered leg (Plummer sphere, adaptive tree, ncrit 128)")
    ap.add_argument("--no-clustered", action="store_true")
    ap.add_argument("--workload", choices=["uniform", "plummer"], default="uniform",
                    help="inputs of --single-path runs: the uniform N=--n workload or the clustered leg's")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run one plain copy kernel of 1 GiB read + 1 GiB written (elementwise x * 1.0), "
                         "the reference point for the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: the evaluation path is single-device (asked for {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import pyfmm3dpp

    if args.workload == "plummer":   # the clustered leg's inputs, for profiles of that workload
        args.n, args.ncrit = args.clustered_n, 128
        x, y, z, q = plummer_workload(args.n, args.seed + 3)
    else:
        x, y, z, q = workload(args.n, args.seed)
    dev = torch.device("cuda", 0)
    if args.copy_calibration:
        src = torch.ones(1 << 27, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    d_x, d_y, d_z = (torch.from_numpy(a).to(dev) for a in (x, y, z))
    d_q = torch.from_numpy(q).to(dev)
    torch.cuda.synchronize()

    if args.single_path:
        prec = pyfmm3dpp.PREC_MIXED if args.path == "mixed" else pyfmm3dpp.PREC_FP64
        r = run_path(pyfmm3dpp, torch, args, prec, d_x, d_y, d_z, d_q)
        print(json.dumps({"path": args.path, "ms_per_step": round(r["ms_per_step"], 3),
                          "particles_per_s": round(args.n / (r["ms_per_step"] / 1e3), 1),
                          "breakdown_ms": r["stages"]}), flush=True)
        return
    # one plan alive at a time: each plan has two HIP streams, and with more streams than hardware
    # queues a plan's second stream may share a queue with its first (no overlap of the near field)
    r64 = finish(run_path(pyfmm3dpp, torch, args, pyfmm3dpp.PREC_FP64, d_x, d_y, d_z, d_q), False)
    log(f"float64 path {r64['ms_per_step']:.3f} ms/step")
    rmx = finish(run_path(pyfmm3dpp, torch, args, pyfmm3dpp.PREC_MIXED, d_x, d_y, d_z, d_q), True)
    log(f"mixed path {rmx['ms_per_step']:.3f} ms/step")
    main_run = rmx if args.path == "mixed" else r64
    value = args.n / (main_run["ms_per_step"] / 1e3)
    v64 = args.n / (r64["ms_per_step"] / 1e3)
    vmx = args.n / (rmx["ms_per_step"] / 1e3)
    err = {"phi_rel_l2_mixed_vs_fp64": rel_l2(torch, rmx["phi"], r64["phi"]),
           "grad_rel_l2_mixed_vs_fp64": rel_l2(torch, rmx["grad"], r64["grad"])}
    # accuracy against a float64 Ewald sum (alpha L = 12, |k_i| <= 23, r_c = L/2) on 4096 sampled targets
    tar = np.sort(np.random.Generator(np.random.PCG64(args.seed + 1)).choice(args.n, 4096, replace=False))
    d_tar = torch.from_numpy(tar.astype(np.int32)).to(dev)
    d_pe = torch.empty(len(tar), dtype=torch.float64, device=dev)
    d_ge = torch.empty((len(tar), 3), dtype=torch.float64, device=dev)
    t0 = time.perf_counter()
    rc = pyfmm3dpp.lib().fmm3dpp_ewald_device(None, args.n, d_x.data_ptr(), d_y.data_ptr(), d_z.data_ptr(),
                                              d_q.data_ptr(), len(tar), d_tar.data_ptr(), 1.0, 23, 12.0, 0.5,
                                              d_pe.data_ptr(), d_ge.data_ptr())
    ewald_s = time.perf_counter() - t0
    log(f"Ewald reference on {len(tar)} targets {ewald_s:.2f} s")
    if rc != 0:
        raise RuntimeError(f"fmm3dpp_ewald_device failed ({rc})")
    idx = torch.from_numpy(tar).to(dev)
    def phi_err(phi):   # periodic potentials agree up to an additive constant: also report with it removed
        d = phi[idx] - d_pe
        return rel_l2(torch, phi[idx], d_pe), rel_l2(torch, phi[idx] - d.mean(), d_pe)

    for name, r in (("fp64", r64), ("mixed", rmx)):
        err[f"phi_rel_l2_{name}_vs_ewald"], err[f"phi_rel_l2_{name}_vs_ewald_mod_const"] = phi_err(r["phi"])
        err[f"grad_rel_l2_{name}_vs_ewald"] = rel_l2(torch, r["grad"][idx], d_ge)
    # the same at 5 and 6 image levels, where the lattice remainder drops below the FMM's own error
    for im in (5, 6):
        for name, prec in (("fp64", pyfmm3dpp.PREC_FP64), ("mixed", pyfmm3dpp.PREC_MIXED)):
            r5 = run_path(pyfmm3dpp, torch, args, prec, d_x, d_y, d_z, d_q, images=im, steps=1, warmup=1)
            r5["plan"].close()
            err[f"phi_rel_l2_{name}_vs_ewald_images{im}"], err[f"phi_rel_l2_{name}_vs_ewald_images{im}_mod_const"] = \
                phi_err(r5["phi"])
            err[f"grad_rel_l2_{name}_vs_ewald_images{im}"] = rel_l2(torch, r5["grad"][idx], d_ge)
            if name == "mixed":
                err[f"mixed_ms_per_step_images{im}"] = round(r5["ms_per_step"], 3)
    err["ewald_seconds"] = round(ewald_s, 3)

    flops, pairs = main_run["had"]
    had_ms = main_run["stages"]["M2L_hadamard"]
    achieved = flops / (had_ms / 1e3) / 1e12
    mixed = args.path == "mixed"
    peak = FP32_PEAK_TFLOPS if mixed else FP64_PEAK_TFLOPS
    roofline = {"kernel": "k_hadamard8_staged<%s> (M2L 8x8 Hadamard on the matrix cores, sources staged in LDS)"
                          % ("float" if mixed else "double"),
                "bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                "frac": round(achieved / peak, 4), "traffic": profiled_traffic(args.path),
                "flops_per_launch": flops, "pairs": pairs,
                "launch_ms": had_ms}

    out = {
        "metric": "particles/sec, N=1M P=8 periodic: mixed vs FP64 GPU at rel L2 err \u22641e-5",
        "value": round(value, 1),
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_run["ms_per_step"], 3),
        "median_ms_per_step": round(main_run["median_ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if mixed else "f64",
        "data": "synthetic",
        "config": {"workload": f"periodic Laplace KIFMM, N={args.n} uniform unit cube (PCG64 seed {args.seed}), "
                               f"zero-sum uniform charges, P={args.P}, ncrit={args.ncrit}, images={args.images}",
                   "path": args.path, "tree_and_lists": "GPU", "inputs": "positions and charges resident in HBM",
                   "parallelism": "single GPU"},
        "fp64_particles_per_s": round(v64, 1),
        "mixed_particles_per_s": round(vmx, 1),
        "mixed_speedup_vs_fp64": round(vmx / v64, 3),
        "accuracy": {k: float("%.3e" % v) for k, v in err.items()},
        "gauge": "infinite-lattice (DIPOLE_PHYSICAL), compared with the tin-foil Ewald sum",
        "breakdown_ms": {"mixed": rmx["stages"], "fp64": r64["stages"]},
        "precompute_s": round(main_run["precompute_s"], 3),
        "roofline": roofline,
        "stage_rates": {"mixed": rmx["rates"], "fp64": r64["rates"]},
    }
    log("accuracy legs done")
    if not args.no_clustered:
        out["clustered"] = run_clustered(pyfmm3dpp, torch, args)
        log(f"clustered leg: {out['clustered']['ms_per_step']} ms/step")
    if not args.no_batch:
        out["batch"] = run_batch(pyfmm3dpp, torch, args)
        log(f"batch leg: {out['batch']['speedup']}x over one plan per system")
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or len(os.sched_getaffinity(0))
        out["cpu_baseline"] = cpu_baseline(x, y, z, q, args.P, args.ncrit, args.images, threads)
        log(f"CPU baseline {out['cpu_baseline']['value']:.3g} particles/s on {threads} threads")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
