"""Batches of independent systems (fmm3dpp_batch_*) against one plan per system.

A batch merges the tables of all systems of equal P and runs each FMM stage as one launch; every
output element is computed by the same arithmetic in the same order as in the system's own plan
(the dipole moments use the same block partition), so the results agree to the last bits.  The
bound written here, 1e-13 relative L2 per system, only absorbs rounding; the printed differences
are expected to be 0.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TOL = 1e-13


@pytest.fixture(scope="module")
def fmm():
    import pyfmm3dpp
    pyfmm3dpp.lib()
    return pyfmm3dpp


def _systems(counts, seed, clustered=False):
    rng = np.random.Generator(np.random.PCG64(seed))
    xs, qs = [], []
    for n in counts:
        if clustered:
            x = np.concatenate([rng.normal(0.1, 0.05, (n // 2, 3)), rng.random((n - n // 2, 3)) - 0.5])
            x = (x + 0.5) % 1.0 - 0.5
        else:
            x = rng.random((n, 3)) - 0.5
        q = rng.uniform(-1, 1, n)
        q -= q.mean()
        xs.append(x)
        qs.append(q)
    return xs, qs


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _per_plan(fmm, systems, xs, qs, images, precision, dipole_mode, divide_4pi=0):
    out = []
    for (P, ncrit), x, q in zip(systems, xs, qs):
        plan = fmm.LaplaceFMM(P, ncrit, images, 1.0, precision=precision, dipole_mode=dipole_mode,
                              divide_4pi=divide_4pi, device=0)
        plan.precompute()
        plan.set_particles(x[:, 0], x[:, 1], x[:, 2])
        out.append(plan.evaluate(q))
        plan.close()
    return out


def _batch_run(fmm, systems, xs, qs, images, precision, dipole_mode, divide_4pi=0, graph=False):
    from pyfmm3dpp.batch import Batch
    b = Batch(systems, images, 1.0, precision=precision, dipole_mode=dipole_mode, divide_4pi=divide_4pi)
    x = np.concatenate(xs)
    b.set_particles([len(v) for v in xs], x[:, 0], x[:, 1], x[:, 2])
    b.set_graph(graph)
    phi, grad = b.evaluate(np.concatenate(qs))
    return b, phi.cpu().numpy(), grad.cpu().numpy()


def _compare(phi, grad, ref, counts):
    off = np.concatenate([[0], np.cumsum(counts)])
    worst = 0.0
    for k, (p_ref, g_ref) in enumerate(ref):
        ep = _rel(phi[off[k]:off[k + 1]], p_ref)
        eg = _rel(grad[off[k]:off[k + 1]], g_ref)
        worst = max(worst, ep, eg)
    print(f"  batch vs per-system plans: worst rel L2 {worst:.2e}", flush=True)
    return worst


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_matches_plans_uneven(fmm, precision):
    """Six systems, N 700..45000, P 4 and 6 interleaved, ncrit 32/64, images=2."""
    counts = [3000, 20000, 700, 45000, 9000, 1500]
    systems = [(4, 32), (6, 64), (6, 32), (4, 64), (6, 64), (4, 32)]
    print(f"case counts={counts} precision={precision}", flush=True)
    xs, qs = _systems(counts, 11 + precision)
    ref = _per_plan(fmm, systems, xs, qs, 2, precision, fmm.DIPOLE_REFERENCE)
    _, phi, grad = _batch_run(fmm, systems, xs, qs, 2, precision, fmm.DIPOLE_REFERENCE)
    assert _compare(phi, grad, ref, counts) <= TOL


@pytest.mark.parametrize("images,dipole,divide", [(0, 0, 0), (1, 2, 1), (3, 1, 0)])
def test_batch_modes(fmm, images, dipole, divide):
    """Free space, the infinite-lattice gauge with 1/(4 pi), and three image levels; clustered bodies."""
    counts = [5000, 12000, 2500]
    systems = [(5, 48), (5, 48), (5, 24)]
    print(f"case images={images} dipole={dipole} divide_4pi={divide}", flush=True)
    xs, qs = _systems(counts, 100 + images, clustered=True)
    ref = _per_plan(fmm, systems, xs, qs, images, 0, dipole, divide)
    _, phi, grad = _batch_run(fmm, systems, xs, qs, images, 0, dipole, divide)
    assert _compare(phi, grad, ref, counts) <= TOL


def test_batch_single_system(fmm):
    counts = [30000]
    systems = [(8, 64)]
    xs, qs = _systems(counts, 5)
    ref = _per_plan(fmm, systems, xs, qs, 3, 1, fmm.DIPOLE_PHYSICAL)
    _, phi, grad = _batch_run(fmm, systems, xs, qs, 3, 1, fmm.DIPOLE_PHYSICAL)
    assert _compare(phi, grad, ref, counts) <= TOL


def test_batch_graph_replays(fmm):
    """Graph capture and three replays with new charge values (same buffers), eager work and
    synchronisations between the replays; every replay against the uncaptured evaluation."""
    import torch
    from pyfmm3dpp.batch import Batch
    counts = [4096, 16000, 9000, 2000]
    systems = [(6, 64), (8, 64), (6, 32), (8, 64)]
    xs, qs = _systems(counts, 21)
    x = np.concatenate(xs)
    dev = torch.device("cuda", 0)
    eager = Batch(systems, 2, 1.0, precision=1, dipole_mode=fmm.DIPOLE_REFERENCE)
    eager.set_particles(counts, x[:, 0], x[:, 1], x[:, 2])
    graphed = Batch(systems, 2, 1.0, precision=1, dipole_mode=fmm.DIPOLE_REFERENCE)
    graphed.set_particles(counts, x[:, 0], x[:, 1], x[:, 2])
    graphed.set_graph(True)
    q = torch.empty(sum(counts), dtype=torch.float64, device=dev)
    phi = torch.empty(sum(counts), dtype=torch.float64, device=dev)
    grad = torch.empty((sum(counts), 3), dtype=torch.float64, device=dev)
    rng = np.random.Generator(np.random.PCG64(77))
    for rep in range(4):
        qn = rng.uniform(-1, 1, sum(counts))
        q.copy_(torch.from_numpy(qn))
        junk = torch.randn(1 << 20, device=dev).sum()   # eager work between replays
        torch.cuda.synchronize()
        p_ref, g_ref = eager.evaluate(q)
        graphed.evaluate(q, phi, grad)
        e = max(_rel(phi.cpu().numpy(), p_ref.cpu().numpy()), _rel(grad.cpu().numpy(), g_ref.cpu().numpy()))
        print(f"  replay {rep}: rel L2 vs uncaptured {e:.2e} ({float(junk):.1f})", flush=True)
        assert e <= TOL


def test_batch_autotune(fmm):
    """The autotuner picks one of the candidates per P group; the batch then equals plans built with
    the chosen leaf sizes."""
    from pyfmm3dpp.batch import Batch
    counts = [20000, 6000, 40000, 3000]
    systems = [(6, 64), (8, 64), (6, 64), (8, 64)]
    xs, qs = _systems(counts, 31)
    x = np.concatenate(xs)
    b = Batch(systems, 2, 1.0, precision=1, dipole_mode=fmm.DIPOLE_REFERENCE)
    b.set_particles(counts, x[:, 0], x[:, 1], x[:, 2])
    q = np.concatenate(qs)
    chosen = b.autotune([32, 64, 128], q)
    print(f"  chosen ncrit {list(chosen)}", flush=True)
    assert set(chosen) <= {32, 64, 128}
    assert chosen[0] == chosen[2] and chosen[1] == chosen[3]   # one choice per P group
    phi, grad = b.evaluate(q)
    tuned = [(P, int(c)) for (P, _), c in zip(systems, chosen)]
    ref = _per_plan(fmm, tuned, xs, qs, 2, 1, fmm.DIPOLE_REFERENCE)
    assert _compare(phi.cpu().numpy(), grad.cpu().numpy(), ref, counts) <= TOL


def test_batch_log_uniform_mixed_p(fmm):
    """Twelve systems, N log-uniform in [4096, 65536], P 6 and 8, images=2, mixed precision."""
    rng = np.random.Generator(np.random.PCG64(48))
    counts = [int(v) for v in np.exp(rng.uniform(np.log(4096), np.log(65536), 12))]
    systems = [(6 if k % 2 == 0 else 8, 64) for k in range(12)]
    xs, qs = _systems(counts, 49)
    ref = _per_plan(fmm, systems, xs, qs, 2, 1, fmm.DIPOLE_REFERENCE)
    _, phi, grad = _batch_run(fmm, systems, xs, qs, 2, 1, fmm.DIPOLE_REFERENCE)
    assert _compare(phi, grad, ref, counts) <= TOL
