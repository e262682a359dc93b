"""GPU Ewald sum (fmm3dpp_ewald) and the infinite-lattice dipole gauge (FMM3DPP_DIPOLE_PHYSICAL).

The Ewald sum is pinned three ways: against the CPU restatement of the reference's EwaldSolver
(per-pair minimum image, same alpha/cutoff/ksize) to 1e-12, against the reference's logged Ewald
energy and sum of phi^2 on its N=24000 drand48 system, and by its independence of the splitting
parameter alpha (a wrong real-space, reciprocal or self term changes with alpha).
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


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _uniform(n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    x = rng.random((n, 3)) - 0.5
    q = rng.uniform(-1, 1, n)
    q -= q.mean()
    return x[:, 0].copy(), x[:, 1].copy(), x[:, 2].copy(), q


def test_ewald_matches_cpu_restatement(fmm, oracle_mod):
    x, y, z, q = _uniform(3000, 5)
    tar = np.arange(0, 3000, 7, dtype=np.int32)
    p, g = fmm.ewald(x, y, z, q, 1.0, 11, targets=tar)
    pr, gr = oracle_mod.ewald(x, y, z, q, 1.0, 11, targets=tar, threads=16)
    print(f"ewald vs cpu: phi {_rel(p, pr):.2e} grad {_rel(g, gr):.2e}", flush=True)
    assert _rel(p, pr) <= 1e-12 and _rel(g, gr) <= 1e-12


def test_ewald_reference_log(fmm, oracle_mod):
    """N=24000 drand48 seed 5 in a 2 pi box, ksize 11: the reference logged sum q phi =
    2.773120969004e+01 and sum phi^2 = 2973740.69023801 (its real space runs over cell pairs, this
    one over body pairs with the same cutoff: 2e-10 apart)."""
    with open(os.path.join(HERE, "golden", "log_energies.json")) as f:
        row = json.load(f)["periodic"][0]
    x, y, z, q = oracle_mod.drand48_bodies(24000, np.pi, seed=5, zero_netcharge=1)
    p, g = fmm.ewald(x, y, z, q, 2 * np.pi, 11)
    e = float(np.sum(q * p))
    print(f"sum q phi {e:.12e} (log {row['energy_ref']:.12e}), sum phi^2 {np.sum(p * p):.8f}", flush=True)
    assert abs(e - row["energy_ref"]) <= 1e-9 * abs(row["energy_ref"])
    assert abs(np.sum(p * p) - row["pnrm"]) <= 1e-9 * row["pnrm"]


def test_ewald_alpha_independent(fmm):
    x, y, z, q = _uniform(20000, 9)
    tar = np.arange(0, 20000, 13, dtype=np.int32)
    p1, g1 = fmm.ewald(x, y, z, q, 1.0, 23, alpha=12.0, cutoff=0.5, targets=tar)
    p2, g2 = fmm.ewald(x, y, z, q, 1.0, 23, alpha=10.0, cutoff=0.5, targets=tar)
    p3, g3 = fmm.ewald(x, y, z, q, 1.0, 20, alpha=12.0, cutoff=0.5, targets=tar)
    print(f"alpha 12 vs 10: phi {_rel(p1, p2):.2e} grad {_rel(g1, g2):.2e}; K 23 vs 20: phi {_rel(p1, p3):.2e}",
          flush=True)
    assert _rel(p1, p2) <= 1e-12 and _rel(g1, g2) <= 1e-12
    assert _rel(p1, p3) <= 1e-13 and _rel(g1, g3) <= 1e-13


def test_physical_gauge_is_analytic_correction(fmm):
    """DIPOLE_PHYSICAL = raw lattice sum (DIPOLE_OFF) - (4 pi/3V) M.x + (2 pi/3V) sum q |x|^2."""
    x, y, z, q = _uniform(8192, 21)
    out = {}
    for mode in (fmm.DIPOLE_OFF, fmm.DIPOLE_PHYSICAL):
        plan = fmm.LaplaceFMM(6, 64, 2, 1.0, dipole_mode=mode, device=0)
        plan.precompute()
        out[mode] = plan.solve(x, y, z, q)
    X = np.stack([x, y, z], 1)
    M = (q[:, None] * X).sum(0)
    c = 4 * np.pi / 3
    p = out[fmm.DIPOLE_OFF][0] - c * (X @ M) + 0.5 * c * np.sum(q * (X * X).sum(1))
    g = out[fmm.DIPOLE_OFF][1] - c * M
    assert _rel(out[fmm.DIPOLE_PHYSICAL][0], p) <= 1e-13 and _rel(out[fmm.DIPOLE_PHYSICAL][1], g) <= 1e-13


@pytest.mark.parametrize("prec", [0, 1])
def test_fmm_physical_gauge_vs_ewald(fmm, prec):
    """The lattice FMM in the infinite-lattice gauge converges to the Ewald sum as images grow (the
    remainder of a (3^i)^3 cube decays like 9^-i); float64 and mixed precision alike.  A periodic
    potential is defined up to an additive constant, which leaves forces and the energy of a neutral
    system unchanged: the FMM carries a constant offset of about 1e-5 of the potential's norm (the
    image levels' approximation error on the constant part of the lattice sum), so the potential is
    compared after removing the mean difference over the targets; the raw figure is printed."""
    x, y, z, q = _uniform(16384, 33)
    tar = np.arange(0, 16384, 4, dtype=np.int32)
    pe, ge = fmm.ewald(x, y, z, q, 1.0, 23, alpha=12.0, cutoff=0.5, targets=tar)
    errs = {}
    for images in (3, 5):
        plan = fmm.LaplaceFMM(8, 64, images, 1.0, dipole_mode=fmm.DIPOLE_PHYSICAL, precision=prec, device=0)
        plan.precompute()
        p, g = plan.solve(x, y, z, q)
        d = p[tar] - pe
        errs[images] = (_rel(p[tar] - np.mean(d), pe), _rel(g[tar], ge))
        print(f"precision {prec} images {images}: phi {errs[images][0]:.2e} (raw {_rel(p[tar], pe):.2e}, "
              f"mean offset {np.mean(d):.3e}) grad {errs[images][1]:.2e}", flush=True)
    assert errs[3][0] <= 1e-3 and errs[3][1] <= 2e-4
    assert errs[5][0] <= 1e-5 and errs[5][1] <= 1e-5
