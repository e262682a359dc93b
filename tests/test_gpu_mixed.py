"""Mixed-precision path (FMM3DPP_PREC_MIXED) against the float64 path on the same plan inputs.

The mixed path runs the M2L spectra and 8x8 Hadamard (matrix cores) and the P2P near field in
float32.  For orders P = 6..8 it also runs the tree operator stages in float32 (P2M check
potentials, the P2M / L2P pseudo-inverse applies at the truncated rank, M2M and L2L on the float32
matrix cores, the L2P evaluation) and sets every tree cell's equivalent-density monopole to the
cell's charge after P2M and M2M; other orders keep those stages in float64.  The image lattice is
float64 in both.  Both paths share the tree, lists and operators.

Bounds: 2e-6 relative L2 for the potential, 1e-5 for the gradient (1e-5 for the potential at
P = 10).  The gradient's share comes from the P2P: float32 leaf-relative coordinates carry an
absolute error of ~6e-8 of the leaf size, a large relative error in the displacement of the closest
pairs, and 1/r^2 weighs those most (a host emulation of the float32 P2P alone gives 4.2e-6 at
N=65536, P=6, ncrit=64 and 8.3e-6 at N=60000, ncrit=128).

In the reference's dipole mode with image levels the root's monopole is not pinned, so the float64
path carries the rounding and truncation monopole of its root equivalent density, amplified by each
image level into a constant potential offset (1.7e-5 at P=6 with 2 levels, measured here and in
the float32 emulation of the restated reference); the mixed path's monopole constraint does not have it.
The potential is compared up to that constant there (periodic potentials are defined up to one);
every other case compares the raw potential.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PHI_TOL, GRAD_TOL = 2e-6, 1e-5


@pytest.fixture(scope="module")
def fmm():
    import pyfmm3dpp
    pyfmm3dpp.lib()
    return pyfmm3dpp


def _rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _uniform(n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    x = rng.random((n, 3)) - 0.5
    q = rng.uniform(-1, 1, n)
    q -= q.mean()
    return x[:, 0].copy(), x[:, 1].copy(), x[:, 2].copy(), q


def _clustered(n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    x = np.concatenate([rng.normal(0.1, 0.05, (n // 2, 3)), rng.random((n - n // 2, 3)) - 0.5])
    x = (x + 0.5) % 1.0 - 0.5
    q = rng.choice([-1.0, 1.0], n)
    q -= q.mean()
    return x[:, 0].copy(), x[:, 1].copy(), x[:, 2].copy(), q


def _both(fmm, x, y, z, q, P, ncrit, images, dipole=1, mixed_applies=1):
    """Potential difference (raw, or up to a constant in the reference dipole mode with images) and
    gradient difference of the mixed path against the float64 path."""
    print(f"case n={len(x)} P={P} ncrit={ncrit} images={images} dipole={dipole}", flush=True)
    out = []
    for prec in (fmm.PREC_FP64, fmm.PREC_MIXED):
        plan = fmm.LaplaceFMM(P, ncrit, images, 1.0, dipole_mode=dipole, precision=prec, device=0)
        if prec == fmm.PREC_MIXED:
            plan.set_mixed_applies(mixed_applies)
        plan.precompute()
        plan.set_particles(x, y, z)
        out.append(plan.evaluate(q))
    (p64, g64), (pmx, gmx) = out
    d = pmx - p64
    ep, epc, eg = _rel_l2(pmx, p64), float(np.linalg.norm(d - d.mean()) / np.linalg.norm(p64)), _rel_l2(gmx, g64)
    print(f"  mixed vs float64: rel L2 phi {ep:.2e} (up to a constant {epc:.2e}) grad {eg:.2e}", flush=True)
    assert np.all(np.isfinite(pmx)) and np.all(np.isfinite(gmx))
    return (epc if (dipole == 1 and images > 0) else ep), eg


@pytest.mark.parametrize("P,images", [(4, 0), (4, 1), (6, 2), (8, 3), (10, 1)])
def test_mixed_vs_fp64_small(fmm, P, images):
    x, y, z, q = _uniform(8192, 2000 + P + images)
    ep, eg = _both(fmm, x, y, z, q, P, 64, images)
    # the float32 M2L's share grows with P (measured potential difference 1.2e-6 at P=8, 7.8e-6 at
    # P=10), still within the 1e-5 bound of the mixed path
    assert ep <= (1e-5 if P >= 10 else PHI_TOL) and eg <= GRAD_TOL


@pytest.mark.parametrize("P,images", [(6, 2), (7, 2), (8, 3)])
def test_mixed_infinite_lattice_gauge(fmm, P, images):
    """Infinite-lattice gauge: both paths pin the root and image monopoles, but the float64 path's
    tree cells keep the monopole truncation error of the float64 operators, which the mixed path's
    constraint removes; at P=6 that leaves a constant of ~5e-6 between the two (measured 5.4e-6 raw,
    1.1e-6 up to a constant; P=7: 2.1e-6 / 1.3e-6; P=8: 1.5e-6 / 1.4e-6).  So: up to a constant
    within 2e-6 of float64, and against the Ewald sum the raw potential is no further off than the
    float64 path's (plus 2e-6)."""
    x, y, z, q = _uniform(16384, 3000 + P)
    print(f"case n={len(x)} P={P} images={images} infinite-lattice gauge", flush=True)
    tar = np.arange(0, len(x), 8, dtype=np.int32)
    pe, _ = fmm.ewald(x, y, z, q, 1.0, 23, alpha=12.0, cutoff=0.5, targets=tar)
    out = []
    for prec in (fmm.PREC_FP64, fmm.PREC_MIXED):
        plan = fmm.LaplaceFMM(P, 64, images, 1.0, dipole_mode=fmm.DIPOLE_PHYSICAL, precision=prec, device=0)
        plan.precompute()
        out.append(plan.solve(x, y, z, q))
        plan.close()
    (p64, g64), (pmx, gmx) = out
    d = pmx - p64
    epc, eg = float(np.linalg.norm(d - d.mean()) / np.linalg.norm(p64)), _rel_l2(gmx, g64)
    e64, emx = _rel_l2(p64[tar], pe), _rel_l2(pmx[tar], pe)
    print(f"  mixed vs float64: phi {_rel_l2(pmx, p64):.2e} (up to a constant {epc:.2e}) grad {eg:.2e}; "
          f"vs Ewald: float64 {e64:.2e} mixed {emx:.2e}", flush=True)
    assert epc <= PHI_TOL and eg <= GRAD_TOL and emx <= e64 + 2e-6


def test_mixed_float64_applies_switch(fmm):
    """fmm3dpp_set_mixed_applies(plan, 0): the tree operator stages in float64 (no monopole
    constraint), float32 M2L and P2P only; the raw potential stays within the float32 M2L's 2e-6."""
    x, y, z, q = _uniform(16384, 77)
    ep, eg = _both(fmm, x, y, z, q, 8, 64, 0, mixed_applies=0)
    assert ep <= PHI_TOL and eg <= GRAD_TOL


def test_mixed_vs_fp64_config1(fmm):
    """N=65536, P=6, ncrit=64, images=2 (the CPU-runnable configuration)."""
    x, y, z, q = _uniform(65536, 42)
    ep, eg = _both(fmm, x, y, z, q, 6, 64, 2)
    assert ep <= PHI_TOL and eg <= GRAD_TOL


def test_mixed_vs_fp64_adaptive(fmm):
    """Clustered bodies: deep adaptive tree, uneven leaves (1..ncrit bodies), partial 8x8 patterns."""
    x, y, z, q = _clustered(30000, 7)
    ep, eg = _both(fmm, x, y, z, q, 6, 32, 2)
    assert ep <= PHI_TOL and eg <= GRAD_TOL


def test_mixed_large_leaves(fmm):
    """ncrit=128: leaves above one wave's 64 target slots take the chunked P2P loop.  The gradient
    difference sits at 8.2e-6 of the 1e-5 bound (deterministic for these inputs): the oracle's
    emulation of the float32 M2L and P2P on the same inputs gives 8.24e-6, the coordinate rounding
    of the P2P's closest pairs in leaves twice the usual size."""
    x, y, z, q = _uniform(60000, 11)
    ep, eg = _both(fmm, x, y, z, q, 4, 128, 1)
    assert ep <= PHI_TOL and eg <= GRAD_TOL
