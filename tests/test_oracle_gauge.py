"""Pins the oracle's infinite-lattice gauge (dipole mode 2, no counterpart in the reference).

Mode 2 is the raw lattice FMM with two changes: the equivalent densities of the root and of every
image cell are shifted so that their monopole equals their charge (the bodies' total, times 27^m for
the image cell c(-m)), removing the rounding-level monopole the pseudo-inverse leaves, and the potential and gradient
receive the surface term of the cubic lattice sum, phi -= (4 pi/3V) M.x - (2 pi/3V) sum q |x|^2,
grad -= (4 pi/3V) M (de Leeuw, Perram and Smith 1980).  What pins it:

* the tin-foil Ewald sum (the oracle's own, pinned against the reference's logged Ewald energy):
  the raw potential error falls with the number of image levels and is below 1e-5 from 5 levels on
  (P=8); the gradient falls below 1e-6;
* the same lattice without the monopole shift (mode 0 plus the surface term, computed here) drifts
  away from Ewald: its constant offset grows ~9x per image level, so at 7 levels it is >1e-3 while
  mode 2 stays <1e-6 -- the drift the shift removes;
* at 2 and 3 levels, where the drift is still small, mode 2 equals mode 0 plus the surface term up to
  an additive constant, to 1e-9 relative (a wrong sign or factor in the surface term fails this).
"""
import numpy as np
import pytest


def _uniform(n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    x = rng.random((n, 3)) - 0.5
    q = rng.uniform(-1, 1, n)
    q -= q.mean()
    q[-1] -= q.sum()
    return x[:, 0].copy(), x[:, 1].copy(), x[:, 2].copy(), q


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def lattice(oracle_mod):
    x, y, z, q = _uniform(4096, 11)
    pe, ge = oracle_mod.ewald(x, y, z, q, 1.0, 23, alpha=12.0, cutoff=0.5, threads=8)
    X = np.stack([x, y, z], 1)
    M = (q[:, None] * X).sum(0)
    c = 4 * np.pi / 3
    surface_phi = -(c * (X @ M) - 0.5 * c * np.sum(q * (X * X).sum(1)))
    out = {}
    for images in (2, 3, 4, 5, 6, 7):
        for mode in (0, 2):
            f = oracle_mod.OracleFMM(8, 64, images, 1.0, dipole=mode, threads=8)
            f.precompute()
            f.set_particles(x, y, z)
            p, g = f.evaluate(q)
            if mode == 0:
                p, g = p + surface_phi, g - c * M
            out[images, mode] = (p, g)
    return pe, ge, out


def test_gauge_converges_to_ewald(lattice):
    pe, ge, out = lattice
    raw = {i: _rel(out[i, 2][0], pe) for i in range(2, 8)}
    grad = {i: _rel(out[i, 2][1], ge) for i in range(2, 8)}
    print("raw phi error vs Ewald by images:", {k: f"{v:.2e}" for k, v in raw.items()})
    print("grad error vs Ewald by images:", {k: f"{v:.2e}" for k, v in grad.items()})
    for i in range(2, 6):
        assert raw[i + 1] < raw[i] / 4, (i, raw)      # the lattice remainder decays ~9x per level
    assert raw[5] <= 1e-5 and raw[6] <= 1e-5 and raw[7] <= 1e-5
    assert grad[5] <= 1e-6 and grad[6] <= 1e-6


def test_without_monopole_shift_the_offset_grows(lattice):
    pe, ge, out = lattice
    d6, d7 = out[6, 0][0] - pe, out[7, 0][0] - pe
    print(f"mode 0 + surface term: mean offset {d6.mean():.2e} (6 levels), {d7.mean():.2e} (7 levels)")
    assert abs(d7.mean()) > 5 * abs(d6.mean())
    assert _rel(out[7, 0][0], pe) > 1e-3 and _rel(out[7, 2][0], pe) < 1e-6


@pytest.mark.parametrize("images", [2, 3])
def test_gauge_is_surface_term_up_to_constant(lattice, images):
    pe, ge, out = lattice
    p2, g2 = out[images, 2]
    p0, g0 = out[images, 0]
    d = p2 - p0
    assert _rel(p2 - d.mean(), p0) <= 1e-9
    assert _rel(g2, g0) <= 1e-9
