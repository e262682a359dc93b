"""Which operator applies of the mixed-precision path may run in float32: the restated reference with
each operator stage in an emulated arithmetic, against its own float64 result (infinite-lattice
gauge, P=8, the pseudo-inverses at the mixed path's truncated rank 192).

Stages: P2M check potentials, P2M pseudo-inverse, M2M, L2L, L2P pseudo-inverse, L2P evaluation.
Arithmetics: float32 (operands and sums), split bfloat16 with 2 pieces and 3 products ("BF16x3")
and with 3 pieces and 6 products, each product exact and summed in float32.  Findings pinned here:

* float32 everywhere leaves the gradient at the truncated rank's own level (~1e-7) but shifts the
  potential by a constant of a few 1e-6: the equivalent densities' monopoles pick up float32
  rounding, and every cell's M2L turns its monopole error into a potential offset;
* constraining each tree cell's equivalent-density monopole to the cell's charge (uniform shift of
  the density after P2M and M2M) removes the offset: float32 in all six stages stays below 1e-6;
* BF16x3 under the same constraint is an order of magnitude worse than float32 (~1e-5 in the
  potential, which is the whole error budget), so the matrix-core operator applies use float32;
* the float32 stages hold from P = 6 to 8 only: at P = 4 the constraint moves the potential by the
  float64 operators' own monopole truncation error (towards the direct sum), at P = 10 the float32
  downward pass loses 2e-5 in the gradient of a periodic system;
* the float32 near field's gradient error (4e-6) is the rounding of cell-relative coordinates to
  float32, not its float32 sums: a perfectly compensated sum leaves it unchanged, a displacement
  formed in float64 and rounded once brings it to 1e-7, and so do two-float (hi + lo) coordinates
  with float32 arithmetic throughout;
* with the float32 M2L emulated as well (EMUL_M2L), the whole mixed path stays within its bounds,
  its gradient error is the P2P's and its potential error the far field's.
"""
import numpy as np
import pytest


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _relc(a, b):
    d = a - b
    return float(np.linalg.norm(d - d.mean()) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def emul(oracle_mod):
    rng = np.random.Generator(np.random.PCG64(42))
    n = 32768
    pts = rng.random((n, 3)) - 0.5
    q = rng.uniform(-1, 1, n)
    q -= q.mean()
    f = oracle_mod.OracleFMM(8, 64, 2, 1.0, dipole=2, threads=8)
    f.precompute()
    f.set_particles(pts[:, 0].copy(), pts[:, 1].copy(), pts[:, 2].copy())
    p0, g0 = f.evaluate(q)

    def run(mask, kind=0, rank=192):
        f.set_emulation(mask, kind, rank)
        p, g = f.evaluate(q)
        return _rel(p, p0), _relc(p, p0), _rel(g, g0)

    yield f, run, (p0, g0, q)
    f.set_emulation(0, 0, 0)


def test_emulation_off_is_the_plain_path(emul):
    f, run, (p0, g0, q) = emul
    f.set_emulation(0, 0, 0)
    p, g = f.evaluate(q)
    assert np.array_equal(p, p0) and np.array_equal(g, g0)


def test_float32_upward_pass_shifts_the_potential_by_a_constant(emul):
    f, run, _ = emul
    O = type(f)
    up = O.EMUL_P2M_PINV | O.EMUL_M2M
    raw, modc, grad = run(up)
    print(f"float32 P2M pinv + M2M, no monopole constraint: phi {raw:.2e} (mod const {modc:.2e}) grad {grad:.2e}")
    assert raw > 1e-6 and raw > 3 * modc and grad < 3e-7


def test_float32_everywhere_with_the_monopole_constraint(emul):
    f, run, _ = emul
    O = type(f)
    base = run(0)
    allx = (O.EMUL_P2M_CHECK | O.EMUL_P2M_PINV | O.EMUL_M2M | O.EMUL_L2L | O.EMUL_L2P_PINV | O.EMUL_L2P_EVAL
            | O.EMUL_MONOPOLE)
    raw, modc, grad = run(allx)
    print(f"truncated rank only: phi {base[0]:.2e} grad {base[2]:.2e}; float32 in all six stages + monopole "
          f"constraint: phi {raw:.2e} (mod const {modc:.2e}) grad {grad:.2e}")
    assert raw < 1e-6 and grad < 3e-7


def test_bf16x3_is_much_worse_than_float32(emul):
    f, run, _ = emul
    O = type(f)
    ops = O.EMUL_P2M_PINV | O.EMUL_M2M | O.EMUL_L2L | O.EMUL_L2P_PINV | O.EMUL_MONOPOLE
    f32 = run(ops, O.EMUL_F32)
    b3 = run(ops, O.EMUL_BF16X3)
    b6 = run(ops, O.EMUL_BF16X6)
    print(f"operator GEMMs + monopole constraint: float32 phi {f32[0]:.2e} grad {f32[2]:.2e}; "
          f"bf16x3 phi {b3[0]:.2e} grad {b3[2]:.2e}; bf16 3 pieces/6 products phi {b6[0]:.2e} grad {b6[2]:.2e}")
    assert b3[0] > 5 * f32[0] and b3[0] > 2e-6
    assert b6[0] < 2 * f32[0] + 2e-7


def _case(oracle_mod, P, images, dipole, n=8192, seed=None):
    rng = np.random.Generator(np.random.PCG64(2000 + P + images if seed is None else seed))
    pts = rng.random((n, 3)) - 0.5
    q = rng.uniform(-1, 1, n)
    q -= q.mean()
    f = oracle_mod.OracleFMM(P, 64, images, 1.0, dipole=dipole, threads=8)
    f.precompute()
    xyz = [pts[:, k].copy() for k in range(3)]
    f.set_particles(*xyz)
    return f, xyz, q


def _rank(f, P):
    """The mixed path's truncated pseudo-inverse rank: singular values above 1e-7 of the largest,
    rounded up to a multiple of 16."""
    S = 6 * (P - 1) ** 2 + 2
    n2 = (f.get_operator("VSinv").reshape(S, S) ** 2).sum(0)
    s = np.where(n2 > 0, 1.0 / np.sqrt(np.where(n2 > 0, n2, 1.0)), 0.0)   # zero columns: dropped
    r = int(np.nonzero(s > 1e-7 * s.max())[0].max() + 1)
    return min(S, (r + 15) // 16 * 16)


def test_float32_tree_stages_by_order(oracle_mod):
    """The orders at which the mixed path runs its tree operator stages in float32 (P = 6..8):
    * P = 6, 8: float32 in all six stages + the monopole constraint stays within 1e-6 of float64 (up
      to a constant in the reference dipole mode with images, whose float64 root monopole the image
      levels turn into an offset);
    * P = 4: the monopole constraint moves the potential by the float64 operators' own monopole
      truncation error (>5e-6, no images), towards the direct sum: no longer float64 in float32;
    * P = 10: float32 L2L / L2P lose 1e-5 in the gradient of a periodic system."""
    O = oracle_mod.OracleFMM
    allx = (O.EMUL_P2M_CHECK | O.EMUL_P2M_PINV | O.EMUL_M2M | O.EMUL_L2L | O.EMUL_L2P_PINV | O.EMUL_L2P_EVAL
            | O.EMUL_MONOPOLE)
    down = O.EMUL_L2L | O.EMUL_L2P_PINV | O.EMUL_L2P_EVAL
    for P, images in ((6, 2), (8, 3)):
        f, _, q = _case(oracle_mod, P, images, 1)
        p0, g0 = f.evaluate(q)
        f.set_emulation(allx, O.EMUL_F32, _rank(f, P))
        p, g = f.evaluate(q)
        print(f"P={P} images={images}: phi up to a constant {_relc(p, p0):.2e} grad {_rel(g, g0):.2e}")
        assert _relc(p, p0) < 1e-6 and _rel(g, g0) < 1e-6
    f, (x, y, z), q = _case(oracle_mod, 4, 0, 0)
    p0, g0 = f.evaluate(q)
    f.set_emulation(allx, O.EMUL_F32, _rank(f, 4))
    p, g = f.evaluate(q)
    tar = np.arange(0, len(q), 4, dtype=np.int32)
    pd, gd = oracle_mod.direct(x, y, z, q, 0, 1.0, targets=tar, threads=8)
    print(f"P=4: float32 vs float64 phi {_rel(p, p0):.2e}; vs direct: float32 {_rel(p[tar], pd):.2e}, "
          f"float64 {_rel(p0[tar], pd):.2e}")
    assert _rel(p, p0) > 5e-6 and _rel(p[tar], pd) <= _rel(p0[tar], pd)
    f, _, q = _case(oracle_mod, 10, 1, 1)
    p0, g0 = f.evaluate(q)
    f.set_emulation(down, O.EMUL_F32, _rank(f, 10))
    p, g = f.evaluate(q)
    print(f"P=10 images=1, float32 downward pass: grad {_rel(g, g0):.2e}")
    assert _rel(g, g0) > 1e-5


def test_float32_near_field_error_is_the_coordinate_rounding(oracle_mod):
    """The mixed path's gradient error comes from its float32 P2P, and there from rounding the
    cell-relative coordinates to float32 (an absolute error of ~3e-8 of the cell size, a large
    relative error in the displacement of the closest pairs, which 1/r^2 weighs most), not from
    float32 summation: summing every pair in float64 (a perfectly compensated sum) leaves the error
    where it is, while rounding only the float64-formed displacement removes it."""
    O = oracle_mod.OracleFMM
    f, _, q = _case(oracle_mod, 6, 1, 1, n=32768, seed=7)
    p0, g0 = f.evaluate(q)
    err = {}
    for form in (O.EMUL_P2P_CELLREL, O.EMUL_P2P_CELLREL_F64SUM, O.EMUL_P2P_DISP, O.EMUL_P2P_TWOFLOAT):
        f.set_p2p_emulation(form)
        p, g = f.evaluate(q)
        err[form] = (_rel(p, p0), _rel(g, g0))
    f.set_p2p_emulation(O.EMUL_P2P_OFF)
    print("float32 near field, phi / grad vs float64: cell-relative coordinates {:.2e} / {:.2e}; the same with "
          "float64 sums {:.2e} / {:.2e}; float64 displacement rounded once {:.2e} / {:.2e}; two-float coordinates "
          "{:.2e} / {:.2e}".format(*err[O.EMUL_P2P_CELLREL], *err[O.EMUL_P2P_CELLREL_F64SUM], *err[O.EMUL_P2P_DISP],
                                   *err[O.EMUL_P2P_TWOFLOAT]))
    cell, summed, disp = err[O.EMUL_P2P_CELLREL][1], err[O.EMUL_P2P_CELLREL_F64SUM][1], err[O.EMUL_P2P_DISP][1]
    assert cell > 1e-6 and abs(summed - cell) < 0.1 * cell and disp < 0.1 * cell
    # hi + lo float32 coordinates (two-sum per target and list entry, two differences per pair) recover
    # the displacement's accuracy with float32 arithmetic throughout
    assert err[O.EMUL_P2P_TWOFLOAT][1] < 0.05 * cell and err[O.EMUL_P2P_TWOFLOAT][0] < err[O.EMUL_P2P_CELLREL][0]


def test_whole_mixed_path_emulated(oracle_mod):
    """Every float32 stage of the mixed path at once (M2L of the tree parents, the six tree stages with
    the monopole constraint at the truncated rank, the cell-relative float32 P2P) against float64, in
    the infinite-lattice gauge, P=8, 3 image levels: within the mixed path's bounds (2e-6 potential,
    1e-5 gradient), the gradient set by the P2P and the potential by the M2L and the tree stages.
    At the bench's N=2^20 inputs the same emulation gives 9.1e-7 / 3.96e-6 against 1.7e-6 / 3.93e-6
    measured on the GPU (whose float32 transforms, matrix-op sums and v_rsq_f32 round differently)."""
    O = oracle_mod.OracleFMM
    f, _, q = _case(oracle_mod, 8, 3, 2, n=32768, seed=11)
    p0, g0 = f.evaluate(q)
    tree = O.EMUL_P2M_CHECK | O.EMUL_P2M_PINV | O.EMUL_M2M | O.EMUL_L2L | O.EMUL_L2P_PINV | O.EMUL_L2P_EVAL | O.EMUL_MONOPOLE
    err = {}
    for name, mask, p2p in (("m2l", O.EMUL_M2L, 0), ("tree", tree, 0), ("p2p", 0, O.EMUL_P2P_CELLREL),
                            ("all", tree | O.EMUL_M2L, O.EMUL_P2P_CELLREL)):
        f.set_emulation(mask, O.EMUL_F32, 192 if mask & tree else 0)
        f.set_p2p_emulation(p2p)
        p, g = f.evaluate(q)
        err[name] = (_rel(p, p0), _rel(g, g0))
    f.set_emulation(0, 0, 0)
    f.set_p2p_emulation(O.EMUL_P2P_OFF)
    print("float32 stages vs float64 (phi / grad): " + ", ".join(f"{k} {v[0]:.2e} / {v[1]:.2e}" for k, v in err.items()))
    assert err["all"][0] < 2e-6 and err["all"][1] < 1e-5
    assert err["p2p"][1] > 0.8 * err["all"][1]                      # the gradient is the P2P's
    assert max(err["m2l"][0], err["tree"][0]) > 2 * err["p2p"][0]   # the potential is the far field's
    assert err["m2l"][1] < 0.3 * err["all"][1] and err["tree"][1] < 0.3 * err["all"][1]


def test_hadamard_split_bfloat16(oracle_mod):
    """Which matrix-core arithmetic the 8x8 M2L Hadamard could use instead of float32: two bfloat16
    pieces with 3 products ("BF16x3") put the potential 2e-5 off (measured 1.9e-5 here, 6e-5 at
    N=2^20), beyond the whole 1e-5 budget; three pieces with 6 products are as accurate as float32
    (the products are exact, only the float32 sums round)."""
    O = oracle_mod.OracleFMM
    f, _, q = _case(oracle_mod, 8, 3, 2, n=32768, seed=42)
    p0, g0 = f.evaluate(q)
    err = {}
    for kind in (O.EMUL_F32, O.EMUL_BF16X3, O.EMUL_BF16X6):
        f.set_emulation(O.EMUL_M2L, kind, 0)
        p, g = f.evaluate(q)
        err[kind] = (_rel(p, p0), _rel(g, g0))
    f.set_emulation(0, 0, 0)
    print("M2L Hadamard, phi / grad vs float64: float32 {:.2e} / {:.2e}, bf16x3 {:.2e} / {:.2e}, bf16 3 pieces / "
          "6 products {:.2e} / {:.2e}".format(*err[O.EMUL_F32], *err[O.EMUL_BF16X3], *err[O.EMUL_BF16X6]))
    assert err[O.EMUL_BF16X3][0] > 1e-5 and err[O.EMUL_BF16X3][0] > 10 * err[O.EMUL_F32][0]
    assert err[O.EMUL_BF16X6][0] < 1.5 * err[O.EMUL_F32][0] and err[O.EMUL_BF16X6][1] < 1.5 * err[O.EMUL_F32][1]
