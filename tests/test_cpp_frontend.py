"""The C++ drop-in front end's Tree and Traverser (exafmm-rd_amd/cpp/rtfmm_fmm3dpp.hpp, mirroring
3dpp/include/tree.h:9-75 and include/traverser.h:9-56) on the host: a small program built against
the header and libfmm3dpp.so builds the tree of the reference's drand48 bodies, checks the tree's
nesting invariants (children at Tree::get_child_cell_x of their parent, half radius, nested body
ranges, leaves within ncrit, bodies inside their cells), and prints its lists, which must equal the
structure queries of the Python binding (the same structure-only plan path the list tests pin)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "exafmm-rd_amd")


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    import pyfmm3dpp
    pyfmm3dpp.lib()
    exe = str(tmp_path_factory.mktemp("cpp") / "tree_traverser")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(PKG, "cpp"),
                    os.path.join(ROOT, "tests", "cpp", "tree_traverser.cpp"), "-o", exe, "-L", PKG, "-lfmm3dpp",
                    f"-Wl,-rpath,{PKG}"], check=True)
    return exe


@pytest.mark.parametrize("n,ncrit,images", [(3000, 32, 2), (2000, 16, 0), (1500, 64, 4)])
def test_tree_and_traverser_match_binding(prog, oracle_mod, n, ncrit, images):
    import pyfmm3dpp
    r = subprocess.run([prog, str(n), str(ncrit), str(images)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[:2000] + r.stderr
    lines = r.stdout.splitlines()
    assert " bad 0 " in lines[0]
    P = [l.split()[1:] for l in lines if l.startswith("P ")]
    M = [l.split()[1:] for l in lines if l.startswith("M ")]
    B = [int(l.split()[1]) for l in lines if l.startswith("B ")]
    x, y, z, q = oracle_mod.drand48_bodies(n, np.pi, seed=5, zero_netcharge=1)
    plan = pyfmm3dpp.LaplaceFMM(4, ncrit, images, 2 * np.pi, device=pyfmm3dpp.DEVICE_NONE)
    plan.set_particles(x, y, z)
    assert B == list(plan.body_order())
    tar, src, off = plan.p2p_list()
    assert [(int(a), int(b)) for a, b, *_ in P] == list(zip(tar.tolist(), src.tolist()))
    assert np.array_equal(np.array([[float(v) for v in p[2:]] for p in P]), off)
    mt, ms, mo, mi = plan.m2l_list()
    want = sorted(zip(mt.tolist(), ms.tolist(), map(tuple, mo.tolist()), mi.tolist()))
    got = sorted((int(a), int(b), (float(c), float(d), float(e)), int(f)) for a, b, c, d, e, f in M)
    assert got == want
    c = plan.counts()
    assert f"cells {c['n_cells']} p2p {c['n_p2p']}" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("n,P,ncrit,images", [(3000, 6, 32, 2), (2500, 4, 16, 3)])
def test_kernel_operator_api_equals_solve(n, P, ncrit, images):
    """The stage loops of LaplaceFMM::solve (3dpp/src/fmm.cpp:83-300) over the drop-in LaplaceKernel's
    per-cell operators (p2m_precompute, m2m(_img)_precompute, m2l_fft_precompute_t,
    l2l(_img)_precompute, l2p_precompute, p2p_1toN_256) equal the drop-in LaplaceFMM::solve: the
    operators run the same kernels per cell; the P2P is a direct sum over the listed sources (a
    different summation order) and the dipole term is summed on the host, so 1e-12."""
    exe = os.path.join(ROOT, "tests", "cpp", "kernel_api")
    assert os.path.exists(exe), "build with make -C exafmm-rd_amd/csrc"
    print(f"kernel API n={n} P={P} ncrit={ncrit} images={images}", flush=True)
    r = subprocess.run([exe, str(n), str(P), str(ncrit), str(images)], capture_output=True, text=True, timeout=300)
    print(r.stdout.strip(), flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    f = r.stdout.split()
    assert float(f[1]) <= 1e-12 and float(f[3]) <= 1e-12
