"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU restatement in fmm_oracle.cpp.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module.
It restates the 3dpp periodic KIFMM path of mumairsadiq/exafmm-rd (3dpp/src/fmm.cpp:13-81 and the
files it calls); see fmm_oracle.cpp for the per-function citations.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liborc_fmm3dpp.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile the oracle with its Makefile (g++/OpenMP)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = ctypes.CDLL(_LIB)
        L.orc_plan_create.restype = ctypes.c_void_p
        L.orc_plan_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_plan_destroy.argtypes = [ctypes.c_void_p]
        L.orc_set_option.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_last_error.argtypes = [ctypes.c_void_p]
        L.orc_precompute.argtypes = [ctypes.c_void_p]
        L.orc_set_particles.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]
        L.orc_evaluate.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.orc_timer.restype = ctypes.c_double
        L.orc_timer.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        for f in ("orc_num_cells", "orc_num_tree_cells", "orc_lost_pairs", "orc_surface_num", "orc_num_freq"):
            getattr(L, f).argtypes = [ctypes.c_void_p]
        L.orc_get_cells.argtypes = [ctypes.c_void_p, _ip, _dp]
        L.orc_get_body_order.argtypes = [ctypes.c_void_p, _ip]
        L.orc_p2p_size.restype = ctypes.c_long
        L.orc_p2p_size.argtypes = [ctypes.c_void_p]
        L.orc_m2l_size.restype = ctypes.c_long
        L.orc_m2l_size.argtypes = [ctypes.c_void_p]
        L.orc_get_p2p.argtypes = [ctypes.c_void_p, _ip, _ip, _dp]
        L.orc_get_m2l.argtypes = [ctypes.c_void_p, _ip, _ip, _dp, _ip]
        L.orc_get_operator.restype = ctypes.c_long
        L.orc_get_operator.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _dp, ctypes.c_long]
        L.orc_set_operator.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _dp, ctypes.c_long]
        L.orc_get_stage_vectors.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_set_stage_vectors.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_stage.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _dp]
        L.orc_stage_finalize.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_get_body_fields.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_drand48_bodies.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_direct.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int, _ip, ctypes.c_int,
                                 ctypes.c_double, _dp, _dp, _dp, _dp, ctypes.c_int]
        L.orc_ewald.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int, _ip, ctypes.c_double,
                                ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def drand48_bodies(n, r, seed=5, zero_netcharge=1, offset=(0.0, 0.0, 0.0)):
    """generate_random_bodies (3dpp/src/body.cpp:98-129): returns x, y, z, q."""
    x, y, z, q = (np.empty(n) for _ in range(4))
    lib().orc_drand48_bodies(n, r, offset[0], offset[1], offset[2], seed, zero_netcharge, _d(x), _d(y), _d(z), _d(q))
    return x, y, z, q


class OracleFMM:
    """LaplaceFMM(bodies, args).solve() restated (3dpp/src/fmm.cpp:13-81)."""

    def __init__(self, P, ncrit, images, cycle, dipole=1, divide_4pi=0, threads=0, m2l_bruteforce=False):
        self.h = lib().orc_plan_create(P, ncrit, images, cycle, dipole, divide_4pi, threads)
        if not self.h:
            raise ValueError("invalid oracle configuration")
        if m2l_bruteforce:
            lib().orc_set_option(self.h, b"m2l_bruteforce", 1)
        self.P, self.images, self.cycle = P, images, cycle
        self._ops = False

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_plan_destroy(self.h)
            self.h = None

    def precompute(self):
        lib().orc_precompute(self.h)
        self._ops = True

    # Stage bits and arithmetic of the reduced-precision emulation (orc_set_option "emulate").
    EMUL_P2M_CHECK, EMUL_P2M_PINV, EMUL_M2M, EMUL_L2L, EMUL_L2P_PINV, EMUL_L2P_EVAL, EMUL_MONOPOLE = 1, 2, 4, 8, 16, 32, 64
    EMUL_M2L = 128   # float32 M2L of the tree parents (up transform, 8x8 Hadamard, down transform)
    EMUL_F32, EMUL_BF16X3, EMUL_BF16X6 = 0, 1, 2
    # Forms of the float32 near field (orc_set_option "emulate_p2p"), see stage_p2p_emul.
    EMUL_P2P_OFF, EMUL_P2P_CELLREL, EMUL_P2P_CELLREL_F64SUM, EMUL_P2P_DISP, EMUL_P2P_TWOFLOAT = 0, 1, 2, 3, 4

    def set_emulation(self, mask, kind=0, pinv_rank=0):
        """Run the stages in `mask` in the emulated arithmetic `kind` (float32, or split bfloat16
        products accumulated in float32), and the P2M/L2P pseudo-inverses at `pinv_rank` (0 = full)."""
        for name, v in ((b"emulate", mask), (b"emulate_kind", kind), (b"pinv_rank", pinv_rank)):
            if lib().orc_set_option(self.h, name, int(v)) != 0:
                raise ValueError(f"bad emulation option {name.decode()}={v}")

    def set_p2p_emulation(self, form):
        """Run the near field in float32 in `form` (EMUL_P2P_*; 0 = float64)."""
        if lib().orc_set_option(self.h, b"emulate_p2p", int(form)) != 0:
            raise ValueError(f"bad near-field emulation form {form}")

    def set_particles(self, x, y, z):
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        self.n = len(x)
        if lib().orc_set_particles(self.h, self.n, _d(x), _d(y), _d(z)) != 0:
            raise RuntimeError(lib().orc_last_error(self.h).decode())

    def evaluate(self, q):
        if not self._ops:
            self.precompute()
        q = np.ascontiguousarray(q, dtype=np.float64)
        p, gx, gy, gz = (np.empty(self.n) for _ in range(4))
        if lib().orc_evaluate(self.h, _d(q), _d(p), _d(gx), _d(gy), _d(gz)) != 0:
            raise RuntimeError(lib().orc_last_error(self.h).decode())
        return p, np.stack([gx, gy, gz], axis=1)

    def timer(self, name):
        return lib().orc_timer(self.h, name.encode())

    # --- structure -------------------------------------------------------
    def cells(self):
        n = lib().orc_num_cells(self.h)
        ints = np.empty((n, 7), dtype=np.int32)
        dbl = np.empty((n, 4))
        lib().orc_get_cells(self.h, _i(ints), _d(dbl))
        return ints, dbl

    def num_tree_cells(self):
        return lib().orc_num_tree_cells(self.h)

    def lost_pairs(self):
        return lib().orc_lost_pairs(self.h)

    def body_order(self):
        perm = np.empty(self.n, dtype=np.int32)
        lib().orc_get_body_order(self.h, _i(perm))
        return perm

    def p2p_list(self):
        m = lib().orc_p2p_size(self.h)
        tar = np.empty(m, dtype=np.int32)
        src = np.empty(m, dtype=np.int32)
        off = np.empty((m, 3))
        lib().orc_get_p2p(self.h, _i(tar), _i(src), _d(off))
        return tar, src, off

    def m2l_list(self):
        m = lib().orc_m2l_size(self.h)
        tar = np.empty(m, dtype=np.int32)
        src = np.empty(m, dtype=np.int32)
        img = np.empty(m, dtype=np.int32)
        off = np.empty((m, 3))
        lib().orc_get_m2l(self.h, _i(tar), _i(src), _d(off), _i(img))
        return tar, src, off, img

    # --- operators -------------------------------------------------------
    def get_operator(self, name):
        n = lib().orc_get_operator(self.h, name.encode(), None, 0)
        if n < 0:
            raise KeyError(name)
        out = np.empty(n)
        lib().orc_get_operator(self.h, name.encode(), _d(out), n)
        return out

    def set_operator(self, name, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        if lib().orc_set_operator(self.h, name.encode(), _d(arr), arr.size) != 0:
            raise KeyError(name)

    def stage(self, name, q=None):
        """One stage of evaluate on the state the previous calls left: "charges" (q, input order),
        "p2m", "m2m", "m2l", "l2l", "l2p" (body fields = far field), "p2p" (+= near field),
        "p2p_only" (body fields = near field)."""
        if not self._ops:
            self.precompute()
        qa = None if q is None else np.ascontiguousarray(q, dtype=np.float64)
        if lib().orc_stage(self.h, name.encode(), None if qa is None else _d(qa)) != 0:
            raise RuntimeError(lib().orc_last_error(self.h).decode())

    def stage_finalize(self):
        p, gx, gy, gz = (np.empty(self.n) for _ in range(4))
        if lib().orc_stage_finalize(self.h, _d(p), _d(gx), _d(gy), _d(gz)) != 0:
            raise RuntimeError("charges not set")
        return p, np.stack([gx, gy, gz], axis=1)

    def body_fields(self):
        """Potential and gradient of the bodies in tree order before the dipole term and scaling."""
        p, gx, gy, gz = (np.empty(self.n) for _ in range(4))
        if lib().orc_get_body_fields(self.h, _d(p), _d(gx), _d(gy), _d(gz)) != 0:
            raise RuntimeError("charges not set")
        return p, np.stack([gx, gy, gz], axis=1)

    def set_stage_vectors(self, qeq=None, pck=None):
        a = None if qeq is None else np.ascontiguousarray(qeq, dtype=np.float64)
        b = None if pck is None else np.ascontiguousarray(pck, dtype=np.float64)
        lib().orc_set_stage_vectors(self.h, None if a is None else _d(a), None if b is None else _d(b))

    def stage_vectors(self):
        S = lib().orc_surface_num(self.h)
        C = lib().orc_num_cells(self.h)
        qeq = np.empty((C, S))
        pck = np.empty((C, S))
        lib().orc_get_stage_vectors(self.h, _d(qeq), _d(pck))
        return qeq, pck


def direct(x, y, z, q, images, cycle, targets=None, threads=0):
    """LaplaceKernel::direct (3dpp/src/kernel.cpp:11-77) at the given target indices."""
    n = len(x)
    targets = np.arange(n, dtype=np.int32) if targets is None else np.ascontiguousarray(targets, dtype=np.int32)
    x, y, z, q = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, q))
    nt = len(targets)
    p, gx, gy, gz = (np.empty(nt) for _ in range(4))
    lib().orc_direct(n, _d(x), _d(y), _d(z), _d(q), nt, _i(targets), images, cycle, _d(p), _d(gx), _d(gy), _d(gz), threads)
    return p, np.stack([gx, gy, gz], axis=1)


def ewald(x, y, z, q, cycle, ksize, alpha=None, cutoff=None, targets=None, threads=0, ref_ncrit=0):
    """EwaldSolver (3dpp/src/ewald.cpp:6-203); defaults alpha = ksize/cycle, cutoff = cycle/2.

    ref_ncrit > 0 reproduces the reference's cell-pair real-space schedule (all bodies are targets);
    otherwise the real-space sum uses the per-body minimum image."""
    n = len(x)
    alpha = ksize / cycle if alpha is None else alpha
    cutoff = cycle / 2 if cutoff is None else cutoff
    targets = np.arange(n, dtype=np.int32) if targets is None else np.ascontiguousarray(targets, dtype=np.int32)
    x, y, z, q = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, q))
    nt = len(targets)
    p, gx, gy, gz = (np.empty(nt) for _ in range(4))
    lib().orc_ewald(n, _d(x), _d(y), _d(z), _d(q), nt, _i(targets), cycle, ksize, alpha, cutoff,
                    _d(p), _d(gx), _d(gy), _d(gz), threads, ref_ncrit)
    return p, np.stack([gx, gy, gz], axis=1)


def compare(p1, f1, q, p2, f2):
    """compare() of 3dpp/src/body.cpp:131-175 (second argument is the reference)."""
    pdif = np.sum((p1 - p2) ** 2)
    pnrm = np.sum(p2 ** 2)
    fdif = np.sum((f1 - f2) ** 2)
    fnrm = np.sum(f2 ** 2)
    e1 = np.sum(p1 * q)
    e2 = np.sum(p2 * q)
    return dict(l2p=np.sqrt(pdif / pnrm), l2f=np.sqrt(fdif / fnrm), e1=e1, e2=e2,
                l2e=np.sqrt((e1 - e2) ** 2 / e2 ** 2), pnrm=pnrm)


def test_r2c(g):
    """The restatement's 3-D r2c DFT (stands in for FFTW r2c; x fastest)."""
    N = g.shape[0]
    g = np.ascontiguousarray(g, dtype=np.float64)
    out = np.empty((N, N, N // 2 + 1), dtype=np.complex128)
    lib().orc_test_r2c(N, _d(g), out.ctypes.data_as(_dp))
    return out


def test_c2r(G):
    N = G.shape[0]
    G = np.ascontiguousarray(G, dtype=np.complex128)
    g = np.empty((N, N, N))
    lib().orc_test_c2r(N, G.ctypes.data_as(_dp), _d(g))
    return g


def test_svd(A):
    n = A.shape[0]
    A = np.ascontiguousarray(A, dtype=np.float64)
    U = np.empty((n, n)); s = np.empty(n); V = np.empty((n, n))
    lib().orc_test_svd(n, _d(A), _d(U), _d(s), _d(V))
    return U, s, V
