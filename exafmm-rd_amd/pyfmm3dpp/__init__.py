"""ctypes binding of libfmm3dpp.so (include/fmm3dpp.h), the MI355X periodic KIFMM.

`LaplaceFMM` mirrors the reference's evaluate API (rtfmm::LaplaceFMM(bodies, args).solve(),
3dpp/include/fmm.h:15-17): a plan is built once per configuration, `set_particles` builds the tree
and lists, `solve`/`evaluate` return potential and gradient in input order.  The library must be
built (`make -C exafmm-rd_amd/csrc`); nothing here falls back to another implementation.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libfmm3dpp.so")

OK = 0
ERR_ARG, ERR_DEVICE, ERR_STATE, ERR_TREE, ERR_ALLOC, ERR_HIP = -1, -2, -3, -4, -5, -6
PREC_FP64, PREC_MIXED = 0, 1
DIPOLE_OFF, DIPOLE_REFERENCE, DIPOLE_PHYSICAL = 0, 1, 2
DEVICE_NONE = -1

# Every entry point include/fmm3dpp.h declares.
EXPORTS = (
    "fmm3dpp_plan_create", "fmm3dpp_plan_destroy", "fmm3dpp_last_error", "fmm3dpp_precompute",
    "fmm3dpp_set_particles", "fmm3dpp_set_particles_device", "fmm3dpp_evaluate", "fmm3dpp_evaluate_device", "fmm3dpp_set_stream",
    "fmm3dpp_solve_device", "fmm3dpp_set_hadamard_staging", "fmm3dpp_set_near_field_stream",
    "fmm3dpp_set_mixed_applies",
    "fmm3dpp_get_counts", "fmm3dpp_get_cells", "fmm3dpp_get_body_order", "fmm3dpp_get_p2p", "fmm3dpp_get_m2l",
    "fmm3dpp_get_operator", "fmm3dpp_set_operator", "fmm3dpp_get_stage_vectors", "fmm3dpp_timing_ms",
    "fmm3dpp_ewald", "fmm3dpp_ewald_device", "fmm3dpp_evaluate_adjoint", "fmm3dpp_evaluate_adjoint_device",
    "fmm3dpp_save_operators", "fmm3dpp_load_operators", "fmm3dpp_direct", "fmm3dpp_direct_device",
    "fmm3dpp_direct_device_noalloc",
    "fmm3dpp_batch_create", "fmm3dpp_batch_destroy", "fmm3dpp_batch_last_error",
    "fmm3dpp_batch_set_particles_device", "fmm3dpp_batch_evaluate_device", "fmm3dpp_batch_set_graph",
    "fmm3dpp_batch_autotune", "fmm3dpp_batch_timing_ms",
    "fmm3dpp_stage_charges", "fmm3dpp_stage_p2m", "fmm3dpp_stage_m2m", "fmm3dpp_stage_m2l", "fmm3dpp_stage_l2l",
    "fmm3dpp_stage_l2p", "fmm3dpp_stage_p2p", "fmm3dpp_stage_finalize", "fmm3dpp_get_cell_vectors",
    "fmm3dpp_set_cell_vectors", "fmm3dpp_get_body_fields", "fmm3dpp_set_body_fields",
)


class Config(ctypes.Structure):
    _fields_ = [("P", ctypes.c_int), ("ncrit", ctypes.c_int), ("images", ctypes.c_int), ("cycle", ctypes.c_double),
                ("center", ctypes.c_double * 3), ("precision", ctypes.c_int), ("dipole_mode", ctypes.c_int),
                ("divide_4pi", ctypes.c_int)]


class Counts(ctypes.Structure):
    _fields_ = [("n_bodies", ctypes.c_int64), ("n_cells", ctypes.c_int64), ("n_tree_cells", ctypes.c_int64),
                ("n_p2p", ctypes.c_int64), ("n_m2l", ctypes.c_int64), ("n_lost", ctypes.c_int64),
                ("S", ctypes.c_int32), ("n_freq", ctypes.c_int32)]


class FMMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fmm3dpp error {code}: {msg}")
        self.code = code


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


def lib():
    """Load libfmm3dpp.so; raises if it was not built."""
    global _lib
    if _lib is None:
        try:   # torch bundles its own HIP runtime (same soname): load it first so that the process has one
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: build it with `make -C exafmm-rd_amd/csrc`")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.fmm3dpp_plan_create.argtypes = [ctypes.POINTER(Config), ctypes.c_int, ctypes.POINTER(vp)]
        L.fmm3dpp_plan_destroy.argtypes = [vp]
        L.fmm3dpp_last_error.argtypes = [vp]
        L.fmm3dpp_last_error.restype = ctypes.c_char_p
        L.fmm3dpp_precompute.argtypes = [vp]
        L.fmm3dpp_set_particles.argtypes = [vp, ctypes.c_int64, _dp, _dp, _dp]
        L.fmm3dpp_set_particles_device.argtypes = [vp, ctypes.c_int64, vp, vp, vp]
        L.fmm3dpp_evaluate.argtypes = [vp, _dp, _dp, _dp]
        L.fmm3dpp_evaluate_device.argtypes = [vp, vp, vp, vp]
        L.fmm3dpp_set_stream.argtypes = [vp, vp]
        L.fmm3dpp_set_hadamard_staging.argtypes = [vp, ctypes.c_int]
        L.fmm3dpp_set_near_field_stream.argtypes = [vp, ctypes.c_int]
        L.fmm3dpp_set_mixed_applies.argtypes = [vp, ctypes.c_int]
        L.fmm3dpp_solve_device.argtypes = [vp, ctypes.c_int64, vp, vp, vp, vp, vp, vp]
        L.fmm3dpp_get_counts.argtypes = [vp, ctypes.POINTER(Counts)]
        L.fmm3dpp_get_cells.argtypes = [vp, _ip, _dp]
        L.fmm3dpp_get_body_order.argtypes = [vp, _ip]
        L.fmm3dpp_get_p2p.argtypes = [vp, _ip, _ip, _dp]
        L.fmm3dpp_get_m2l.argtypes = [vp, _ip, _ip, _dp, _ip]
        L.fmm3dpp_get_operator.argtypes = [vp, ctypes.c_char_p, _dp, ctypes.c_int64]
        L.fmm3dpp_get_operator.restype = ctypes.c_int64
        L.fmm3dpp_set_operator.argtypes = [vp, ctypes.c_char_p, _dp, ctypes.c_int64]
        L.fmm3dpp_get_stage_vectors.argtypes = [vp, _dp, _dp]
        L.fmm3dpp_timing_ms.argtypes = [vp, ctypes.c_char_p]
        L.fmm3dpp_timing_ms.restype = ctypes.c_double
        L.fmm3dpp_direct.argtypes = [ctypes.c_int, ctypes.c_int64, _dp, _dp, _dp, _dp, ctypes.c_int32, _ip,
                                     ctypes.c_double, ctypes.c_int, _dp, _dp]
        L.fmm3dpp_direct_device.argtypes = [vp, ctypes.c_int64, vp, vp, vp, vp, ctypes.c_int32, vp, ctypes.c_double,
                                            ctypes.c_int, vp, vp]
        L.fmm3dpp_direct_device_noalloc.argtypes = [vp, ctypes.c_int64, vp, vp, vp, vp, ctypes.c_int32, vp, ctypes.c_double,
                                            ctypes.c_int, vp, vp]
        L.fmm3dpp_save_operators.argtypes = [vp, ctypes.c_char_p]
        L.fmm3dpp_load_operators.argtypes = [vp, ctypes.c_char_p]
        L.fmm3dpp_evaluate_adjoint.argtypes = [vp, _dp, _dp, _dp, _dp, _dp]
        L.fmm3dpp_evaluate_adjoint_device.argtypes = [vp, vp, vp, vp, vp, vp]
        L.fmm3dpp_ewald.argtypes = [ctypes.c_int, ctypes.c_int64, _dp, _dp, _dp, _dp, ctypes.c_int32, _ip,
                                    ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.fmm3dpp_ewald_device.argtypes = [vp, ctypes.c_int64, vp, vp, vp, vp, ctypes.c_int32, vp, ctypes.c_double,
                                           ctypes.c_int, ctypes.c_double, ctypes.c_double, vp, vp]
        L.fmm3dpp_batch_create.argtypes = [ctypes.POINTER(Config), ctypes.c_int32, ctypes.c_int, ctypes.POINTER(vp)]
        L.fmm3dpp_batch_destroy.argtypes = [vp]
        L.fmm3dpp_batch_last_error.argtypes = [vp]
        L.fmm3dpp_batch_last_error.restype = ctypes.c_char_p
        L.fmm3dpp_batch_set_particles_device.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), vp, vp, vp]
        L.fmm3dpp_batch_evaluate_device.argtypes = [vp, vp, vp, vp]
        L.fmm3dpp_batch_set_graph.argtypes = [vp, ctypes.c_int]
        L.fmm3dpp_batch_autotune.argtypes = [vp, _ip, ctypes.c_int32, vp, vp, vp, _ip]
        L.fmm3dpp_batch_timing_ms.argtypes = [vp, ctypes.c_char_p]
        L.fmm3dpp_batch_timing_ms.restype = ctypes.c_double
        L.fmm3dpp_stage_charges.argtypes = [vp, _dp]
        for f in ("fmm3dpp_stage_p2m", "fmm3dpp_stage_m2m", "fmm3dpp_stage_l2l", "fmm3dpp_stage_l2p"):
            getattr(L, f).argtypes = [vp, ctypes.c_int64, _ip]
        L.fmm3dpp_stage_m2l.argtypes = [vp]
        L.fmm3dpp_stage_p2p.argtypes = [vp, ctypes.c_int64, _ip, ctypes.c_int]
        L.fmm3dpp_stage_finalize.argtypes = [vp, _dp, _dp]
        L.fmm3dpp_get_cell_vectors.argtypes = [vp, ctypes.c_int, ctypes.c_int64, _ip, _dp]
        L.fmm3dpp_set_cell_vectors.argtypes = [vp, ctypes.c_int, ctypes.c_int64, _ip, _dp]
        L.fmm3dpp_get_body_fields.argtypes = [vp, ctypes.c_int64, ctypes.c_int64, _dp, _dp]
        L.fmm3dpp_set_body_fields.argtypes = [vp, ctypes.c_int64, ctypes.c_int64, _dp, _dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


class LaplaceFMM:
    """Plan of the periodic Laplace KIFMM on one GPU (or structure-only with device=DEVICE_NONE).

    Arguments follow the reference's Argument fields (3dpp/include/argument.h:14-44): P, ncrit,
    images, cycle, dipole_correction, divide_4pi; the box is centred at `center` with half size cycle/2.
    `dipole_mode` (DIPOLE_OFF / DIPOLE_REFERENCE / DIPOLE_PHYSICAL) overrides dipole_correction.
    """

    def __init__(self, P, ncrit, images, cycle, dipole_correction=1, divide_4pi=0, precision=PREC_FP64,
                 device=0, center=(0.0, 0.0, 0.0), dipole_mode=None):
        cfg = Config()
        cfg.P, cfg.ncrit, cfg.images, cfg.cycle = P, ncrit, images, cycle
        for k in range(3):
            cfg.center[k] = center[k]
        cfg.precision = precision
        if dipole_mode is None:
            dipole_mode = DIPOLE_REFERENCE if dipole_correction else DIPOLE_OFF
        cfg.dipole_mode = dipole_mode
        cfg.divide_4pi = int(bool(divide_4pi))
        h = ctypes.c_void_p()
        rc = lib().fmm3dpp_plan_create(ctypes.byref(cfg), device, ctypes.byref(h))
        if rc != OK:
            raise FMMError(rc, "plan_create failed (invalid configuration or no device)")
        self.h = h
        self.P, self.ncrit, self.images, self.cycle = P, ncrit, images, cycle
        self.device = device
        self.n = 0

    def close(self):
        if getattr(self, "h", None):
            lib().fmm3dpp_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != OK:
            raise FMMError(rc, lib().fmm3dpp_last_error(self.h).decode())

    def precompute(self):
        self._check(lib().fmm3dpp_precompute(self.h))

    def save_operators(self, path):
        """Write the precomputed operators to `path` (fmm3dpp_save_operators)."""
        self._check(lib().fmm3dpp_save_operators(self.h, os.fsencode(path)))

    def load_operators(self, path):
        """Use operators saved by a plan of the same P, images and cycle instead of the precompute."""
        self._check(lib().fmm3dpp_load_operators(self.h, os.fsencode(path)))

    def set_particles(self, x, y, z):
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        self.n = len(x)
        self._check(lib().fmm3dpp_set_particles(self.h, self.n, _d(x), _d(y), _d(z)))

    def set_particles_device(self, n, x_ptr, y_ptr, z_ptr):
        """Positions already in HBM (device pointers as ints); the tree and lists are built on the GPU."""
        self.n = int(n)
        self._check(lib().fmm3dpp_set_particles_device(self.h, self.n, ctypes.c_void_p(x_ptr),
                                                       ctypes.c_void_p(y_ptr), ctypes.c_void_p(z_ptr)))

    def evaluate(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        if len(q) != self.n:
            raise ValueError("charge count differs from the particle count")
        phi = np.empty(self.n)
        grad = np.empty((self.n, 3))
        self._check(lib().fmm3dpp_evaluate(self.h, _d(q), _d(phi), _d(grad)))
        return phi, grad

    def solve(self, x, y, z, q):
        """set_particles + evaluate, like LaplaceFMM::solve (3dpp/src/fmm.cpp:13-81)."""
        self.set_particles(x, y, z)
        return self.evaluate(q)

    def evaluate_device(self, q_ptr, phi_ptr, grad_ptr):
        """Device pointers (ints), ordered on the plan's stream."""
        self._check(lib().fmm3dpp_evaluate_device(self.h, ctypes.c_void_p(q_ptr), ctypes.c_void_p(phi_ptr),
                                                   ctypes.c_void_p(grad_ptr)))

    def solve_device(self, n, x_ptr, y_ptr, z_ptr, q_ptr, phi_ptr, grad_ptr):
        """set_particles_device + evaluate_device in one call, the near-field list build overlapping the
        far-field stages (fmm3dpp_solve_device); device pointers as ints."""
        self.n = int(n)
        self._check(lib().fmm3dpp_solve_device(self.h, self.n, *(ctypes.c_void_p(p) for p in
                                                               (x_ptr, y_ptr, z_ptr, q_ptr, phi_ptr, grad_ptr))))

    def evaluate_adjoint(self, q, w, grad):
        """Backward of phi = evaluate(q) for w = dL/dphi and the forward gradient `grad`:
        returns (dL/dq, dL/dx) (fmm3dpp_evaluate_adjoint)."""
        q, w = (np.ascontiguousarray(a, dtype=np.float64) for a in (q, w))
        grad = np.ascontiguousarray(grad, dtype=np.float64)
        dq = np.empty(self.n)
        dx = np.empty((self.n, 3))
        self._check(lib().fmm3dpp_evaluate_adjoint(self.h, _d(q), _d(w), _d(grad), _d(dq), _d(dx)))
        return dq, dx

    def evaluate_adjoint_device(self, q_ptr, w_ptr, grad_ptr, dq_ptr, dx_ptr):
        """Device pointers (ints), ordered on the plan's stream."""
        self._check(lib().fmm3dpp_evaluate_adjoint_device(self.h, *(ctypes.c_void_p(p) for p in
                                                                    (q_ptr, w_ptr, grad_ptr, dq_ptr, dx_ptr))))

    def set_stream(self, stream_ptr):
        self._check(lib().fmm3dpp_set_stream(self.h, ctypes.c_void_p(stream_ptr)))

    def set_near_field_stream(self, enable):
        """1 (default): the near field runs on the plan's second stream beside the far field; 0: on the
        plan's stream after it.  Results are bit-identical."""
        self._check(lib().fmm3dpp_set_near_field_stream(self.h, int(bool(enable))))

    def set_hadamard_staging(self, enable):
        """1 (default): the 8x8 Hadamard stages each chunk's distinct sources through LDS; 0: it reads
        them from HBM / L2.  Results are bit-identical."""
        self._check(lib().fmm3dpp_set_hadamard_staging(self.h, int(bool(enable))))

    def set_mixed_applies(self, float32):
        """Mixed path: 1 (default) runs the tree operator stages (P2M check, pseudo-inverses, M2M, L2L,
        L2P evaluation) in float32 with a per-cell monopole constraint; 0 runs them in float64."""
        self._check(lib().fmm3dpp_set_mixed_applies(self.h, int(bool(float32))))

    def timing_ms(self, name):
        return lib().fmm3dpp_timing_ms(self.h, name.encode())

    # --- structure ---------------------------------------------------------
    def counts(self):
        c = Counts()
        self._check(lib().fmm3dpp_get_counts(self.h, ctypes.byref(c)))
        return {k: getattr(c, k) for k, _ in Counts._fields_}

    def cells(self):
        n = self.counts()["n_cells"]
        ints = np.empty((n, 7), dtype=np.int32)
        dbl = np.empty((n, 4))
        self._check(lib().fmm3dpp_get_cells(self.h, _i(ints), _d(dbl)))
        return ints, dbl

    def body_order(self):
        perm = np.empty(self.n, dtype=np.int32)
        self._check(lib().fmm3dpp_get_body_order(self.h, _i(perm)))
        return perm

    def p2p_list(self):
        m = self.counts()["n_p2p"]
        tar, src = np.empty(m, dtype=np.int32), np.empty(m, dtype=np.int32)
        off = np.empty((m, 3))
        self._check(lib().fmm3dpp_get_p2p(self.h, _i(tar), _i(src), _d(off)))
        return tar, src, off

    def m2l_list(self):
        m = self.counts()["n_m2l"]
        tar, src, img = (np.empty(m, dtype=np.int32) for _ in range(3))
        off = np.empty((m, 3))
        self._check(lib().fmm3dpp_get_m2l(self.h, _i(tar), _i(src), _d(off), _i(img)))
        return tar, src, off, img

    # --- operators -----------------------------------------------------------
    def get_operator(self, name):
        n = lib().fmm3dpp_get_operator(self.h, name.encode(), None, 0)
        if n < 0:
            self._check(int(n))
        out = np.empty(n)
        lib().fmm3dpp_get_operator(self.h, name.encode(), _d(out), n)
        return out

    def set_operator(self, name, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        self._check(lib().fmm3dpp_set_operator(self.h, name.encode(), _d(arr), arr.size))

    # --- per-stage entry points (LaplaceKernel's operator API) ---------------
    def _cells(self, cells):
        if cells is None:
            return 0, None
        a = np.ascontiguousarray(cells, dtype=np.int32)
        self._keep = a
        return len(a), _i(a)

    def stage_charges(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        self._check(lib().fmm3dpp_stage_charges(self.h, _d(q)))

    def stage(self, name, cells=None, accumulate=1):
        """One far- or near-field stage: "p2m", "m2m", "m2l", "l2l", "l2p", "p2p" over the listed cells
        (None: every cell the stage covers, as evaluate)."""
        L = lib()
        if name == "m2l":
            self._check(L.fmm3dpp_stage_m2l(self.h))
            return
        n, p = self._cells(cells)
        if name == "p2p":
            self._check(L.fmm3dpp_stage_p2p(self.h, n, p, int(accumulate)))
        else:
            self._check(getattr(L, f"fmm3dpp_stage_{name}")(self.h, n, p))

    def stage_finalize(self):
        phi = np.empty(self.n)
        grad = np.empty((self.n, 3))
        self._check(lib().fmm3dpp_stage_finalize(self.h, _d(phi), _d(grad)))
        return phi, grad

    def cell_vectors(self, which, cells=None):
        """which 0: q_equiv, 1: p_check of the listed cells (None: all), shape (ncells, S)."""
        c = self.counts()
        n, p = self._cells(cells)
        out = np.empty((c["n_cells"] if cells is None else n, c["S"]))
        self._check(lib().fmm3dpp_get_cell_vectors(self.h, which, n, p, _d(out)))
        return out

    def set_cell_vectors(self, which, values, cells=None):
        v = np.ascontiguousarray(values, dtype=np.float64)
        n, p = self._cells(cells)
        self._check(lib().fmm3dpp_set_cell_vectors(self.h, which, n, p, _d(v)))

    def body_fields(self, first=0, count=None):
        """Potential and gradient of bodies [first, first + count) in tree order, before the dipole
        term and scaling."""
        count = self.n - first if count is None else count
        phi = np.empty(count)
        grad = np.empty((count, 3))
        self._check(lib().fmm3dpp_get_body_fields(self.h, first, count, _d(phi), _d(grad)))
        return phi, grad

    def set_body_fields(self, phi, grad, first=0):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        grad = np.ascontiguousarray(grad, dtype=np.float64)
        self._check(lib().fmm3dpp_set_body_fields(self.h, first, len(phi), _d(phi), _d(grad)))

    def stage_vectors(self):
        c = self.counts()
        qeq = np.empty((c["n_cells"], c["S"]))
        pck = np.empty((c["n_cells"], c["S"]))
        self._check(lib().fmm3dpp_get_stage_vectors(self.h, _d(qeq), _d(pck)))
        return qeq, pck


def ewald(x, y, z, q, cycle, ksize, alpha=None, cutoff=None, targets=None, device=0):
    """Tin-foil Ewald sum on the GPU (fmm3dpp_ewald); the reference's defaults alpha = ksize / cycle,
    cutoff = cycle / 2 (3dpp/src/ewald.cpp:9-14).  Returns phi[ntar], grad[ntar, 3]."""
    x, y, z, q = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, q))
    n = len(x)
    alpha = ksize / cycle if alpha is None else alpha
    cutoff = cycle / 2 if cutoff is None else cutoff
    targets = np.arange(n, dtype=np.int32) if targets is None else np.ascontiguousarray(targets, dtype=np.int32)
    phi = np.empty(len(targets))
    grad = np.empty((len(targets), 3))
    rc = lib().fmm3dpp_ewald(device, n, _d(x), _d(y), _d(z), _d(q), len(targets), _i(targets), cycle, ksize,
                             alpha, cutoff, _d(phi), _d(grad))
    if rc != OK:
        raise FMMError(rc, "fmm3dpp_ewald failed")
    return phi, grad


def direct(x, y, z, q, images, cycle, targets=None, device=0):
    """Direct lattice sum on the GPU (fmm3dpp_direct): (3^images)^3 copies, no dipole term.
    Returns phi[ntar], grad[ntar, 3]."""
    x, y, z, q = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, q))
    n = len(x)
    targets = np.arange(n, dtype=np.int32) if targets is None else np.ascontiguousarray(targets, dtype=np.int32)
    phi = np.empty(len(targets))
    grad = np.empty((len(targets), 3))
    rc = lib().fmm3dpp_direct(device, n, _d(x), _d(y), _d(z), _d(q), len(targets), _i(targets), cycle, images,
                              _d(phi), _d(grad))
    if rc != OK:
        raise FMMError(rc, "fmm3dpp_direct failed")
    return phi, grad


OPERATOR_NAMES = (["UT", "VSinv", "UT_l2p", "VSinv_l2p", "GK8"] + [f"M2M{o}" for o in range(8)]
                  + [f"L2L{o}" for o in range(8)])
IMAGE_OPERATOR_NAMES = [f"M2M_IMG{o}" for o in range(27)] + ["L2L_IMG"]
