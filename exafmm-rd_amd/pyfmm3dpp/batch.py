"""Batches of independent periodic systems (fmm3dpp_batch_* of include/fmm3dpp.h).

The reference solves one system per rtfmm::LaplaceFMM (3dpp/include/fmm.h:15-17); a `Batch` holds
many systems of uneven size that share images, cycle, precision and dipole mode, and evaluates them
with one launch per FMM stage for every group of equal P.  Arrays are torch tensors on the batch's
GPU (float64), all systems concatenated in system order.
"""
import ctypes

import numpy as np

from . import FMMError, OK, PREC_FP64, DIPOLE_REFERENCE, DIPOLE_OFF, Config, lib


class Batch:
    """`systems` is a sequence of (P, ncrit) pairs, one per system; the other fields are shared."""

    def __init__(self, systems, images, cycle, dipole_correction=1, divide_4pi=0, precision=PREC_FP64,
                 device=0, dipole_mode=None):
        import torch
        if dipole_mode is None:
            dipole_mode = DIPOLE_REFERENCE if dipole_correction else DIPOLE_OFF
        cfgs = (Config * len(systems))()
        for k, (P, ncrit) in enumerate(systems):
            c = cfgs[k]
            c.P, c.ncrit, c.images, c.cycle = int(P), int(ncrit), int(images), float(cycle)
            c.precision, c.dipole_mode, c.divide_4pi = precision, dipole_mode, int(bool(divide_4pi))
        h = ctypes.c_void_p()
        rc = lib().fmm3dpp_batch_create(cfgs, len(systems), device, ctypes.byref(h))
        if rc != OK:
            raise FMMError(rc, "batch_create failed (invalid configurations or no device)")
        self.h = h
        self.nsys = len(systems)
        self.systems = [tuple(s) for s in systems]
        self.device = torch.device("cuda", device)
        self.counts = None

    def close(self):
        if getattr(self, "h", None):
            lib().fmm3dpp_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != OK:
            raise FMMError(rc, lib().fmm3dpp_batch_last_error(self.h).decode())

    def _dev(self, a, rows=None):
        import torch
        t = torch.as_tensor(a, dtype=torch.float64, device=self.device).contiguous()
        if rows is not None and t.shape[0] != rows:
            raise ValueError("array length differs from the total body count")
        return t

    def set_particles(self, counts, x, y, z):
        """counts[k] bodies of system k; x, y, z concatenated (tensors or arrays)."""
        import torch
        counts = np.ascontiguousarray(counts, dtype=np.int64)
        if counts.shape != (self.nsys,):
            raise ValueError("one count per system")
        n = int(counts.sum())
        x, y, z = (self._dev(a, n) for a in (x, y, z))
        torch.cuda.synchronize(self.device)
        self._check(lib().fmm3dpp_batch_set_particles_device(
            self.h, counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), ctypes.c_void_p(x.data_ptr()),
            ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(z.data_ptr())))
        self.counts = counts
        self.n = n

    def evaluate(self, q, phi=None, grad=None):
        """Potentials [n] and gradients [n, 3] (torch, on the batch's device) for charges q."""
        import torch
        if self.counts is None:
            raise RuntimeError("set_particles must precede evaluate")
        q = self._dev(q, self.n)
        phi = torch.empty(self.n, dtype=torch.float64, device=self.device) if phi is None else phi
        grad = torch.empty((self.n, 3), dtype=torch.float64, device=self.device) if grad is None else grad
        torch.cuda.synchronize(self.device)
        self._check(lib().fmm3dpp_batch_evaluate_device(self.h, ctypes.c_void_p(q.data_ptr()),
                                                        ctypes.c_void_p(phi.data_ptr()),
                                                        ctypes.c_void_p(grad.data_ptr())))
        return phi, grad

    def set_graph(self, enable=True):
        """Capture the evaluation into a HIP graph on the next call and replay it afterwards."""
        self._check(lib().fmm3dpp_batch_set_graph(self.h, int(bool(enable))))

    def autotune(self, candidates, q):
        """Per-P-group leaf size from timed evaluations over `candidates`; returns each system's ncrit."""
        import torch
        cand = np.ascontiguousarray(candidates, dtype=np.int32)
        q = self._dev(q, self.n)
        phi = torch.empty(self.n, dtype=torch.float64, device=self.device)
        grad = torch.empty((self.n, 3), dtype=torch.float64, device=self.device)
        chosen = np.empty(self.nsys, dtype=np.int32)
        torch.cuda.synchronize(self.device)
        self._check(lib().fmm3dpp_batch_autotune(
            self.h, cand.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(cand), ctypes.c_void_p(q.data_ptr()),
            ctypes.c_void_p(phi.data_ptr()), ctypes.c_void_p(grad.data_ptr()),
            chosen.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return chosen

    def timing_ms(self, name):
        return lib().fmm3dpp_batch_timing_ms(self.h, name.encode())
