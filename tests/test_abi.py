"""The C-ABI boundary (include/fmm3dpp.h) on the CPU: exports, status codes, error paths.

The reference exits the process on every error (assert_exit, 3dpp/include/type.h:42-49); the
C-ABI returns a status instead.  These run without a GPU: device plans must fail cleanly when no
device exists, structure-only plans must refuse to evaluate.
"""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "fmm3dpp.h")).read()
    return sorted(set(re.findall(r"\b(fmm3dpp_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    import pyfmm3dpp
    lib = ctypes.CDLL(pyfmm3dpp.LIB_PATH)
    names = _declared()
    assert len(names) >= 17
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(pyfmm3dpp.EXPORTS) == names


def _cfg(**kw):
    import pyfmm3dpp
    c = pyfmm3dpp.Config()
    c.P, c.ncrit, c.images, c.cycle = 4, 16, 1, 1.0
    c.precision, c.dipole_mode, c.divide_4pi = 0, 1, 0
    for k, v in kw.items():
        setattr(c, k, v)
    return c


@pytest.mark.parametrize("bad", [dict(P=1), dict(P=13), dict(ncrit=0), dict(images=-1), dict(images=21),
                                 dict(cycle=0.0), dict(cycle=float("nan")), dict(precision=7), dict(dipole_mode=5)])
def test_invalid_config_rejected(bad):
    import pyfmm3dpp
    h = ctypes.c_void_p()
    rc = pyfmm3dpp.lib().fmm3dpp_plan_create(ctypes.byref(_cfg(**bad)), pyfmm3dpp.DEVICE_NONE, ctypes.byref(h))
    assert rc == pyfmm3dpp.ERR_ARG and not h.value


@pytest.mark.parametrize("prec", [0, 1])
def test_both_precisions_accepted(prec):
    """FMM3DPP_PREC_FP64 and FMM3DPP_PREC_MIXED both validate; the structure is the same for both."""
    import pyfmm3dpp
    h = ctypes.c_void_p()
    rc = pyfmm3dpp.lib().fmm3dpp_plan_create(ctypes.byref(_cfg(precision=prec)), pyfmm3dpp.DEVICE_NONE,
                                             ctypes.byref(h))
    assert rc == pyfmm3dpp.OK and h.value
    assert pyfmm3dpp.lib().fmm3dpp_plan_destroy(h) == pyfmm3dpp.OK


def test_null_arguments():
    import pyfmm3dpp
    L = pyfmm3dpp.lib()
    assert L.fmm3dpp_plan_create(None, -1, None) == pyfmm3dpp.ERR_ARG
    assert L.fmm3dpp_plan_destroy(None) == pyfmm3dpp.ERR_ARG
    assert L.fmm3dpp_precompute(None) == pyfmm3dpp.ERR_ARG


def test_no_device_fails_cleanly():
    """This container has no GPU: a device plan must return FMM3DPP_ERR_DEVICE, not crash."""
    import pyfmm3dpp
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is visible")
    h = ctypes.c_void_p()
    rc = pyfmm3dpp.lib().fmm3dpp_plan_create(ctypes.byref(_cfg()), 0, ctypes.byref(h))
    assert rc == pyfmm3dpp.ERR_DEVICE


def test_structure_only_plan_refuses_to_evaluate():
    import pyfmm3dpp
    plan = pyfmm3dpp.LaplaceFMM(4, 16, 1, 1.0, device=pyfmm3dpp.DEVICE_NONE)
    with pytest.raises(pyfmm3dpp.FMMError) as e:
        plan.evaluate(np.zeros(0))
    rng = np.random.default_rng(0)
    p = rng.random((100, 3)) - 0.5
    plan.set_particles(p[:, 0], p[:, 1], p[:, 2])
    with pytest.raises(pyfmm3dpp.FMMError) as e:
        plan.evaluate(np.zeros(100))
    assert e.value.code == pyfmm3dpp.ERR_DEVICE


def test_bad_particles():
    import pyfmm3dpp
    plan = pyfmm3dpp.LaplaceFMM(4, 16, 1, 1.0, device=pyfmm3dpp.DEVICE_NONE)
    with pytest.raises(pyfmm3dpp.FMMError) as e:
        plan.set_particles(np.array([0.1, np.nan]), np.zeros(2), np.zeros(2))
    assert e.value.code == pyfmm3dpp.ERR_ARG
    with pytest.raises(pyfmm3dpp.FMMError) as e:
        plan.set_particles(np.zeros(0), np.zeros(0), np.zeros(0))
    assert e.value.code == pyfmm3dpp.ERR_ARG
    # more than ncrit coincident bodies: the reference recurses forever (tree.cpp:125-185)
    with pytest.raises(pyfmm3dpp.FMMError) as e:
        plan.set_particles(np.full(40, 0.1), np.full(40, 0.2), np.full(40, -0.3))
    assert e.value.code == pyfmm3dpp.ERR_TREE
    # exactly ncrit coincident bodies are one leaf
    plan.set_particles(np.full(16, 0.1), np.full(16, 0.2), np.full(16, -0.3))
    assert plan.counts()["n_tree_cells"] == 1


def test_operator_queries_on_host_plan():
    import pyfmm3dpp
    plan = pyfmm3dpp.LaplaceFMM(4, 16, 2, 1.0, device=pyfmm3dpp.DEVICE_NONE)
    S = 6 * 3 * 3 + 2
    assert plan.get_operator("UT").size == S * S
    assert plan.get_operator("GK27").size == 26 * 7 * 7 * 4 * 56
    with pytest.raises(pyfmm3dpp.FMMError):
        plan.get_operator("NOPE")
    with pytest.raises(pyfmm3dpp.FMMError):
        plan.set_operator("UT", np.zeros(3))
    with pytest.raises(pyfmm3dpp.FMMError):
        plan.set_operator("M2M_IMG_SUM", plan.get_operator("M2M_IMG_SUM"))
    # L2L = M2M^T (kernel.cpp:1686-1690); L2L_IMG = M2M_IMG13^T (1693-1696)
    m2m = plan.get_operator("M2M3").reshape(S, S)
    assert np.array_equal(plan.get_operator("L2L3").reshape(S, S), m2m.T)
    assert np.array_equal(plan.get_operator("L2L_IMG").reshape(S, S), plan.get_operator("M2M_IMG13").reshape(S, S).T)


def test_batch_config_checks():
    """Batch creation refuses invalid or disagreeing configurations before it looks for a device."""
    import pyfmm3dpp
    L = pyfmm3dpp.lib()
    h = ctypes.c_void_p()
    assert L.fmm3dpp_batch_create(None, 1, 0, ctypes.byref(h)) == pyfmm3dpp.ERR_ARG
    for bad in (dict(images=2), dict(cycle=2.0), dict(precision=1), dict(dipole_mode=2), dict(divide_4pi=1),
                dict(P=13)):
        cfgs = (pyfmm3dpp.Config * 2)(_cfg(), _cfg(**bad))
        assert L.fmm3dpp_batch_create(cfgs, 2, 0, ctypes.byref(h)) == pyfmm3dpp.ERR_ARG and not h.value
    cfgs = (pyfmm3dpp.Config * 2)(_cfg(), _cfg(P=6, ncrit=64))   # P and ncrit may differ
    rc = L.fmm3dpp_batch_create(cfgs, 2, 0, ctypes.byref(h))
    assert rc in (pyfmm3dpp.OK, pyfmm3dpp.ERR_DEVICE)
    if rc == pyfmm3dpp.OK:
        assert L.fmm3dpp_batch_destroy(h) == pyfmm3dpp.OK
    assert L.fmm3dpp_batch_destroy(None) == pyfmm3dpp.ERR_ARG
    assert L.fmm3dpp_batch_timing_ms(None, b"evaluate") == -1


def test_stage_entry_points_need_a_device_plan():
    """The per-stage entry points act on device state: a structure-only plan refuses them."""
    import pyfmm3dpp
    L = pyfmm3dpp.lib()
    plan = pyfmm3dpp.LaplaceFMM(4, 16, 1, 1.0, device=pyfmm3dpp.DEVICE_NONE)
    rng = np.random.default_rng(1)
    p = rng.random((200, 3)) - 0.5
    plan.set_particles(p[:, 0], p[:, 1], p[:, 2])
    q = np.zeros(200)
    assert L.fmm3dpp_stage_charges(plan.h, q.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == pyfmm3dpp.ERR_DEVICE
    for f in ("fmm3dpp_stage_p2m", "fmm3dpp_stage_m2m", "fmm3dpp_stage_l2l", "fmm3dpp_stage_l2p"):
        assert getattr(L, f)(plan.h, 0, None) == pyfmm3dpp.ERR_DEVICE
    assert L.fmm3dpp_stage_m2l(plan.h) == pyfmm3dpp.ERR_DEVICE
    assert L.fmm3dpp_stage_p2p(plan.h, 0, None, 1) == pyfmm3dpp.ERR_DEVICE
    assert L.fmm3dpp_stage_m2l(None) == pyfmm3dpp.ERR_ARG
