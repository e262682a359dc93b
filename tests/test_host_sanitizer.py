"""The host-sanitizer build (`make -C exafmm-rd_amd/csrc asan`: AddressSanitizer and
UndefinedBehaviorSanitizer on the library's host code, device code unchanged) driven through the
C-ABI by tests/cpp/host_asan.cpp: argument checks, structure-only plans, operator precompute, get,
set and cache round trip, a corrupted and a truncated cache on the CPU; plans of both precisions,
every per-stage entry point (all and listed cells), the adjoint, Ewald, direct (with and without
allocation) and a batch with graph replay on the GPU.  A sanitizer finding aborts the driver.
The operator precompute's thread pool is also run under ThreadSanitizer (tests/cpp/precompute_tsan.cpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "host_asan")


def _run(args, timeout):
    if not os.path.exists(EXE):
        pytest.skip("host-sanitizer build absent (make -C exafmm-rd_amd/csrc asan)")
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"   # HIP runtime allocations live to exit
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    r = subprocess.run([EXE] + args, capture_output=True, text=True, timeout=timeout, env=env, cwd="/tmp")
    print(r.stdout[-3000:], r.stderr[-3000:])
    return r


def test_host_paths_under_sanitizers():
    r = _run([], 600)
    assert r.returncode == 0 and "OK (0 failures)" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr


@pytest.mark.gpu
@pytest.mark.skipif(os.environ.get("FMM3DPP_HOST_SANITIZER_GPU") != "1",
                    reason="the GPU leg of the sanitizer driver runs on request (FMM3DPP_HOST_SANITIZER_GPU=1)")
def test_device_paths_under_host_sanitizers():
    r = _run(["gpu"], 300)
    assert r.returncode == 0 and "OK (0 failures)" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr


def test_precompute_pool_under_thread_sanitizer(tmp_path):
    """The operator precompute's host thread pool under ThreadSanitizer (operators.cpp built with
    -fsanitize=thread beside tests/cpp/precompute_tsan.cpp): no data race is reported, and repeated
    builds of the same operators give the same bits."""
    exe = str(tmp_path / "precompute_tsan")
    src = os.path.join(ROOT, "exafmm-rd_amd", "csrc")
    b = subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=thread", "-I", src, "-I",
                        os.path.join(ROOT, "include"), os.path.join(src, "operators.cpp"),
                        os.path.join(ROOT, "tests", "cpp", "precompute_tsan.cpp"), "-o", exe, "-lpthread"],
                       capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-3000:]
    env = dict(os.environ)
    env["TSAN_OPTIONS"] = "halt_on_error=1"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0 and "OK" in r.stdout and "WARNING: ThreadSanitizer" not in r.stderr
