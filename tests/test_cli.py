"""The test_pbc driver over the drop-in C++ API: the reference's flags and printed blocks
(3dpp/src/argument.cpp:7-101, 3dpp/test/test_pbc.cpp, 3dpp/src/body.cpp:192-205)."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(os.path.dirname(HERE), "exafmm-rd_amd", "cpp", "test_pbc")


def _run(args, timeout=600):
    if not os.path.exists(EXE):
        pytest.skip("exafmm-rd_amd/cpp/test_pbc not built")
    return subprocess.run([EXE] + args, capture_output=True, text=True, timeout=timeout)


def test_parameter_block_and_flags():
    """No algorithm selected: only the title and the parameter block, with the reference's formats."""
    r = _run(["-P", "6", "-n", "100", "-i", "2", "-a", "x", "--ewald_ksize=9", "-m32", "--seed", "7"])
    assert r.returncode == 0, r.stdout + r.stderr
    out = r.stdout
    assert out.startswith("<rtfmm_3dpp_test_pbc>\n\ninput parameters=[\n")
    for key, val in [("P", "6"), ("n", "100"), ("ncrit", "32"), ("images", "2"), ("cycle", "6.2832"),
                     ("ewald_ksize", "9"), ("(f,d,e)", "(0,0,0)"), ("num_compare", "100"), ("seed", "7"),
                     ("dipole_correction", "1"), ("divide_4pi", "0")]:
        assert ("  %-20s = %s\n" % (key, val)) in out, key


def test_setting_t_and_bad_flags():
    r = _run(["-a", "x", "--setting_t", "1"])
    assert "  dipole_correction    = 0" in r.stdout and "  divide_4pi           = 1" in r.stdout
    bad = _run(["--no_such_flag", "1"])
    assert bad.returncode == 1 and bad.stdout.startswith("<rtfmm_3dpp_test_pbc>") and "error : " in bad.stdout


@pytest.mark.gpu
def test_reference_log_blocks_reproduced():
    """P=6, images=2, N=24000, ncrit=128 (the reference's logged run): the FMM vs Ewald block
    carries the logged energies (FMM 2.639499940725e+01 to 2e-10, Ewald 2.773120969004e+01 to 1e-9)."""
    with open(os.path.join(HERE, "golden", "log_energies.json")) as f:
        rows = json.load(f)["periodic"]
    row = [r for r in rows if r["P"] == 6 and r["images"] == 2][0]
    r = _run(["-P", "6", "-n", "24000", "-m", "128", "-i", "2", "-a", "fe"])
    assert r.returncode == 0, r.stdout + r.stderr
    print(r.stdout, flush=True)
    assert "[FMM time measured : " in r.stdout and "[EWALD time measured : " in r.stdout
    block = r.stdout[r.stdout.index("FMM vs Ewald"):]
    e1 = float(re.search(r"p-energy1 : (\S+)", block).group(1))
    e2 = float(re.search(r"p-energy2 : (\S+)", block).group(1))
    assert abs(e1 - row["energy_fmm"]) <= 2e-10 * abs(row["energy_fmm"])
    assert abs(e2 - row["energy_ref"]) <= 1e-9 * abs(row["energy_ref"])
    l2f = float(re.search(r"L2  \(f\)  : (\S+)", block).group(1))
    assert abs(l2f - row["l2f"]) <= 1e-4 * row["l2f"]


@pytest.mark.gpu
def test_fmm_driver_nonperiodic_log():
    """test_fmm (3dpp/test/test_fmm.cpp) with the logged free-space setting (setting_t: no dipole term,
    no net-charge removal, 1/(4 pi)), P=6, N=24000, ncrit=128: FMM energy and direct energy as logged
    (slurm-152673.out), L2(p) and L2(f) to their 6 printed digits."""
    exe = os.path.join(os.path.dirname(HERE), "exafmm-rd_amd", "cpp", "test_fmm")
    if not os.path.exists(exe):
        pytest.skip("exafmm-rd_amd/cpp/test_fmm not built")
    with open(os.path.join(HERE, "golden", "log_energies.json")) as f:
        row = [r for r in json.load(f)["nonperiodic"] if r["P"] == 6][0]
    r = subprocess.run([exe, "-P", "6", "-n", "24000", "-m", "128", "-i", "0", "--setting_t", "1"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    print(r.stdout, flush=True)
    assert r.stdout.startswith("rtfmm_3dp_test_fmm\n")
    block = r.stdout[r.stdout.index("FMM vs Direct"):]
    e1 = float(re.search(r"p-energy1 : (\S+)", block).group(1))
    e2 = float(re.search(r"p-energy2 : (\S+)", block).group(1))
    assert abs(e1 - row["energy_fmm"]) <= 2e-12 * abs(row["energy_fmm"])
    assert abs(e2 - row["energy_ref"]) <= 1e-12 * abs(row["energy_ref"])
    l2p = float(re.search(r"L2  \(p\)  : (\S+)", block).group(1))
    l2f = float(re.search(r"L2  \(f\)  : (\S+)", block).group(1))
    assert abs(l2p - row["l2p"]) <= 1e-4 * row["l2p"] and abs(l2f - row["l2f"]) <= 1e-4 * row["l2f"]
