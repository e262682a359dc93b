"""Extract the reference's logged end-to-end energies into tests/golden/log_energies.json.

Reads the slurm logs the reference ships under 3dpp/exp/ (run of its own test_pbc / test_fmm
executables, N=24000, drand48 seed 5, cycle 2*pi, ncrit 128).  Each row records the run
parameters and the printed `p-energy1` (FMM, sum q*p) and `p-energy2` (Ewald or direct) with the
printed L2 errors.  Run from the repository root with the reference mounted at /root/reference.
"""
import json
import re
import sys

LOGS = {
    "periodic": "/root/reference/3dpp/exp/slurm-153086.out",     # test_pbc -a fe, P=4..8, i=0..9
    "nonperiodic": "/root/reference/3dpp/exp/slurm-152673.out",  # test_fmm --setting_t 1, P=2..20
}


def parse(path):
    rows, cur = [], {}
    keys = {"P": int, "n": int, "ncrit": int, "images": int, "cycle": float, "seed": int,
            "dipole_correction": int, "zero_netcharge": int, "divide_4pi": int, "setting_t": int}
    for line in open(path, errors="replace"):
        m = re.match(r"^\s+(\w+)\s+= (\S+)", line)
        if m and m.group(1) in keys:
            cur[m.group(1)] = keys[m.group(1)](m.group(2))
            continue
        m = re.match(r"L2  \(p\)\s+: (\S+)\s+L2  \(f\)\s+: (\S+)\s+L2  \(e\)\s+: (\S+)", line)
        if m:
            cur["l2p"], cur["l2f"], cur["l2e"] = map(float, m.groups())
        m = re.match(r"p-energy1 : (\S+)", line)
        if m:
            cur["energy_fmm"] = float(m.group(1))
        m = re.match(r"p-energy2 : (\S+)", line)
        if m:
            cur["energy_ref"] = float(m.group(1))
            rows.append(dict(cur))
        m = re.match(r"pnrm = (\S+)", line)
        if m:
            cur["pnrm"] = float(m.group(1))
    return rows


if __name__ == "__main__":
    out = {k: parse(v) for k, v in LOGS.items()}
    out["source"] = {k: v.replace("/root/reference/", "") for k, v in LOGS.items()}
    json.dump(out, open(sys.argv[1] if len(sys.argv) > 1 else "tests/golden/log_energies.json", "w"), indent=1)
    print({k: len(v) for k, v in out.items() if isinstance(v, list)})
