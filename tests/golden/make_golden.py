"""Extract the golden numbers of the reference's shipped outputs into
tests/golden/reference_outputs.json.

The reference ships captured stdout of its own runs (data/*_out).  This
script reads those text files (data, not code) and records the energies,
iteration counts and basis-set counts that pin the oracle.  It only runs in
a checkout that has the reference tree:

    python tests/golden/make_golden.py /root/reference/data
"""
import json
import os
import re
import sys

CASES = {
    "H2O_ccpVDZ_spherical": "H2O_CCSD_ccpVDZ_out",
    "BH_STO3G": "BH_EOM_out",
    "FH_STO3G_zeroT": "FH_finiteT_MBPT_zeroT_HF_ref_out",
    "FH_STO3G_T1e5": "FH_finiteT_MBPT_finiteT_HF_ref_out",
}

PATTERNS = {
    "e_nuc": r"^NUCLEAR REPULSION ENERGY\s+=\s+(\S+) HARTREE",
    "e_kin": r"^KINETIC ENERGY\s+=\s+(\S+) HARTREE",
    "e_nuc_att": r"^NUCLEAR ATTRACTION ENERGY\s+=\s+(\S+) HARTREE",
    "e_coul": r"^COULOMB ENERGY\s+=\s+(\S+) HARTREE",
    "e_hfx": r"^HF-EXCHANGE ENERGY\s+=\s+(\S+) HARTREE",
    "e_total": r"^TOTAL SCF ENERGY\s+=\s+(\S+) HARTREE",
    "npgs": r"^NUMBER OF PRIMITIVE GAUSSIANS  \(INCLUDING GHOST FUNCTIONS\) =\s+(\d+)",
    "ncgs": r"^NUMBER OF CONTRACTED GAUSSIANS \(INCLUDING GHOST FUNCTIONS\) =\s+(\d+)",
    "npshell": r"^NUMBER OF SHELLS\s+\(INCLUDING GHOST FUNCTIONS\) =\s+(\d+)",
    "e2_alg1": r"^E\(2\)\s+=\s+(\S+) HARTREE",
    "e2_cc": r"^E\(2\) =\s+(\S+) \(",
    "e2_zero_t": r"^ZERO-TEMPERATURE MP2 ENERGY =\s+(\S+) HARTREE",
    "e3_cc": r"^E\(3\) =\s+(\S+) \(",
    "e_cc_singles": r"^E\(SINGLES\)\s+=\s+(\S+) HARTREE",
    "e_cc_doubles": r"^E\(DOUBLES\)\s+=\s+(\S+) HARTREE",
    "e_cc_corr": r"^E\(CORRELATION\)\s+=\s+(\S+) HARTREE",
    "e_ccsd": r"^E\(CCSD\)\s+=\s+(\S+) HARTREE",
}


def parse(path):
    out = {"source": os.path.basename(path)}
    lines = open(path).read().splitlines()
    for key, pat in PATTERNS.items():
        rx = re.compile(pat)
        for n, line in enumerate(lines, 1):
            m = rx.match(line)
            if m:
                v = m.group(1)
                out[key] = int(v) if v.isdigit() else float(v)
                out[key + "_line"] = n
                if key != "e_nuc":
                    break
    # SCF iterations: the last "ITR" table row before "SCF CONVERGED"
    it = None
    trace = []
    for line in lines:
        if line.startswith("SCF CONVERGED"):
            break
        m = re.match(r"^\s*(\d+)\s+(\d+)\s+\S+E[-+]\d+\s+(\S+)\s+(\S+)", line)
        if m:
            it = int(m.group(1))
            trace.append([it, int(m.group(2)), float(m.group(3)), float(m.group(4))])
    out["iterations"] = it
    for n, line in enumerate(lines):
        if line.strip() == "X   MAXIMUM OVERLAP":
            out["max_overlap"] = float(lines[n + 1].split()[1])  # F18.13
            out["max_overlap_line"] = n + 2
            break
    out["trace"] = trace  # ITR, DIIS, DENSITY CHANGE, TOTAL ENERGY
    # band tables (PRINT_ENERGYBANDS): the KX = 0 block after the HF / MP2
    # quasi-particle headers, rows "label value" from the top band down
    for key, head in (("hf_bands", "HF/DFT ENERGY BANDS"),
                      ("qp_bands", "MP2 QUASI-PARTICLE ENERGY BANDS (DIAGONAL, FREQ-INDEP)")):
        for n, line in enumerate(lines):
            if line.startswith(head):
                rows, started = [], False
                for line2 in lines[n + 1:]:
                    if line2.startswith("KX ="):
                        started = True
                        continue
                    if not started or line2.startswith("-----"):
                        if started and rows:
                            break
                        continue
                    m = re.match(r"^\s*(\S+)\s+(-?\d+\.\d+)\s*$", line2)
                    if m:
                        rows.append([m.group(1), float(m.group(2))])
                out[key] = rows
                out[key + "_line"] = n + 1
                break
    # coupled-cluster iteration table: rows "ITER RESIDUUM CORRELATION CPU"
    for n, line in enumerate(lines):
        if line.startswith("CCSD ITERATION HAS STARTED"):
            rows = []
            for line2 in lines[n + 3:]:
                if line2.startswith("-----"):
                    break
                m = re.match(r"^\s*(\d+)\s+(\S+)\s+(\S+)\s+\S+\s*$", line2)
                if m:
                    rows.append([int(m.group(1)), float(m.group(2)), float(m.group(3))])
            out["cc_trace"] = rows
            out["cc_trace_line"] = n + 1
            break
    return out


if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/data"
    res = {k: parse(os.path.join(src, v)) for k, v in CASES.items()}
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "reference_outputs.json"), "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)
    print(json.dumps(res, indent=1, sort_keys=True))
