"""Polyethylene on one GPU: the periodic direct-SCF and k-point MP2 path at
the size of a production polymer run.

    python -m polymer_amd.run_polyethylene [--cutoff1 5] [--cutoff2 10] [--kpoints 16] [--mp2]

Unit cell C2H4 (C-C 1.53 A zig-zag, period 2.55 A along the chain, which is
the x axis of the periodic code; helix angle 0), 6-31G**, CUTOFF1 / CUTOFF2
lattice cells (src/pselfconsistentfield.f90, src/pfourindexdirect.f90),
2 * kpoints k-points, direct RHF to RMS density change 1e-8, then (--mp2)
MP2_ENERGY_ALG1 with the two carbon 1s bands frozen, transformed once with
the FP64 MFMA kernel and once with the BF16x3 kernel.  Prints one JSON line:
energy per cell, iterations, timings, REDUCED_CEL1X, band gap, the largest
|e(k) - e(-k)| (time reversal), E(2) of both transforms and their
difference.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import Context, read_basis

ANG = 1.0 / 0.529177249


def cell():
    h = 0.4229
    return [(6, 0.0, 0.0, h * ANG), (6, 1.275 * ANG, 0.0, -h * ANG),
            (1, 0.0, 0.876 * ANG, 1.071 * ANG), (1, 0.0, -0.876 * ANG, 1.071 * ANG),
            (1, 1.275 * ANG, 0.876 * ANG, -1.071 * ANG), (1, 1.275 * ANG, -0.876 * ANG, -1.071 * ANG)]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--cutoff1", type=int, default=5)
    ap.add_argument("--cutoff2", type=int, default=10)
    ap.add_argument("--kpoints", type=int, default=16, help="KPOINTS: 2 * kpoints k-points")
    ap.add_argument("--basis", default="6-31Gss")
    ap.add_argument("--mp2", action="store_true")
    args = ap.parse_args()
    atoms = cell()
    b = read_basis(args.basis, atoms)
    t0 = time.perf_counter()
    ctx = Context(b, schwarz_tol=1e-12, period=2.55 * ANG, cel1=args.cutoff1, cel2=args.cutoff2)
    t1 = time.perf_counter()
    r = ctx.self_consistent_field_periodic(kpoints=args.kpoints, scfconv=1e-8, diis=5, scfcycles=150)
    t2 = time.perf_counter()
    kvc, iocc, eps = args.kpoints, r["iocc"], r["eps"]
    print(f"scf: E {r['e_total']:.10f} per cell, {r['iterations']} iterations, {t2 - t1:.1f} s", flush=True)
    out = {"workload": f"polyethylene C2H4 {args.basis}, CUTOFF1 {args.cutoff1} / CUTOFF2 {args.cutoff2}, "
                       f"{2 * kvc if kvc else 1} k-points",
           "ncgs": b.ncgs, "reduced_cel1": ctx.reduced_cel1, "e_total_per_cell": r["e_total"],
           "iterations": r["iterations"], "context_s": t1 - t0, "scf_s": t2 - t1, "jk_s": r["jk_seconds"],
           "quartets": r["quartets"],
           "band_gap": float(np.min(eps[:, iocc]) - np.max(eps[:, :iocc])),
           "time_reversal_max": float(max([np.max(np.abs(eps[kvc + i, :iocc + 4] - eps[kvc - i, :iocc + 4]))
                                           for i in range(1, kvc)] or [0.0]))}
    if args.mp2:
        mp2 = {}
        for tr in ("f64", "bf16x3"):
            print(f"mp2 {tr}", flush=True)
            s, t, sec = ctx.mp2_energy_periodic(kvc, r["co"], eps, r["nall"], 2, iocc, transform=tr, timing=True)
            mp2[tr] = {"e2_singlet": s, "e2_triplet": t, "e2": s + t, "seconds": sec}
        out["mp2"] = mp2
        out["e2_bf16x3_minus_f64"] = mp2["bf16x3"]["e2"] - mp2["f64"]["e2"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
