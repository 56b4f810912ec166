"""Train and evaluate the learned SCF guess for water clusters.

    python -m polymer_amd.guess_train --n 2048 --nfeat 48 --mode train --out polymer_amd/guess_water_631gss.npz
    python -m polymer_amd.guess_train --n 2048 --nfeat 48 --mode eval --out polymer_amd/guess_water_631gss.npz
    python -m polymer_amd.guess_train --basis cc-pVDZ --n 2048 --nfeat 64 --hidden 128,128 --eval-limit 60 \
        --core-subset 10 --out polymer_amd/guess_water_ccpvdz.npz

Data: n seeded geometries of (H2O)4 .. (H2O)8 (seed s has 4 + s % 5
waters), each built like the benchmark clusters (rigid placement, random
orientations, O-O >= 2.6 A, O inside a ball of radius 2.9 (n/4)^(1/3) A)
with 0.02 A bond and 2 degree angle jitter, 6-31G** (or --basis cc-pVDZ).  Converged densities
come from this library's FP64 SCF (RMS density change 1e-8).  The first
80 % of the seeds train; the last 20 % are held out.

Evaluation on the held-out geometries, every SCF converged to RMS density
change 1e-12 so that all final energies agree well within 1e-8 Eh: the reference
energy is that of the reference's diagonal guess; for each guess (diagonal,
superposition of atomic densities, core Hamiltonian on a subset, learned)
the iteration from which every energy stays within 1e-8 Eh of the reference
energy, the iteration at which the reference's convergence test (RMS
density change 1e-8) first passes, and the final energy difference.  The
weights and a JSON report are written next to --out.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import Context, PolyeriError, read_basis
from . import guess as G
from .geometry import water_cluster


def geometry(seed):
    n = 4 + seed % 5
    return water_cluster(n, 10_000 + seed, radius=2.9 * (n / 4) ** (1 / 3), bond_sigma=0.02, angle_sigma_deg=2.0)


def log(msg):
    print(msg, flush=True)


def energy_iterations(trace, e_ref, tol=1e-8):
    """First iteration (1-based) from which every SCF energy stays within tol
    of e_ref; None if the last one is not within."""
    ok = np.abs(np.asarray(trace)[:, 1] - e_ref) < tol
    i = len(ok)
    while i > 0 and ok[i - 1]:
        i -= 1
    return i + 1 if i < len(ok) else None


def rms_iterations(trace, conv=1e-8):
    """First iteration whose RMS density change is below conv (the reference's test)."""
    d = np.asarray(trace)[:, 0]
    hit = np.nonzero(d < conv)[0]
    return int(hit[0]) + 1 if len(hit) else None


def train(args, basis, ntrain):
    model = G.GuessModel(basis, nfeat=args.nfeat, hidden=tuple(int(h) for h in args.hidden.split(",")))
    full = G.GuessModel(basis, nfeat=G.NFEAT, hidden=(16,))      # all feature columns, for the data
    t0 = time.perf_counter()
    if args.data and os.path.exists(args.data):
        z = np.load(args.data, allow_pickle=False)
        data = {t: (z[f"{t}_X"], z[f"{t}_Y"]) for t in G.TYPES}
        log(f"data from {args.data}")
    else:
        data = {t: ([], []) for t in G.TYPES}
        for s in range(ntrain):
            atoms = geometry(s)
            b = read_basis(basis, atoms)
            ctx = Context(b, schwarz_tol=1e-12)
            # the converged density does not depend on the solver: purification
            # and commutator DIIS reach it in about half the iterations
            r = ctx.self_consistent_field(scfconv=1e-8, diis=2, scfcycles=100, solver=1, diis_mode=1)
            if not r["converged"]:
                continue
            smp, _, _ = full.samples(atoms, b, r["P"])
            for t in G.TYPES:
                data[t][0].extend(smp[t][0])
                data[t][1].extend(smp[t][1])
            if s % 100 == 0:
                log(f"data {s}/{ntrain} ({time.perf_counter() - t0:.0f} s)")
        if args.data:
            np.savez(args.data, **{f"{t}_{k}": np.array(data[t][i], np.float32)
                                   for t in G.TYPES for i, k in enumerate("XY")})
    data = {t: (np.asarray(data[t][0], np.float32)[:, :model.nfeat], data[t][1]) for t in G.TYPES}
    arrays = {t: (np.array(data[t][0], np.float32), np.array(data[t][1], np.float32)) for t in G.TYPES}
    log("pairs per type: " + ", ".join(f"{t} {len(arrays[t][0])}" for t in G.TYPES))
    t1 = time.perf_counter()
    ctx0 = Context(read_basis(basis, geometry(0)), schwarz_tol=1e-12)
    losses = model.fit(arrays, ctx0, epochs=args.epochs, log=log)
    train_s = time.perf_counter() - t1
    model.save(args.out)
    log(f"trained in {train_s:.1f} s: " + ", ".join(f"{t} {v:.3e}" for t, v in losses.items()))
    rep = {"basis": basis, "n_geometries": args.n, "n_train": ntrain, "nfeat": model.nfeat,
           "hidden": list(model.hidden),
           "train_pairs": {t: int(len(arrays[t][0])) for t in G.TYPES}, "train_loss": losses,
           "train_seconds": train_s, "data_seconds": t1 - t0, "epochs": args.epochs}
    with open(args.out.replace(".npz", "_train.json"), "w") as f:
        json.dump(rep, f, indent=1)


def evaluate(args, basis, ntrain):
    model = G.GuessModel.load(args.out)
    seeds = list(range(ntrain, args.n))[:args.eval_limit]
    rows = []
    t0 = time.perf_counter()
    for k, s in enumerate(seeds):
        atoms = geometry(s)
        b = read_basis(basis, atoms)
        ctx = Context(b, schwarz_tol=1e-12)
        S = ctx.one_electron()[0]
        tg = time.perf_counter()
        P = model.predict(ctx, b, atoms, S=S)
        tg = time.perf_counter() - tg
        guesses = {"diag": None, "sad": G.sad_guess(b, basis), "learned": P}
        if k < args.core_subset:
            guesses["core"] = G.core_guess(ctx, b)
        runs = {}
        for g, v in guesses.items():
            try:
                runs[g] = ctx.self_consistent_field(scfconv=1e-12, diis=2, scfcycles=200, guess=v)
            except PolyeriError as e:          # SCF CONVERGENCE NOT MET
                log(f"seed {s} guess {g}: {e}")
        row = {"seed": s, "nwater": len(atoms) // 3, "guess_s": tg, "conv": len(runs) == len(guesses)}
        if "diag" not in runs:
            rows.append(row)
            continue
        e_ref = runs["diag"]["e_total"]
        for g, r in runs.items():
            row[f"itE_{g}"] = energy_iterations(r["trace"], e_ref)
            row[f"itR_{g}"] = rms_iterations(r["trace"])
            row[f"dE_{g}"] = r["e_total"] - e_ref
        # the same guesses with purification + commutator DIIS to RMS 1e-8
        for g, v in guesses.items():
            if g == "core":
                continue
            try:
                rc_ = ctx.self_consistent_field(scfconv=1e-8, diis=2, scfcycles=100, guess=v, solver=1, diis_mode=1)
                row[f"itC_{g}"] = rc_["iterations"]
                row[f"dEC_{g}"] = rc_["e_total"] - e_ref
            except PolyeriError as e:
                log(f"seed {s} guess {g} (commutator DIIS): {e}")
        rows.append(row)
        if k % 25 == 0:
            log(f"eval {k}/{len(seeds)} ({time.perf_counter() - t0:.0f} s): iterations to 1e-8 Eh "
                + ", ".join(f"{g} {row[f'itE_{g}']}" for g in runs))
    conv = [r for r in rows if r["conv"]]

    def mean(key):
        v = [r[key] for r in conv if r.get(key) is not None]
        return float(np.mean(v)) if v else None

    rep = {"basis": basis, "n_geometries": args.n, "n_train": ntrain, "n_test": len(rows),
           "not_converged": len(rows) - len(conv), "energy_tolerance": 1e-8}
    for g in ("diag", "sad", "learned", "core"):
        rep[f"mean_iterations_to_1e-8Eh_{g}"] = mean(f"itE_{g}")
        rep[f"mean_iterations_rms1e-8_{g}"] = mean(f"itR_{g}")
        if g != "core":
            rep[f"mean_iterations_cdiis_rms1e-8_{g}"] = mean(f"itC_{g}")
            d = [abs(r[f"dEC_{g}"]) for r in conv if f"dEC_{g}" in r]
            rep[f"max_abs_dE_cdiis_{g}"] = float(max(d)) if d else None
        d = [abs(r[f"dE_{g}"]) for r in conv if f"dE_{g}" in r]
        rep[f"max_abs_final_dE_{g}"] = float(max(d)) if d else None
    rep["mean_iterations_to_1e-8Eh_diag_core_subset"] = float(np.mean(
        [r["itE_diag"] for r in conv if "itE_core" in r and r["itE_diag"] is not None])) if any(
        "itE_core" in r for r in conv) else None
    rep["mean_guess_seconds"] = float(np.mean([r["guess_s"] for r in rows]))
    rep["rows"] = rows
    with open(args.out.replace(".npz", ".json"), "w") as f:
        json.dump(rep, f, indent=1)
    log(json.dumps({k: v for k, v in rep.items() if k != "rows"}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--out", default="polymer_amd/guess_water_631gss.npz")
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--core-subset", type=int, default=100)
    ap.add_argument("--eval-limit", type=int, default=None)
    ap.add_argument("--mode", choices=("all", "train", "eval"), default="all")
    ap.add_argument("--basis", default="6-31Gss", choices=("6-31Gss", "cc-pVDZ"))
    ap.add_argument("--nfeat", type=int, default=G.NFEAT,
                    help=f"feature columns ({G.NFEAT_GEOM}: geometry only; {G.NFEAT}: with the field features)")
    ap.add_argument("--hidden", default="128,128", help="hidden layer widths")
    ap.add_argument("--data", default=None, help="cache of the training pairs (all feature columns): read if it "
                                                "exists, else written after the SCFs")
    args = ap.parse_args()
    basis = args.basis
    ntrain = int(round(0.8 * args.n))
    if args.mode in ("all", "train"):
        train(args, basis, ntrain)
    if args.mode in ("all", "eval"):
        evaluate(args, basis, ntrain)
    return 0


if __name__ == "__main__":
    sys.exit(main())
