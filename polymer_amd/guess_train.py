"""Train and evaluate the learned SCF guess for water clusters.

    python -m polymer_amd.guess_train --n 2048 --out polymer_amd/guess_water_631gss.npz

Data: n seeded geometries of (H2O)4 .. (H2O)8 (seed s has 4 + s % 5
waters), each built like the benchmark clusters (rigid placement, random
orientations, O-O >= 2.6 A, O inside a ball of radius 2.9 (n/4)^(1/3) A)
with 0.02 A bond and 2 degree angle jitter, 6-31G**.  Converged densities
come from this library's FP64 SCF (RMS density change 1e-8).  Seeds with
s % 5 == 0 .. 3 of the first 80 % train; the last 20 % of seeds are held out.

Evaluation on the held-out geometries: SCF iterations from the reference's
diagonal guess, the core-Hamiltonian guess (on a subset) and the learned
guess, all converged to 1e-8, and the energy differences.  The weights and
a JSON report are written next to --out.
"""
import argparse
import json
import math
import sys
import time

import numpy as np

from . import Context, read_basis
from . import guess as G
from .geometry import water_cluster


def geometry(seed):
    n = 4 + seed % 5
    return water_cluster(n, 10_000 + seed, radius=2.9 * (n / 4) ** (1 / 3), bond_sigma=0.02, angle_sigma_deg=2.0)


def log(msg):
    print(msg, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--out", default="polymer_amd/guess_water_631gss.npz")
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--core-subset", type=int, default=100)
    args = ap.parse_args()
    basis = "6-31Gss"
    ntrain = int(round(0.8 * args.n))
    model = G.GuessModel(basis)
    data = {t: ([], []) for t in G.TYPES}
    t0 = time.perf_counter()
    iters_train = []
    for s in range(ntrain):
        atoms = geometry(s)
        b = read_basis(basis, atoms)
        ctx = Context(b, schwarz_tol=1e-12)
        r = ctx.self_consistent_field(scfconv=1e-8, diis=2, scfcycles=100)
        if not r["converged"]:
            continue
        iters_train.append(r["iterations"])
        smp, _, _ = model.samples(atoms, b, r["P"])
        for t in G.TYPES:
            data[t][0].extend(smp[t][0])
            data[t][1].extend(smp[t][1])
        if s % 100 == 0:
            log(f"data {s}/{ntrain} ({time.perf_counter() - t0:.0f} s)")
    arrays = {t: (np.array(data[t][0], np.float32), np.array(data[t][1], np.float32)) for t in G.TYPES}
    log("pairs per type: " + ", ".join(f"{t} {len(arrays[t][0])}" for t in G.TYPES))
    t1 = time.perf_counter()
    ctx0 = Context(read_basis(basis, geometry(0)), schwarz_tol=1e-12)
    losses = model.fit(arrays, ctx0, epochs=args.epochs, log=log)
    train_s = time.perf_counter() - t1
    model.save(args.out)
    log(f"trained in {train_s:.1f} s: " + ", ".join(f"{t} {v:.3e}" for t, v in losses.items()))
    # held-out evaluation
    rows = []
    for k, s in enumerate(range(ntrain, args.n)):
        atoms = geometry(s)
        b = read_basis(basis, atoms)
        ctx = Context(b, schwarz_tol=1e-12)
        S = ctx.one_electron()[0]
        rd = ctx.self_consistent_field(scfconv=1e-8, diis=2, scfcycles=100)
        tg = time.perf_counter()
        P = model.predict(ctx, b, atoms, S=S)
        tg = time.perf_counter() - tg
        rl = ctx.self_consistent_field(scfconv=1e-8, diis=2, scfcycles=100, guess=P)
        row = {"seed": s, "nwater": len(atoms) // 3, "it_diag": rd["iterations"], "it_learned": rl["iterations"],
               "e_diag": rd["e_total"], "e_learned": rl["e_total"], "dE": rl["e_total"] - rd["e_total"],
               "conv": bool(rd["converged"] and rl["converged"]), "guess_s": tg}
        if k < args.core_subset:
            rc = ctx.self_consistent_field(scfconv=1e-8, diis=2, scfcycles=100, guess=G.core_guess(ctx, b))
            row["it_core"] = rc["iterations"]
            row["dE_core"] = rc["e_total"] - rd["e_total"]
        rows.append(row)
        if k % 50 == 0:
            log(f"eval {k}/{args.n - ntrain}: diag {row['it_diag']} learned {row['it_learned']} dE {row['dE']:.1e}")
    conv = [r for r in rows if r["conv"]]
    rep = {
        "basis": basis, "n_geometries": args.n, "n_train": ntrain, "n_test": len(rows),
        "train_pairs": {t: int(len(arrays[t][0])) for t in G.TYPES}, "train_loss": losses,
        "train_seconds": train_s, "epochs": args.epochs,
        "mean_iterations_diag": float(np.mean([r["it_diag"] for r in conv])),
        "mean_iterations_learned": float(np.mean([r["it_learned"] for r in conv])),
        "mean_iterations_core": float(np.mean([r["it_core"] for r in conv if "it_core" in r])),
        "mean_iterations_diag_core_subset": float(np.mean([r["it_diag"] for r in conv if "it_core" in r])),
        "max_abs_dE_learned": float(max(abs(r["dE"]) for r in conv)),
        "max_abs_dE_core": float(max(abs(r["dE_core"]) for r in conv if "dE_core" in r)),
        "not_converged": len(rows) - len(conv),
        "rows": rows,
    }
    with open(args.out.replace(".npz", ".json"), "w") as f:
        json.dump(rep, f, indent=1)
    log(json.dumps({k: v for k, v in rep.items() if k != "rows"}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
