"""The J/K paths the benchmark times, checked against the oracle.

The benchmark's FP64 build runs autotuned launch shapes (segment, generally
contracted and wavefront kernels with tuned blocks, grids and primitive
batches) replayed from a captured HIP graph; the mixed build evaluates the
wavefront classes' quartets below Q_ab Q_cd = 1e-5 with an FP32 recursion.
Here both are compared with the oracle (the C restatement of
FOURINDEX_ERI, src/pfourindexdirect.f90:1-814) on the same density:

* autotuned graph replays at Schwarz 0 against the oracle at INTTOL 0,
  1e-10 absolute, on every replay;
* the mixed build against the oracle, 1e-6 absolute (FP32 recursion for
  quartets whose integrals are below 1e-5);
* at the benchmark's (H2O)16 cc-pVDZ size, where the oracle is too slow, the
  autotuned graph path against the default-shape eager path, 1e-10.
"""
import os

import numpy as np
import pytest

import oracle
import polymer_amd as pa
from polymer_amd import geometry as G

pytestmark = pytest.mark.gpu


def _density(b, seed, scale=0.05):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((b.ncgs, 20))
    ghost = [i for i in range(b.ncgs) if i not in set(b.real_functions)]
    x[ghost] = 0.0
    return scale * x @ x.T


def _oracle_jk(atoms, p, inttol=0.0):
    ob = oracle.Basis(os.path.join(pa.BASIS_DIR, "cc-pVDZ"), atoms)
    j, k, _ = ob.fourindex_eri(p, inttol)
    return j, k


def test_autotuned_graph_replays_vs_oracle():
    atoms = G.water_cluster(2, 5)
    print("case: autotuned graph (H2O)2/cc-pVDZ Schwarz 0 vs oracle", flush=True)
    b = pa.read_basis("cc-pVDZ", atoms)
    p = _density(b, 3)
    jr, kr = _oracle_jk(atoms, p)
    ctx = pa.Context(b, schwarz_tol=0.0)
    ctx.autotune(reps=1)
    kinds = sorted({(c["cls"], c["kernel"], c["block"], c["grid"]) for c in ctx.stats()["classes"] if c["quartets"]})
    print("tuned shapes:", kinds, flush=True)
    q = _density(b, 4)
    for rep in range(3):
        j, k = ctx.fourindex_eri(p)
        dj, dk = np.max(np.abs(j - jr)), np.max(np.abs(k - kr))
        print(f"replay {rep}: max|dJ| {dj:.2e} max|dK| {dk:.2e}", flush=True)
        assert dj < 1e-10 and dk < 1e-10, rep
        ctx.fourindex_eri(q)               # another density's graph in between


def test_mixed_jk_vs_oracle():
    atoms = G.water_cluster(2, 5)
    print("case: mixed J/K (H2O)2/cc-pVDZ vs oracle", flush=True)
    b = pa.read_basis("cc-pVDZ", atoms)
    p = _density(b, 5)
    jr, kr = _oracle_jk(atoms, p)
    ctx = pa.Context(b, schwarz_tol=0.0, precision=pa.MIXED, mixed_cut=1e-5)
    st = ctx.stats()
    n32 = sum(c["quartets"] for c in st["classes"] if c["prec"] == "fp32")
    assert n32 > 0, "no quartet below the cut"
    j, k = ctx.fourindex_eri(p)
    dj, dk = np.max(np.abs(j - jr)), np.max(np.abs(k - kr))
    print(f"fp32 quartets {n32} of {st['quartets']}: max|dJ| {dj:.2e} max|dK| {dk:.2e}", flush=True)
    assert dj < 1e-6 and dk < 1e-6
    # the FP32 quartets are a strict subset: the FP64 build of the same
    # context type at the same screen agrees with the oracle to 1e-10
    j64, k64 = pa.Context(b, schwarz_tol=0.0).fourindex_eri(p)
    assert np.max(np.abs(j64 - jr)) < 1e-10 and np.max(np.abs(k64 - kr)) < 1e-10


def test_water16_autotuned_graph_vs_default_shapes():
    atoms = G.water_cluster(16, 0)
    print("case: (H2O)16/cc-pVDZ autotuned graph vs default eager", flush=True)
    b = pa.read_basis("cc-pVDZ", atoms)
    p = _density(b, 6, scale=0.02)
    ref = pa.Context(b, schwarz_tol=1e-12)
    ref.set_graph(False)
    j0, k0 = ref.fourindex_eri(p)
    ref.close()
    ctx = pa.Context(b, schwarz_tol=1e-12)
    ctx.autotune(reps=1)
    for rep in range(2):
        j, k = ctx.fourindex_eri(p)
        dj, dk = np.max(np.abs(j - j0)), np.max(np.abs(k - k0))
        print(f"replay {rep}: max|dJ| {dj:.2e} max|dK| {dk:.2e}", flush=True)
        assert dj < 1e-10 and dk < 1e-10
    # the mixed build of the same cluster, against the FP64 one
    mx = pa.Context(b, schwarz_tol=1e-12, precision=pa.MIXED, mixed_cut=1e-5)
    jm, km = mx.fourindex_eri(p)
    dj, dk = np.max(np.abs(jm - j0)), np.max(np.abs(km - k0))
    print(f"mixed: max|dJ| {dj:.2e} max|dK| {dk:.2e}", flush=True)
    assert dj < 1e-6 and dk < 1e-6
