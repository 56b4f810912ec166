"""Captured-graph J/K builds and the launch-shape autotuner.

Without per-class profiling a J/K build is captured once into a HIP graph
and replayed; with profiling it runs eagerly.  The graph must give the eager
result on every replay, also with eager work (other builds, host syncs)
between the replays, and the autotuned launch shapes must give the same
J/K as the default ones.  Bars: 1e-12 absolute (the J/K sums are atomics,
so runs differ only in the order of FP64 additions)."""
import numpy as np
import pytest

import polymer_amd as pa
from polymer_amd import geometry as G

pytestmark = pytest.mark.gpu


def _density(b, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((b.ncgs, 20))
    ghost = [i for i in range(b.ncgs) if i not in set(b.real_functions)]
    x[ghost] = 0.0
    return 0.05 * x @ x.T


@pytest.mark.parametrize("dens_tol", [0.0, 1e-10])
def test_graph_replays_match_eager(dens_tol):
    atoms = G.water_cluster(4, 7)
    print(f"case: graph replay (H2O)4 dens_tol {dens_tol}", flush=True)
    b = pa.read_basis("cc-pVDZ", atoms)
    ctx = pa.Context(b, schwarz_tol=1e-12, precision=pa.MIXED)
    p = _density(b, 0)
    q = _density(b, 1)
    ctx.set_profiling(True)                      # eager
    je, ke = ctx.fourindex_eri(p, dens_tol=dens_tol)
    jq, kq = ctx.fourindex_eri(q, dens_tol=dens_tol)
    ctx.set_profiling(False)                     # captured graph
    for rep in range(4):
        j, k = ctx.fourindex_eri(p, dens_tol=dens_tol)
        assert np.max(np.abs(j - je)) < 1e-12 and np.max(np.abs(k - ke)) < 1e-12, rep
        # eager work in between: another density through the eager path
        ctx.set_profiling(True)
        j2, k2 = ctx.fourindex_eri(q, dens_tol=dens_tol)
        ctx.set_profiling(False)
        assert np.max(np.abs(j2 - jq)) < 1e-12 and np.max(np.abs(k2 - kq)) < 1e-12, rep


def test_autotune_keeps_results():
    atoms = G.water_cluster(4, 8)
    print("case: autotune (H2O)4", flush=True)
    b = pa.read_basis("cc-pVDZ", atoms)
    ctx = pa.Context(b, schwarz_tol=1e-12)
    p = _density(b, 2)
    j0, k0 = ctx.fourindex_eri(p)
    ctx.autotune(reps=1)
    st = ctx.stats()
    shapes = {(c["kernel"], c["block"], c["grid"]) for c in st["classes"] if c["quartets"]}
    print("shapes", sorted(shapes), flush=True)
    j1, k1 = ctx.fourindex_eri(p)
    assert np.max(np.abs(j1 - j0)) < 1e-12 and np.max(np.abs(k1 - k0)) < 1e-12
    s1 = ctx.self_consistent_field(scfconv=1e-8, diis=2)
    s0 = pa.Context(b, schwarz_tol=1e-12).self_consistent_field(scfconv=1e-8, diis=2)
    assert abs(s1["e_total"] - s0["e_total"]) < 1e-10
    assert s1["iterations"] == s0["iterations"]
