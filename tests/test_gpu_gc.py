"""Generally contracted (ls|l's) lists against the shell-pair lists.

cc-pVDZ's O 1s/2s and H 1s/2s shells share exponents, so the classes
(ss|ss) (ps|ss) (ps|ps) (ds|ss) run by default over pairs of shell groups:
one primitive loop per group quartet, each member quartet weighted by its
own coefficients and digested when it passes the same Schwarz test as in
the shell-pair lists.  (pp|ss) (pp|ps) (ds|ps) (dp|ss) can run with
shell-pair bras against group-pair kets (on where the autotuner finds it
faster; switched on here for all).  Both forms evaluate the same quartets, so J/K agree
to the rounding of the FP64 atomics (1e-12 absolute); against the oracle
the bar is the usual 1e-10."""
import numpy as np
import pytest

import oracle
import polymer_amd as pa
from polymer_amd import geometry as G
from tests.conftest import basis_path

pytestmark = pytest.mark.gpu

GC_CLASSES = {"0000", "1000", "1010", "2000"}
GCK_CLASSES = {"1100", "1110", "2010", "2100"}     # shell-pair bra, group-pair kets


def _density(b, seed, rank=20):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((b.ncgs, rank))
    ghost = [i for i in range(b.ncgs) if i not in set(b.real_functions)]
    x[ghost] = 0.0
    return 0.05 * x @ x.T


@pytest.mark.parametrize("tol", [0.0, 1e-12, 1e-9])
def test_gc_matches_pair_lists(tol):
    atoms = G.water_cluster(4, 3)
    print(f"case: generally contracted vs pair lists (H2O)4 cc-pVDZ Schwarz {tol}", flush=True)
    b = pa.read_basis("cc-pVDZ", atoms)
    ctx = pa.Context(b, schwarz_tol=tol)
    p = _density(b, 5)
    kinds = {c["cls"]: c["kernel"] for c in ctx.stats()["classes"] if c["quartets"]}
    assert all(kinds[c] == "gc" for c in GC_CLASSES if c in kinds), kinds
    j1, k1 = ctx.fourindex_eri(p)
    q1 = ctx.stats()["quartets"]
    ctx.set_general_contraction(True)        # also the shell-pair-bra lists
    kinds = {c["cls"]: c["kernel"] for c in ctx.stats()["classes"] if c["quartets"]}
    assert all(kinds[c] == "gc" for c in GC_CLASSES | GCK_CLASSES if c in kinds), kinds
    j2, k2 = ctx.fourindex_eri(p)
    ctx.set_general_contraction(False)
    kinds = {c["cls"]: c["kernel"] for c in ctx.stats()["classes"] if c["quartets"]}
    assert "gc" not in kinds.values(), kinds
    j0, k0 = ctx.fourindex_eri(p)
    print(f"max |dJ| {np.max(np.abs(j1 - j0)):.2e} |dK| {np.max(np.abs(k1 - k0)):.2e}; all lists: "
          f"{np.max(np.abs(j2 - j0)):.2e} {np.max(np.abs(k2 - k0)):.2e}", flush=True)
    assert np.max(np.abs(j1 - j0)) < 1e-12 and np.max(np.abs(k1 - k0)) < 1e-12
    assert np.max(np.abs(j2 - j0)) < 1e-12 and np.max(np.abs(k2 - k0)) < 1e-12
    assert ctx.stats()["quartets"] == q1


@pytest.mark.parametrize("dens_tol", [1e-10, 1e-7])
def test_gc_density_screened_matches_pair_lists(dens_tol):
    """Density-screened builds: the generally contracted kernel skips group
    quartets by a group bound and digests the member quartets that pass the
    shell-pair lists' own test, so both forms keep the same quartets and
    give the same J/K (1e-12)."""
    atoms = G.water_cluster(4, 6)
    print(f"case: screened generally contracted vs pair lists (H2O)4 dens_tol {dens_tol}", flush=True)
    b = pa.read_basis("cc-pVDZ", atoms)
    ctx = pa.Context(b, schwarz_tol=1e-12)
    p = _density(b, 8)
    p[np.abs(p) < 1e-3] *= 1e-6            # a sparse-ish density, so the screen drops quartets
    ctx.set_general_contraction(True)
    j1, k1 = ctx.fourindex_eri(p, dens_tol=dens_tol)
    kept1 = ctx.stats()["last_quartets"]
    ctx.set_general_contraction(False)
    j0, k0 = ctx.fourindex_eri(p, dens_tol=dens_tol)
    kept0 = ctx.stats()["last_quartets"]
    print(f"kept {kept1} vs {kept0} of {ctx.stats()['quartets']}; max |dJ| {np.max(np.abs(j1 - j0)):.2e} "
          f"|dK| {np.max(np.abs(k1 - k0)):.2e}", flush=True)
    assert kept1 == kept0 < ctx.stats()["quartets"]
    assert np.max(np.abs(j1 - j0)) < 1e-12 and np.max(np.abs(k1 - k0)) < 1e-12


def test_gc_h2o_vs_oracle():
    atoms = G.water()
    print("case: generally contracted J/K H2O cc-pVDZ vs oracle", flush=True)
    b = pa.read_basis("cc-pVDZ", atoms)
    ctx = pa.Context(b, schwarz_tol=0.0)
    assert any(c["kernel"] == "gc" for c in ctx.stats()["classes"])
    rng = np.random.default_rng(11)
    p = rng.standard_normal((b.ncgs, b.ncgs))
    p = 0.05 * (p + p.T)
    j, k = ctx.fourindex_eri(p)
    jr, kr, _ = oracle.Basis(basis_path("cc-pVDZ"), atoms).fourindex_eri(p, 0.0)
    assert np.max(np.abs(j - jr)) < 1e-10 and np.max(np.abs(k - kr)) < 1e-10


def test_gc_scf_same_energy():
    atoms = G.water_cluster(3, 2)
    print("case: SCF (H2O)3 cc-pVDZ with and without generally contracted lists", flush=True)
    b = pa.read_basis("cc-pVDZ", atoms)
    on = pa.Context(b, schwarz_tol=1e-12)
    off = pa.Context(b, schwarz_tol=1e-12)
    off.set_general_contraction(False)
    r1 = on.self_consistent_field(scfconv=1e-9, diis=2)
    r0 = off.self_consistent_field(scfconv=1e-9, diis=2)
    print(f"E {r1['e_total']:.12f} vs {r0['e_total']:.12f}, iterations {r1['iterations']} {r0['iterations']}",
          flush=True)
    assert abs(r1["e_total"] - r0["e_total"]) < 1e-10
    assert r1["iterations"] == r0["iterations"]


def test_basis_without_groups_has_no_gc():
    """6-31G** shells share no exponents on a centre: no generally
    contracted lists are built."""
    atoms = G.water_cluster(2, 1)
    ctx = pa.Context(pa.read_basis("6-31Gss", atoms), schwarz_tol=1e-12)
    assert all(c["kernel"] != "gc" for c in ctx.stats()["classes"])
