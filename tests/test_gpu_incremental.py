"""Density-screened and incremental Fock builds.

A quartet is skipped when Q_ab Q_cd times the largest |P| over the six
density blocks it reads is below dens_tol, so each J/K element moves by at
most (number of skipped quartets touching it) x dens_tol.  Incremental SCF
builds J/K(P_i) = J/K(P_prev) + J/K(P_i - P_prev) with that screen applied to
the difference, and a full build every n iterations.

The plain FP64 path (pinned to the oracle and the goldens elsewhere) is the
reference.  Bars: screened J/K of a full density within 1e-8 of the
unscreened ones; a zero density evaluates no quartet; the incremental SCF
energy within 1 microhartree of the plain SCF energy with fewer quartets
evaluated in total."""
import numpy as np
import pytest

import polymer_amd as pa
from polymer_amd import geometry as G

pytestmark = pytest.mark.gpu


def _density(b, seed, scale=0.02):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((b.ncgs, 30))
    ghost = [i for i in range(b.ncgs) if i not in set(b.real_functions)]
    x[ghost] = 0.0
    return scale * x @ x.T


def test_screened_jk_water8():
    atoms = G.water_cluster(8, 2)
    print("case: density-screened J/K (H2O)8/cc-pVDZ", flush=True)
    b = pa.read_basis("cc-pVDZ", atoms)
    ctx = pa.Context(b, schwarz_tol=1e-12)
    p = _density(b, 3)
    j0, k0 = ctx.fourindex_eri(p)
    n_all = ctx.stats()["last_quartets"]
    j1, k1 = ctx.fourindex_eri(p, dens_tol=1e-10)
    n_scr = ctx.stats()["last_quartets"]
    dj, dk = np.max(np.abs(j1 - j0)), np.max(np.abs(k1 - k0))
    print(f"quartets {n_scr} of {n_all}; max |dJ| {dj:.2e} |dK| {dk:.2e}", flush=True)
    assert 0 < n_scr < n_all
    assert dj < 1e-8 and dk < 1e-8
    # a tiny density: the bound scales with |P|, most quartets drop out
    j2, k2 = ctx.fourindex_eri(1e-6 * p, dens_tol=1e-10)
    assert ctx.stats()["last_quartets"] < n_scr
    assert np.max(np.abs(j2 - 1e-6 * j0)) < 1e-8 and np.max(np.abs(k2 - 1e-6 * k0)) < 1e-8
    # zero density: nothing survives, J = K = 0
    j3, k3 = ctx.fourindex_eri(0.0 * p, dens_tol=1e-10)
    assert ctx.stats()["last_quartets"] == 0
    assert not np.any(j3) and not np.any(k3)


@pytest.mark.parametrize("precision", [pa.FP64, pa.MIXED])
def test_incremental_scf_water8(precision):
    atoms = G.water_cluster(8, 4)
    print(f"case: incremental SCF (H2O)8/cc-pVDZ precision {precision}", flush=True)
    b = pa.read_basis("cc-pVDZ", atoms)
    ref = pa.Context(b, schwarz_tol=1e-12).self_consistent_field(scfconv=1e-8, diis=2, scfcycles=80)
    ctx = pa.Context(b, schwarz_tol=1e-12, precision=precision, mixed_cut=1e-5)
    r = ctx.self_consistent_field(scfconv=1e-8, diis=2, scfcycles=80, incremental=8, dens_tol=1e-11)
    de = r["e_total"] - ref["e_total"]
    print(f"E plain {ref['e_total']:.10f} incremental {r['e_total']:.10f} dE {de:.2e}; iterations "
          f"{ref['iterations']} / {r['iterations']}; quartets {ref['quartets']} / {r['quartets']}", flush=True)
    assert ref["converged"] and r["converged"]
    assert abs(de) < 1e-6
    assert r["quartets"] < ref["quartets"]


@pytest.mark.parametrize("incremental", [0, 80])
def test_screened_builds_converged_energy_exact(incremental):
    """Density-screened builds (full or incremental, never a periodic full
    rebuild) with purification and commutator DIIS, the combined solver's
    settings: the converged energy comes from one unscreened build of the
    last input density, so it matches a tightly converged FP64 reference far
    below 1 uEh."""
    atoms = G.water_cluster(8, 4)
    print(f"case: screened SCF (H2O)8/cc-pVDZ incremental {incremental} dens_tol 2e-9", flush=True)
    b = pa.read_basis("cc-pVDZ", atoms)
    ctx = pa.Context(b, schwarz_tol=1e-12)
    ref = ctx.self_consistent_field(scfconv=1e-10, diis=2, scfcycles=80, solver=1, diis_mode=1)
    r = ctx.self_consistent_field(scfconv=1e-8, diis=2, scfcycles=80, solver=1, diis_mode=1,
                                  incremental=incremental, dens_tol=2e-9)
    de = r["e_total"] - ref["e_total"]
    print(f"E reference {ref['e_total']:.10f} screened {r['e_total']:.10f} dE {de:.2e}; iterations "
          f"{ref['iterations']} / {r['iterations']}", flush=True)
    assert ref["converged"] and r["converged"]
    assert abs(de) < 1e-8
