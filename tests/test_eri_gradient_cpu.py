"""Two-electron nuclear gradient (FOURINDEX_ERI_GRADIENT,
src/pfourindexgradient.f90:1334-1420) through the library's host replay of
the device code (polyeri_eri_gradient_host), pinned by what the mathematics
fixes:

* central finite differences of the oracle's two-electron energy
  E2(R) = 1/2 sum P_C C_C - x/4 sum P_C X_C at a fixed contracted density
  (the basis functions move with their atoms), with the oracle's J/K
  (itself pinned against McMurchie-Davidson integrals);
* translational invariance: the gradient sums to zero over the atoms;
* linearity in the exchange scale x.

The density is a random symmetric matrix (E2 is a quadratic form in P, so
any P exercises every term)."""
import os

import numpy as np
import pytest

import oracle
import polymer_amd as pa
from tests import molecules as M


def _e2(basis, atoms, P, x):
    ob = oracle.Basis(os.path.join(pa.BASIS_DIR, basis), atoms)
    c, k, _ = ob.fourindex_eri(P, 0.0)
    return 0.5 * np.sum(P * c) - 0.25 * x * np.sum(P * k)


def _fd(basis, atoms, P, x, h=1e-4):
    g = np.zeros((len(atoms), 3))
    for a in range(len(atoms)):
        for e in range(3):
            def moved(d):
                at = [list(t) for t in atoms]
                at[a][1 + e] += d
                return [tuple(t) for t in at]
            g[a, e] = (_e2(basis, moved(h), P, x) - _e2(basis, moved(-h), P, x)) / (2 * h)
    return g


@pytest.mark.parametrize("basis", ["STO-3G", "6-31Gss", "cc-pVDZ"])
def test_eri_gradient_matches_finite_differences(basis):
    atoms = M.H2O_GOLDEN
    b = pa.read_basis(basis, atoms)
    rng = np.random.default_rng(3)
    P = rng.standard_normal((b.ncgs, b.ncgs))
    P = 0.1 * (P + P.T)
    x = 1.0
    g = pa.eri_gradient_host(b, P, x)
    fd = _fd(basis, atoms, P, x)
    err = np.max(np.abs(g - fd))
    print(f"{basis}: max |g - fd| = {err:.2e}, max |g| = {np.max(np.abs(g)):.3e}")
    assert err < 1e-6 * max(1.0, np.max(np.abs(g)))
    assert np.max(np.abs(g.sum(axis=0))) < 1e-10 * max(1.0, np.max(np.abs(g)))
    # linear in the exchange scale
    g0 = pa.eri_gradient_host(b, P, 0.0)
    g2 = pa.eri_gradient_host(b, P, 2.0)
    assert np.max(np.abs(g2 - 2 * g + g0)) < 1e-10 * max(1.0, np.max(np.abs(g)))


def _e1(basis, atoms, P, W):
    ob = oracle.Basis(os.path.join(pa.BASIS_DIR, basis), atoms)
    _, s, t, n = ob.one_electron()
    return np.sum(P * (t + n)) - np.sum(W * s) + ob.nuclear_repulsion()


@pytest.mark.parametrize("basis", ["STO-3G", "cc-pVDZ"])
def test_one_electron_gradient_matches_finite_differences(basis):
    """sum P (T + N) - sum W S + V_nn at fixed P, W, differentiated by central
    differences of the oracle's one-electron matrices and nuclear repulsion."""
    atoms = M.H2O_GOLDEN
    b = pa.read_basis(basis, atoms)
    rng = np.random.default_rng(5)
    P = rng.standard_normal((b.ncgs, b.ncgs)); P = 0.1 * (P + P.T)
    W = rng.standard_normal((b.ncgs, b.ncgs)); W = 0.1 * (W + W.T)
    g = pa.one_electron_gradient_host(b, P, W)
    h = 1e-4
    fd = np.zeros_like(g)
    for a in range(len(atoms)):
        for e in range(3):
            def moved(d):
                at = [list(t) for t in atoms]
                at[a][1 + e] += d
                return [tuple(t) for t in at]
            fd[a, e] = (_e1(basis, moved(h), P, W) - _e1(basis, moved(-h), P, W)) / (2 * h)
    err = np.max(np.abs(g - fd))
    print(f"{basis}: one-electron max |g - fd| = {err:.2e}, max |g| = {np.max(np.abs(g)):.3e}")
    assert err < 1e-6 * max(1.0, np.max(np.abs(g)))


def test_scf_gradient_matches_energy_differences():
    """The total analytic RHF gradient (one-electron part with the
    energy-weighted density + two-electron part) against central differences
    of the oracle's converged SCF energy, H2O/6-31G**, bent off symmetry."""
    basis = "6-31Gss"
    atoms = [(1, 0.1, 1.079252144093028, 1.474611055780858), (8, 0.0, 0.0, 0.0),
             (1, -0.05, 1.1, -1.4)]

    def scf(at):
        ob = oracle.Basis(os.path.join(pa.BASIS_DIR, basis), at)
        return ob.scf(inttol=0.0, scfconv=1e-11, scfcycles=200, diis=2)

    r = scf(atoms)
    b = pa.read_basis(basis, atoms)
    W = pa.energy_weighted_density(r["co"], r["eps"], r["iocc"])
    g = pa.one_electron_gradient_host(b, r["P"], W) + pa.eri_gradient_host(b, r["P"], 1.0)
    h = 5e-4
    fd = np.zeros_like(g)
    for a in range(len(atoms)):
        for e in range(3):
            def moved(d):
                at = [list(t) for t in atoms]
                at[a][1 + e] += d
                return [tuple(t) for t in at]
            fd[a, e] = (scf(moved(h))["e_total"] - scf(moved(-h))["e_total"]) / (2 * h)
    err = np.max(np.abs(g - fd))
    print(f"SCF gradient max |g - fd| = {err:.2e}, max |g| = {np.max(np.abs(g)):.3e}, "
          f"sum over atoms {np.max(np.abs(g.sum(axis=0))):.1e}")
    assert err < 2e-6
    assert np.max(np.abs(g.sum(axis=0))) < 1e-9


def test_eri_gradient_screened_enumeration():
    """With a Schwarz threshold only the quartets with Q_ab Q_cd >= tol are
    enumerated (pairs ordered by bound inside each primitive count, the
    survivors of each bra taken as prefixes).  Three waters 8 bohr apart drop
    most cross quartets; a dropped prefix element, a lost group or a wrong
    diagonal (pair ab == pair cd) changes the gradient by far more than the
    screening error against the unscreened sum, and the survivors' sum stays
    translation invariant."""
    atoms = [(z, x + 8.0 * k, y, w + 3.0 * (k % 2)) for k in range(3) for (z, x, y, w) in M.H2O_GOLDEN]
    b = pa.read_basis("6-31Gss", atoms)
    rng = np.random.default_rng(5)
    P = rng.standard_normal((b.ncgs, b.ncgs))
    P = 0.1 * (P + P.T)
    g0 = pa.eri_gradient_host(b, P, 1.0, schwarz_tol=0.0)
    for tol in (1e-12, 1e-9, 1e-6):
        g = pa.eri_gradient_host(b, P, 1.0, schwarz_tol=tol)
        err = np.max(np.abs(g - g0))
        print(f"tol {tol:.0e}: max |g - g(tol 0)| = {err:.2e}")
        assert err < 100 * tol * max(1.0, np.max(np.abs(g0)))
        assert np.max(np.abs(g.sum(axis=0))) < 1e-10 * max(1.0, np.max(np.abs(g0)))
    assert np.max(np.abs(g - g0)) > 0.0          # tol 1e-6 does drop quartets
