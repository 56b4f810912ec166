"""Pins the oracle's periodic direct J/K (src/pfourindexdirect.f90:222-750,
HELIX = 0, periodic axis x) without the reference.

The lattice sums are rebuilt from a supercell *molecule* -- the unit cell
repeated over every cell the sums touch -- whose contracted ERI tensor comes
from the oracle's molecular engine (pinned by tests/test_oracle_eri.py):

    J(Q1)[i,j]    = sum_{Q2,Q3} (i_0 j_Q1 | k_Q3 l_Q3+Q2) P(Q2)[k,l]
    K(Q3+Q2)[i,l] = sum (i_0 j_Q1 | k_Q3 l_Q3+Q2) P(Q1-Q3)[k,j]
                    over |Q3+Q2|, |Q1-Q3| <= CEL1

with Q1, Q2 in [-RC1, RC1], Q3 in [-CEL2, CEL2].  A wrong cell index, sign or
shift in the cell loops fails this; the screen is off (INTTOL 0).  Also: a
period long enough that neighbour cells do not overlap leaves J(0), K(0) of
the lattice equal to the molecule's J, K plus the (small, analytic)
long-range Coulomb of the image cells."""
import numpy as np
import pytest

import oracle
from tests.conftest import basis_path


def _chain(ncell_lo, ncell_hi, period, d=1.4):
    """H2 units (H at x = 0 and x = d bohr) in cells ncell_lo..ncell_hi."""
    atoms = []
    for c in range(ncell_lo, ncell_hi + 1):
        atoms += [(1, c * period, 0.0, 0.0), (1, c * period + d, 0.1, -0.05)]
    return atoms


def _supercell_map(unit, sup, period, lo):
    """Index of function f of the unit cell, translated to cell c, in the
    supercell basis (matched by centre and position within the atom)."""
    def per_atom(b, atoms):
        out = [[] for _ in atoms]
        for f in range(b.ncgs):
            if not np.any(b.cc[f]):
                continue
            cen = np.array([b.cgsx[f], b.cgsy[f], b.cgsz[f]])
            a = int(np.argmin([np.linalg.norm(cen - np.array(at[1:])) for at in atoms]))
            out[a].append(f)
        return out
    fu = per_atom(unit, unit.atoms_list)
    fs = per_atom(sup, sup.atoms_list)
    nat = len(unit.atoms_list)

    def idx(c, f):
        for a, fl in enumerate(fu):
            if f in fl:
                return fs[(c - lo) * nat + a][fl.index(f)]
        raise KeyError(f)
    return idx


@pytest.mark.parametrize("basis", ["STO-3G", "6-31Gss"])
def test_periodic_jk_against_supercell_tensor(basis):
    period, cel1, cel2, rc1 = 3.2, 1, 1, 1
    lo, hi = -(cel2 + rc1), cel2 + rc1
    unit_atoms = _chain(0, 0, period)
    sup_atoms = _chain(lo, hi, period)
    ub = oracle.Basis(basis_path(basis), unit_atoms)
    sb = oracle.Basis(basis_path(basis), sup_atoms)
    ub.atoms_list, sb.atoms_list = unit_atoms, sup_atoms
    g = sb.eri_tensor()
    idx = _supercell_map(ub, sb, period, lo)
    nc = ub.ncgs
    real = [f for f in range(nc) if np.any(ub.cc[f])]
    rng = np.random.default_rng(3)
    P = {}
    a = rng.standard_normal((nc, nc))
    P[0] = 0.1 * (a + a.T)
    for q in range(1, cel1 + 1):
        P[q] = 0.05 * rng.standard_normal((nc, nc))
        P[-q] = P[q].T.copy()
    for q in P:
        mask = np.zeros((nc, nc))
        mask[np.ix_(real, real)] = 1.0
        P[q] *= mask
    pc = np.array([P[q] for q in range(-cel1, cel1 + 1)])
    J, K, info = ub.fourindex_eri_cells(pc, period, rc1, cel1, cel2, 0.0)
    Jr = np.zeros_like(J)
    Kr = np.zeros_like(K)
    for q1 in range(-rc1, rc1 + 1):
        for q2 in range(-rc1, rc1 + 1):
            for q3 in range(-cel2, cel2 + 1):
                lx = abs(q3 + q2) <= cel1 and abs(q1 - q3) <= cel1
                for i in real:
                    for j in real:
                        I, Jx = idx(0, i), idx(q1, j)
                        for k in real:
                            Kk = idx(q3, k)
                            for l in real:
                                v = g[I, Jx, Kk, idx(q3 + q2, l)]
                                Jr[q1 + cel1, i, j] += v * P[q2][k, l]
                                if lx:
                                    Kr[q3 + q2 + cel1, i, l] += v * P[q1 - q3][k, j]
    assert info["composite_quartets"] > 0
    assert np.max(np.abs(J - Jr)) < 1e-12 * max(1.0, np.max(np.abs(Jr)))
    assert np.max(np.abs(K - Kr)) < 1e-12 * max(1.0, np.max(np.abs(Kr)))


def test_periodic_isolated_cell_limit():
    """Period 200 bohr: images do not overlap, so K(0) is the molecule's K,
    K(Q != 0) = 0, and J(0) exceeds the molecule's J only by the image
    cells' Coulomb potential, sum_Q N_el(P) / (Q * 200) per overlap."""
    period, cel1, cel2 = 200.0, 1, 2
    atoms = _chain(0, 0, period)
    ub = oracle.Basis(basis_path("6-31Gss"), atoms)
    nc = ub.ncgs
    rng = np.random.default_rng(4)
    a = rng.standard_normal((nc, nc))
    p0 = 0.1 * (a + a.T)
    real = [f for f in range(nc) if np.any(ub.cc[f])]
    m = np.zeros((nc, nc))
    m[np.ix_(real, real)] = 1
    p0 *= m
    pc = np.zeros((2 * cel1 + 1, nc, nc))
    pc[cel1] = p0
    J, K, _ = ub.fourindex_eri_cells(pc, period, 0, cel1, cel2, 0.0)
    jm, km, _ = ub.fourindex_eri(np.asfortranarray(p0), 0.0)
    assert np.max(np.abs(K[cel1] - km)) < 1e-12
    assert np.max(np.abs(K[0])) == 0.0 and np.max(np.abs(K[2])) == 0.0
    # image Coulomb: (ij|k_Q l_Q) ~ S_ij * S_kl / |Q| period for distant
    # cells (monopole term; dipole and higher terms are O((extent/R)^2) of it)
    s = ub.one_electron()[1]
    q_el = float(np.sum(p0 * s))
    img = sum(2.0 / (q * period) for q in range(1, cel2 + 1)) * q_el * s
    rel = np.max(np.abs(J[cel1] - jm - img)) / np.max(np.abs(img))
    assert rel < 1e-3
