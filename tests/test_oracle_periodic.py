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


def test_lattice_tensor_against_supercell():
    """The oracle's direct lattice integrals (or_eri_tensor_cells, the values
    STORE_FOURINDEX_ERI writes) equal the supercell molecule's ERIs of the
    same functions: (i_0 j_Q1 | k_Q3 l_Q3+Q2) for every cell triple."""
    from oracle import periodic
    atoms = _chain(0, 0, 3.2)
    ub = oracle.Basis(basis_path("6-31Gss"), atoms)
    T = ub.eri_tensor_cells(3.2, 1, 1)
    T2 = periodic.ao_tensor_cells(basis_path("6-31Gss"), atoms, 3.2, 1, 1)
    assert np.max(np.abs(T2)) > 0.5
    assert np.max(np.abs(T - T2)) < 1e-13


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


# ---------------------------------------------------------------------------
# periodic SCF (oracle/periodic.py)
# ---------------------------------------------------------------------------
def test_periodic_scf_isolated_limit_is_molecular():
    """Period 100 bohr, CUTOFF1 = CUTOFF2 = 0, Gamma only: the periodic SCF
    is the molecular SCF (same energy, same iteration count)."""
    from oracle import periodic
    atoms = [(1, 0.0, 0.0, 0.0), (1, 1.4, 0.0, 0.0), (1, 0.3, 2.8, 0.1), (1, 0.4, 4.2, 0.0)]
    r = periodic.scf(basis_path("6-31Gss"), atoms, 100.0, 0, 0, 0, inttol=1e-12, scfconv=1e-9, diis=2,
                     scfcycles=60)
    ob = oracle.Basis(basis_path("6-31Gss"), atoms)
    m = ob.scf(inttol=1e-12, scfconv=1e-9, diis=2, scfcycles=60, direct=True)
    assert r["converged"] and m["converged"]
    assert abs(r["e_total"] - m["e_total"]) < 1e-10
    assert r["iterations"] == m["iterations"]


def test_band_folding_identity():
    """Bands of a chain at 2K k-points of its primitive cell equal the bands
    of the doubled cell (built from the same cell matrices) at K k-points:
    checks the Bloch phase convention, the Hermitian part and the k grid."""
    from oracle import periodic
    rng = np.random.default_rng(8)
    n, cel1 = 3, 2
    F = rng.standard_normal((2 * cel1 + 1, n, n)) * np.array([1.0, 0.5, 1.0, 0.5, 0.2])[:, None, None]
    F[cel1] = 0.5 * (F[cel1] + F[cel1].T)
    for q in range(1, cel1 + 1):
        F[cel1 - q] = F[cel1 + q].T
    S = np.zeros_like(F)
    S[cel1] = np.eye(n)
    # doubled cell: index (i, a), a in {0, 1} sub-cell; F2(Q)[(i,a),(j,b)] = F(2Q + b - a)[i, j]
    c2 = (2 * cel1 + 1) // 2 + 1
    F2 = np.zeros((2 * c2 + 1, 2 * n, 2 * n))
    for Q in range(-c2, c2 + 1):
        for a in range(2):
            for b in range(2):
                q = 2 * Q + b - a
                if abs(q) <= cel1:
                    F2[Q + c2][a * n:(a + 1) * n, b * n:(b + 1) * n] = F[q + cel1]
    S2 = np.zeros_like(F2)
    S2[c2] = np.eye(2 * n)
    K = 2
    e1, _, _ = periodic.bands(F, S, np.eye(n), cel1, 2 * K, 1e-8)
    e2, _, _ = periodic.bands(F2, S2, np.eye(2 * n), c2, K, 1e-8)
    assert np.allclose(np.sort(e1.ravel()), np.sort(e2.ravel()), atol=1e-12)


# ---------------------------------------------------------------------------
# k-point MP2 (oracle/periodic.py mp2_kpoints)
# ---------------------------------------------------------------------------
def test_mp2_kpoints_isolated_limit_is_molecular_golden(golden):
    """CUTOFF1 = CUTOFF2 = 0, Gamma only: the lattice tensor is the molecular
    ERI tensor and the k-point MP2 is MP2_ENERGY_ALG1 of the molecule: the
    shipped BH/STO-3G frozen-core E(2) (data/BH_EOM_out:286)."""
    from oracle import periodic
    from tests import molecules as M
    g = golden["BH_STO3G"]
    b = oracle.Basis(basis_path("STO-3G"), M.BH_GOLDEN)
    r = b.scf(inttol=0.0, scfconv=1e-11)
    T = periodic.ao_tensor_cells(basis_path("STO-3G"), M.BH_GOLDEN, 100.0, 0, 0)
    assert np.max(np.abs(T[0, 0, 0] - b.eri_tensor())) < 1e-14
    co = np.asarray(r["co"], dtype=complex)[None]
    eps = np.asarray(r["eps"])[None]
    s, t = periodic.mp2_kpoints(T, co, eps, 0, 0, 0, 1, r["iocc"], [r["n_indep"]])
    assert abs(s + t - g["e2_alg1"]) < 1e-10
    s0, t0 = oracle.mp2(b.eri_tensor(), r["co"], r["eps"], 1, r["iocc"], r["n_indep"])
    assert abs(s - s0) < 1e-12 and abs(t - t0) < 1e-12


def test_mp2_kpoints_qp_isolated_limit_is_molecular_golden(golden):
    """CUTOFF 0, Gamma only, IOPTN(44) = 2: the k-point quasi-particle bands
    (src/pmp2energy.f90:445-522, 549-558) are the molecule's -- the shipped
    BH/STO-3G bands (data/BH_EOM_out:289-297, 4 decimals) and the molecular
    oracle (1e-12) -- and E(2) is unchanged by the wider j, i and a ranges."""
    from oracle import periodic
    from tests import molecules as M
    g = golden["BH_STO3G"]
    b = oracle.Basis(basis_path("STO-3G"), M.BH_GOLDEN)
    r = b.scf(inttol=0.0, scfconv=1e-11)
    T = periodic.ao_tensor_cells(basis_path("STO-3G"), M.BH_GOLDEN, 100.0, 0, 0)
    co = np.asarray(r["co"], dtype=complex)[None]
    eps = np.asarray(r["eps"])[None]
    s, t, q = periodic.mp2_kpoints(T, co, eps, 0, 0, 0, 1, r["iocc"], [r["n_indep"]], nqp=2)
    s0, t0 = periodic.mp2_kpoints(T, co, eps, 0, 0, 0, 1, r["iocc"], [r["n_indep"]])
    assert abs(s - s0) < 1e-14 and abs(t - t0) < 1e-14
    bands, qm = oracle.mp2_qp(b.eri_tensor(), r["co"], r["eps"], 1, r["iocc"], r["n_indep"], 2)
    assert np.max(np.abs(q[0][bands] - qm)) < 1e-12
    gold = [v for _, v in g["qp_bands"]][::-1]
    assert np.max(np.abs(q[0][bands] - np.array(gold))) <= 5.001e-5
    rest = np.setdiff1d(np.arange(len(eps[0])), bands)
    assert np.array_equal(q[0][rest], eps[0][rest])


def _fold_case(seed):
    """Random translation-invariant integrals G(i_0 j_dj | k_dk l_dl),
    |dj|, |dk|, |dl| <= 1, on a 2-function primitive cell; primitive-cell
    orbitals at 4 k-points (one occupied, one virtual band, gapped)."""
    rng = np.random.default_rng(seed)
    n = 2
    G = rng.standard_normal((n, n, n, n, 3, 3, 3))
    kvc_p = 2
    nk_p = 2 * kvc_p
    c = rng.standard_normal((nk_p, n, n)) + 1j * rng.standard_normal((nk_p, n, n))
    e = np.stack([np.array([-1.0 - rng.random(), 1.0 + rng.random()]) for _ in range(nk_p)])
    return n, G, kvc_p, c, e


def _tensor_prim(n, G, rc1, cel2):
    T = np.zeros((2 * rc1 + 1, 2 * rc1 + 1, 2 * cel2 + 1, n, n, n, n))
    for q1 in range(-rc1, rc1 + 1):
        for q2 in range(-rc1, rc1 + 1):
            for q3 in range(-cel2, cel2 + 1):
                dj, dk, dl = q1, q3, q3 + q2
                if max(abs(dj), abs(dk), abs(dl)) <= 1:
                    T[q1 + rc1, q2 + rc1, q3 + cel2] = G[..., dj + 1, dk + 1, dl + 1]
    return T


def _tensor_doubled(n, G, rc1, cel2):
    """Doubled cell: function (s, f) = primitive function f in primitive cell
    2 Q + s of doubled cell Q."""
    T = np.zeros((2 * rc1 + 1, 2 * rc1 + 1, 2 * cel2 + 1, 2 * n, 2 * n, 2 * n, 2 * n))
    for q1 in range(-rc1, rc1 + 1):
        for q2 in range(-rc1, rc1 + 1):
            for q3 in range(-cel2, cel2 + 1):
                for si in range(2):
                    for sj in range(2):
                        for sk in range(2):
                            for sl in range(2):
                                dj = 2 * q1 + sj - si
                                dk = 2 * q3 + sk - si
                                dl = 2 * (q3 + q2) + sl - si
                                if max(abs(dj), abs(dk), abs(dl)) <= 1:
                                    T[q1 + rc1, q2 + rc1, q3 + cel2, si * n:(si + 1) * n, sj * n:(sj + 1) * n,
                                      sk * n:(sk + 1) * n, sl * n:(sl + 1) * n] = G[..., dj + 1, dk + 1, dl + 1]
    return T


@pytest.mark.parametrize("seed", [1, 2])
def test_mp2_kpoints_band_folding_identity(seed):
    """The same crystal described by its primitive cell at 4 k-points and by
    the doubled cell at 2 k-points (Bloch states folded: k -> K = 2k, the
    coefficients c e^{i k s} / sqrt(2) on sub-cell s): E(2) per doubled cell
    is twice E(2) per primitive cell, singlet and triplet separately.  Fails
    for a wrong Bloch phase sign, momentum-conservation index, swap-term
    index or k normalisation."""
    from oracle import periodic
    n, G, kvc_p, c, e = _fold_case(seed)
    nk_p = 2 * kvc_p
    sp, tp = periodic.mp2_kpoints(_tensor_prim(n, G, 2, 1), c, e, 2, 1, kvc_p, 0, 1, [n] * nk_p)
    kvc_d = kvc_p // 2
    nk_d = 2 * kvc_d
    cd = np.zeros((nk_d, 2 * n, 2 * n), dtype=complex)
    ed = np.zeros((nk_d, 2 * n))
    for K in range(nk_d):
        IXd = K - kvc_d
        cols, vals = [], []
        for kp in range(nk_p):
            IX = kp - kvc_p
            if (IX - IXd) % nk_d:
                continue
            k = np.pi * IX / kvc_p
            for band in range(n):
                v = np.concatenate([c[kp][:, band] * np.exp(1j * k * s) for s in range(2)]) / np.sqrt(2.0)
                cols.append(v)
                vals.append(e[kp][band])
        order = np.argsort(vals)
        cd[K] = np.array(cols).T[:, order]
        ed[K] = np.array(vals)[order]
    sd, td = periodic.mp2_kpoints(_tensor_doubled(n, G, 2, 1), cd, ed, 2, 1, kvc_d, 0, 2, [2 * n] * nk_d)
    assert abs(sp) > 1e-3 and abs(tp) > 1e-3
    assert abs(sd - 2 * sp) < 1e-11 * max(1.0, abs(sd))
    assert abs(td - 2 * tp) < 1e-11 * max(1.0, abs(td))


def _fold_bands(n, kvc_p, c, e):
    """The primitive bands folded onto the doubled cell's k grid: returns
    (cd, ed, src) with src[K][band] = (primitive k index, primitive band)."""
    nk_p = 2 * kvc_p
    kvc_d = kvc_p // 2
    nk_d = 2 * kvc_d
    cd = np.zeros((nk_d, 2 * n, 2 * n), dtype=complex)
    ed = np.zeros((nk_d, 2 * n))
    src = []
    for K in range(nk_d):
        IXd = K - kvc_d
        cols, vals, who = [], [], []
        for kp in range(nk_p):
            IX = kp - kvc_p
            if (IX - IXd) % nk_d:
                continue
            k = np.pi * IX / kvc_p
            for band in range(n):
                cols.append(np.concatenate([c[kp][:, band] * np.exp(1j * k * s) for s in range(2)]) / np.sqrt(2.0))
                vals.append(e[kp][band])
                who.append((kp, band))
        order = np.argsort(vals)
        cd[K] = np.array(cols).T[:, order]
        ed[K] = np.array(vals)[order]
        src.append([who[o] for o in order])
    return cd, ed, src


@pytest.mark.parametrize("seed", [1, 2])
def test_mp2_kpoints_qp_band_folding_identity(seed):
    """Quasi-particle energies are intensive: the doubled cell's QP band at K
    equals the primitive QP band it folds from (every band, both sums), and
    E(2) is unchanged by nqp.  Fails for a wrong kg = ka - kj + kb or
    ki = kg - kj + kb, a swap term at the wrong k-point, or a QP
    normalisation other than (2 kvc)^2."""
    from oracle import periodic
    n, G, kvc_p, c, e = _fold_case(seed)
    nk_p = 2 * kvc_p
    Tp = _tensor_prim(n, G, 2, 1)
    sp, tp, qp = periodic.mp2_kpoints(Tp, c, e, 2, 1, kvc_p, 0, 1, [n] * nk_p, nqp=1)
    sp0, tp0 = periodic.mp2_kpoints(Tp, c, e, 2, 1, kvc_p, 0, 1, [n] * nk_p)
    assert abs(sp - sp0) < 1e-13 and abs(tp - tp0) < 1e-13
    cd, ed, src = _fold_bands(n, kvc_p, c, e)
    sd, td, qd = periodic.mp2_kpoints(_tensor_doubled(n, G, 2, 1), cd, ed, 2, 1, kvc_p // 2, 0, 2,
                                      [2 * n] * (nk_p // 2), nqp=2)
    assert abs(sd - 2 * sp) < 1e-11 * max(1.0, abs(sd))
    shift = np.abs(qp - e)
    assert shift.min() > 1e-4                      # every band is corrected
    for K, row in enumerate(src):
        for band, (kp, bp) in enumerate(row):
            assert abs(qd[K][band] - qp[kp][bp]) < 1e-12 * max(1.0, abs(qp[kp][bp]))
