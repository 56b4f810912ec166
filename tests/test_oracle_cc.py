"""Pins the oracle's coupled-cluster restatement (oracle/cc.py: the full MO
transform of INTEGRAL_TRANSFORM_DIRECTACCESS, MP2/MP3ENERGY_DISK and the
spin-restricted CCSD of COUPLED_CLUSTER_ENERGY, src/pcc.f90) against the
reference's shipped H2O/cc-pVDZ CCSD run and against two-electron
exactness (CCSD = FCI)."""
import numpy as np
import pytest

import oracle
from oracle import cc
from tests import molecules as M
from tests.conftest import basis_path


@pytest.fixture(scope="module")
def h2o_mo():
    b = oracle.Basis(basis_path("cc-pVDZ"), M.H2O_GOLDEN)
    r = b.scf(inttol=0.0, scfconv=1e-11, lindep=1e-12, diis=2, spherical=True, direct=False)
    g = cc.mo_integrals(b.eri_tensor(), r["co"], 0, r["n_indep"])
    return r, g


def test_mo_integrals_symmetry_and_mp2(golden, h2o_mo):
    """The MO tensor has the 8-fold permutational symmetry and gives the
    golden E(2) (data/H2O_CCSD_ccpVDZ_out:391) and the AO-path MP2."""
    r, g = h2o_mo
    assert np.max(np.abs(g - g.transpose(1, 0, 2, 3))) < 1e-12
    assert np.max(np.abs(g - g.transpose(2, 3, 0, 1))) < 1e-12
    e2 = cc.mp2(g, r["eps"][:r["n_indep"]], r["iocc"])
    assert abs(e2 - golden["H2O_ccpVDZ_spherical"]["e2_cc"]) < 1e-12


def test_mp3_golden(golden, h2o_mo):
    """E(3) of the shipped run (data/H2O_CCSD_ccpVDZ_out:394).  A dropped
    ladder or ring term, or a wrong exchange sign, moves it by > 1e-4."""
    r, g = h2o_mo
    e3 = cc.mp3(g, r["eps"][:r["n_indep"]], r["iocc"])
    assert abs(e3 - golden["H2O_ccpVDZ_spherical"]["e3_cc"]) < 1e-11


def test_ccsd_golden(golden, h2o_mo):
    """E(SINGLES), E(DOUBLES) and E(CORRELATION) of the converged CCSD of the
    shipped run (data/H2O_CCSD_ccpVDZ_out:448-450, residuum 6e-13)."""
    gd = golden["H2O_ccpVDZ_spherical"]
    r, g = h2o_mo
    res = cc.ccsd(g, r["eps"][:r["n_indep"]], r["iocc"], conv=1e-12)
    print(res["iterations"], res["e_corr"], gd["e_cc_corr"])
    assert abs(res["e_corr"] - gd["e_cc_corr"]) < 1e-10
    assert abs(res["e_singles"] - gd["e_cc_singles"]) < 1e-10
    assert abs(res["e_doubles"] - gd["e_cc_doubles"]) < 1e-10
    assert abs(r["e_total"] + res["e_corr"] - gd["e_ccsd"]) < 1e-9


@pytest.mark.parametrize("basis", ["6-31Gss", "cc-pVDZ"])
def test_ccsd_two_electron_exact(basis):
    """For two electrons CCSD is exact: its correlation energy equals the
    lowest singlet eigenvalue of the full CI Hamiltonian in the MO basis
    (brute force over symmetric pair functions), H2 at 1.4 bohr and a
    stretched 2.2 bohr bond where singles and doubles both matter."""
    for r_hh in (1.4, 2.2):
        atoms = [(1, 0.0, 0.0, 0.0), (1, 0.0, 0.3, r_hh)]
        b = oracle.Basis(basis_path(basis), atoms)
        r = b.scf(inttol=0.0, scfconv=1e-12, lindep=1e-12, diis=2, direct=False)
        n = r["n_indep"]
        g = cc.mo_integrals(b.eri_tensor(), r["co"], 0, n)
        res = cc.ccsd(g, r["eps"][:n], 1, conv=1e-12)
        efci = cc.fci_two_electron(g, r["eps"][:n])
        assert efci < -1e-3
        assert abs(res["e_corr"] - efci) < 1e-10, (res["e_corr"], efci)


def test_residual_components_reassemble():
    """The degree components of the CCSD residual sum back to the residual
    at (t1, t2) and at (2 t1, -t2) (exact polynomial of degree 4 / 2)."""
    rng = np.random.default_rng(3)
    no, nv = 2, 4
    n = no + nv
    g = rng.standard_normal((n, n, n, n))
    g = g + g.transpose(1, 0, 2, 3)
    g = g + g.transpose(0, 1, 3, 2)
    g = g + g.transpose(2, 3, 0, 1)
    I = cc._Ints(g, no)
    t1 = 0.3 * rng.standard_normal((no, nv))
    t2 = rng.standard_normal((no, no, nv, nv))
    t2 = 0.3 * (t2 + t2.transpose(1, 0, 3, 2))
    C1, C2 = cc.residual_components(I, t1, t2)
    for lam, mu in ((1.0, 1.0), (2.0, -1.0), (0.5, 0.7)):
        r1, r2 = cc.ccsd_residual(I, lam * t1, mu * t2)
        s1 = sum(lam ** p * mu ** q * C1[p][q] for p in range(5) for q in range(3))
        s2 = sum(lam ** p * mu ** q * C2[p][q] for p in range(5) for q in range(3))
        assert np.max(np.abs(s1 - r1)) < 1e-10 * max(1.0, np.max(np.abs(r1)))
        assert np.max(np.abs(s2 - r2)) < 1e-10 * max(1.0, np.max(np.abs(r2)))
    # no singles residual of degree 0 in t1 beyond the t2-linear one, none above t1^3
    assert np.max(np.abs(C1[4][0])) < 1e-10 and np.max(np.abs(C1[0][2])) < 1e-10


def test_lccd_is_the_linear_solve():
    """LCCD is a linear system: the iterated amplitudes equal the direct
    solution of D t - L_lin t = V (the linear doubles operator, itself
    pinned by the golden E(3)), and E(LCCD) is near E(2) + E(3) (H2O/6-31G)."""
    b = oracle.Basis(basis_path("6-31G"), M.H2O_GOLDEN)
    r = b.scf(inttol=0.0, scfconv=1e-11, lindep=1e-12, diis=2, direct=False)
    g = cc.mo_integrals(b.eri_tensor(), r["co"], 0, r["n_indep"])
    no = r["iocc"]
    e = r["eps"][:r["n_indep"]]
    res = cc.ccsd(g, e, no, conv=1e-12, method="LCCD")
    I = cc._Ints(g, no)
    _, d2 = cc._denoms(e, no)
    nv = d2.shape[2]
    m = no * no * nv * nv
    A = np.zeros((m, m))
    for k in range(m):
        u = np.zeros(m)
        u[k] = 1.0
        A[:, k] = cc._linear_doubles(I, u.reshape(d2.shape)).ravel()
    t = np.linalg.solve(np.diag(d2.ravel()) - A, I.ovov.transpose(0, 2, 1, 3).ravel())
    ed = float(np.einsum("iajb,ijab->", I.L, t.reshape(d2.shape)))
    assert abs(res["e_corr"] - ed) < 1e-10
    e2, e3 = cc.mp2(g, e, no), cc.mp3(g, e, no)
    assert res["e_corr"] < e2 + e3 + 1e-3 and abs(res["e_corr"] - (e2 + e3)) < 5e-3


@pytest.mark.parametrize("method", ["QCISD", "CCSD"])
def test_two_electron_methods_exact(method):
    """QCISD and CCSD are exact for two electrons (T1 T2 and T2^2 vanish),
    H2/6-31G** stretched to 2.2 bohr; LCCSD and CCD are not."""
    atoms = [(1, 0.0, 0.0, 0.0), (1, 0.0, 0.3, 2.2)]
    b = oracle.Basis(basis_path("6-31Gss"), atoms)
    r = b.scf(inttol=0.0, scfconv=1e-12, lindep=1e-12, diis=2, direct=False)
    n = r["n_indep"]
    g = cc.mo_integrals(b.eri_tensor(), r["co"], 0, n)
    efci = cc.fci_two_electron(g, r["eps"][:n])
    res = cc.ccsd(g, r["eps"][:n], 1, conv=1e-12, method=method)
    assert abs(res["e_corr"] - efci) < 1e-10
    if method == "QCISD":
        for other in ("LCCSD", "CCD"):
            assert abs(cc.ccsd(g, r["eps"][:n], 1, conv=1e-12, method=other)["e_corr"] - efci) > 1e-6


def test_ccd_two_electron_no_singles_exact():
    """H2/STO-3G has one occupied and one virtual orbital of different
    symmetry: no singles, so CCD (and LCCD is not) is exact."""
    atoms = [(1, 0.0, 0.0, 0.0), (1, 0.0, 0.0, 1.4)]
    b = oracle.Basis(basis_path("STO-3G"), atoms)
    r = b.scf(inttol=0.0, scfconv=1e-12, lindep=1e-12, diis=2, direct=False)
    g = cc.mo_integrals(b.eri_tensor(), r["co"], 0, 2)
    efci = cc.fci_two_electron(g, r["eps"][:2])
    assert abs(cc.ccsd(g, r["eps"][:2], 1, conv=1e-12, method="CCD")["e_corr"] - efci) < 1e-12
    assert abs(cc.ccsd(g, r["eps"][:2], 1, conv=1e-12, method="LCCD")["e_corr"] - efci) > 1e-6
