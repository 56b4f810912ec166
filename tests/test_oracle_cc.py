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
