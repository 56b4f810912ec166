"""Coupled-cluster path on the device (src/pcc.f90): the full MO transform
(INTEGRAL_TRANSFORM_DIRECTACCESS), E(2), E(3) and spin-restricted CCSD with
every amplitude contraction on the library's FP64 MFMA GEMM.  Bars: the
oracle (oracle/cc.py) on the device's own orbitals within 1e-10 Eh, the
shipped H2O/cc-pVDZ CCSD run within 1e-8 Eh, CCSD = FCI for two electrons."""
import numpy as np
import pytest

import oracle
import polymer_amd as pa
from oracle import cc
from tests import molecules as M
from tests.conftest import basis_path

pytestmark = pytest.mark.gpu
EH = 1e-8


@pytest.fixture(scope="module")
def h2o():
    ctx = pa.Context(pa.read_basis("cc-pVDZ", M.H2O_GOLDEN), schwarz_tol=0.0)
    r = ctx.self_consistent_field(scfconv=1e-11, lindep=1e-12, diis=2, spherical=True)
    feri = oracle.Basis(basis_path("cc-pVDZ"), M.H2O_GOLDEN).eri_tensor()
    return ctx, r, feri


def test_mo_integrals_match_oracle(h2o):
    ctx, r, feri = h2o
    n = r["n_indep"]
    print(f"case: MO integrals H2O/cc-pVDZ nmo {n}", flush=True)
    g = ctx.mo_integrals(r["co"], 0, n)
    gr = cc.mo_integrals(feri, r["co"], 0, n)
    print(f"max |g - oracle| {np.max(np.abs(g - gr)):.2e}", flush=True)
    assert np.max(np.abs(g - gr)) < 1e-12


@pytest.mark.parametrize("icore", [0, 1])
def test_ccsd_h2o_matches_oracle_and_golden(golden, h2o, icore):
    """All-electron: E(2), E(3), E(SINGLES), E(DOUBLES), E(CORRELATION) of
    the shipped run (data/H2O_CCSD_ccpVDZ_out:391-451) within 1e-8 Eh; both
    all-electron and frozen-core against the oracle on the device's bands
    within 1e-10 Eh."""
    ctx, r, feri = h2o
    n, iocc = r["n_indep"], r["iocc"]
    print(f"case: CCSD H2O/cc-pVDZ icore {icore}", flush=True)
    res = ctx.ccsd(r["co"], r["eps"], icore, iocc, n, conv=1e-11, maxiter=100)
    g = cc.mo_integrals(feri, r["co"], icore, n)
    e = r["eps"][icore:n]
    ref = cc.ccsd(g, e, iocc - icore, conv=1e-12)
    print(f"CCSD: corr {res['e_corr']:.12f} oracle {ref['e_corr']:.12f} singles {res['e_singles']:.3e} "
          f"E2 {res['e2']:.12f} E3 {res['e3']:.12f} iterations {res['iterations']} "
          f"({res['iter_ms']:.2f} ms/iteration, transform {res['transform_ms']:.2f} ms)", flush=True)
    assert res["converged"] == 1
    assert abs(res["e2"] - cc.mp2(g, e, iocc - icore)) < 1e-10
    assert abs(res["e3"] - cc.mp3(g, e, iocc - icore)) < 1e-10
    assert abs(res["e_corr"] - ref["e_corr"]) < 1e-10
    assert abs(res["e_singles"] - ref["e_singles"]) < 1e-10
    assert abs(res["e_doubles"] - ref["e_doubles"]) < 1e-10
    if icore == 0:
        gd = golden["H2O_ccpVDZ_spherical"]
        assert abs(res["e2"] - gd["e2_cc"]) < EH
        assert abs(res["e3"] - gd["e3_cc"]) < EH
        assert abs(res["e_singles"] - gd["e_cc_singles"]) < EH
        assert abs(res["e_doubles"] - gd["e_cc_doubles"]) < EH
        assert abs(res["e_corr"] - gd["e_cc_corr"]) < EH
        assert abs(r["e_total"] + res["e_corr"] - gd["e_ccsd"]) < EH


def test_ccsd_two_electron_exact():
    """H2/6-31G** at a stretched 2.2 bohr bond: the device CCSD correlation
    energy equals the full CI of the device's own MO integrals (1e-9)."""
    atoms = [(1, 0.0, 0.0, 0.0), (1, 0.0, 0.3, 2.2)]
    print("case: CCSD = FCI, H2/6-31G**", flush=True)
    ctx = pa.Context(pa.read_basis("6-31Gss", atoms), schwarz_tol=0.0)
    r = ctx.self_consistent_field(scfconv=1e-12, lindep=1e-12, diis=2)
    n = r["n_indep"]
    res = ctx.ccsd(r["co"], r["eps"], 0, 1, n, conv=1e-11)
    g = ctx.mo_integrals(r["co"], 0, n)
    efci = cc.fci_two_electron(g, r["eps"][:n])
    print(f"CCSD {res['e_corr']:.12f} FCI {efci:.12f}", flush=True)
    assert abs(res["e_corr"] - efci) < 1e-9


def test_ccsd_rejects_bad_ranges(h2o):
    ctx, r, _ = h2o
    with pytest.raises(pa.PolyeriError, match="empty occupied"):
        ctx.ccsd(r["co"], r["eps"], 5, 5, r["n_indep"])
    with pytest.raises(pa.PolyeriError, match="CC ITERATION DID NOT CONVERGE"):
        ctx.ccsd(r["co"], r["eps"], 0, r["iocc"], r["n_indep"], conv=1e-14, maxiter=3)


@pytest.mark.parametrize("method", ["CCD", "LCCD", "LCCSD", "QCISD"])
def test_cc_methods_match_oracle(h2o, method):
    """COPTN(70) = CCD, LCCD, LCCSD, QCISD (src/pcc.f90:39-66) on the device
    vs the oracle's degree components of the CCSD residual, H2O/cc-pVDZ
    frozen core, E(CORRELATION) (the doubles energy) within 1e-10 Eh."""
    ctx, r, feri = h2o
    n, iocc, icore = r["n_indep"], r["iocc"], 1
    print(f"case: {method} H2O/cc-pVDZ icore {icore}", flush=True)
    res = ctx.ccsd(r["co"], r["eps"], icore, iocc, n, conv=1e-11, maxiter=100, method=method)
    g = cc.mo_integrals(feri, r["co"], icore, n)
    ref = cc.ccsd(g, r["eps"][icore:n], iocc - icore, conv=1e-12, method=method)
    print(f"{method}: corr {res['e_corr']:.12f} oracle {ref['e_corr']:.12f} iterations {res['iterations']}",
          flush=True)
    assert res["converged"] == 1
    assert abs(res["e_corr"] - ref["e_corr"]) < 1e-10
    assert abs(res["e_doubles"] - ref["e_doubles"]) < 1e-10
    if method in ("CCD", "LCCD"):
        assert res["e_singles"] == 0.0


def test_qcisd_two_electron_exact():
    """QCISD on the device is exact for two electrons (H2/6-31G**, 2.2 bohr)."""
    atoms = [(1, 0.0, 0.0, 0.0), (1, 0.0, 0.3, 2.2)]
    print("case: QCISD = FCI, H2/6-31G**", flush=True)
    ctx = pa.Context(pa.read_basis("6-31Gss", atoms), schwarz_tol=0.0)
    r = ctx.self_consistent_field(scfconv=1e-12, lindep=1e-12, diis=2)
    n = r["n_indep"]
    res = ctx.ccsd(r["co"], r["eps"], 0, 1, n, conv=1e-11, method="QCISD")
    efci = cc.fci_two_electron(ctx.mo_integrals(r["co"], 0, n), r["eps"][:n])
    assert abs(res["e_corr"] - efci) < 1e-9
    with pytest.raises(pa.PolyeriError, match="UNKNOWN OR UNIMPLEMENTED"):
        ctx.ccsd(r["co"], r["eps"], 0, 1, n, method="D45")
