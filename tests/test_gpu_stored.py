"""DIRECT .F.: J/K, SCF and MP2 from the stored .ao integrals on the device
(RESTORE_FOURINDEX_ERI, src/pfourindexstore.f90:805-1004, called from
src/pselfconsistentfield.f90:45 when DIRECT is false).

Bars: J/K from the records equal the direct build's to 1e-12 (the file
holds every nonzero integral at INTTOL 0); the H2O/cc-pVDZ spherical SCF
from the file reproduces the shipped energy components (1e-8 Eh) in the
shipped 30 iterations and the shipped E(2) with the MP2 tensor taken from
the same records; periodic J/K cells and the periodic SCF from a lattice
file equal the direct ones (J/K 1e-10, energy per cell 1e-8 Eh, same
iteration count)."""
import numpy as np
import pytest

import polymer_amd as pa
from tests import molecules as M
from tests.test_gpu_periodic import _case, _densities

pytestmark = pytest.mark.gpu
EH = 1e-8


def test_stored_h2o_jk_scf_mp2_ccsd_golden(golden, tmp_path):
    g = golden["H2O_ccpVDZ_spherical"]
    print("case: DIRECT .F. H2O/cc-pVDZ spherical golden", flush=True)
    b = pa.read_basis("cc-pVDZ", M.H2O_GOLDEN)
    direct = pa.Context(b, schwarz_tol=0.0)
    path = tmp_path / "h2o.ao.00000"
    n = direct.store_fourindex_eri(path, 0.0)
    ctx = pa.Context(b, schwarz_tol=0.0)
    assert ctx.load_fourindex_eri(path) == n
    rng = np.random.default_rng(7)
    p = rng.standard_normal((b.ncgs, b.ncgs))
    p = 0.05 * (p + p.T)
    j, k = ctx.fourindex_eri(p)
    jd, kd = direct.fourindex_eri(p)
    dj, dk = float(np.max(np.abs(j - jd))), float(np.max(np.abs(k - kd)))
    print(f"stored {n} records: max|dJ| {dj:.2e} max|dK| {dk:.2e}", flush=True)
    assert np.max(np.abs(jd)) > 1e-2 and dj < 1e-12 and dk < 1e-12
    r = ctx.self_consistent_field(scfconv=1e-11, lindep=1e-12, diis=2, spherical=True)
    for key in ["e_nuc", "e_kin", "e_nuc_att", "e_coul", "e_hfx", "e_total"]:
        assert abs(r[key] - g[key]) < EH, key
    assert r["iterations"] == g["iterations"] == 30
    assert r["quartets"] == n * r["iterations"]       # every build digested the records
    e2s, e2t = ctx.mp2_energy(r["co"], r["eps"], 0, r["iocc"], r["n_indep"])
    print(f"E {r['e_total']:.12f} (shipped {g['e_total']:.12f}), E(2) {e2s + e2t:.15f} "
          f"(shipped {g['e2_cc']:.15f})", flush=True)
    assert abs(e2s + e2t - g["e2_cc"]) < EH
    # the rest of the shipped run: E(3) and CCSD from the same stored integrals
    c = ctx.ccsd(r["co"], r["eps"], 0, r["iocc"], r["n_indep"], conv=1e-11)
    print(f"E(3) {c['e3']:.15f} (shipped {g['e3_cc']:.15f}), E(CORRELATION) {c['e_corr']:.15f} "
          f"(shipped {g['e_cc_corr']:.15f})", flush=True)
    assert abs(c["e3"] - g["e3_cc"]) < EH and abs(c["e_corr"] - g["e_cc_corr"]) < EH
    assert abs(r["e_total"] + c["e_corr"] - g["e_ccsd"]) < EH
    ctx.set_direct(True)                               # back to the direct build
    j2, _ = ctx.fourindex_eri(p)
    assert np.max(np.abs(j2 - jd)) < 1e-12


def test_stored_periodic_jk_and_scf(tmp_path):
    basis, atoms, period, cel1, cel2, _, _ = _case("h2_631gss")
    print("case: DIRECT .F. periodic H2 chain 6-31G**", flush=True)
    b = pa.read_basis(basis, atoms)
    direct = pa.Context(b, schwarz_tol=0.0, period=period, cel1=cel1, cel2=cel2)
    path = tmp_path / "h2.ao.00000"
    n = direct.store_fourindex_eri(path, inttol=0.0)
    ctx = pa.Context(b, schwarz_tol=0.0, period=period, cel1=cel1, cel2=cel2)
    assert ctx.load_fourindex_eri(path) == n
    pc = _densities(b.ncgs, list(b.real_functions), cel1, 3)
    J, K = ctx.fourindex_eri_cells(pc)
    Jd, Kd = direct.fourindex_eri_cells(pc)
    dj, dk = float(np.max(np.abs(J - Jd))), float(np.max(np.abs(K - Kd)))
    print(f"stored {n} records: max|dJ| {dj:.2e} max|dK| {dk:.2e}", flush=True)
    assert np.max(np.abs(Jd)) > 1e-2 and dj < 1e-10 and dk < 1e-10
    r = ctx.self_consistent_field_periodic(kpoints=2, scfconv=1e-8, diis=2, scfcycles=60)
    rd = direct.self_consistent_field_periodic(kpoints=2, scfconv=1e-8, diis=2, scfcycles=60)
    print(f"E stored {r['e_total']:.12f} direct {rd['e_total']:.12f} iterations {r['iterations']} / "
          f"{rd['iterations']}", flush=True)
    assert r["converged"] and abs(r["e_total"] - rd["e_total"]) < EH
    assert r["iterations"] == rd["iterations"]


def test_stored_file_rejected_beyond_cells(tmp_path):
    """A lattice file read by a context with a smaller CUTOFF1 unpacks cells
    outside its range: "INTEGRAL FILE HAS BEEN DEGRADED", as the reference."""
    basis, atoms, period, cel1, cel2, _, _ = _case("h2_631gss")
    b = pa.read_basis(basis, atoms)
    big = pa.Context(b, schwarz_tol=0.0, period=period, cel1=cel1 + 1, cel2=cel2 + 1)
    path = tmp_path / "h2big.ao.00000"
    big.store_fourindex_eri(path, inttol=1e-12)
    small = pa.Context(b, schwarz_tol=0.0, period=period, cel1=cel1, cel2=cel2)
    with pytest.raises(pa.PolyeriError, match="DEGRADED"):
        small.load_fourindex_eri(path)
