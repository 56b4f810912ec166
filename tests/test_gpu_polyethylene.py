"""The helical-polymer configuration at its real size: polyethylene C2H4,
6-31G**, CUTOFF1 5 / CUTOFF2 10 (REDUCED_CEL1X 4), helix angle 0, 2 x 16
k-points (src/pfourindexdirect.f90:222-781, src/pfock.f90:180-470,
src/pmp2energy.f90:1-548).

* Lattice J/K cells of a seeded density against the oracle's cells at the
  same cut-offs (tests/golden/pe_jk_cutoff5_10.npz, written by
  tests/golden/make_pe_fixture.py), 1e-10 absolute.  The library screens
  contracted shell pairs at Schwarz 1e-12; the oracle's composite-primitive
  screen is taken at INTTOL 1e-14, because at 1e-12 it drops lattice-sum
  terms worth 3e-10 (J) and 4e-10 (K) that the contracted screen keeps.
* The J/K cells are accumulated in fixed point (integer atomics), so two
  builds of one density are bit-identical, and J/K are linear in P.
* Two SCF runs give bit-identical energies and iteration counts (every
  step after the J/K build is deterministic too); band energies satisfy
  e(k) = e(-k) (time reversal of a real Hamiltonian).
* The k-point MP2 with the BF16x3 transforms is within 1e-6 Eh of the FP64
  MFMA transforms.
"""
import os
import sys

import numpy as np
import pytest

import polymer_amd as pa

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import pe_case as C  # noqa: E402

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(HERE, "golden", "pe_jk_cutoff5_10.npz")


@pytest.fixture(scope="module")
def pe():
    atoms = C.cell_atoms()
    b = pa.read_basis(C.BASIS, atoms)
    ctx = pa.Context(b, schwarz_tol=1e-12, period=C.PERIOD, cel1=C.CEL1, cel2=C.CEL2)
    yield b, ctx
    ctx.close()


def _density(b):
    xyz = np.stack([b.cgsx, b.cgsy, b.cgsz], axis=1)
    return C.density_cells(xyz, b.real_functions)


def test_pe_jk_cells_vs_oracle_fixture(pe):
    b, ctx = pe
    print("case: polyethylene 6-31G** CUTOFF 5/10 lattice J/K vs oracle fixture", flush=True)
    fx = np.load(FIXTURE)
    assert np.allclose(np.array(C.cell_atoms()), fx["atoms"], rtol=0, atol=1e-14)
    assert ctx.reduced_cel1 == int(fx["rc1"])
    P = _density(b)
    assert abs(np.abs(P).sum() - float(fx["p_abs_sum"])) < 1e-9
    J, K = ctx.fourindex_eri_cells(P)
    dj, dk = np.max(np.abs(J - fx["J"])), np.max(np.abs(K - fx["K"]))
    print(f"max|dJ| {dj:.2e} (max|J| {np.max(np.abs(J)):.2f}), max|dK| {dk:.2e}", flush=True)
    assert dj < 1e-10 and dk < 1e-10
    J2, K2 = ctx.fourindex_eri_cells(P)
    assert np.array_equal(J, J2) and np.array_equal(K, K2), "J/K cells differ between two builds"
    # linearity: J(2 P + Q) = 2 J(P) + J(Q), same for K
    Q = C.density_cells(np.stack([b.cgsx, b.cgsy, b.cgsz], axis=1), b.real_functions, seed=C.SEED + 1)
    JQ, KQ = ctx.fourindex_eri_cells(Q)
    JS, KS = ctx.fourindex_eri_cells(2.0 * P + Q)
    assert np.max(np.abs(JS - (2.0 * J + JQ))) < 1e-11
    assert np.max(np.abs(KS - (2.0 * K + KQ))) < 1e-11


def test_pe_scf_reproducible_and_kpoint_mp2(pe):
    b, ctx = pe
    kvc = 16
    print("case: polyethylene 6-31G** CUTOFF 5/10, 32 k-points: SCF twice, MP2 FP64 and BF16x3", flush=True)
    r1 = ctx.self_consistent_field_periodic(kpoints=kvc, scfconv=1e-8, diis=5, scfcycles=150)
    r2 = ctx.self_consistent_field_periodic(kpoints=kvc, scfconv=1e-8, diis=5, scfcycles=150)
    print(f"E {r1['e_total']:.12f} / {r2['e_total']:.12f} per cell, {r1['iterations']} / {r2['iterations']} "
          f"iterations", flush=True)
    assert r1["converged"] and r2["converged"]
    assert r1["iterations"] == r2["iterations"]
    assert abs(r1["e_total"] - r2["e_total"]) < 1e-10
    eps = r1["eps"]
    iocc = r1["iocc"]
    tr = max(np.max(np.abs(eps[kvc + i, :iocc + 4] - eps[kvc - i, :iocc + 4])) for i in range(1, kvc))
    print(f"time reversal max |e(k) - e(-k)| {tr:.2e}, gap {np.min(eps[:, iocc]) - np.max(eps[:, :iocc]):.4f}",
          flush=True)
    assert tr < 1e-10
    icore = 2                      # the two carbon 1s bands
    e64 = sum(ctx.mp2_energy_periodic(kvc, r1["co"], eps, r1["nall"], icore, iocc, transform="f64"))
    e16 = sum(ctx.mp2_energy_periodic(kvc, r1["co"], eps, r1["nall"], icore, iocc, transform="bf16x3"))
    d = e16 - e64
    print(f"E(2) FP64 {e64:.10f}, BF16x3 {e16:.10f}, difference {d:.2e}", flush=True)
    assert -0.5 < e64 < 0.0
    assert abs(d) < 1e-6
