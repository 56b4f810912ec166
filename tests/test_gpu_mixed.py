"""Mixed-precision direct J/K: quartets whose Schwarz bound Q_ab Q_cd falls
below the cut (1e-5) are evaluated with an FP32 recursion, the rest in FP64;
Boys function, accumulation into J/K, DIIS and diagonalisation stay FP64.

The FP64 path of the same library is the reference here (it is pinned to the
oracle and the shipped goldens by test_gpu_parity.py).  Bars: the SCF energy
of the mixed path within 1 microhartree of the FP64 energy (the tolerance
the mixed mode is specified for), J/K within 1e-6 elementwise, and a mixed
context with the cut at or below the Schwarz threshold reproducing FP64 to
the last bits of the FP64 tolerance."""
import numpy as np
import pytest

import polymer_amd as pa
from polymer_amd import geometry as G

pytestmark = pytest.mark.gpu

MICRO_EH = 1e-6


def _density(b, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((b.ncgs, 30))
    ghost = [i for i in range(b.ncgs) if i not in set(b.real_functions)]
    x[ghost] = 0.0
    return 0.02 * x @ x.T


def test_mixed_jk_water4():
    atoms = G.water_cluster(4, 3)
    print("case: mixed J/K (H2O)4/cc-pVDZ", flush=True)
    b = pa.read_basis("cc-pVDZ", atoms)
    p = _density(b, 1)
    c64 = pa.Context(b, schwarz_tol=1e-12)
    cmx = pa.Context(b, schwarz_tol=1e-12, precision=pa.MIXED, mixed_cut=1e-5)
    j0, k0 = c64.fourindex_eri(p)
    j1, k1 = cmx.fourindex_eri(p)
    st = cmx.stats()
    n32 = sum(c["quartets"] for c in st["classes"] if c["prec"] == "fp32")
    assert n32 > 0 and st["quartets"] == c64.stats()["quartets"]
    dj, dk = np.max(np.abs(j1 - j0)), np.max(np.abs(k1 - k0))
    print(f"fp32 quartets {n32} of {st['quartets']}; max |dJ| {dj:.2e} |dK| {dk:.2e}", flush=True)
    assert dj < 1e-6 and dk < 1e-6
    assert dj > 0.0 or dk > 0.0          # the FP32 lists really ran


def test_mixed_cut_below_threshold_is_fp64():
    atoms = G.water_cluster(2, 5)
    print("case: mixed with cut <= Schwarz threshold", flush=True)
    b = pa.read_basis("cc-pVDZ", atoms)
    p = _density(b, 2)
    j0, k0 = pa.Context(b, schwarz_tol=1e-12).fourindex_eri(p)
    cmx = pa.Context(b, schwarz_tol=1e-12, precision=pa.MIXED, mixed_cut=1e-12)
    j1, k1 = cmx.fourindex_eri(p)
    assert all(c["prec"] == "fp64" for c in cmx.stats()["classes"])
    assert np.max(np.abs(j1 - j0)) < 1e-12 and np.max(np.abs(k1 - k0)) < 1e-12


@pytest.mark.parametrize("seed", [0, 1])
def test_mixed_scf_alkane_energy(seed):
    """C16H34 / 6-31G** with 0.05 A coordinate noise: the mixed SCF energy
    within 1 microhartree of the FP64 one, same iteration count +- 1."""
    atoms = G.alkane(16, seed=seed)
    print(f"case: mixed SCF C16H34/6-31G** seed {seed}", flush=True)
    b = pa.read_basis("6-31Gss", atoms)
    r64 = pa.Context(b, schwarz_tol=1e-12).self_consistent_field(scfconv=1e-8, diis=2, scfcycles=60)
    rmx = pa.Context(b, schwarz_tol=1e-12, precision=pa.MIXED, mixed_cut=1e-5).self_consistent_field(
        scfconv=1e-8, diis=2, scfcycles=60)
    de = rmx["e_total"] - r64["e_total"]
    print(f"E(fp64) {r64['e_total']:.10f} E(mixed) {rmx['e_total']:.10f} dE {de:.2e} "
          f"iterations {r64['iterations']} / {rmx['iterations']}", flush=True)
    assert r64["converged"] and rmx["converged"]
    assert abs(de) < MICRO_EH
    assert abs(rmx["iterations"] - r64["iterations"]) <= 1
