"""MP2 AO->MO transform on this library's MFMA kernels.

The four quarter transforms of polyeri_mp2_energy run on the library's FP64
MFMA kernel (v_mfma_f64_16x16x4f64, the default), on rocBLAS dgemm or on its
BF16x3-split kernel (three BF16 pieces per FP64 operand, six MFMA products,
FP32 accumulation).  Bars: the shipped goldens (H2O/cc-pVDZ E(2), BH
frozen-core E(2)) within 1e-8 Eh for FP64 MFMA; the BF16x3 error against
FP64 stated and below 1e-6 Eh; the singlet/triplet split of both FP64 paths
equal to 1e-11."""
import numpy as np
import pytest

import polymer_amd as pa
from polymer_amd import geometry as Gm
from tests import molecules as M

pytestmark = pytest.mark.gpu
EH = 1e-8


@pytest.fixture(scope="module")
def h2o_scf():
    ctx = pa.Context(pa.read_basis("cc-pVDZ", M.H2O_GOLDEN), schwarz_tol=0.0)
    r = ctx.self_consistent_field(scfconv=1e-11, lindep=1e-12, diis=2, spherical=True)
    return ctx, r


@pytest.mark.parametrize("transform", ["f64", "bf16x3"])
def test_mp2_h2o_golden(golden, h2o_scf, transform):
    ctx, r = h2o_scf
    print(f"case: MP2 H2O/cc-pVDZ transform {transform}", flush=True)
    s0, t0 = ctx.mp2_energy(r["co"], r["eps"], 0, r["iocc"], r["n_indep"], transform="rocblas")
    s1, t1 = ctx.mp2_energy(r["co"], r["eps"], 0, r["iocc"], r["n_indep"], transform=transform)
    e = golden["H2O_ccpVDZ_spherical"]["e2_cc"]
    print(f"E2 rocblas {s0 + t0:.12f} {transform} {s1 + t1:.12f} golden {e:.12f} "
          f"diff {s1 + t1 - (s0 + t0):.2e}", flush=True)
    if transform == "f64":
        assert abs(s1 + t1 - e) < EH
        assert abs(s1 - s0) < 1e-11 and abs(t1 - t0) < 1e-11
    else:
        assert abs(s1 + t1 - (s0 + t0)) < 1e-6


def test_mp2_bh_frozen_core_mfma(golden):
    g = golden["BH_STO3G"]
    ctx = pa.Context(pa.read_basis("STO-3G", M.BH_GOLDEN), schwarz_tol=0.0)
    r = ctx.self_consistent_field(scfconv=1e-11)
    s, t = ctx.mp2_energy(r["co"], r["eps"], 1, r["iocc"], r["n_indep"], transform="f64")
    assert abs(s + t - g["e2_alg1"]) < EH


def test_mp2_water4_transforms_agree():
    """A size where every quarter transform has ragged 16 x 16 tiles:
    (H2O)4 / 6-31G** (100 functions, 20 occupied)."""
    atoms = Gm.water_cluster(4, 11, radius=3.2)
    ctx = pa.Context(pa.read_basis("6-31Gss", atoms), schwarz_tol=1e-12)
    r = ctx.self_consistent_field(scfconv=1e-9, diis=2)
    print("case: MP2 (H2O)4 transforms", flush=True)
    e = {k: sum(ctx.mp2_energy(r["co"], r["eps"], 4, r["iocc"], r["n_indep"], transform=k))
         for k in (None, "rocblas", "f64", "bf16x3")}
    print({k: f"{v:.12f}" for k, v in e.items()}, flush=True)
    assert abs(e[None] - e["f64"]) < 1e-12     # the default transform is the FP64 MFMA kernel
    assert abs(e["f64"] - e["rocblas"]) < 1e-10
    assert abs(e["bf16x3"] - e["rocblas"]) < 1e-6


@pytest.mark.parametrize("transform", ["rocblas", "f64"])
def test_mp2_quasiparticle_bh_golden(golden, transform):
    """MP2 quasi-particle bands (IOPTN(44) = 2, frozen core) of BH/STO-3G:
    the shipped bands 2..5 (data/BH_EOM_out:289-297, 4 decimals), the
    oracle's restatement to 1e-10, and E(2) still the golden."""
    import oracle
    from tests.conftest import basis_path
    g = golden["BH_STO3G"]
    ctx = pa.Context(pa.read_basis("STO-3G", M.BH_GOLDEN), schwarz_tol=0.0)
    r = ctx.self_consistent_field(scfconv=1e-11)
    print(f"case: MP2 quasi-particle bands BH/STO-3G transform {transform}", flush=True)
    s, t, q = ctx.mp2_quasiparticle(r["co"], r["eps"], 1, r["iocc"], r["n_indep"], 2, transform=transform)
    ob = oracle.Basis(basis_path("STO-3G"), M.BH_GOLDEN)
    bands, qr = oracle.mp2_qp(ob.eri_tensor(), r["co"], r["eps"], 1, r["iocc"], r["n_indep"], 2)
    print("QEPSILON", q[bands], "oracle", qr, flush=True)
    assert abs(s + t - g["e2_alg1"]) < EH
    assert np.max(np.abs(q[bands] - qr)) < 1e-10
    gold = [v for _, v in g["qp_bands"]][::-1]
    assert np.max(np.abs(q[bands] - np.array(gold))) <= 5.001e-5
    rest = [k for k in range(len(q)) if k not in bands]
    assert np.array_equal(q[rest], r["eps"][rest])


def test_mp2_quasiparticle_water_transforms():
    """H2O/cc-pVDZ, three quasi-particle bands either side of the gap: the
    three transforms against the oracle (FP64 1e-10; BF16x3 error printed
    and below 1e-6 Eh)."""
    import oracle
    from tests.conftest import basis_path
    ctx = pa.Context(pa.read_basis("cc-pVDZ", M.H2O_GOLDEN), schwarz_tol=0.0)
    r = ctx.self_consistent_field(scfconv=1e-11, diis=2)
    ob = oracle.Basis(basis_path("cc-pVDZ"), M.H2O_GOLDEN)
    bands, qr = oracle.mp2_qp(ob.eri_tensor(), r["co"], r["eps"], 1, r["iocc"], r["n_indep"], 3)
    for tr in ("rocblas", "f64", "bf16x3"):
        print(f"case: MP2 quasi-particle bands H2O/cc-pVDZ transform {tr}", flush=True)
        _, _, q = ctx.mp2_quasiparticle(r["co"], r["eps"], 1, r["iocc"], r["n_indep"], 3, transform=tr)
        d = np.max(np.abs(q[bands] - qr))
        print(f"{tr}: max |dQ| {d:.2e}", flush=True)
        assert d < (1e-6 if tr == "bf16x3" else 1e-10)
