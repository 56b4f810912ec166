"""Periodic direct J/K (lattice sums along x, HELIX = 0) against the oracle.

polyeri_jk_build_cells vs the oracle's restatement of the periodic loop nest
of src/pfourindexdirect.f90:222-750 (itself pinned against a supercell ERI
tensor by tests/test_oracle_periodic.py), unscreened (threshold 0) so both
sides evaluate every quartet: 1e-10 absolute on J and K of every cell.
Cases: an H2 chain in 6-31G** (s and p shells), a Ne chain in cc-pVDZ
(d shells: the wavefront-per-quartet path) and the polyethylene unit cell
C2H4 (period 2.55 A along x) in STO-3G.  Cells within CUTOFF1 overlap, so
the reduced cell range equals CUTOFF1 on both sides."""
import numpy as np
import pytest

import oracle
import polymer_amd as pa
from tests.conftest import basis_path

pytestmark = pytest.mark.gpu

ANG = 1.0 / 0.529177249


def _h2_chain(period):
    return [(1, 0.0, 0.0, 0.0), (1, 1.4, 0.1, -0.05)]


def _ne_chain(period):
    return [(10, 0.0, 0.2, 0.0)]


def _pe_cell():
    h = 0.4229
    return [(6, 0.0, 0.0, h * ANG), (6, 1.275 * ANG, 0.0, -h * ANG),
            (1, 0.0, 0.876 * ANG, 1.071 * ANG), (1, 0.0, -0.876 * ANG, 1.071 * ANG),
            (1, 1.275 * ANG, 0.876 * ANG, -1.071 * ANG), (1, 1.275 * ANG, -0.876 * ANG, -1.071 * ANG)]


def _densities(nc, real, cel1, seed):
    rng = np.random.default_rng(seed)
    P = {}
    a = rng.standard_normal((nc, nc))
    P[0] = 0.1 * (a + a.T)
    for q in range(1, cel1 + 1):
        P[q] = 0.05 * rng.standard_normal((nc, nc)) / q
        P[-q] = P[q].T.copy()
    m = np.zeros((nc, nc))
    m[np.ix_(real, real)] = 1.0
    return np.array([P[q] * m for q in range(-cel1, cel1 + 1)])


@pytest.mark.parametrize("case", ["h2_631gss", "ne_ccpvdz", "pe_sto3g"])
def test_periodic_jk_matches_oracle(case):
    if case == "h2_631gss":
        basis, atoms, period, cel1, cel2 = "6-31Gss", _h2_chain(3.2), 3.2, 1, 2
    elif case == "ne_ccpvdz":
        basis, atoms, period, cel1, cel2 = "cc-pVDZ", _ne_chain(5.0), 5.0, 1, 1
    else:
        basis, atoms, period, cel1, cel2 = "STO-3G", _pe_cell(), 2.55 * ANG, 1, 2
    print(f"case: periodic J/K {case} cel1 {cel1} cel2 {cel2}", flush=True)
    b = pa.read_basis(basis, atoms)
    real = list(b.real_functions)
    pc = _densities(b.ncgs, real, cel1, 5)
    ctx = pa.Context(b, schwarz_tol=0.0, period=period, cel1=cel1, cel2=cel2)
    J, K = ctx.fourindex_eri_cells(pc)
    ob = oracle.Basis(basis_path(basis), atoms)
    Jr, Kr, _ = ob.fourindex_eri_cells(pc, period, cel1, cel1, cel2, 0.0)
    dj, dk = np.max(np.abs(J - Jr)), np.max(np.abs(K - Kr))
    print(f"max |dJ| {dj:.2e} |dK| {dk:.2e}  (|J| {np.max(np.abs(Jr)):.2e})", flush=True)
    assert np.max(np.abs(Jr)) > 0 and np.max(np.abs(Kr)) > 0
    assert dj < 1e-10 and dk < 1e-10


@pytest.mark.parametrize("case", ["h2_chain_631gss", "pe_sto3g"])
def test_periodic_scf_matches_oracle(case):
    """polyeri_scf_periodic (k-point FOCK_BUILD on the device) vs the
    oracle's numpy restatement (oracle/periodic.py, pinned by
    tests/test_oracle_periodic.py): energy per cell 1e-8 Eh, same
    iteration count, band energies 1e-7."""
    from oracle import periodic
    if case == "h2_chain_631gss":
        basis, atoms, period, cel1, cel2, kvc = "6-31Gss", _h2_chain(4.0), 4.0, 2, 3, 2
    else:
        basis, atoms, period, cel1, cel2, kvc = "STO-3G", _pe_cell(), 2.55 * ANG, 1, 1, 2
    print(f"case: periodic SCF {case} cel1 {cel1} cel2 {cel2} kpoints {2 * kvc}", flush=True)
    b = pa.read_basis(basis, atoms)
    ctx = pa.Context(b, schwarz_tol=0.0, period=period, cel1=cel1, cel2=cel2)
    r = ctx.self_consistent_field_periodic(kpoints=kvc, scfconv=1e-8, diis=2, scfcycles=60)
    ref = periodic.scf(basis_path(basis), atoms, period, cel1, cel2, kvc, inttol=0.0, scfconv=1e-8, diis=2,
                       scfcycles=60)
    print(f"E gpu {r['e_total']:.12f} oracle {ref['e_total']:.12f} iterations {r['iterations']} / "
          f"{ref['iterations']}", flush=True)
    assert r["converged"] and ref["converged"]
    assert abs(r["e_total"] - ref["e_total"]) < 1e-8
    assert r["iterations"] == ref["iterations"]
    ea, eb = np.sort(r["eps"].ravel()), np.sort(ref["eps"].ravel())
    fin = eb < 1e90
    assert np.max(np.abs(ea[fin] - eb[fin])) < 1e-7


def _case(case):
    if case == "h2_631gss":
        return "6-31Gss", _h2_chain(3.2), 3.2, 1, 2, 2, 0
    if case == "ne_ccpvdz":
        return "cc-pVDZ", _ne_chain(5.0), 5.0, 1, 1, 2, 1
    return "STO-3G", _pe_cell(), 2.55 * ANG, 1, 1, 2, 2


@pytest.mark.parametrize("case", ["h2_631gss", "ne_ccpvdz", "pe_sto3g"])
def test_periodic_lattice_tensor_matches_oracle(case):
    """The lattice integrals STORE_FOURINDEX_ERI writes (i_0 j_Q1 | k_Q3
    l_Q3+Q2), computed by the periodic quartet kernels in their tensor-store
    form, vs the oracle's direct restatement (pinned against a supercell
    ERI tensor): 1e-12 absolute, same REDUCED_CEL1X."""
    from oracle import periodic
    basis, atoms, period, cel1, cel2, _, _ = _case(case)
    print(f"case: lattice tensor {case}", flush=True)
    b = pa.read_basis(basis, atoms)
    ctx = pa.Context(b, schwarz_tol=0.0, period=period, cel1=cel1, cel2=cel2)
    _, S, _, _, _ = periodic.one_electron_cells(basis_path(basis), atoms, period, cel1, cel2)
    rc1 = periodic.reduced_cel1(S, cel1)
    assert ctx.reduced_cel1 == rc1
    T = ctx.eri_tensor_cells()
    Tr = oracle.Basis(basis_path(basis), atoms).eri_tensor_cells(period, rc1, cel2)
    print(f"max |dT| {np.max(np.abs(T - Tr)):.2e} (|T| {np.max(np.abs(Tr)):.2e})", flush=True)
    assert T.shape == Tr.shape
    assert np.max(np.abs(T - Tr)) < 1e-12


@pytest.mark.parametrize("case", ["h2_631gss", "ne_ccpvdz", "pe_sto3g"])
def test_periodic_mp2_matches_oracle(case):
    """k-point MP2 (MP2_ENERGY_ALG1, src/pmp2energy.f90:187-548) on the
    device: lattice tensor plus four complex quarter transforms on the FP64
    MFMA kernel vs the oracle's restatement on the same bands (the device
    SCF's CO, eps, IALL): singlet and triplet within 1e-10 Eh.  The BF16x3
    split transform's error against FP64 is printed and below 1e-6 Eh.
    Frozen core: Ne 1s, polyethylene's two C 1s bands."""
    from oracle import periodic
    basis, atoms, period, cel1, cel2, kvc, icore = _case(case)
    print(f"case: k-point MP2 {case} kpoints {2 * kvc} icore {icore}", flush=True)
    b = pa.read_basis(basis, atoms)
    ctx = pa.Context(b, schwarz_tol=0.0, period=period, cel1=cel1, cel2=cel2)
    r = ctx.self_consistent_field_periodic(kpoints=kvc, scfconv=1e-8, diis=2, scfcycles=60)
    assert r["converged"]
    iocc = r["iocc"]
    s, t, sec = ctx.mp2_energy_periodic(kvc, r["co"], r["eps"], r["nall"], icore, iocc, transform="f64",
                                        timing=True)
    rc1 = ctx.reduced_cel1
    Tr = oracle.Basis(basis_path(basis), atoms).eri_tensor_cells(period, rc1, cel2)
    sr, tr = periodic.mp2_kpoints(Tr, r["co"], r["eps"], rc1, cel2, kvc, icore, iocc, list(r["nall"]))
    sb, tb = ctx.mp2_energy_periodic(kvc, r["co"], r["eps"], r["nall"], icore, iocc, transform="bf16x3")
    print(f"E2 singlet {s:.12f} oracle {sr:.12f}  triplet {t:.12f} oracle {tr:.12f}  ({sec:.3f} s)", flush=True)
    print(f"E2 bf16x3 {sb + tb:.12f} - f64 {s + t:.12f} = {sb + tb - s - t:.2e}", flush=True)
    assert abs(s + t) > 1e-3
    assert abs(s - sr) < 1e-10 and abs(t - tr) < 1e-10
    assert abs(sb + tb - s - t) < 1e-6


def test_periodic_store_restore_roundtrip(tmp_path):
    """STORE_FOURINDEX_ERI of a periodic context: the records written
    (packed cells, src/pfourindexstore.f90:735-738) read back by the cells
    reader equal the lattice tensor above the threshold, and the count is the
    number of such integrals."""
    basis, atoms, period, cel1, cel2, _, _ = _case("h2_631gss")
    print("case: periodic .ao round trip", flush=True)
    ctx = pa.Context(pa.read_basis(basis, atoms), schwarz_tol=0.0, period=period, cel1=cel1, cel2=cel2)
    T = ctx.eri_tensor_cells()
    tol = 1e-9
    n = ctx.store_fourindex_eri(tmp_path / "h2.ao.00000", inttol=tol)
    assert n == int(np.sum(np.abs(T) > tol))
    R, m = pa.restore_fourindex_eri_cells(tmp_path / "h2.ao.00000", ctx.ncgs, cel1, cel2)
    rc1 = ctx.reduced_cel1
    sl = slice(cel1 - rc1, cel1 + rc1 + 1)
    assert m == n
    assert np.array_equal(R[sl, sl], np.where(np.abs(T) > tol, T, 0.0))


def test_periodic_mp2_modulo():
    """MODULO (IOPTN(99), src/pmp2energy.f90:64-80): the k-point MP2 of an
    SCF on 8 k-points with MODULO 2 is the k-point MP2 of the bands at every
    second k-point (k = kb * 2, kb in [-2, 1]) with the 4-point grid's phases
    and normalisation: equal to the device's 4-point MP2 of that subset and
    to the oracle's restatement on it (1e-10 Eh); MODULO 3 of KPOINTS 4 is
    rejected with the reference's message."""
    from oracle import periodic
    basis, atoms, period, cel1, cel2, _, icore = _case("h2_631gss")
    kvc, nmod = 4, 2
    print(f"case: k-point MP2 MODULO {nmod} of {2 * kvc} k-points", flush=True)
    b = pa.read_basis(basis, atoms)
    ctx = pa.Context(b, schwarz_tol=0.0, period=period, cel1=cel1, cel2=cel2)
    r = ctx.self_consistent_field_periodic(kpoints=kvc, scfconv=1e-8, diis=2, scfcycles=60)
    assert r["converged"]
    iocc = r["iocc"]
    s, t = ctx.mp2_energy_periodic(kvc, r["co"], r["eps"], r["nall"], icore, iocc, modulo=nmod)
    idx = [kb * nmod + kvc for kb in range(-kvc // nmod, kvc // nmod)]
    co2, eps2, nall2 = r["co"][idx], r["eps"][idx], np.asarray(r["nall"])[idx]
    s2, t2 = ctx.mp2_energy_periodic(kvc // nmod, co2, eps2, nall2, icore, iocc)
    rc1 = ctx.reduced_cel1
    Tr = oracle.Basis(basis_path(basis), atoms).eri_tensor_cells(period, rc1, cel2)
    sr, tr = periodic.mp2_kpoints(Tr, co2, eps2, rc1, cel2, kvc // nmod, icore, iocc, list(nall2))
    print(f"MODULO {nmod}: E2 singlet {s:.12f} subset {s2:.12f} oracle {sr:.12f}; triplet {t:.12f} "
          f"{t2:.12f} {tr:.12f}", flush=True)
    assert abs(s - s2) < 1e-12 and abs(t - t2) < 1e-12
    assert abs(s - sr) < 1e-10 and abs(t - tr) < 1e-10
    with pytest.raises(pa.PolyeriError, match="MODULO MUST EVENLY DIVIDE"):
        ctx.mp2_energy_periodic(2, r["co"][2:6], r["eps"][2:6], np.asarray(r["nall"])[2:6], icore, iocc, modulo=3)


@pytest.mark.parametrize("case", ["h2_631gss", "ne_ccpvdz", "pe_sto3g"])
def test_periodic_mp2_qp_matches_oracle(case):
    """k-point MP2 quasi-particle bands (IOPTN(44) = 2, src/pmp2energy.f90:
    445-522, 549-558): the device's wider transforms and its k_mp2k_qp sums
    vs the oracle's restatement on the same bands (1e-10 Eh); E(2) equals
    the nqp = 0 call (1e-12); bands outside the corrected range copy eps."""
    from oracle import periodic
    basis, atoms, period, cel1, cel2, kvc, icore = _case(case)
    nqp = 2
    print(f"case: k-point MP2 QP {case} kpoints {2 * kvc} icore {icore} nqp {nqp}", flush=True)
    b = pa.read_basis(basis, atoms)
    ctx = pa.Context(b, schwarz_tol=0.0, period=period, cel1=cel1, cel2=cel2)
    r = ctx.self_consistent_field_periodic(kpoints=kvc, scfconv=1e-8, diis=2, scfcycles=60)
    assert r["converged"]
    iocc = r["iocc"]
    s, t, q = ctx.mp2_quasiparticle_periodic(kvc, r["co"], r["eps"], r["nall"], icore, iocc, nqp)
    s0, t0 = ctx.mp2_energy_periodic(kvc, r["co"], r["eps"], r["nall"], icore, iocc)
    rc1 = ctx.reduced_cel1
    Tr = oracle.Basis(basis_path(basis), atoms).eri_tensor_cells(period, rc1, cel2)
    sr, tr, qr = periodic.mp2_kpoints(Tr, r["co"], r["eps"], rc1, cel2, kvc, icore, iocc, list(r["nall"]), nqp=nqp)
    lo, hi = max(iocc - nqp, icore), iocc + nqp
    print(f"E2 {s + t:.12f} (nqp 0: {s0 + t0:.12f}, oracle {sr + tr:.12f}); QP bands k=0: "
          f"{np.array2string(q[0][lo:hi], precision=8)} max |dq| {np.max(np.abs(q - qr)):.2e}", flush=True)
    assert abs(s - s0) < 1e-12 and abs(t - t0) < 1e-12
    assert abs(s - sr) < 1e-10 and abs(t - tr) < 1e-10
    assert np.max(np.abs(q - qr)) < 1e-10
    assert np.min(np.abs(q[:, lo:hi] - r["eps"][:, lo:hi])) > 1e-5
    keep = np.ones(q.shape[1], dtype=bool)
    keep[lo:hi] = False
    assert np.array_equal(q[:, keep], np.asarray(r["eps"])[:, keep])


def test_periodic_mp2_qp_modulo():
    """Quasi-particle bands with MODULO 2 on an 8-k-point SCF: the bands at
    k = kb * 2 equal the 4-k-point QP bands of that subset (1e-12), the
    other k-points copy eps."""
    basis, atoms, period, cel1, cel2, _, icore = _case("h2_631gss")
    kvc, nmod, nqp = 4, 2, 1
    print(f"case: k-point MP2 QP MODULO {nmod} of {2 * kvc} k-points", flush=True)
    ctx = pa.Context(pa.read_basis(basis, atoms), schwarz_tol=0.0, period=period, cel1=cel1, cel2=cel2)
    r = ctx.self_consistent_field_periodic(kpoints=kvc, scfconv=1e-8, diis=2, scfcycles=60)
    assert r["converged"]
    iocc = r["iocc"]
    s, t, q = ctx.mp2_quasiparticle_periodic(kvc, r["co"], r["eps"], r["nall"], icore, iocc, nqp, modulo=nmod)
    idx = [kb * nmod + kvc for kb in range(-kvc // nmod, kvc // nmod)]
    s2, t2, q2 = ctx.mp2_quasiparticle_periodic(kvc // nmod, r["co"][idx], r["eps"][idx],
                                                np.asarray(r["nall"])[idx], icore, iocc, nqp)
    print(f"MODULO QP: max |q - q_subset| {np.max(np.abs(q[idx] - q2)):.2e}", flush=True)
    assert abs(s - s2) < 1e-12 and abs(t - t2) < 1e-12
    assert np.max(np.abs(q[idx] - q2)) < 1e-12
    rest = [k for k in range(2 * kvc) if k not in idx]
    assert np.array_equal(q[rest], np.asarray(r["eps"])[rest])
