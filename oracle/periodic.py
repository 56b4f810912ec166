"""Periodic SCF of the reference (1-D, HELIX = 0, axis x), restated in numpy.

TEST INFRASTRUCTURE ONLY (the oracle; the product never imports it).

* one_electron_cells: S_C, T_C, N_C of cells Q in [-CEL1, CEL1] as blocks
  (cell 0, cell Q) of the molecular one-electron matrices of a supercell
  molecule spanning cells -CEL2..CEL2, so the nuclear attraction runs over
  the nuclei of exactly those cells (src/ponee.f90:486-560), and the
  nuclear repulsion per unit cell (src/pselfconsistentfield.f90:652-702);
* reduced_cel1: REDUCED_CEL1X, the last cell whose contracted overlap with
  cell 0 exceeds 1e-15 (src/ponee.f90:359-371);
* bands / density: the k-point part of FOCK_BUILD (src/pfock.f90:180-470):
  Bloch sums over the cells with phase exp(i pi IX Q / KVC), IX = -KVC ..
  KVC-1, Hermitian part, C2S, canonical orthogonalisation with LINDEP,
  diagonalisation, CO = C2S^T X C', one Fermi level over all k with weight
  1/(2 KVC), P_out(L,K,Q) = Re sum_k w sum_l occ conj(CO(K,l)) CO(L,l) e^{ikQ};
* scf: SELF_CONSISTENT_FIELD with DIIS every other iteration from IOPTN(16)
  on the cell-stacked densities (src/pselfconsistentfield.f90:1-162,
  src/pdensitymatrix.f90:131-339) and ENERGY_CALC summed over the cells
  (src/pselfconsistentfield.f90:487-574); J/K from the oracle's periodic
  FOURINDEX_ERI (oracle.c or_fourindex_eri_cells);
* ao_tensor_cells / mp2_kpoints: the stored lattice integrals
  (i_0 j_Q1 | k_Q3 l_Q3+Q2) of STORE_FOURINDEX_ERI (src/pfourindexstore.f90:
  258-747, from a supercell ERI tensor) and the k-point MP2_ENERGY_ALG1
  (src/pmp2energy.f90:187-548, HELIX = 0, MODULO 1), pinned by the
  isolated-molecule limit and the band-folding identity
  (tests/test_oracle_periodic.py).
"""
import numpy as np

import oracle

BOLTZMANN = 1.0 / 3.1577464e5      # src/pmodules.f90:12


def _supercell(atoms, period, cel2):
    out = []
    for c in range(-cel2, cel2 + 1):
        out += [(z, x + c * period, y, w) for z, x, y, w in atoms]
    return out


def _real(b):
    return [f for f in range(b.ncgs) if np.any(b.cc[f])]


def one_electron_cells(basis_path, atoms, period, cel1, cel2):
    if cel1 > cel2:
        raise ValueError("CUTOFF1 > CUTOFF2 is not supported here")
    unit = oracle.Basis(basis_path, atoms)
    sup_atoms = _supercell(atoms, period, cel2)
    sup = oracle.Basis(basis_path, sup_atoms)
    _, s_c, t_c, n_c = sup.one_electron()
    nat = len(atoms)
    nc = unit.ncgs
    # unit-cell function f of atom a -> supercell index in cell c
    def per_atom(b, ats):
        out = [[] for _ in ats]
        for f in _real(b):
            cen = np.array([b.cgsx[f], b.cgsy[f], b.cgsz[f]])
            a = int(np.argmin([np.linalg.norm(cen - np.array(t[1:])) for t in ats]))
            out[a].append(f)
        return out
    fu, fs = per_atom(unit, atoms), per_atom(sup, sup_atoms)
    idx = {}
    for a, fl in enumerate(fu):
        for k, f in enumerate(fl):
            for c in range(-cel2, cel2 + 1):
                idx[c, f] = fs[(c + cel2) * nat + a][k]
    S = np.zeros((2 * cel1 + 1, nc, nc))
    T = np.zeros_like(S)
    N = np.zeros_like(S)
    real = _real(unit)
    for q in range(-cel1, cel1 + 1):
        for i in real:
            for j in real:
                I, J = idx[0, i], idx[q, j]
                S[q + cel1, i, j] = s_c[I, J]
                T[q + cel1, i, j] = t_c[I, J]
                N[q + cel1, i, j] = n_c[I, J]
    enuc = 0.0
    for i, (zi, xi, yi, wi) in enumerate(atoms):
        for j, (zj, xj, yj, wj) in enumerate(atoms):
            for k in range(-cel2, cel2 + 1):
                if k == 0 and i == j:
                    continue
                r = np.sqrt((xj + k * period - xi) ** 2 + (yj - yi) ** 2 + (wj - wi) ** 2)
                enuc += 0.5 * zi * zj / r
    return unit, S, T, N, enuc


def reduced_cel1(S, cel1):
    rc = 0
    for q in range(0, cel1 + 1):
        if np.max(np.abs(S[q + cel1])) > 1e-15:
            rc = q
    return rc


def kphases(kvc):
    return [np.pi * ix / kvc for ix in range(-kvc, max(0, kvc - 1) + 1)] if kvc > 0 else [0.0]


def bloch(M, cel1, phi):
    """Hermitian part of sum_Q M(Q) e^{i phi Q} (M[Q + cel1] = M(Q))."""
    W = sum(M[q + cel1] * np.exp(1j * phi * q) for q in range(-cel1, cel1 + 1))
    return 0.5 * (W + W.conj().T)


def bands(F, S, c2s, cel1, kvc, lindep):
    """Band energies (nk, nc) and Cartesian CO (nk, nc, nc) per k."""
    nc = F.shape[1]
    eps, cos, nind = [], [], []
    for phi in kphases(kvc):
        ws = c2s @ bloch(S, cel1, phi) @ c2s.T
        wf = c2s @ bloch(F, cel1, phi) @ c2s.T
        e, u = np.linalg.eigh(ws)
        keep = e > lindep
        X = u[:, keep] / np.sqrt(e[keep])
        ek, ck = np.linalg.eigh(X.conj().T @ wf @ X)
        co = np.zeros((nc, nc), dtype=complex)
        co[:, :keep.sum()] = c2s.T @ (X @ ck)
        ee = np.full(nc, 1.0e99)
        ee[:keep.sum()] = ek
        eps.append(ee)
        cos.append(co)
        nind.append(int(keep.sum()))
    return np.array(eps), np.array(cos), nind


def fermi_level(eps, nel, temphf):
    nk = eps.shape[0]

    def count(mu):
        return np.sum(2.0 / (np.exp(np.clip((eps - mu) / BOLTZMANN / temphf, -700, 700)) + 1.0)) / nk - nel
    mine, maxe = np.min(eps), np.max(eps)
    f1, f2, f12 = mine, maxe, 0.5 * (mine + maxe)
    for _ in range(100000):
        n1, n2, n12 = count(f1), count(f2), count(f12)
        if abs(n12) < 1e-10:
            return f12
        if n1 > 0 and n2 > 0:
            f2, f1 = f1, f1 - 0.1
        elif n1 < 0 and n2 < 0:
            f1, f2 = f2, f2 + 0.1
        elif n1 < 0 and n12 >= 0:
            f2 = f12
        elif n12 <= 0 and n2 > 0:
            f1 = f12
        else:
            break
        f12 = 0.5 * (f1 + f2)
    return f12


def density(eps, co, cel1, kvc, fermi, temphf):
    nk, nc = eps.shape
    occ = 2.0 / (np.exp(np.clip((eps - fermi) / BOLTZMANN / temphf, -700, 700)) + 1.0)
    P = np.zeros((2 * cel1 + 1, nc, nc))
    w = 1.0 / nk
    for k, phi in enumerate(kphases(kvc)):
        ws = (co[k].conj() * occ[k]) @ co[k].T          # WS(K,J) = sum_l occ conj(CO(K,l)) CO(J,l)
        for q in range(-cel1, cel1 + 1):
            # P_C_OUT(L,K,Q) += WS(L,K) e^{i phi Q} w
            P[q + cel1] += np.real(ws * np.exp(1j * phi * q)) * w
    return P


def _lu_solve(B, rhs):
    x = np.linalg.solve(B, rhs)
    return x - np.linalg.solve(B, B @ x - rhs)


def scf(basis_path, atoms, period, cel1, cel2, kvc, inttol=1e-12, scfconv=1e-8, scfrelax=1.0,
        hfexchange=1.0, lindep=1e-6, temphf=0.0, scfcycles=50, diis=5, charge=0, spherical=False):
    unit, S, T, N, enuc = one_electron_cells(basis_path, atoms, period, cel1, cel2)
    rc1 = reduced_cel1(S, cel1)
    nc = unit.ncgs
    c2s = unit.c2s(spherical)
    nel = sum(a[0] for a in atoms) - charge
    temphf = temphf if temphf > 0 else 1.0e-4
    # GUESS_DENSITYMATRIX in cell 0
    P = np.zeros((2 * cel1 + 1, nc, nc))
    for a, (z, x, y, w) in enumerate(atoms):
        fl = [f for f in _real(unit) if (unit.cgsx[f], unit.cgsy[f], unit.cgsz[f]) == (x, y, w)]
        for f in fl:
            P[cel1, f, f] = z / len(fl)
    hp, hd, pulay = {}, {}, np.zeros((scfcycles + 2, scfcycles + 2))
    trace = []
    res = None
    for it in range(1, scfcycles + 1):
        J, K, _ = unit.fourindex_eri_cells(P, period, rc1, cel1, cel2, inttol)
        F = T + N + J - 0.5 * hfexchange * K
        eps, co, nind = bands(F, S, c2s, cel1, kvc, lindep)
        mu = fermi_level(eps, nel, temphf)
        Pout = density(eps, co, cel1, kvc, mu, temphf)
        hp[it], hd[it] = P.copy(), Pout - P
        for i in range(max(1, it - 8), it + 1):
            pulay[i, it] = pulay[it, i] = float(np.sum(hd[i] * hd[it]))
        dchg = np.sqrt(pulay[it, it] / P.size)
        if diis != 0 and it >= diis and (it - diis) % 2 == 0:
            nt = min(8, it)
            n = nt + 1
            B = np.zeros((n, n))
            rhs = np.zeros(n)
            for i in range(nt):
                for j in range(nt):
                    B[i, j] = pulay[it - j, it - i]
                B[nt, i] = B[i, nt] = -1.0
            rhs[nt] = -1.0
            cv = _lu_solve(B, rhs)
            P = sum(cv[it - i] * hp[i] for i in range(it - nt + 1, it + 1))
        else:
            P = scfrelax * Pout + (1.0 - scfrelax) * P
        ek, ev = float(np.sum(P * T)), float(np.sum(P * N))
        ej, ex = 0.5 * float(np.sum(P * J)), -0.25 * hfexchange * float(np.sum(P * K))
        etot = enuc + ek + ev + ej + ex
        trace.append((dchg, etot))
        res = dict(e_total=etot, e_nuc=enuc, e_kin=ek, e_nuc_att=ev, e_coul=ej, e_hfx=ex, fermi=mu,
                   iterations=it, converged=dchg < scfconv, P=P, eps=eps, co=co, rc1=rc1, nind=nind)
        if dchg < scfconv:
            break
    res["trace"] = np.array(trace)
    return res


# ---------------------------------------------------------------------------
# k-point MP2: MP2_ENERGY_ALG1 with HELIX = 0 and MODULO 1
# (src/pmp2energy.f90:187-548)
# ---------------------------------------------------------------------------
def ao_tensor_cells(basis_path, atoms, period, rc1, cel2):
    """(i_0 j_Q1 | k_Q3 l_Q3+Q2) as T[Q1+rc1, Q2+rc1, Q3+cel2, i, j, k, l]
    (contracted, reference layout) from the molecular ERI tensor of a
    supercell spanning every cell the indices reach."""
    lo = -(cel2 + rc1)
    sup_atoms = _supercell(atoms, period, cel2 + rc1)
    unit = oracle.Basis(basis_path, atoms)
    sup = oracle.Basis(basis_path, sup_atoms)
    g = sup.eri_tensor()
    nat, nc = len(atoms), unit.ncgs

    def per_atom(b, ats):
        out = [[] for _ in ats]
        for f in _real(b):
            cen = np.array([b.cgsx[f], b.cgsy[f], b.cgsz[f]])
            a = int(np.argmin([np.linalg.norm(cen - np.array(t[1:])) for t in ats]))
            out[a].append(f)
        return out
    fu, fs = per_atom(unit, atoms), per_atom(sup, sup_atoms)
    real = _real(unit)
    idx = np.zeros((2 * (cel2 + rc1) + 1, nc), dtype=int)
    for a, fl in enumerate(fu):
        for k, f in enumerate(fl):
            for c in range(lo, -lo + 1):
                idx[c - lo, f] = fs[(c - lo) * nat + a][k]
    n1, n3 = 2 * rc1 + 1, 2 * cel2 + 1
    T = np.zeros((n1, n1, n3, nc, nc, nc, nc))
    R = np.array(real)
    for q1 in range(-rc1, rc1 + 1):
        for q2 in range(-rc1, rc1 + 1):
            for q3 in range(-cel2, cel2 + 1):
                I = idx[0 - lo, R]
                J = idx[q1 - lo, R]
                K = idx[q3 - lo, R]
                L = idx[q3 + q2 - lo, R]
                blk = g[np.ix_(I, J, K, L)]
                T[q1 + rc1, q2 + rc1, q3 + cel2][np.ix_(R, R, R, R)] = blk
    return T


def _kwrap(k, kvc):
    """K1 - K2 + K3 folded back into -KVC .. KVC-1 (src/pmp2energy.f90:366-374)."""
    if kvc == 0:
        return 0
    if k >= kvc:
        k -= 2 * kvc
    if k < -kvc:
        k += 2 * kvc
    return k


def mp2_kpoints(T, co, eps, rc1, cel2, kvc, icore, iocc, nall, ivirtcore=0, nqp=0):
    """E(2) singlet and triplet per unit cell (src/pmp2energy.f90:187-548).

    T[Q1 + rc1, Q2 + rc1, Q3 + cel2, i, j, k, l] = (i_0 j_Q1 | k_Q3 l_Q3+Q2)
    (the records STORE_FOURINDEX_ERI writes, src/pfourindexstore.f90:727-747);
    co[k] (ncgs x ncgs, columns = bands) and eps[k] for the k grid
    IX = -kvc .. kvc-1 (kvc 0: Gamma only); nall[k] = IALL(k), the bands
    kept by LINDEP.  Per (k_b, b) the four quarter transforms of the
    reference, here for every virtual b of one k_b at once:
      TNNNKK(I,J,K,Q1,Q3) = sum_{L,Q2} T CO(L,b,kb) e^{i kb (Q2+Q3)}      :253-258
      TOK(j,kj)           = sum_{K,Q3} TNNNKK conj(CO(K,j,kj)) e^{-i kj Q3} :294-299
      TVOKK(a,j,ka,kj)    = sum_{J,Q1} TOK CO(J,a,ka) e^{i ka Q1}           :323-329
      TOVOKK(i,a,j,ka,kj) = sum_I TVOKK conj(CO(I,i,ki)), ki = ka-kj+kb     :350-386
    then E2S/E2T over i, j (occupied), a, b (virtual) divided by (2 kvc)^3
    (:400-433, :542-543).

    nqp > 0 (IOPTN(44)): the j and i ranges reach band min(iocc + nqp,
    IALL - IVIRTCORE) - 1 and the a range starts at max(iocc - nqp, icore)
    (:137-153), and the quasi-particle bands g in max(iocc - nqp, icore) ..
    min(iocc + nqp, IALL(kg) - IVIRTCORE) - 1 at every kg collect (:445-522)
      sum_{kj,ka; kg = ka-kj+kb} sum_{j,a} [2|T(g,a,j)|^2 - Re T(g,a,j;ka,kj) T*(j,a,g;ka,kg)]
                                            / (e_g + e_j - e_a - e_b)
      sum_{kj,kg; ki = kg-kj+kb} sum_{j,i} [2|T(i,g,j)|^2 - Re T(i,g,j;kg,kj) T*(j,g,i;kg,ki)]
                                            / (e_g + e_b - e_i - e_j)
    and QEPSILON = eps + that / (2 kvc)^2 (:549-558).  Returns (es, et) for
    nqp = 0, else (es, et, qeps) with qeps[k] a copy of eps[k] outside the
    corrected bands."""
    nk = 2 * kvc if kvc > 0 else 1
    ks = list(range(-kvc, kvc)) if kvc > 0 else [0]

    def ph(n):
        return np.exp(1j * np.pi * n / kvc) if kvc > 0 else 1.0 + 0j
    q1s = np.arange(-rc1, rc1 + 1)
    q3s = np.arange(-cel2, cel2 + 1)
    for k in range(nk):
        if nall[k] - ivirtcore < iocc:
            raise ValueError("fewer bands than occupied orbitals at a k-point")
    plim = [nall[k] - ivirtcore for k in range(nk)]
    pmax = max(plim)
    no = iocc - icore
    op = np.arange(icore, min(iocc + nqp, pmax))         # j and i bands (occupied + QP)
    vlo = max(iocc - nqp, icore)
    vv = np.arange(vlo, pmax)                            # a bands (QP + virtual)
    aoff = iocc - vlo
    nop, nvp = len(op), len(vv)

    def bands(k_i, idx):
        """CO columns idx at k, zero past IALL(k) - IVIRTCORE."""
        c = co[k_i][:, idx].copy()
        c[:, idx >= plim[k_i]] = 0.0
        return c
    es, et = 0.0, 0.0
    dq = np.zeros((nk, pmax))
    for kb_i, kb in enumerate(ks):
        vb = np.arange(iocc, plim[kb_i])
        if len(vb) == 0:
            continue
        # TNNNKK[q1, q3, i, j, k, b]
        cb = co[kb_i][:, vb]
        tn = np.zeros((len(q1s), len(q3s)) + T.shape[3:6] + (len(vb),), dtype=complex)
        for a3, q3 in enumerate(q3s):
            for a2, q2 in enumerate(q1s):
                tn[:, a3] += np.einsum("qijkl,lb->qijkb", T[:, a2, a3], cb) * ph(kb * (q2 + q3))
        # TOK[q1, i, j, kj, jo, b]
        tok = np.zeros((len(q1s),) + T.shape[3:5] + (nk, nop, len(vb)), dtype=complex)
        for kj_i, kj in enumerate(ks):
            cj = bands(kj_i, op).conj()
            for a3, q3 in enumerate(q3s):
                tok[:, :, :, kj_i] += np.einsum("qijkb,ko->qijob", tn[:, a3], cj) * ph(-kj * q3)
        # TVOKK[i, ka, a, kj, jo, b]
        tv = np.zeros((T.shape[3], nk, nvp, nk, nop, len(vb)), dtype=complex)
        for ka_i, ka in enumerate(ks):
            ca = bands(ka_i, vv)
            for a1, q1 in enumerate(q1s):
                tv[:, ka_i] += np.einsum("ijkob,ja->iakob", tok[a1], ca) * ph(ka * q1)
        # TOVOKK[io, ka, a, kj, jo, b]
        tov = np.zeros((nop, nk, nvp, nk, nop, len(vb)), dtype=complex)
        for ka_i, ka in enumerate(ks):
            for kj_i, kj in enumerate(ks):
                ki_i = ks.index(_kwrap(ka - kj + kb, kvc))
                tov[:, ka_i, :, kj_i] = np.einsum("iaob,im->maob", tv[:, ka_i, :, kj_i], bands(ki_i, op).conj())
        eb = eps[kb_i][vb]
        for ka_i, ka in enumerate(ks):
            nva = plim[ka_i] - iocc
            ea = eps[ka_i][iocc:plim[ka_i]]
            for kj_i, kj in enumerate(ks):
                ki_i = ks.index(_kwrap(ka - kj + kb, kvc))
                t1 = tov[:no, ka_i, aoff:aoff + nva, kj_i, :no]                # (i, a, j, b)
                t2 = np.transpose(tov[:no, ka_i, aoff:aoff + nva, ki_i, :no], (2, 1, 0, 3))   # TOVOKK(j, a, i) at (ka, ki)
                d = (eps[ki_i][icore:iocc][:, None, None, None] + eps[kj_i][icore:iocc][None, None, :, None]
                     - ea[None, :, None, None] - eb[None, None, None, :])
                aa = np.real(t1 * t1.conj())
                ab = np.real(t1 * t2.conj())
                es += float(np.sum((0.5 * aa + 0.5 * ab) / d))
                et += float(np.sum((1.5 * aa - 1.5 * ab) / d))
        if nqp <= 0:
            continue
        ej = [eps[k][icore:iocc] for k in range(nk)]
        for kj_i, kj in enumerate(ks):
            for ka_i, ka in enumerate(ks):              # first sum: g in the i slot (:446-483)
                kg_i = ks.index(_kwrap(ka - kj + kb, kvc))
                ea = eps[ka_i][iocc:plim[ka_i]]
                nva = len(ea)
                for g in range(vlo, min(iocc + nqp, plim[kg_i])):
                    t1 = tov[g - icore, ka_i, aoff:aoff + nva, kj_i, :no]           # T(g,a,j)[a, j, b]
                    t2 = tov[:no, ka_i, aoff:aoff + nva, kg_i, g - icore]           # T(j,a,g)[j, a, b]
                    t2 = np.transpose(t2, (1, 0, 2))
                    d = eps[kg_i][g] + ej[kj_i][None, :, None] - ea[:, None, None] - eb[None, None, :]
                    dq[kg_i, g] += float(np.sum((2.0 * np.real(t1 * t1.conj()) - np.real(t1 * t2.conj())) / d))
            for kg_i, kg in enumerate(ks):              # second sum: g in the a slot (:484-520)
                ki_i = ks.index(_kwrap(kg - kj + kb, kvc))
                for g in range(vlo, min(iocc + nqp, plim[kg_i])):
                    t1 = tov[:no, kg_i, g - vlo, kj_i, :no]                         # T(i,g,j)[i, j, b]
                    t2 = np.transpose(tov[:no, kg_i, g - vlo, ki_i, :no], (1, 0, 2))  # T(j,g,i)[i, j, b]
                    d = eps[kg_i][g] + eb[None, None, :] - ej[ki_i][:, None, None] - ej[kj_i][None, :, None]
                    dq[kg_i, g] += float(np.sum((2.0 * np.real(t1 * t1.conj()) - np.real(t1 * t2.conj())) / d))
    norm = float(nk) ** 3
    if nqp <= 0:
        return es / norm, et / norm
    qeps = np.array([np.asarray(eps[k], dtype=np.float64).copy() for k in range(nk)])
    for k in range(nk):
        for g in range(vlo, min(iocc + nqp, plim[k])):
            qeps[k][g] = eps[k][g] + dq[k, g] / float(nk) ** 2
    return es / norm, et / norm, qeps
