"""Closed-shell coupled-cluster path of the reference, restated in numpy.

TEST INFRASTRUCTURE ONLY (the oracle; the product never imports it).

COUPLED_CLUSTER_ENERGY (src/pcc.f90:1-243) at one k-point with
COPTN(70) = 'CCSD':

* mo_integrals: INTEGRAL_TRANSFORM_DIRECTACCESS (src/pcc.f90:247-436), the
  full MO integral set (pq|rs) over the active orbitals ICORE .. IALL -
  IVIRTCORE - 1 from the AO tensor and CO;
* mp2 / mp3: MP2ENERGY_DISK and MP3ENERGY_DISK (src/pcc.f90:503-807): E(2)
  from the first-order doubles, E(3) = sum L t2(2) with t2(2) the linear
  doubles residual of t2(1) over the denominators (singles enter at fourth
  order for canonical orbitals);
* ccsd: the spin-restricted CC iteration (src/pcc.f90:120-240, the kappa
  and amplitude subroutines :1170-3094), here as the spin-adapted closed-shell
  equations with the usual one- and two-body intermediates (F_oo, F_vv, F_ov,
  L_oo, L_vv, W_oooo, W_vvvv, W_voov, W_vovo) for canonical HF orbitals, and
  commutator-free DIIS over the amplitude vectors; E(SINGLES) and
  E(DOUBLES) as ENERGY_T1_CONTRIBUTION / ENERGY_T2_CONTRIBUTION
  (src/pcc.f90:1041-1166): sum L_ijab t_i^a t_j^b and sum L_ijab t_ij^ab with
  L_ijab = 2 (ia|jb) - (ib|ja);
* CCD, LCCD, LCCSD, QCISD (COPTN(70), src/pcc.f90:39-66): the components of
  the CCSD residual of given degrees (p, q) in (t1, t2), extracted exactly
  from CCSD residuals at integer-scaled amplitudes (residual_components):
  CCD = CCSD at t1 = 0; LCCD its linear part; LCCSD the linear part in
  (t1, t2); QCISD singles linear + t1 t2, doubles linear + t2^2.  E(CORRELATION)
  is the doubles energy for all but CCSD (src/pcc.f90:202-207).

Pinned by the shipped H2O/cc-pVDZ CCSD run (E(2), E(3), E(SINGLES),
E(DOUBLES), E(CORRELATION); data/H2O_CCSD_ccpVDZ_out:391-451) and by
two-electron exactness (CCSD = FCI for H2, brute-force FCI in the MO basis)
in tests/test_oracle_cc.py.
"""
import numpy as np


def mo_integrals(feri, co, lo, hi):
    """(pq|rs) for MOs lo .. hi-1 (columns of co) from the AO tensor feri."""
    c = np.asarray(co, dtype=np.float64)[:, lo:hi]
    t = np.tensordot(np.asarray(feri, dtype=np.float64), c, axes=([3], [0]))        # (pqr|s')
    t = np.tensordot(t, c, axes=([2], [0]))                                          # p q s' r'
    t = np.tensordot(t, c, axes=([1], [0]))                                          # p s' r' q'
    t = np.tensordot(t, c, axes=([0], [0]))                                          # s' r' q' p'
    return np.ascontiguousarray(t.transpose(3, 2, 1, 0))


class _Ints:
    """Occupied/virtual blocks of the MO integrals (chemists' notation)."""

    def __init__(self, g, no):
        o, v = slice(0, no), slice(no, g.shape[0])
        self.ovov = g[o, v, o, v]
        self.oovv = g[o, o, v, v]
        self.ovvo = g[o, v, v, o]
        self.ovoo = g[o, v, o, o]
        self.ovvv = g[o, v, v, v]
        self.oooo = g[o, o, o, o]
        self.vvvv = g[v, v, v, v]
        self.L = 2.0 * self.ovov - self.ovov.transpose(0, 3, 2, 1)     # L[i,a,j,b] = 2(ia|jb) - (ib|ja)


def _denoms(eps, no):
    eo, ev = eps[:no], eps[no:]
    d1 = eo[:, None] - ev[None, :]
    d2 = d1[:, None, :, None] + d1[None, :, None, :]
    return d1, d2


def _pair_sym(x):
    """P(ia, jb) X_ijab = X_ijab + X_jiba."""
    return x + x.transpose(1, 0, 3, 2)


def _linear_doubles(I, t2):
    """Doubles residual linear in t2 with no singles and no Fock terms (the
    MP3 / linear CCD part)."""
    ein = np.einsum
    r = ein("kilj,klab->ijab", I.oooo, t2, optimize=True)
    r += ein("acbd,ijcd->ijab", I.vvvv, t2, optimize=True)
    x = 2.0 * ein("kcai,kjcb->ijab", I.ovvo, t2, optimize=True)
    x -= ein("kiac,kjcb->ijab", I.oovv, t2, optimize=True)
    x -= ein("kcai,kjbc->ijab", I.ovvo, t2, optimize=True)
    x -= ein("kibc,kjac->ijab", I.oovv, t2, optimize=True)
    return r + _pair_sym(x)


def _energy_parts(I, t1, t2):
    es = float(np.einsum("iajb,ia,jb->", I.L, t1, t1, optimize=True))
    ed = float(np.einsum("iajb,ijab->", I.L, t2, optimize=True))
    return es, ed


def mp2(g, eps, no):
    """E(2) from the MO integrals over the active space (occupied first)."""
    I = _Ints(g, no)
    _, d2 = _denoms(eps, no)
    t2 = I.ovov.transpose(0, 2, 1, 3) / d2
    return _energy_parts(I, np.zeros(d2.shape[0:1] + d2.shape[2:3]), t2)[1]


def mp3(g, eps, no):
    """E(3): sum L t2(2), t2(2) = linear doubles residual of t2(1) / D."""
    I = _Ints(g, no)
    _, d2 = _denoms(eps, no)
    t1 = I.ovov.transpose(0, 2, 1, 3) / d2
    t2 = _linear_doubles(I, t1) / d2
    return float(np.einsum("iajb,ijab->", I.L, t2, optimize=True))


def ccsd_residual(I, t1, t2):
    """The CCSD amplitude equations without the diagonal Fock terms, for
    canonical orbitals (f_ia = 0, f diagonal): returns (r1, r2) with the new
    amplitudes r / D."""
    ein = lambda s, *a: np.einsum(s, *a, optimize=True)
    tau = t2 + ein("ia,jb->ijab", t1, t1)
    Lov = I.L                                                        # [k,c,l,d]
    # one-body intermediates (t-dependent parts)
    Foo = ein("kcld,ilcd->ki", Lov, tau)
    Fvv = -ein("kcld,klad->ac", Lov, tau)
    Fov = ein("kcld,ld->kc", Lov, t1)
    Loo = Foo + 2.0 * ein("lcki,lc->ki", I.ovoo, t1) - ein("kcli,lc->ki", I.ovoo, t1)
    Lvv = Fvv + 2.0 * ein("kdac,kd->ac", I.ovvv, t1) - ein("kcad,kd->ac", I.ovvv, t1)
    # singles
    r1 = ein("ac,ic->ia", Fvv, t1) - ein("ki,ka->ia", Foo, t1)
    r1 += 2.0 * ein("kc,kica->ia", Fov, t2) - ein("kc,ikca->ia", Fov, t2) + ein("kc,ic,ka->ia", Fov, t1, t1)
    r1 += 2.0 * ein("kcai,kc->ia", I.ovvo, t1) - ein("kiac,kc->ia", I.oovv, t1)
    tk = t2 + ein("ic,kd->ikcd", t1, t1)
    r1 += 2.0 * ein("kdac,ikcd->ia", I.ovvv, tk) - ein("kcad,ikcd->ia", I.ovvv, tk)
    tl = t2 + ein("ka,lc->klac", t1, t1)
    r1 -= 2.0 * ein("lcki,klac->ia", I.ovoo, tl) - ein("kcli,klac->ia", I.ovoo, tl)
    # two-body intermediates
    Woooo = I.oooo.transpose(0, 2, 1, 3).copy()                      # (ki|lj) at [k,l,i,j]
    Woooo += ein("lcki,jc->klij", I.ovoo, t1) + ein("kclj,ic->klij", I.ovoo, t1)
    Woooo += ein("kcld,ijcd->klij", I.ovov, tau)
    Wvvvv = I.vvvv.transpose(0, 2, 1, 3).copy()                      # (ac|bd) at [a,b,c,d]
    Wvvvv -= ein("kdac,kb->abcd", I.ovvv, t1) + ein("kcbd,ka->abcd", I.ovvv, t1)
    Wvoov = I.ovvo.transpose(2, 0, 3, 1).copy()                      # (kc|ai) at [a,k,i,c]
    Wvoov += ein("kcad,id->akic", I.ovvv, t1) - ein("kcli,la->akic", I.ovoo, t1)
    Wvoov -= 0.5 * ein("ldkc,ilda->akic", I.ovov, t2) + 0.5 * ein("lckd,ilad->akic", I.ovov, t2)
    Wvoov -= ein("ldkc,id,la->akic", I.ovov, t1, t1)
    Wvoov += ein("ldkc,ilad->akic", I.ovov, t2)
    Wvovo = I.oovv.transpose(2, 0, 3, 1).copy()                      # (ki|ac) at [a,k,c,i]
    Wvovo += ein("kdac,id->akci", I.ovvv, t1) - ein("lcki,la->akci", I.ovoo, t1)
    Wvovo -= 0.5 * ein("lckd,ilda->akci", I.ovov, t2) + ein("lckd,id,la->akci", I.ovov, t1, t1)
    # doubles
    r2 = I.ovov.transpose(0, 2, 1, 3).copy()
    x = ein("iacb,jc->ijab", I.ovvv, t1) - ein("kibc,ka,jc->ijab", I.oovv, t1, t1)
    r2 += _pair_sym(x)
    x = ein("iajk,kb->ijab", I.ovoo, t1) + ein("kcai,jc,kb->ijab", I.ovvo, t1, t1)
    r2 -= _pair_sym(x)
    r2 += ein("klij,klab->ijab", Woooo, tau) + ein("abcd,ijcd->ijab", Wvvvv, tau)
    x = ein("ac,ijcb->ijab", Lvv, t2) - ein("ki,kjab->ijab", Loo, t2)
    x += 2.0 * ein("akic,kjcb->ijab", Wvoov, t2) - ein("akci,kjcb->ijab", Wvovo, t2)
    x -= ein("akic,kjbc->ijab", Wvoov, t2) + ein("bkci,kjac->ijab", Wvovo, t2)
    r2 += _pair_sym(x)
    return r1, r2


# Degree components of the CCSD residual kept by each method of COPTN(70)
# (src/pcc.f90:39-66): (p, q) = powers of (t1, t2); None = no singles.
METHODS = {
    "CCSD": (None, None),
    "CCD": ([], [(0, 0), (0, 1), (0, 2)]),
    "LCCD": ([], [(0, 0), (0, 1)]),
    "LCCSD": ([(0, 0), (1, 0), (0, 1)], [(0, 0), (1, 0), (0, 1)]),
    "QCISD": ([(0, 0), (1, 0), (0, 1), (1, 1)], [(0, 0), (1, 0), (0, 1), (0, 2)]),
}


def residual_components(I, t1, t2):
    """The CCSD residual as a polynomial in (lambda t1, mu t2): returns
    C1[p][q], C2[p][q] (p <= 4, q <= 2) with R(lambda t1, mu t2) =
    sum lambda^p mu^q C[p][q], from 5 x 3 exact integer nodes."""
    lam = np.array([-2.0, -1.0, 0.0, 1.0, 2.0])
    mu = np.array([-1.0, 0.0, 1.0])
    Vl = np.vander(lam, 5, increasing=True)
    Vm = np.vander(mu, 3, increasing=True)
    Il, Im = np.linalg.inv(Vl), np.linalg.inv(Vm)
    R1 = [[None] * 3 for _ in range(5)]
    R2 = [[None] * 3 for _ in range(5)]
    for a, x in enumerate(lam):
        for b, y in enumerate(mu):
            R1[a][b], R2[a][b] = ccsd_residual(I, x * t1, y * t2)
    C1 = [[sum(Il[p, a] * Im[q, b] * R1[a][b] for a in range(5) for b in range(3)) for q in range(3)]
          for p in range(5)]
    C2 = [[sum(Il[p, a] * Im[q, b] * R2[a][b] for a in range(5) for b in range(3)) for q in range(3)]
          for p in range(5)]
    return C1, C2


def method_residual(I, t1, t2, method):
    """The residual of COPTN(70) = method: CCSD in full, the others as the
    (p, q) components of the CCSD residual listed in METHODS."""
    if method == "CCSD":
        return ccsd_residual(I, t1, t2)
    keep1, keep2 = METHODS[method]
    C1, C2 = residual_components(I, t1, t2)
    r1 = sum((C1[p][q] for p, q in keep1), np.zeros_like(t1))
    r2 = sum((C2[p][q] for p, q in keep2), np.zeros_like(t2))
    return r1, r2


def ccsd(g, eps, no, conv=1e-11, maxiter=200, ndiis=8, method="CCSD"):
    """Closed-shell coupled cluster (method: CCSD, CCD, LCCD, LCCSD, QCISD)
    on the active MO integrals g (occupied first, no occupied) and orbital
    energies eps.  Returns a dict with e_singles, e_doubles, e_corr (CCSD:
    singles + doubles; the others: doubles, src/pcc.f90:202-207),
    iterations, t1, t2."""
    if method not in METHODS:
        raise ValueError("UNKNOWN OR UNIMPLEMENTED COUPLED-CLUSTER METHOD")
    I = _Ints(np.asarray(g, dtype=np.float64), no)
    eps = np.asarray(eps, dtype=np.float64)
    d1, d2 = _denoms(eps, no)
    t1 = np.zeros_like(d1)
    t2 = np.zeros_like(d2)
    vecs, errs = [], []
    singles = METHODS[method][0] is None or len(METHODS[method][0]) > 0
    for it in range(1, maxiter + 1):
        r1, r2 = method_residual(I, t1, t2, method)
        if not singles:
            r1 = np.zeros_like(t1)
        n1, n2 = r1 / d1, r2 / d2
        dt = np.concatenate([(n1 - t1).ravel(), (n2 - t2).ravel()])
        vec = np.concatenate([n1.ravel(), n2.ravel()])
        vecs.append(vec)
        errs.append(dt)
        if len(vecs) > ndiis:
            vecs.pop(0)
            errs.pop(0)
        if len(vecs) > 1:
            m = len(vecs)
            B = -np.ones((m + 1, m + 1))
            B[m, m] = 0.0
            for a in range(m):
                for b in range(m):
                    B[a, b] = errs[a] @ errs[b]
            rhs = np.zeros(m + 1)
            rhs[m] = -1.0
            w = np.linalg.solve(B, rhs)[:m]
            vec = sum(wi * v for wi, v in zip(w, vecs))
        t1 = vec[:d1.size].reshape(d1.shape)
        t2 = vec[d1.size:].reshape(d2.shape)
        if np.max(np.abs(dt)) < conv:
            break
    es, ed = _energy_parts(I, t1, t2)
    ec = es + ed if method == "CCSD" else ed
    return {"e_singles": es, "e_doubles": ed, "e_corr": ec, "iterations": it, "t1": t1, "t2": t2}


def fci_two_electron(g, eps):
    """Exact singlet ground-state correlation energy of a two-electron
    closed-shell system in the MO basis (one occupied orbital, index 0):
    h = f - (2 J_0 - K_0) from the canonical Fock matrix, H on the symmetric
    spatial products phi_p(1) phi_q(2), lowest eigenvalue minus E_HF."""
    n = g.shape[0]
    h = np.diag(np.asarray(eps, dtype=np.float64)) - (2.0 * g[:, :, 0, 0] - g[:, 0, 0, :])
    H = (np.einsum("pr,qs->pqrs", h, np.eye(n)) + np.einsum("pr,qs->pqrs", np.eye(n), h)
         + g.transpose(0, 2, 1, 3)).reshape(n * n, n * n)
    cols = []
    for p in range(n):
        for q in range(p, n):
            v = np.zeros((n, n))
            v[p, q] += 1.0
            v[q, p] += 1.0
            cols.append(v.ravel() / np.linalg.norm(v))
    B = np.array(cols).T
    e = np.linalg.eigvalsh(B.T @ H @ B)[0]
    ehf = 2.0 * h[0, 0] + g[0, 0, 0, 0]
    return float(e - ehf)
