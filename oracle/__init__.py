"""CPU oracle for polymer's direct two-electron path.

TEST INFRASTRUCTURE ONLY.  This package wraps oracle/liboracle.so, a C
restatement of the reference algorithm (src/pfourindexdirect.f90,
src/pread.f90 READ3, src/ponee.f90, src/pselfconsistentfield.f90,
src/pfock.f90, src/pdensitymatrix.f90, src/pmp2energy.f90).  It is the
checker the MI355X product is compared against and the CPU baseline that
bench.py times.  The product (polymer_amd) never imports it.

Parity status of each restated function is listed in the C header
(oracle/oracle.h); every function is pinned by tests/test_oracle_*.py.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ANG_BOHR = 0.529177249  # src/pmodules.f90:9


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        d, i, p = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        ll = ctypes.c_longlong
        L.or_basis_build.argtypes = [ctypes.c_char_p, i, p, p, ctypes.POINTER(p)]
        L.or_basis_build.restype = i
        L.or_basis_free.argtypes = [p]
        L.or_basis_dims.argtypes = [p, p, p, p]
        L.or_basis_export.argtypes = [p] * 15
        L.or_boys_table.argtypes = [p]
        L.or_boys.argtypes = [d, i]
        L.or_boys.restype = d
        L.or_one_electron.argtypes = [p, p, p, p, p]
        L.or_nuclear_repulsion.argtypes = [p]
        L.or_nuclear_repulsion.restype = d
        L.or_fourindex_eri.argtypes = [p, p, p, d, i, p, p, p, p]
        L.or_fourindex_eri.restype = ll
        L.or_fourindex_eri_range.argtypes = [p, p, p, d, i, ll, ll, p, p, p, p]
        L.or_fourindex_eri_range.restype = ll
        L.or_fourindex_eri_rows.argtypes = [p, p, p, d, i, p, ll, p, p, p, p, p, p]
        L.or_fourindex_eri_rows.restype = ll
        L.or_fourindex_eri_rows_mt.argtypes = [p, p, d, i, p, ll, i, p, p, p, p, p]
        L.or_fourindex_eri_rows_mt.restype = ll
        L.or_eri_tensor.argtypes = [p, p, p]
        L.or_eri_tensor_cells.argtypes = [p, d, i, i, p]
        L.or_fourindex_eri_cells_c.argtypes = [p, d, i, i, i, p, d, p, p]
        L.or_fourindex_eri_cells_c.restype = ll
        L.or_decontract.argtypes = [p, p, p]
        L.or_form_c2s.argtypes = [p, i, p]
        L.or_scf.argtypes = [p, p, p, p, p, p, p]
        L.or_scf.restype = i
        L.or_mp2.argtypes = [i, p, p, p, i, i, i, i, p, p]
        L.or_flop_model.argtypes = [i, i, i, i]
        L.or_flop_model.restype = d
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class ScfOpts(ctypes.Structure):
    _fields_ = [("inttol", ctypes.c_double), ("scfconv", ctypes.c_double),
                ("scfrelax", ctypes.c_double), ("hfexchange", ctypes.c_double),
                ("lindep", ctypes.c_double), ("temphf", ctypes.c_double),
                ("scfcycles", ctypes.c_int), ("diis", ctypes.c_int),
                ("charge", ctypes.c_int), ("spherical", ctypes.c_int),
                ("direct", ctypes.c_int), ("guess", ctypes.c_int)]


class ScfResult(ctypes.Structure):
    _fields_ = [("e_total", ctypes.c_double), ("e_nuc", ctypes.c_double),
                ("e_kin", ctypes.c_double), ("e_nuc_att", ctypes.c_double),
                ("e_coul", ctypes.c_double), ("e_hfx", ctypes.c_double),
                ("fermi", ctypes.c_double), ("e_grand", ctypes.c_double),
                ("iterations", ctypes.c_int),
                ("converged", ctypes.c_int), ("n_indep", ctypes.c_int),
                ("iocc", ctypes.c_int)]


class Basis:
    """Reference basis layout (READ3, src/pread.f90:584-1482) for a molecule.

    atoms: list of (Z, x, y, z) in bohr.  Arrays are 0-based copies of the
    reference's module BASISSET (src/pmodules.f90:64-81)."""

    def __init__(self, basis_path, atoms):
        L = lib()
        z = np.ascontiguousarray([a[0] for a in atoms], dtype=np.int32)
        xyz = np.ascontiguousarray([[a[1], a[2], a[3]] for a in atoms], dtype=np.float64)
        h = ctypes.c_void_p()
        rc = L.or_basis_build(basis_path.encode(), len(atoms), _ptr(z), _ptr(xyz), ctypes.byref(h))
        if rc != 0:
            raise RuntimeError(f"or_basis_build failed rc={rc}")
        self.h = h
        self.atoms = [(int(a[0]), float(a[1]), float(a[2]), float(a[3])) for a in atoms]
        n = [ctypes.c_int(), ctypes.c_int(), ctypes.c_int()]
        L.or_basis_dims(h, *[ctypes.byref(x) for x in n])
        self.npgs, self.ncgs, self.npshell = [x.value for x in n]
        np_, nc, ns = self.npgs, self.ncgs, self.npshell
        self.pax = np.zeros(np_, np.int32); self.pay = np.zeros(np_, np.int32)
        self.paz = np.zeros(np_, np.int32); self.pgs_atom = np.zeros(np_, np.int32)
        self.pgsx = np.zeros(np_); self.pgsy = np.zeros(np_); self.pgsz = np.zeros(np_)
        self.zt = np.zeros(np_)
        self.cc = np.zeros((nc, np_), order="F")
        self.p_shell = np.zeros((ns, 5, 5, 5), np.int32)
        self.p_shell_ang = np.zeros(ns, np.int32)
        self.cgsx = np.zeros(nc); self.cgsy = np.zeros(nc); self.cgsz = np.zeros(nc)
        L.or_basis_export(h, _ptr(self.pax), _ptr(self.pay), _ptr(self.paz), _ptr(self.pgs_atom),
                          _ptr(self.pgsx), _ptr(self.pgsy), _ptr(self.pgsz), _ptr(self.zt),
                          _ptr(self.cc), _ptr(self.p_shell), _ptr(self.p_shell_ang),
                          _ptr(self.cgsx), _ptr(self.cgsy), _ptr(self.cgsz))

    def __del__(self):
        try:
            lib().or_basis_free(self.h)
        except Exception:
            pass

    def one_electron(self):
        n, m = self.npgs, self.ncgs
        s_p = np.zeros((n, n), order="F")
        s_c = np.zeros((m, m), order="F"); t_c = np.zeros((m, m), order="F")
        n_c = np.zeros((m, m), order="F")
        lib().or_one_electron(self.h, _ptr(s_p), _ptr(s_c), _ptr(t_c), _ptr(n_c))
        return s_p, s_c, t_c, n_c

    def nuclear_repulsion(self):
        return lib().or_nuclear_repulsion(self.h)

    def decontract(self, p_c):
        p_c = np.asfortranarray(p_c, dtype=np.float64)
        p_p = np.zeros((self.npgs, self.npgs), order="F")
        lib().or_decontract(self.h, _ptr(p_c), _ptr(p_p))
        return p_p

    def fourindex_eri(self, p_c, inttol, recompute_kl=False):
        """J_C, K_C for contracted density p_c (src/pfourindexdirect.f90)."""
        s_p = np.zeros((self.npgs, self.npgs), order="F")
        p_p = self.decontract(p_c)
        c_c = np.zeros((self.ncgs, self.ncgs), order="F")
        x_c = np.zeros((self.ncgs, self.ncgs), order="F")
        nreal = ctypes.c_longlong()
        fl = ctypes.c_double()
        nq = lib().or_fourindex_eri(self.h, _ptr(s_p), _ptr(p_p), float(inttol), int(recompute_kl),
                                    _ptr(c_c), _ptr(x_c), ctypes.byref(nreal), ctypes.byref(fl))
        return c_c, x_c, {"composite_quartets": nq, "real_quartets": nreal.value, "flops": fl.value}

    def fourindex_eri_cells(self, p_cells, period, rc1, cel1, cel2, inttol):
        """Periodic J_C, K_C per cell (src/pfourindexdirect.f90:222-750,
        HELIX = 0, axis x).  p_cells: (2 cel1 + 1, ncgs, ncgs), cell Q at
        index Q + cel1, P(Q)[k, l] = P(k in cell 0, l in cell Q)."""
        nc, ncell = self.ncgs, 2 * cel1 + 1
        pc = np.zeros((ncell, nc, nc))
        for c in range(ncell):
            pc[c] = np.asarray(p_cells[c], dtype=np.float64).T    # column-major blocks
        jc = np.zeros((ncell, nc, nc))
        kc = np.zeros((ncell, nc, nc))
        nq = lib().or_fourindex_eri_cells_c(self.h, float(period), int(rc1), int(cel1), int(cel2),
                                            _ptr(pc), float(inttol), _ptr(jc), _ptr(kc))
        return (np.ascontiguousarray(jc.transpose(0, 2, 1)), np.ascontiguousarray(kc.transpose(0, 2, 1)),
                {"composite_quartets": nq})

    def fourindex_eri_range(self, p_c, inttol, ij_begin, ij_end, recompute_kl=True):
        s_p = np.zeros((self.npgs, self.npgs), order="F")
        p_p = self.decontract(p_c)
        c_p = np.zeros((self.npgs, self.npgs), order="F")
        x_p = np.zeros((self.npgs, self.npgs), order="F")
        nreal = ctypes.c_longlong()
        fl = ctypes.c_double()
        nq = lib().or_fourindex_eri_range(self.h, _ptr(s_p), _ptr(p_p), float(inttol),
                                          int(recompute_kl), int(ij_begin), int(ij_end),
                                          _ptr(c_p), _ptr(x_p), ctypes.byref(nreal), ctypes.byref(fl))
        return nq, nreal.value, fl.value

    def fourindex_eri_rows(self, p_c, inttol, rows, recompute_kl=True):
        """Times the reference loop nest on selected bra rows (I*npshell+J):
        returns (quartets, real quartets, setup seconds, row seconds)."""
        s_p = np.zeros((self.npgs, self.npgs), order="F")
        p_p = self.decontract(p_c)
        c_p = np.zeros((self.npgs, self.npgs), order="F")
        x_p = np.zeros((self.npgs, self.npgs), order="F")
        r = np.ascontiguousarray(rows, dtype=np.int64)
        nreal = ctypes.c_longlong()
        fl = ctypes.c_double()
        ts, tr = ctypes.c_double(), ctypes.c_double()
        nq = lib().or_fourindex_eri_rows(self.h, _ptr(s_p), _ptr(p_p), float(inttol), int(recompute_kl),
                                         _ptr(r), len(r), _ptr(c_p), _ptr(x_p), ctypes.byref(nreal),
                                         ctypes.byref(fl), ctypes.byref(ts), ctypes.byref(tr))
        return nq, nreal.value, ts.value, tr.value

    def fourindex_eri_rows_mt(self, p_c, inttol, rows, nthreads, recompute_kl=True):
        """fourindex_eri_rows over nthreads threads: (quartets, real quartets,
        set-up seconds, wall seconds of the rows, C_P, X_P)."""
        p_p = self.decontract(p_c)
        c_p = np.zeros((self.npgs, self.npgs), order="F")
        x_p = np.zeros((self.npgs, self.npgs), order="F")
        r = np.ascontiguousarray(rows, dtype=np.int64)
        nreal = ctypes.c_longlong()
        ts, tr = ctypes.c_double(), ctypes.c_double()
        nq = lib().or_fourindex_eri_rows_mt(self.h, _ptr(p_p), float(inttol), int(recompute_kl), _ptr(r), len(r),
                                            int(nthreads), _ptr(c_p), _ptr(x_p), ctypes.byref(nreal),
                                            ctypes.byref(ts), ctypes.byref(tr))
        return nq, nreal.value, ts.value, tr.value, c_p, x_p

    def eri_tensor(self):
        m = self.ncgs
        s_p = np.zeros((self.npgs, self.npgs), order="F")
        feri = np.zeros((m, m, m, m), order="F")
        lib().or_eri_tensor(self.h, _ptr(s_p), _ptr(feri))
        return feri

    def eri_tensor_cells(self, period, rc1, cel2):
        """T[Q1 + rc1, Q2 + rc1, Q3 + cel2, i, j, k, l] = (i_0 j_Q1 | k_Q3 l_Q3+Q2),
        the lattice integrals of STORE_FOURINDEX_ERI (unscreened)."""
        m, n1, n3 = self.ncgs, 2 * rc1 + 1, 2 * cel2 + 1
        T = np.zeros((n1, n1, n3, m, m, m, m))
        lib().or_eri_tensor_cells(self.h, float(period), int(rc1), int(cel2), _ptr(T))
        return T

    def c2s(self, spherical):
        m = self.ncgs
        c = np.zeros((m, m), order="F")
        lib().or_form_c2s(self.h, int(spherical), _ptr(c))
        return c

    def scf(self, inttol=1e-10, scfconv=1e-8, scfrelax=1.0, hfexchange=1.0, lindep=1e-6,
            temphf=0.0, scfcycles=50, diis=5, charge=0, spherical=False, direct=True,
            guess=None):
        """SELF_CONSISTENT_FIELD (src/pselfconsistentfield.f90:1-162).  Defaults are
        the reference's keyword defaults (src/pread.f90:65-172)."""
        o = ScfOpts(inttol, scfconv, scfrelax, hfexchange, lindep, temphf, scfcycles, diis,
                    charge, int(spherical), int(direct), 0 if guess is None else 1)
        m = self.ncgs
        p = np.zeros((m, m), order="F") if guess is None else np.asfortranarray(guess, dtype=np.float64).copy()
        eps = np.zeros(m)
        co = np.zeros((m, m), order="F")
        trace = np.zeros(2 * scfcycles)
        r = ScfResult()
        rc = lib().or_scf(self.h, ctypes.byref(o), _ptr(p), _ptr(eps), _ptr(co), _ptr(trace),
                          ctypes.byref(r))
        res = {f: getattr(r, f) for f, _ in ScfResult._fields_}
        res.update(P=p, eps=eps, co=co, trace=trace[:2 * r.iterations].reshape(-1, 2), rc=rc)
        return res


def boys_table():
    t = np.zeros((1701, 28), order="F")
    lib().or_boys_table(_ptr(t))
    return t


def boys(t, m):
    return lib().or_boys(float(t), int(m))


def flop_model(la, lb, lc, ld):
    return lib().or_flop_model(la, lb, lc, ld)


def mp2(feri, co, eps, icore, iocc, nall, ivirtcore=0):
    """Closed-shell molecular MP2 (src/pmp2energy.f90:395-444 at one k-point)."""
    nc = co.shape[0]
    feri = np.asfortranarray(feri, dtype=np.float64)
    co = np.asfortranarray(co, dtype=np.float64)
    eps = np.ascontiguousarray(eps, dtype=np.float64)
    s, t = ctypes.c_double(), ctypes.c_double()
    lib().or_mp2(nc, _ptr(feri), _ptr(co), _ptr(eps), icore, iocc, nall, ivirtcore,
                 ctypes.byref(s), ctypes.byref(t))
    return s.value, t.value


def mp2_qp(feri, co, eps, icore, iocc, nall, nqp, ivirtcore=0):
    """MP2 quasi-particle band energies of MP2_ENERGY_ALG1 at one k-point
    (IOPTN(44) = nqp > 0; src/pmp2energy.f90:445-522, normalisation :549-558):
    for every band g from max(iocc - nqp, icore) to min(iocc + nqp, nall -
    ivirtcore) - 1 (0-based),

      QEPSILON(g) = eps(g) + sum_{b,j,a} [2 (ga|jb)^2 - (ga|jb)(ja|gb)] / (e_g + e_j - e_a - e_b)
                           + sum_{b,j,i} [2 (ig|jb)^2 - (ig|jb)(jg|ib)] / (e_g + e_b - e_i - e_j)

    i, j over icore..iocc-1, a, b over iocc..nall-ivirtcore-1; (pq|rs) in
    chemists' notation from the AO tensor feri(I,J,K,L) and the real
    orbitals co (columns).  Returns (band indices, QEPSILON)."""
    feri = np.asarray(feri, dtype=np.float64)
    co = np.asarray(co, dtype=np.float64)
    eps = np.asarray(eps, dtype=np.float64)
    hi = nall - ivirtcore
    g0, g1 = max(iocc - nqp, icore), min(iocc + nqp, hi)
    occ, vir = np.arange(icore, iocc), np.arange(iocc, hi)
    mo = np.arange(icore, hi)
    c = co[:, mo]
    t = np.einsum("pqrs,pi,qa,rj,sb->iajb", feri, c, c, c, c, optimize=True)   # (ia|jb) over icore..hi-1
    o, v = occ - icore, vir - icore
    e = eps[mo]
    out = []
    for g in range(g0, g1):
        x = g - icore
        tg = t[x][np.ix_(v, o, v)]                       # (ga|jb)[a, j, b]
        ts = t[np.ix_(o, v)][:, :, x][:, :, v]           # (ja|gb)[j, a, b]
        d1 = e[x] + e[o][None, :, None] - e[v][:, None, None] - e[v][None, None, :]
        q = np.sum((2.0 * tg ** 2 - tg * np.transpose(ts, (1, 0, 2))) / d1)
        ti = t[np.ix_(o, [x], o, v)][:, 0]               # (ig|jb)[i, j, b]
        tj = np.transpose(ti, (1, 0, 2))                 # (jg|ib)[i, j, b] = ti[j, i, b]
        d2 = e[x] + e[v][None, None, :] - e[o][:, None, None] - e[o][None, :, None]
        q += np.sum((2.0 * ti ** 2 - ti * tj) / d2)
        out.append(eps[g] + q)
    return list(range(g0, g1)), np.array(out)
