"""polymer_amd -- MI355X-native direct two-electron path for polymer.

Host-side mirror of the reference's interface for this path, over the C ABI
of libpolyeri (include/polyeri.h):

    read_basis          READ3            src/pread.f90:584-1482
    Context.fourindex_eri   FOURINDEX_ERI    src/pfourindexdirect.f90:1-814
    Context.self_consistent_field  SELF_CONSISTENT_FIELD + FOCK_BUILD
                                     src/pselfconsistentfield.f90:1-162,
                                     src/pfock.f90:1-506
    Context.mp2_energy  MP2_ENERGY_ALG1  src/pmp2energy.f90:1-571

Matrices are (ncgs, ncgs) numpy arrays in the reference's contracted layout
(ghost functions included, Fortran order).  Every compute call runs on the
GPU; errors raise PolyeriError with the reference's message where it has one
(e.g. 'SCF CONVERGENCE NOT MET').
"""
import ctypes
import os

import numpy as np

from ._lib import (LIB_PATH, Opts, PolyeriError, ScfOpts, ScfResult, Stats, check, lib)

BASIS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "basis")
ANG_BOHR = 0.529177249  # src/pmodules.f90:9
FP64, MIXED = 0, 1

__all__ = ["read_basis", "Context", "PolyeriError", "BASIS_DIR", "ANG_BOHR", "FP64", "MIXED"]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def basis_path(name):
    return name if os.path.sep in name else os.path.join(BASIS_DIR, name)


class RefBasis:
    """The reference's basis arrays for a molecule (module BASISSET)."""

    def __init__(self, atoms, arrays):
        self.atoms = atoms
        self.__dict__.update(arrays)

    @property
    def real_functions(self):
        return np.nonzero(np.any(self.cc != 0.0, axis=1))[0]


def eri_gradient_host(basis, P, hfexchange=1.0, schwarz_tol=0.0):
    """The two-electron gradient of Context.eri_gradient evaluated on the
    host by the same code, without a context or a GPU (slow; for checks)."""
    b = basis
    z = np.ascontiguousarray([a[0] for a in b.atoms], dtype=np.int32)
    xyz = np.ascontiguousarray([[a[1], a[2], a[3]] for a in b.atoms], dtype=np.float64)
    P = np.asfortranarray(P, dtype=np.float64)
    g = np.zeros((len(b.atoms), 3))
    check(lib().polyeri_eri_gradient_host(b.npgs, b.ncgs, _p(b.pax), _p(b.pay), _p(b.paz), _p(b.pgsx), _p(b.pgsy),
                                          _p(b.pgsz), _p(b.zt), _p(b.cc), len(b.atoms), _p(z), _p(xyz),
                                          float(schwarz_tol), _p(P), float(hfexchange), _p(g)))
    return g


def _basis_args(b):
    z = np.ascontiguousarray([a[0] for a in b.atoms], dtype=np.int32)
    xyz = np.ascontiguousarray([[a[1], a[2], a[3]] for a in b.atoms], dtype=np.float64)
    return z, xyz


def energy_weighted_density(co, eps, iocc):
    """W_C = sum_l f_l e_l C_l C_l^T with f = 2 for the iocc occupied
    orbitals (src/pfock.f90:478-497 at zero temperature)."""
    c = np.asarray(co)[:, :iocc]
    return 2.0 * (c * np.asarray(eps)[:iocc]) @ c.T


def one_electron_gradient_host(basis, P, W):
    """The one-electron and nuclear parts of the gradient of
    Context.one_electron_gradient, without a context or a GPU."""
    b = basis
    z, xyz = _basis_args(b)
    P = np.asfortranarray(P, dtype=np.float64)
    W = np.asfortranarray(W, dtype=np.float64)
    g = np.zeros((len(b.atoms), 3))
    check(lib().polyeri_one_electron_gradient_host(b.npgs, b.ncgs, _p(b.pax), _p(b.pay), _p(b.paz), _p(b.pgsx),
                                                   _p(b.pgsy), _p(b.pgsz), _p(b.zt), _p(b.cc), len(b.atoms), _p(z),
                                                   _p(xyz), _p(P), _p(W), _p(g)))
    return g


def restore_fourindex_eri(path, ncgs):
    """Read a molecular .ao file (RESTORE_FOURINDEX_ERI) into an (ncgs,)*4
    Fortran-order tensor; returns (tensor, number of integrals read)."""
    t = np.zeros((ncgs,) * 4, order="F")
    n = ctypes.c_longlong()
    check(lib().polyeri_restore_fourindex_eri(str(path).encode(), int(ncgs), _p(t), ctypes.byref(n)))
    return t, n.value


def restore_fourindex_eri_cells(path, ncgs, cel1, cel2):
    """Read a periodic .ao file into T[Q1 + cel1, Q2 + cel1, Q3 + cel2, i, j,
    k, l]; returns (tensor, number of integrals read)."""
    n1, n3 = 2 * cel1 + 1, 2 * cel2 + 1
    t = np.zeros((n1, n1, n3) + (ncgs,) * 4)
    n = ctypes.c_longlong()
    check(lib().polyeri_restore_fourindex_eri_cells(str(path).encode(), int(ncgs), int(cel1), int(cel2), _p(t),
                                                    ctypes.byref(n)))
    # blocks are column-major FERI(I,J,K,L)
    return np.ascontiguousarray(t.transpose(0, 1, 2, 6, 5, 4, 3)), n.value


def rys_nodes(n, T):
    """Nodes u (ascending) and weights w of the n-point Rys rule at T
    (polyeri_rys_nodes): sum_r w_r u_r^m = F_m(T) for m < 2n."""
    u = np.zeros(n)
    w = np.zeros(n)
    check(lib().polyeri_rys_nodes(int(n), ctypes.c_double(T), _p(u), _p(w)))
    return u, w


def rys_block(la, lb, lc, ld, xyz, ex):
    """(ab|cd) of one primitive quartet of unnormalised Cartesian Gaussians
    (centres xyz (4, 3), exponents ex (4)) from the Rys kernels' per-lane
    code on the host (polyeri_rys_block); shape (ncart(la), .., ncart(ld))."""
    nc = [(l + 1) * (l + 2) // 2 for l in (la, lb, lc, ld)]
    out = np.zeros(nc)
    x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(12)
    e = np.ascontiguousarray(ex, dtype=np.float64).reshape(4)
    check(lib().polyeri_rys_block(la, lb, lc, ld, _p(x), _p(e), _p(out)))
    return out


def read_basis(basis, atoms):
    """READ3 (src/pread.f90:584-1482): atoms = [(Z, x, y, z)] in bohr."""
    L = lib()
    z = np.ascontiguousarray([a[0] for a in atoms], dtype=np.int32)
    xyz = np.ascontiguousarray([[a[1], a[2], a[3]] for a in atoms], dtype=np.float64)
    n = [ctypes.c_int() for _ in range(3)]
    path = basis_path(basis).encode()
    check(L.polyeri_basis_read(path, len(atoms), _p(z), _p(xyz), *[ctypes.byref(x) for x in n],
                               *([None] * 14)))
    npgs, ncgs, nsh = [x.value for x in n]
    ar = dict(npgs=npgs, ncgs=ncgs, npshell=nsh,
              pax=np.zeros(npgs, np.int32), pay=np.zeros(npgs, np.int32), paz=np.zeros(npgs, np.int32),
              pgs_atom=np.zeros(npgs, np.int32), pgsx=np.zeros(npgs), pgsy=np.zeros(npgs),
              pgsz=np.zeros(npgs), zt=np.zeros(npgs), cc=np.zeros((ncgs, npgs), order="F"),
              cgsx=np.zeros(ncgs), cgsy=np.zeros(ncgs), cgsz=np.zeros(ncgs),
              p_shell=np.zeros((nsh, 5, 5, 5), np.int32), p_shell_ang=np.zeros(nsh, np.int32))
    check(L.polyeri_basis_read(path, len(atoms), _p(z), _p(xyz), *[ctypes.byref(x) for x in n],
                               _p(ar["pax"]), _p(ar["pay"]), _p(ar["paz"]), _p(ar["pgs_atom"]),
                               _p(ar["pgsx"]), _p(ar["pgsy"]), _p(ar["pgsz"]), _p(ar["zt"]),
                               _p(ar["cc"]), _p(ar["cgsx"]), _p(ar["cgsy"]), _p(ar["cgsz"]),
                               _p(ar["p_shell"]), _p(ar["p_shell_ang"])))
    return RefBasis([tuple(a) for a in atoms], ar)


class Context:
    """A libpolyeri context for one geometry and basis (device resident)."""

    def __init__(self, basis, schwarz_tol=1e-12, prim_tol=1e-20, precision=FP64, mixed_cut=1e-5,
                 device=0, period=None, cel1=0, cel2=0):
        if not isinstance(basis, RefBasis):
            raise TypeError("Context(basis): basis must come from read_basis()")
        self.basis = basis
        b = basis
        z = np.ascontiguousarray([a[0] for a in b.atoms], dtype=np.int32)
        xyz = np.ascontiguousarray([[a[1], a[2], a[3]] for a in b.atoms], dtype=np.float64)
        o = Opts(schwarz_tol, prim_tol, precision, mixed_cut, device)
        h = ctypes.c_void_p()
        self._keep = (z, xyz)
        if period is None:
            check(lib().polyeri_ctx_create(b.npgs, b.ncgs, _p(b.pax), _p(b.pay), _p(b.paz), _p(b.pgsx),
                                           _p(b.pgsy), _p(b.pgsz), _p(b.zt), _p(b.cc), len(b.atoms),
                                           _p(z), _p(xyz), ctypes.byref(o), ctypes.byref(h)))
        else:
            # the reference's periodic axis is x (src/pfourindexdirect.f90:262-270)
            check(lib().polyeri_ctx_create_periodic(b.npgs, b.ncgs, _p(b.pax), _p(b.pay), _p(b.paz),
                                                    _p(b.pgsx), _p(b.pgsy), _p(b.pgsz), _p(b.zt), _p(b.cc),
                                                    len(b.atoms), _p(z), _p(xyz), float(period), int(cel1),
                                                    int(cel2), ctypes.byref(o), ctypes.byref(h)))
        self.h = h
        self.ncgs = b.ncgs
        self.period, self.cel1, self.cel2 = period, cel1, cel2

    def close(self):
        if getattr(self, "h", None):
            lib().polyeri_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _mat(self):
        return np.zeros((self.ncgs, self.ncgs), order="F")

    def fourindex_eri(self, p_c, dens_tol=0.0):
        """J (C_C) and K (X_C) from the contracted density P_C; dens_tol > 0
        skips quartets whose Schwarz bound times the largest |P| they read
        is below dens_tol."""
        p = np.asfortranarray(p_c, dtype=np.float64)
        j, k = self._mat(), self._mat()
        if dens_tol > 0.0:
            check(lib().polyeri_jk_build_screened(self.h, _p(p), float(dens_tol), _p(j), _p(k)))
        else:
            check(lib().polyeri_jk_build(self.h, _p(p), _p(j), _p(k)))
        return j, k

    def fourindex_eri_cells(self, p_cells):
        """Periodic J and K cells (polyeri_jk_build_cells).  p_cells:
        (2 cel1 + 1, ncgs, ncgs), cell Q at index Q + cel1, P(Q)[k, l] =
        P(k in cell 0, l in cell Q); returns J, K in the same layout."""
        if self.period is None:
            raise PolyeriError(-1, "fourindex_eri_cells needs a periodic context")
        ncell, n = 2 * self.cel1 + 1, self.ncgs
        p = np.zeros((ncell, n, n))
        for c in range(ncell):
            p[c] = np.asarray(p_cells[c], dtype=np.float64).T        # column-major blocks
        j = np.zeros((ncell, n, n))
        k = np.zeros((ncell, n, n))
        check(lib().polyeri_jk_build_cells(self.h, _p(p), _p(j), _p(k)))
        return np.ascontiguousarray(j.transpose(0, 2, 1)), np.ascontiguousarray(k.transpose(0, 2, 1))

    def one_electron(self):
        s, t, v = self._mat(), self._mat(), self._mat()
        e = ctypes.c_double()
        check(lib().polyeri_one_electron(self.h, _p(s), _p(t), _p(v), ctypes.byref(e)))
        return s, t, v, e.value

    def eri_tensor(self):
        n = self.ncgs
        f = np.zeros((n, n, n, n), order="F")
        check(lib().polyeri_eri_tensor(self.h, _p(f)))
        return f

    def self_consistent_field(self, scfconv=1e-8, scfrelax=1.0, hfexchange=1.0, lindep=1e-6,
                              temphf=0.0, scfcycles=50, diis=5, charge=0, spherical=False,
                              guess=None, incremental=0, dens_tol=0.0, solver=0, diis_mode=0):
        """Defaults are the reference's keyword defaults (src/pread.f90:65-172).
        incremental = n > 0 builds J/K from the density difference with a
        full build every n iterations; dens_tol > 0 screens quartets by their
        Schwarz bound times the density (difference) they touch; solver = 1
        forms the zero-temperature density by purification instead of
        diagonalisation while the gap is open; diis_mode = 1 uses commutator
        DIIS on the Fock matrices every iteration."""
        o = ScfOpts(scfconv, scfrelax, hfexchange, lindep, temphf, scfcycles, diis, charge,
                    int(spherical), 0 if guess is None else 1, int(incremental), float(dens_tol),
                    int(solver), int(diis_mode))
        p = self._mat() if guess is None else np.asfortranarray(guess, dtype=np.float64).copy()
        eps = np.zeros(self.ncgs)
        co = self._mat()
        trace = np.zeros(2 * scfcycles)
        r = ScfResult()
        rc = lib().polyeri_scf(self.h, ctypes.byref(o), _p(p), _p(eps), _p(co), _p(trace),
                               ctypes.byref(r))
        res = {f: getattr(r, f) for f, _ in ScfResult._fields_}
        res.update(P=p, eps=eps, co=co, trace=trace[:2 * max(r.iterations, 0)].reshape(-1, 2))
        if rc != 0:
            check(rc)
        return res

    @property
    def reduced_cel1(self):
        """REDUCED_CEL1X of a periodic context."""
        r = lib().polyeri_reduced_cel1(self.h)
        check(min(r, 0))
        return r

    def eri_tensor_cells(self):
        """T[Q1 + rc1, Q2 + rc1, Q3 + cel2, i, j, k, l] = (i_0 j_Q1 | k_Q3 l_Q3+Q2)."""
        rc1 = self.reduced_cel1
        n1, n3, n = 2 * rc1 + 1, 2 * self.cel2 + 1, self.ncgs
        t = np.zeros((n1, n1, n3) + (n,) * 4)
        check(lib().polyeri_eri_tensor_cells(self.h, _p(t)))
        return np.ascontiguousarray(t.transpose(0, 1, 2, 6, 5, 4, 3))

    def store_fourindex_eri(self, path, inttol=0.0):
        """Write the contracted ERIs above inttol in the reference's .ao record
        format (STORE_FOURINDEX_ERI); returns the number stored."""
        n = ctypes.c_longlong()
        check(lib().polyeri_store_fourindex_eri(self.h, str(path).encode(), float(inttol), ctypes.byref(n)))
        return n.value

    def self_consistent_field_periodic(self, kpoints=0, scfconv=1e-8, scfrelax=1.0, hfexchange=1.0,
                                       lindep=1e-6, temphf=0.0, scfcycles=50, diis=5, charge=0,
                                       spherical=False, guess=None):
        """Periodic SCF with 2 * kpoints k-points (KPOINTS, src/pfock.f90:186;
        0 = Gamma only).  Returns the result fields plus P (cells, as
        fourindex_eri_cells), eps (nk, ncgs), co (nk, ncgs, ncgs complex,
        columns = bands) and nall (IALL per k)."""
        if self.period is None:
            raise PolyeriError(-1, "self_consistent_field_periodic needs a periodic context")
        o = ScfOpts(scfconv, scfrelax, hfexchange, lindep, temphf, scfcycles, diis, charge,
                    int(spherical), 0 if guess is None else 1, 0, 0.0)
        ncell, n = 2 * self.cel1 + 1, self.ncgs
        nk = 2 * kpoints if kpoints > 0 else 1
        p = np.zeros((ncell, n, n))
        if guess is not None:
            for c in range(ncell):
                p[c] = np.asarray(guess[c], dtype=np.float64).T
        eps = np.zeros((nk, n))
        trace = np.zeros(2 * scfcycles)
        co = np.zeros((nk, n, n), dtype=np.complex128)      # [k][band][function]: column-major per k
        nall = np.zeros(nk, dtype=np.int32)
        r = ScfResult()
        rc = lib().polyeri_scf_periodic(self.h, ctypes.byref(o), int(kpoints), _p(p), _p(eps), _p(trace),
                                        ctypes.byref(r), co.ctypes.data_as(ctypes.c_void_p),
                                        nall.ctypes.data_as(ctypes.c_void_p))
        res = {f: getattr(r, f) for f, _ in ScfResult._fields_}
        res.update(P=np.ascontiguousarray(p.transpose(0, 2, 1)), eps=eps,
                   co=np.ascontiguousarray(co.transpose(0, 2, 1)), nall=nall.astype(int),
                   trace=trace[:2 * max(r.iterations, 0)].reshape(-1, 2))
        if rc != 0:
            check(rc)
        return res

    def _periodic_mp2_args(self, kpoints, co, eps, nall):
        if self.period is None:
            raise PolyeriError(-1, "the k-point MP2 needs a periodic context")
        nk = 2 * kpoints if kpoints > 0 else 1
        n = self.ncgs
        co = np.asarray(co, dtype=np.complex128)
        if co.shape != (nk, n, n):
            raise PolyeriError(-1, f"co must have shape {(nk, n, n)}")
        cf = np.ascontiguousarray(co.transpose(0, 2, 1))      # column-major per k
        eps = np.ascontiguousarray(eps, dtype=np.float64).reshape(nk, n)
        na = np.ascontiguousarray(nall, dtype=np.int32).reshape(nk)
        return cf, eps, na

    def mp2_energy_periodic(self, kpoints, co, eps, nall, icore, iocc, ivirtcore=0, transform="f64",
                            timing=False, modulo=1):
        """k-point E(2) singlet and triplet per unit cell (MP2_ENERGY_ALG1 of a
        periodic run) with the lattice integrals and all four quarter
        transforms on the device.  co: (nk, ncgs, ncgs) complex, columns =
        bands; eps: (nk, ncgs); nall: IALL per k.  transform "f64" (FP64 MFMA)
        or "bf16x3" (BF16-split MFMA, FP32 accumulation).  modulo: MODULO
        (IOPTN(99)), every modulo-th k-point of the SCF grid."""
        cf, eps, na = self._periodic_mp2_args(kpoints, co, eps, nall)
        mode = {"f64": 1, "bf16x3": 2}[transform]
        s, t, sec = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(lib().polyeri_mp2_energy_periodic_mod(self.h, int(kpoints), int(modulo),
                                                    cf.ctypes.data_as(ctypes.c_void_p), _p(eps),
                                                    na.ctypes.data_as(ctypes.c_void_p), icore, iocc, ivirtcore, mode,
                                                    ctypes.byref(s), ctypes.byref(t), ctypes.byref(sec)))
        if timing:
            return s.value, t.value, sec.value
        return s.value, t.value

    def mp2_quasiparticle_periodic(self, kpoints, co, eps, nall, icore, iocc, nqp, ivirtcore=0, transform="f64",
                                   modulo=1):
        """k-point E(2) singlet, triplet and the MP2 quasi-particle bands
        QEPSILON (IOPTN(44) = nqp, with MODULO = modulo): returns
        (e2s, e2t, qeps) with qeps of eps's shape (nk, ncgs); bands and
        k-points outside the corrected range copy eps."""
        cf, eps, na = self._periodic_mp2_args(kpoints, co, eps, nall)
        mode = {"f64": 1, "bf16x3": 2}[transform]
        q = np.zeros_like(eps)
        s, t, sec = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(lib().polyeri_mp2_energy_periodic_qp(self.h, int(kpoints), int(modulo),
                                                   cf.ctypes.data_as(ctypes.c_void_p), _p(eps),
                                                   na.ctypes.data_as(ctypes.c_void_p), icore, iocc, ivirtcore,
                                                   int(nqp), mode, ctypes.byref(s), ctypes.byref(t), _p(q),
                                                   ctypes.byref(sec)))
        return s.value, t.value, q

    def mp2_quasiparticle(self, co, eps, icore, iocc, nall, nqp, ivirtcore=0, transform="f64"):
        """E(2) singlet, triplet and the MP2 quasi-particle band energies
        QEPSILON (IOPTN(44) = nqp): returns (e2s, e2t, qeps) with qeps of
        length ncgs (bands outside the corrected range copy eps)."""
        co = np.asfortranarray(co, dtype=np.float64)
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        q = np.zeros(self.ncgs)
        s, t = ctypes.c_double(), ctypes.c_double()
        mode = {"rocblas": 0, "f64": 1, "bf16x3": 2}[transform]
        check(lib().polyeri_mp2_energy_qp(self.h, _p(co), _p(eps), icore, iocc, nall, ivirtcore, int(nqp), mode,
                                          ctypes.byref(s), ctypes.byref(t), _p(q)))
        return s.value, t.value, q

    def mo_integrals(self, co, lo, hi):
        """Full MO integral set (pq|rs) over MOs lo .. hi-1 (columns of co),
        INTEGRAL_TRANSFORM_DIRECTACCESS on the device: an (m, m, m, m) array."""
        co = np.asfortranarray(co, dtype=np.float64)
        m = hi - lo
        g = np.zeros((m, m, m, m))
        check(lib().polyeri_mo_integrals(self.h, _p(co), int(lo), int(hi), _p(g)))
        return g

    def ccsd(self, co, eps, icore, iocc, nall, ivirtcore=0, conv=1e-10, maxiter=100, diis=8, method="CCSD"):
        """Spin-restricted coupled cluster (COUPLED_CLUSTER_ENERGY, COPTN(70) =
        method: CCSD, CCD, LCCD, LCCSD or QCISD) with E(2) and E(3) on the
        device: returns a dict with e2, e3, e_singles, e_doubles, e_corr,
        residual, iterations, converged and timings."""
        from ._lib import CcResult
        co = np.asfortranarray(co, dtype=np.float64)
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        r = CcResult()
        check(lib().polyeri_cc(self.h, str(method).encode(), _p(co), _p(eps), int(icore), int(iocc), int(nall),
                               int(ivirtcore), float(conv), int(maxiter), int(diis), ctypes.byref(r)))
        return {k: getattr(r, k) for k, _ in CcResult._fields_}

    def mp2_energy(self, co, eps, icore, iocc, nall, ivirtcore=0, transform=None):
        """E(2) singlet and triplet parts.  transform: None (the library's
        default, its FP64 MFMA kernel), "f64" (the same), "rocblas" (rocBLAS
        dgemm) or "bf16x3" (own BF16-split MFMA kernel)."""
        co = np.asfortranarray(co, dtype=np.float64)
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        s, t = ctypes.c_double(), ctypes.c_double()
        if transform is None:
            check(lib().polyeri_mp2_energy(self.h, _p(co), _p(eps), icore, iocc, nall, ivirtcore,
                                           ctypes.byref(s), ctypes.byref(t)))
        else:
            mode = {"rocblas": 0, "f64": 1, "bf16x3": 2}[transform]
            check(lib().polyeri_mp2_energy_ex(self.h, _p(co), _p(eps), icore, iocc, nall, ivirtcore, mode,
                                              ctypes.byref(s), ctypes.byref(t)))
        return s.value, t.value

    # device-resident buffers (used by bench.py: inputs live in HBM)
    def device_alloc(self, nbytes):
        ptr = ctypes.c_void_p()
        check(lib().polyeri_device_alloc(self.h, nbytes, ctypes.byref(ptr)))
        return ptr

    def device_free(self, ptr):
        check(lib().polyeri_device_free(self.h, ptr))

    def to_device(self, dptr, arr):
        a = np.asfortranarray(arr, dtype=np.float64)
        check(lib().polyeri_memcpy(self.h, dptr, _p(a), a.nbytes, 1))

    def from_device(self, dptr, shape):
        a = np.zeros(shape, order="F")
        check(lib().polyeri_memcpy(self.h, _p(a), dptr, a.nbytes, 0))
        return a

    def to_device_raw(self, dptr, arr):
        """Copy the bytes of a C-contiguous array to device memory."""
        a = np.ascontiguousarray(arr)
        check(lib().polyeri_memcpy(self.h, dptr, _p(a), a.nbytes, 1))

    def from_device_raw(self, dptr, out):
        """Fill a C-contiguous array from device memory."""
        assert out.flags["C_CONTIGUOUS"]
        check(lib().polyeri_memcpy(self.h, _p(out), dptr, out.nbytes, 0))
        return out

    def fock_build_device(self, d_p, d_f, d_j, d_k, hfexchange=1.0):
        check(lib().polyeri_fock_build_device(self.h, d_p, hfexchange, d_f, d_j, d_k))

    def device_copy(self, dst, src, nbytes):
        check(lib().polyeri_device_copy(self.h, dst, src, nbytes))

    def sync(self):
        check(lib().polyeri_sync(self.h))

    def set_profiling(self, on=True):
        check(lib().polyeri_set_profiling(self.h, int(on)))

    def set_graph(self, on):
        """J/K builds replayed from captured HIP graphs (default) or launched
        eagerly."""
        check(lib().polyeri_set_graph(self.h, int(bool(on))))

    def set_general_contraction(self, on):
        """Evaluate the (ls|l's) classes over generally contracted s-shell
        groups (default where the basis has them) or shell pair by shell
        pair."""
        check(lib().polyeri_set_general_contraction(self.h, int(bool(on))))

    def load_fourindex_eri(self, path):
        """DIRECT .F.: append the records of an .ao file (store_fourindex_eri)
        and switch this context's J/K builds, SCFs and MP2 tensor to them.
        Returns the number of records read."""
        n = ctypes.c_longlong()
        check(lib().polyeri_load_fourindex_eri(self.h, str(path).encode(), ctypes.byref(n)))
        return n.value

    def set_direct(self, direct):
        """True: evaluate the integrals in every J/K build (default); False:
        use the loaded .ao records."""
        check(lib().polyeri_set_direct(self.h, int(bool(direct))))

    def eri_gradient(self, P, hfexchange=1.0):
        """Two-electron part of the nuclear gradient at fixed density P_C
        ((ncgs, ncgs), the SCF density): dE2/dR as an (natom, 3) array in
        Eh/bohr (FOURINDEX_ERI_GRADIENT)."""
        P = np.asfortranarray(P, dtype=np.float64)
        g = np.zeros((len(self.basis.atoms), 3))
        check(lib().polyeri_eri_gradient(self.h, _p(P), float(hfexchange), _p(g)))
        return g

    def one_electron_gradient(self, P, W):
        """One-electron and nuclear parts of the nuclear gradient, (natom, 3)
        in Eh/bohr: sum P (dT + dV) - W dS, the attraction operator's
        derivative at each nucleus and the nuclear repulsion."""
        P = np.asfortranarray(P, dtype=np.float64)
        W = np.asfortranarray(W, dtype=np.float64)
        g = np.zeros((len(self.basis.atoms), 3))
        check(lib().polyeri_one_electron_gradient(self.h, _p(P), _p(W), _p(g)))
        return g

    def scf_gradient(self, result, hfexchange=1.0):
        """dE_SCF/dR, (natom, 3) in Eh/bohr, of a converged zero-temperature
        self_consistent_field result: the one-electron part with the
        energy-weighted density of its orbitals plus the two-electron part
        (the reference's sequence at src/psinglepoint.f90:358-383)."""
        W = energy_weighted_density(result["co"], result["eps"], result["iocc"])
        return self.one_electron_gradient(result["P"], W) + self.eri_gradient(result["P"], hfexchange)

    def launch_configs(self):
        """Per work list: the 9-int launch record of polyeri_get_launch_config."""
        n = self.stats()["nclass"]
        out = []
        for i in range(n):
            c = (ctypes.c_int * 9)()
            check(lib().polyeri_get_launch_config(self.h, i, c))
            out.append(list(c))
        return out

    def set_launch_configs(self, cfgs):
        """Restore records from launch_configs() (same geometry and basis)."""
        for i, c in enumerate(cfgs):
            a = (ctypes.c_int * 9)(*[int(x) for x in c])
            check(lib().polyeri_set_launch_config(self.h, i, a))

    def autotune(self, reps=2):
        """Time every work list with each candidate launch shape and keep the
        fastest (results do not change)."""
        check(lib().polyeri_autotune(self.h, int(reps)))

    def stats(self):
        st = Stats()
        check(lib().polyeri_get_stats(self.h, ctypes.byref(st)))
        d = {f: getattr(st, f) for f, _ in Stats._fields_
             if f not in ("last_class_ms", "class_prim_quartets", "class_flops", "class_l",
                          "class_quartets", "class_prec", "class_kind", "class_block", "class_grid")}
        n = st.nclass
        d["classes"] = [dict(cls=f"{st.class_l[i]:04d}", ms=st.last_class_ms[i],
                             prim_quartets=st.class_prim_quartets[i], flops=st.class_flops[i],
                             quartets=st.class_quartets[i], prec="fp32" if st.class_prec[i] else "fp64",
                             kernel={2: "wave", 3: "seg", 4: "gc", 5: "rys"}.get(st.class_kind[i], "thread"), block=st.class_block[i],
                             grid=st.class_grid[i])
                        for i in range(n)]
        return d
