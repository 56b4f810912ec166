"""Learned SCF initial guess for water clusters.

A small network predicts the converged AO density matrix block by block:
one block per atom pair (i, j) within a cutoff, P_ij = P[functions of i,
functions of j] in the reference's contracted layout.  Blocks rotate with
the molecules, so the network predicts them in molecular frames: every water
has a frame F (columns: H-O-H bisector, in-plane perpendicular, plane
normal), each atom's functions transform to that frame by a matrix T_atom
(block diagonal over shells: 1 for s, F for p, the induced 6 x 6 map for
Cartesian d with the reference's 1/sqrt(3) scaling of xx, yy, zz), and

    P_ij = T_i  Q_ij  T_j^T,      Q_ij = network(features of the pair),

with features invariant under rotation and translation of the cluster: the
position of j in i's frame, the orientation of j's molecule relative to i's,
radial-basis expansions of the distance, and radial neighbour densities of O
and H around both atoms.  One network per atom-type pair (O-O, O-H, H-H);
the block of (H, O) is the transpose of the (O, H) prediction.  The
predicted matrix is scaled to the electron count (sum P.S = N).

Training and inference run on the device through the library's MLP
(include/polyeri.h polyeri_mlp_*: BF16 MFMA GEMMs, FP32 accumulation and
master weights, Adam); features, frames and block rotations are host-side
numpy.  There is no counterpart in the reference, whose guess is the
diagonal GUESS_DENSITYMATRIX (src/pdensitymatrix.f90:83-127); that guess
and the core-Hamiltonian guess stay available and the converged energy does
not depend on the guess.
"""
import ctypes
import math

import numpy as np

from . import Context, read_basis
from ._lib import check, lib

ANG = 1.0 / 0.529177249          # bohr per angstrom
CUTOFF = 6.0 * ANG                # pairs beyond this distance get a zero block
NFEAT = 64                        # feature columns computed; a model uses the first nfeat of them
NFEAT_GEOM = 48                   # geometric features only (columns 0..47)
TYPES = ("OO", "OH", "HH")
QO, QH = -0.834, 0.417            # point charges of the other molecules for the field features (TIP3P)


# ---------------------------------------------------------------------------
# geometry: molecules, frames, invariant features
# ---------------------------------------------------------------------------
def molecules(atoms):
    """Water molecules as consecutive (O, H, H) triples (the layout of
    geometry.water_cluster)."""
    if len(atoms) % 3:
        raise ValueError("learned guess: atoms must be (O, H, H) triples")
    for k in range(0, len(atoms), 3):
        if (atoms[k][0], atoms[k + 1][0], atoms[k + 2][0]) != (8, 1, 1):
            raise ValueError("learned guess: atoms must be (O, H, H) triples")
    return [(k, k + 1, k + 2) for k in range(0, len(atoms), 3)]


def frames(atoms):
    """One right-handed orthonormal frame per molecule (3 x 3, columns)."""
    xyz = np.array([a[1:] for a in atoms])
    out = []
    for o, h1, h2 in molecules(atoms):
        a, b = xyz[h1] - xyz[o], xyz[h2] - xyz[o]
        ex = a + b
        ex /= np.linalg.norm(ex)
        ez = np.cross(a, b)
        ez /= np.linalg.norm(ez)
        ey = np.cross(ez, ex)
        out.append(np.stack([ex, ey, ez], axis=1))
    return out


def _env(xyz, z):
    """Radial neighbour densities per atom: for O and H neighbours,
    sum exp(-((r - mu) / 0.5 A)^2), mu = 1..6 A (12 numbers)."""
    n = len(z)
    d = np.linalg.norm(xyz[:, None, :] - xyz[None, :, :], axis=2) / ANG
    mus = np.arange(1.0, 7.0)
    out = np.zeros((n, 12))
    for s, zz in enumerate((8, 1)):
        mask = (z[None, :] == zz) & ~np.eye(n, dtype=bool)
        g = np.exp(-((d[:, :, None] - mus[None, None, :]) / 0.5) ** 2) * mask[:, :, None]
        out[:, 6 * s:6 * s + 6] = g.sum(axis=1)
    return out


def pair_list(atoms):
    """Canonical pairs (i, j, type): O before H for mixed pairs, i <= j
    otherwise, distance below CUTOFF (diagonal pairs included)."""
    xyz = np.array([a[1:] for a in atoms])
    z = np.array([a[0] for a in atoms])
    d = np.linalg.norm(xyz[:, None, :] - xyz[None, :, :], axis=2)
    names = {(8, 8): "OO", (8, 1): "OH", (1, 1): "HH"}
    out = []
    for i, j in zip(*np.nonzero(np.triu(d < CUTOFF))):
        i, j = int(i), int(j)
        a, b = (i, j) if not (z[i] == 1 and z[j] == 8) else (j, i)
        out.append((a, b, names[(int(z[a]), int(z[b]))]))
    return out


def _fields(xyz, z):
    """Per atom: the electric field (bohr units) of the point charges of the
    other molecules, in the atom's molecular frame, and their potential."""
    n = len(z)
    q = np.where(z == 8, QO, QH)
    mol = np.arange(n) // 3
    d = xyz[:, None, :] - xyz[None, :, :]                   # r_a - r_b
    r = np.linalg.norm(d, axis=2)
    other = mol[:, None] != mol[None, :]
    inv = np.where(other, 1.0 / np.where(other, r, 1.0), 0.0)
    E = np.einsum("ab,abx->ax", q[None, :] * inv ** 3, d)
    phi = (q[None, :] * inv).sum(axis=1)
    fr = np.array(frames([(int(zz), *x) for zz, x in zip(z, xyz)]))
    El = np.einsum("axk,ax->ak", fr[mol], E)               # frame components
    return El, phi


def pair_features(atoms, pairs):
    """Invariant features of every pair (one row each, NFEAT columns): same
    molecule, same atom, the position of j in i's frame, the distance and
    its radial-basis expansion, the relative orientation of the two frames,
    the neighbour densities of both atoms (columns 0..46), then the field of
    the other molecules' point charges at i and at j in their own frames and
    the potential at both (columns 48..55)."""
    xyz = np.array([a[1:] for a in atoms])
    z = np.array([a[0] for a in atoms])
    fr = np.array(frames(atoms))                       # (molecules, 3, 3)
    env = _env(xyz, z)
    rbf = np.linspace(0.0, 6.0, 8)
    X = np.zeros((len(pairs), NFEAT), dtype=np.float32)
    if not pairs:
        return X
    i = np.array([p[0] for p in pairs])
    j = np.array([p[1] for p in pairs])
    Fi, Fj = fr[i // 3], fr[j // 3]
    d = xyz[j] - xyz[i]
    dl = np.einsum("pxa,px->pa", Fi, d) / ANG
    dist = np.linalg.norm(d, axis=1) / ANG
    X[:, 0] = i // 3 == j // 3
    X[:, 1] = i == j
    X[:, 2:5] = dl / 5.0
    X[:, 5] = dist / 5.0
    X[:, 6:14] = np.exp(-((dist[:, None] - rbf[None, :]) / 0.75) ** 2)
    X[:, 14:23] = np.einsum("pxa,pxb->pab", Fi, Fj).reshape(-1, 9)
    X[:, 23:35] = env[i]
    X[:, 35:47] = env[j]
    El, phi = _fields(xyz, z)
    X[:, 48:51] = 10.0 * El[i]
    X[:, 51:54] = 10.0 * El[j]
    X[:, 54] = phi[i]
    X[:, 55] = phi[j]
    return X


# ---------------------------------------------------------------------------
# basis: functions per atom and their frame transformation
# ---------------------------------------------------------------------------
def _dfact(n):
    return 1 if n <= 0 else n * _dfact(n - 2)


def atom_functions(basis):
    """Per atom: reference indices of its real functions and their Cartesian
    exponents, in the reference order."""
    real = np.asarray(basis.real_functions, dtype=np.int64)
    cen = np.stack([basis.cgsx, basis.cgsy, basis.cgsz], axis=1)[real]
    axyz = np.array([at[1:] for at in basis.atoms], dtype=np.float64)
    owner = np.argmin(np.linalg.norm(cen[:, None, :] - axyz[None, :, :], axis=2), axis=1)
    first = np.argmax(np.asarray(basis.cc)[real] != 0, axis=1)      # first primitive of each function
    lx, ly, lz = (np.asarray(a)[first] for a in (basis.pax, basis.pay, basis.paz))
    funcs = [[] for _ in basis.atoms]
    for c, a, x, y, z in zip(real.tolist(), owner.tolist(), lx.tolist(), ly.tolist(), lz.tolist()):
        funcs[a].append((c, (x, y, z)))
    return funcs


def _scale(l):
    return 1.0 / math.sqrt(_dfact(2 * l[0] - 1) * _dfact(2 * l[1] - 1) * _dfact(2 * l[2] - 1))


def frame_transform(F, fl):
    """T with phi_local = phi_global T for one atom's functions fl (list of
    (index, (lx, ly, lz))); shells are runs of equal L (s, p, or the six
    Cartesian d), the mapping follows from u = F^T x."""
    n = len(fl)
    T = np.zeros((n, n))
    k = 0
    while k < n:
        L = sum(fl[k][1])
        m = (L + 1) * (L + 2) // 2
        comps = [fl[k + q][1] for q in range(m)]
        if L == 0:
            T[k, k] = 1.0
        else:
            pos = {c: q for q, c in enumerate(comps)}
            for ql, cl in enumerate(comps):
                # local monomial: product over its axis factors u_a = sum_x F[x, a] x
                axes = [a for a in range(3) for _ in range(cl[a])]
                poly = {(0, 0, 0): 1.0}
                for a in axes:
                    nxt = {}
                    for e, v in poly.items():
                        for x in range(3):
                            e2 = list(e)
                            e2[x] += 1
                            e2 = tuple(e2)
                            nxt[e2] = nxt.get(e2, 0.0) + v * F[x, a]
                    poly = nxt
                for e, v in poly.items():
                    qg = pos[e]
                    T[k + qg, k + ql] += _scale(cl) / _scale(e) * v
        k += m
    return T


# ---------------------------------------------------------------------------
# the device MLP
# ---------------------------------------------------------------------------
class MLP:
    """Handle on a polyeri_mlp (device weights, BF16 MFMA GEMMs)."""

    def __init__(self, dims, seed=0):
        self.dims = [int(d) for d in dims]
        arr = (ctypes.c_int * len(dims))(*self.dims)
        h = ctypes.c_void_p()
        check(lib().polyeri_mlp_create(len(dims) - 1, arr, ctypes.c_ulonglong(seed), ctypes.byref(h)))
        self.h = h
        self.nparam = lib().polyeri_mlp_nparam(self.h)
        self._ctx = None

    def __del__(self):
        if getattr(self, "h", None):
            lib().polyeri_mlp_destroy(self.h)
            self.h = None

    def params(self):
        p = np.zeros(self.nparam, np.float32)
        check(lib().polyeri_mlp_params(self.h, p.ctypes.data_as(ctypes.c_void_p), 0))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        check(lib().polyeri_mlp_params(self.h, p.ctypes.data_as(ctypes.c_void_p), 1))

    def grads(self):
        g = np.zeros(self.nparam, np.float32)
        check(lib().polyeri_mlp_grads(self.h, g.ctypes.data_as(ctypes.c_void_p)))
        return g


class DeviceArrays:
    """Device copies of float32 host arrays through a library context."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        d = self.ctx.device_alloc(max(a.nbytes, 4))
        self.ctx.to_device_raw(d, a)
        self.ptrs.append(d)
        return d

    def empty(self, nbytes):
        d = self.ctx.device_alloc(max(nbytes, 4))
        self.ptrs.append(d)
        return d

    def free(self):
        for d in self.ptrs:
            self.ctx.device_free(d)
        self.ptrs = []


def mlp_forward(mlp, ctx, X):
    X = np.ascontiguousarray(X, dtype=np.float32)
    dev = DeviceArrays(ctx)
    try:
        dx = dev.put(X)
        dy = dev.empty(X.shape[0] * mlp.dims[-1] * 4)
        check(lib().polyeri_mlp_forward(mlp.h, dx, X.shape[0], dy))
        y = np.zeros((X.shape[0], mlp.dims[-1]), np.float32)
        ctx.from_device_raw(dy, y)
        return y
    finally:
        dev.free()


def mlp_train_step(mlp, ctx, X, T, W, lr, update=True):
    dev = DeviceArrays(ctx)
    try:
        dx, dt, dw = dev.put(X), dev.put(T), dev.put(W)
        loss = ctypes.c_float()
        check(lib().polyeri_mlp_train_step(mlp.h, dx, dt, dw, X.shape[0], ctypes.c_float(lr), int(update),
                                           ctypes.byref(loss)))
        return loss.value
    finally:
        dev.free()


# ---------------------------------------------------------------------------
# the guess model
# ---------------------------------------------------------------------------
def _block_dims(basis_name):
    """Functions per O and per H atom of a basis (from a single water)."""
    from .geometry import water
    b = read_basis(basis_name, water())
    fl = atom_functions(b)
    return {"O": len(fl[0]), "H": len(fl[1])}


class GuessModel:
    HIDDEN = (128, 128)

    def __init__(self, basis_name="6-31Gss", seed=0, nfeat=NFEAT, hidden=None):
        self.basis_name = basis_name
        self.nfeat = int(nfeat)
        self.hidden = tuple(int(h) for h in (hidden or self.HIDDEN))
        self.nf = _block_dims(basis_name)
        self.out = {"OO": self.nf["O"] ** 2, "OH": self.nf["O"] * self.nf["H"], "HH": self.nf["H"] ** 2}
        self.nets = {t: MLP([self.nfeat, *self.hidden, self.out[t]], seed=seed + k) for k, t in enumerate(TYPES)}
        self.xmu = {t: np.zeros(self.nfeat, np.float32) for t in TYPES}
        self.xsd = {t: np.ones(self.nfeat, np.float32) for t in TYPES}
        self.ymu = {t: np.zeros(self.out[t], np.float32) for t in TYPES}
        self.ysd = {t: np.ones(self.out[t], np.float32) for t in TYPES}

    # -- data -------------------------------------------------------------
    def samples(self, atoms, basis, P=None):
        """Features per type, and (with P) the frame-local target blocks."""
        pairs = pair_list(atoms)
        X = pair_features(atoms, pairs)[:, :self.nfeat]
        fl = atom_functions(basis)
        fr = frames(atoms)
        Tm = [frame_transform(fr[a // 3], fl[a]) for a in range(len(atoms))]
        out = {t: ([], [], []) for t in TYPES}
        for r, (i, j, t) in enumerate(pairs):
            out[t][0].append(X[r])
            out[t][2].append((i, j))
            if P is not None:
                ii = [c for c, _ in fl[i]]
                jj = [c for c, _ in fl[j]]
                blk = P[np.ix_(ii, jj)]
                q = np.linalg.solve(Tm[i], np.linalg.solve(Tm[j], blk.T).T)
                out[t][1].append(q.ravel().astype(np.float32))
        return out, fl, Tm

    # -- inference --------------------------------------------------------
    def predict(self, ctx, basis, atoms, S=None):
        """Guess density (ncgs x ncgs, reference layout, Fortran order).
        Per type, the blocks of all pairs are rotated back and scattered in
        batched array operations."""
        pairs = pair_list(atoms)
        X = pair_features(atoms, pairs)[:, :self.nfeat]
        fl = atom_functions(basis)
        fr = frames(atoms)
        Tm = [frame_transform(fr[a // 3], fl[a]) for a in range(len(atoms))]
        idx = [np.array([c for c, _ in f], dtype=np.int64) for f in fl]
        P = np.zeros((basis.ncgs, basis.ncgs))
        for t in TYPES:
            rows = [r for r, p in enumerate(pairs) if p[2] == t]
            if not rows:
                continue
            Xn = (X[rows] - self.xmu[t]) / self.xsd[t]
            Y = mlp_forward(self.nets[t], ctx, Xn) * self.ysd[t] + self.ymu[t]
            i = np.array([pairs[r][0] for r in rows])
            j = np.array([pairs[r][1] for r in rows])
            ni, nj = len(fl[i[0]]), len(fl[j[0]])
            Ti = np.stack([Tm[a] for a in i])
            Tj = np.stack([Tm[a] for a in j])
            blk = Ti @ Y.reshape(-1, ni, nj).astype(np.float64) @ np.transpose(Tj, (0, 2, 1))
            same = i == j
            if same.any():    # diagonal blocks (types OO, HH only)
                blk[same] = 0.5 * (blk[same] + np.transpose(blk[same], (0, 2, 1)))
            I = np.stack([idx[a] for a in i])[:, :, None]
            J = np.stack([idx[a] for a in j])[:, None, :]
            P[I, J] = blk
            P[J.transpose(0, 2, 1), I.transpose(0, 2, 1)] = np.transpose(blk, (0, 2, 1))
        P = np.asfortranarray(P)
        if S is not None:
            nel = sum(a[0] for a in atoms)
            tr = float(np.sum(P * S))
            if tr > 0:
                P *= nel / tr
        return P

    # -- training ---------------------------------------------------------
    def fit(self, data, ctx, epochs=40, batch=512, lr=2e-3, seed=0, log=None):
        """data[t] = (X, Y) float32 arrays.  Returns final mean losses."""
        rng = np.random.default_rng(seed)
        losses = {}
        for t in TYPES:
            X, Y = data[t]
            if len(X) == 0:
                continue
            self.xmu[t] = X.mean(0)
            self.xsd[t] = np.maximum(X.std(0), 1e-3).astype(np.float32)
            self.ymu[t] = Y.mean(0)
            self.ysd[t] = np.maximum(Y.std(0), 1e-4).astype(np.float32)
            Xn = ((X - self.xmu[t]) / self.xsd[t]).astype(np.float32)
            Yn = ((Y - self.ymu[t]) / self.ysd[t]).astype(np.float32)
            W = np.ones_like(Yn)
            n = len(Xn)
            dev = DeviceArrays(ctx)
            try:
                dx, dy, dw = dev.put(Xn), dev.put(Yn), dev.put(W)
                net = self.nets[t]
                for ep in range(epochs):
                    order = rng.permutation(n)
                    Xs, Ys = Xn[order], Yn[order]
                    ctx.to_device_raw(dx, Xs)
                    ctx.to_device_raw(dy, Ys)
                    tot, cnt = 0.0, 0
                    rate = lr * 0.5 * (1 + math.cos(math.pi * ep / epochs))
                    for s0 in range(0, n, batch):
                        bs = min(batch, n - s0)
                        loss = ctypes.c_float()
                        check(lib().polyeri_mlp_train_step(
                            net.h, ctypes.c_void_p(dx.value + 4 * s0 * self.nfeat),
                            ctypes.c_void_p(dy.value + 4 * s0 * Yn.shape[1]),
                            ctypes.c_void_p(dw.value + 4 * s0 * Yn.shape[1]), bs, ctypes.c_float(rate), 1,
                            ctypes.byref(loss)))
                        tot += loss.value * bs
                        cnt += bs
                    losses[t] = tot / cnt
                    if log and (ep % 10 == 0 or ep == epochs - 1):
                        log(f"train {t}: epoch {ep} loss {losses[t]:.4e}")
            finally:
                dev.free()
        return losses

    # -- persistence ------------------------------------------------------
    def save(self, path):
        d = {"basis": np.array(self.basis_name), "hidden": np.array(self.hidden), "nfeat": np.array(self.nfeat)}
        for t in TYPES:
            d[f"{t}_w"] = self.nets[t].params()
            d[f"{t}_xmu"], d[f"{t}_xsd"] = self.xmu[t], self.xsd[t]
            d[f"{t}_ymu"], d[f"{t}_ysd"] = self.ymu[t], self.ysd[t]
        np.savez_compressed(path, **d)

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        # models saved before the field features use the 48 geometric columns
        nfeat = int(z["nfeat"]) if "nfeat" in z.files else NFEAT_GEOM
        m = cls(str(z["basis"]), nfeat=nfeat, hidden=tuple(int(h) for h in z["hidden"]))
        for t in TYPES:
            m.nets[t].set_params(z[f"{t}_w"])
            m.xmu[t], m.xsd[t] = z[f"{t}_xmu"], z[f"{t}_xsd"]
            m.ymu[t], m.ysd[t] = z[f"{t}_ymu"], z[f"{t}_ysd"]
        return m


def core_guess(ctx, basis, lindep=1e-6):
    """Density of the core Hamiltonian T + V (canonical orthogonalisation,
    aufbau, doubly occupied) in the reference layout."""
    S, T, V, _ = ctx.one_electron()
    real = basis.real_functions
    s = S[np.ix_(real, real)]
    h = (T + V)[np.ix_(real, real)]
    e, U = np.linalg.eigh(s)
    keep = e > lindep
    X = U[:, keep] / np.sqrt(e[keep])
    _, C = np.linalg.eigh(X.T @ h @ X)
    C = X @ C
    nocc = sum(a[0] for a in basis.atoms) // 2
    P = np.zeros((basis.ncgs, basis.ncgs), order="F")
    P[np.ix_(real, real)] = 2.0 * C[:, :nocc] @ C[:, :nocc].T
    return P


_ATOM_CACHE = {}


def atomic_density(basis_name, z, temphf=1000.0):
    """Spherically averaged density of a free atom of element z in its own
    basis (reference layout of the one-atom basis): closed-shell SCF with
    Fermi-smeared occupations (TEMPHF, equal occupation of degenerate
    levels) for even z; for hydrogen the exact one-electron ground state."""
    key = (basis_name, z, temphf)
    if key in _ATOM_CACHE:
        return _ATOM_CACHE[key]
    b = read_basis(basis_name, [(z, 0.0, 0.0, 0.0)])
    ctx = Context(b, schwarz_tol=0.0)
    if z % 2 == 0:
        r = ctx.self_consistent_field(scfconv=1e-9, diis=2, scfcycles=200, temphf=temphf)
        if not r["converged"]:
            raise RuntimeError(f"atomic SCF of Z={z} did not converge")
        P = np.asarray(r["P"])
    elif z == 1:
        S, T, V, _ = ctx.one_electron()
        real = b.real_functions
        s = S[np.ix_(real, real)]
        e, U = np.linalg.eigh(s)
        X = U / np.sqrt(e)
        _, C = np.linalg.eigh(X.T @ (T + V)[np.ix_(real, real)] @ X)
        c = X @ C[:, 0]
        P = np.zeros((b.ncgs, b.ncgs))
        P[np.ix_(real, real)] = np.outer(c, c)
    else:
        raise ValueError("atomic densities: even Z or hydrogen only")
    fl = [f for f, _ in atom_functions(b)[0]]
    _ATOM_CACHE[key] = P[np.ix_(fl, fl)]
    return _ATOM_CACHE[key]


def sad_guess(basis, basis_name):
    """Superposition of atomic densities: the block of each atom is its free
    atom's spherically averaged density, every other block zero."""
    P = np.zeros((basis.ncgs, basis.ncgs), order="F")
    for a, fl in enumerate(atom_functions(basis)):
        idx = [f for f, _ in fl]
        P[np.ix_(idx, idx)] = atomic_density(basis_name, int(basis.atoms[a][0]))
    return P

