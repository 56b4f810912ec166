"""Rys quadrature nodes and weights (polyeri_rys_nodes, the tables the Rys
kernels evaluate), pinned by what the mathematics fixes:

* Gauss exactness: the n-point rule integrates u^m exactly for m < 2n, so
  sum_r w_r u_r^m must equal the Boys function
  F_m(T) = Gamma(m + 1/2) P(m + 1/2, T) / (2 T^(m + 1/2)) (regularised lower
  incomplete gamma P, scipy), F_m(0) = 1 / (2m + 1);
* T = 0: the nodes are the squared positive nodes of the 2n-point
  Gauss-Legendre rule on [-1, 1] and the weights half its weights (numpy);
* nodes ascending in (0, 1), weights positive, and the series continuous
  across the unit intervals and the switch to the asymptotic form.

A wrong node, weight, interval index or Chebyshev variable fails the moment
test at the affected T."""
import numpy as np
import pytest
from scipy.special import gamma, gammainc

import polymer_amd as pa


def boys_ref(m, T):
    if T == 0.0:
        return 1.0 / (2 * m + 1)
    a = m + 0.5
    return gamma(a) * gammainc(a, T) / (2.0 * T ** a)


def _grid():
    rng = np.random.default_rng(7)
    pts = list(np.linspace(0.0, 100.0, 801)) + list(rng.uniform(0.0, 80.0, 400))
    pts += [i + e for i in range(0, 80) for e in (-1e-12, 0.0, 1e-12) if i + e >= 0.0]
    pts += [1e-9, 1e-6, 1e-3, 150.0, 400.0, 1e4]
    return pts


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_rys_moments_match_boys(n):
    worst = 0.0
    for T in _grid():
        u, w = pa.rys_nodes(n, T)
        assert np.all(np.diff(u) > 0) and u[0] > 0.0 and u[-1] < 1.0, (n, T, u)
        assert np.all(w > 0.0), (n, T, w)
        for m in range(2 * n):
            ref = boys_ref(m, T)
            got = float(np.sum(w * u ** m))
            worst = max(worst, abs(got - ref) / ref)
    print(f"n = {n}: largest relative moment error {worst:.2e}")
    assert worst < 1e-13   # u^9 carries nine times the ~3e-15 node error


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_rys_t0_is_gauss_legendre(n):
    x, wl = np.polynomial.legendre.leggauss(2 * n)
    pos = x > 0
    u0 = np.sort(x[pos] ** 2)
    w0 = wl[pos][np.argsort(x[pos] ** 2)]
    u, w = pa.rys_nodes(n, 0.0)
    assert np.max(np.abs(u - u0) / u0) < 1e-14
    assert np.max(np.abs(w - w0) / w0) < 1e-14


@pytest.mark.parametrize("n", [3, 4, 5])
def test_rys_continuity(n):
    tmax = 32 + 8 * n
    for T in list(range(1, tmax)) + [tmax]:
        ua, wa = pa.rys_nodes(n, T - 1e-13)
        ub, wb = pa.rys_nodes(n, float(T))
        assert np.max(np.abs(ua - ub) / ub) < 1e-13, (n, T)
        assert np.max(np.abs(wa - wb) / wb) < 1e-13, (n, T)


def test_rys_bad_arguments():
    with pytest.raises(pa.PolyeriError):
        pa.rys_nodes(6, 1.0)
    with pytest.raises(pa.PolyeriError):
        pa.rys_nodes(3, -1.0)


# Cartesian component order of the library (xx xy xz yy yz zz / x y z)
_COMP = {0: [(0, 0, 0)], 1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
         2: [(2, 0, 0), (1, 1, 0), (1, 0, 1), (0, 2, 0), (0, 1, 1), (0, 0, 2)]}
_RYS_CLASSES = [(2, 1, 1, 1), (2, 1, 2, 1), (2, 2, 1, 1), (2, 2, 2, 0), (2, 2, 2, 1), (2, 2, 2, 2)]


@pytest.mark.parametrize("cls", _RYS_CLASSES)
def test_rys_block_vs_mcmurchie_davidson(cls):
    """The kernels' per-lane Rys code (every lane variant and axis-permutation
    member, mapped back to global components, evaluated on the host) against
    the independent McMurchie-Davidson integrals of tests/md_reference.py, on
    random primitive quartets whose Boys arguments span the Chebyshev
    intervals and the asymptotic range.  A wrong recursion coefficient, a
    wrong horizontal transfer, a wrong member permutation or component map
    fails here; the library checks that every integral is produced exactly
    once."""
    from tests.md_reference import eri_prim
    rng = np.random.default_rng(sum(cls) * 17 + cls[3])
    nc = [len(_COMP[l]) for l in cls]
    worst = 0.0
    # compact pairs whose centres lie 0.2 .. 12 bohr apart: Boys arguments
    # from below 1 through the Chebyshev intervals to the asymptotic form
    tmax = 0.0
    for k, sep in enumerate((0.2, 1.5, 4.0, 7.0, 12.0)):
        xyz = rng.uniform(-0.7, 0.7, (4, 3))
        xyz[2:] += sep * np.array([0.8, 0.36, -0.48])
        ex = rng.uniform(0.3, 3.0, 4)
        pz, qz = ex[0] + ex[1], ex[2] + ex[3]
        P = (ex[0] * xyz[0] + ex[1] * xyz[1]) / pz
        Q = (ex[2] * xyz[2] + ex[3] * xyz[3]) / qz
        tmax = max(tmax, pz * qz / (pz + qz) * float(np.sum((P - Q) ** 2)))
        blk = pa.rys_block(*cls, xyz, ex)
        assert blk.shape == tuple(nc)
        idx = [tuple(int(rng.integers(0, n)) for n in nc) for _ in range(24)]
        idx.append(tuple(n - 1 for n in nc))
        idx.append((0, 0, 0, 0))
        for ia, ib, ic, id_ in idx:
            ref = eri_prim(_COMP[cls[0]][ia], ex[0], xyz[0], _COMP[cls[1]][ib], ex[1], xyz[1],
                           _COMP[cls[2]][ic], ex[2], xyz[2], _COMP[cls[3]][id_], ex[3], xyz[3])
            got = blk[ia, ib, ic, id_]
            err = abs(got - ref) / max(abs(ref), np.max(np.abs(blk)) * 1e-3)
            worst = max(worst, err)
            assert err < 1e-11, (cls, sep, (ia, ib, ic, id_), got, ref)
    assert tmax > 80.0       # the asymptotic form was reached
    print(f"{cls}: worst scaled error {worst:.2e}, largest T {tmax:.1f}")


def test_rys_block_rejects_other_classes():
    with pytest.raises(pa.PolyeriError):
        pa.rys_block(1, 1, 1, 1, np.zeros((4, 3)), np.ones(4))
