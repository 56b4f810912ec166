"""The geometry optimiser (OPTIMIZATION, src/poptimization.f90) on the CPU:
its steps on a quadratic surface, where the secant rescaling and GDIIS
have closed forms, its errors, and a whole optimisation of H2O/STO-3G with
the oracle's SCF and the library's host gradient entries."""
import os

import numpy as np
import pytest

import oracle
import polymer_amd as pa
from polymer_amd import optimize as O


def _quadratic(k, xmin):
    k = np.asarray(k, dtype=float)

    def ev(x):
        d = x - xmin
        return 0.5 * float(np.sum(k * d * d)), k * d
    return ev


def test_relax_steps_on_a_quadratic():
    """Cycle 1 moves by -g (RELAX1 = 1); cycle 2 rescales RELAX1 by the
    secant ratio of the largest component of cycle 1, which on a quadratic
    is 1 / k of that component: R2 = R1 - g1 / k_max."""
    k = np.array([[0.4, 0.9, 0.6], [0.5, 0.7, 0.8]])
    xmin = np.array([[0.0, 0.1, -0.2], [1.5, 0.0, 0.3]])
    x0 = xmin + np.array([[0.3, -0.5, 0.2], [0.1, 0.05, -0.1]])
    r = O.optimize(_quadratic(k, xmin), x0, optconv=1e-9, gdiisterm=0, maxcycles=200)
    h = r["history"]
    g0 = k * (x0 - xmin)
    assert np.allclose(h[1]["x"], x0 - g0, atol=1e-15)
    imax = np.unravel_index(np.argmax(np.abs(g0)), g0.shape)
    g1 = k * (h[1]["x"] - xmin)
    assert h[1]["step"][0] == "relax" and abs(h[1]["step"][1] - 1.0 / k[imax]) < 1e-14
    assert np.allclose(h[2]["x"], h[1]["x"] - g1 / k[imax], atol=1e-14)
    assert r["converged"] and np.max(np.abs(r["x"] - xmin)) < 1e-8


def test_gdiis_step_is_the_constrained_least_squares():
    """The 4-term GDIIS geometry equals sum c_j R_j with sum c = 1 and the
    smallest |sum c_j g_j|, found here by eliminating c_m (least squares over
    the differences g_j - g_m), and on a quadratic its gradient is that
    combination of the stored gradients."""
    k = np.array([[0.3, 1.1, 0.6], [0.9, 0.5, 0.75]])
    xmin = np.zeros((2, 3))
    x0 = np.array([[0.4, -0.3, 0.2], [0.1, 0.25, -0.35]])
    with pytest.raises(O.OptimizationError) as ei:
        O.optimize(_quadratic(k, xmin), x0, optconv=1e-14, gdiisterm=4, maxcycles=5)
    h = ei.value.history
    assert h[3]["step"][0] == "gdiis"
    xs = [c["x"] for c in h[:4]]
    gs = [c["gradient"] for c in h[:4]]
    D = np.stack([(g - gs[3]).ravel() for g in gs[:3]], axis=1)
    d, *_ = np.linalg.lstsq(D, -gs[3].ravel(), rcond=None)
    c = np.append(d, 1.0 - d.sum())
    x_ref = sum(ci * xi for ci, xi in zip(c, xs))
    assert np.allclose(h[4]["x"], x_ref, atol=1e-12)
    assert np.allclose(h[4]["gradient"], sum(ci * gi for ci, gi in zip(c, gs)), atol=1e-12)
    assert abs(sum(h[3]["step"][1]) - 1.0) < 1e-12


def test_errors():
    ev = _quadratic(np.ones((2, 3)), np.zeros((2, 3)))
    with pytest.raises(O.OptimizationError, match="INVALID GDIISTERM"):
        O.optimize(ev, np.ones((2, 3)), gdiisterm=1)
    with pytest.raises(O.OptimizationError, match="ZERO INTERNUCLEAR SEPARATION"):
        O.optimize(ev, np.zeros((2, 3)))
    with pytest.raises(O.OptimizationError, match="TOO MANY ITERATIONS"):
        O.optimize(_quadratic(0.5 * np.ones((2, 3)), np.array([[0.0, 0.0, 0.0], [3.0, 0.0, 0.0]])),
                   np.array([[1.0, 0.0, 0.0], [0.0, 2.0, 0.0]]), optconv=1e-14, maxcycles=2)


H2O_START = [(8, 0.0, 0.0, 0.0), (1, 0.0, 1.35, 1.05), (1, 0.1, -1.2, 1.15)]


def oracle_evaluate(basis):
    """Energy of the oracle's converged SCF and the library's host gradient
    entries at its density (no GPU)."""
    z = [a[0] for a in H2O_START]

    def ev(x):
        at = [(zi, *map(float, p)) for zi, p in zip(z, x)]
        r = oracle.Basis(os.path.join(pa.BASIS_DIR, basis), at).scf(inttol=0.0, scfconv=1e-11, scfcycles=200,
                                                                     diis=2)
        b = pa.read_basis(basis, at)
        W = pa.energy_weighted_density(r["co"], r["eps"], r["iocc"])
        return r["e_total"], pa.one_electron_gradient_host(b, r["P"], W) + pa.eri_gradient_host(b, r["P"], 1.0)
    return ev


def test_h2o_sto3g_optimisation_on_cpu():
    """From a distorted H2O the optimiser converges (every component <=
    OPTCONV 4e-4) with the energy lowered, to a stationary point: a central
    difference of the oracle energy along every coordinate is below 1e-3."""
    x0 = np.array([a[1:] for a in H2O_START])
    ev = oracle_evaluate("STO-3G")
    r = O.optimize(ev, x0)
    print(f"cycles {r['cycles']} E {r['energy']:.10f} start {r['history'][0]['energy']:.10f}")
    assert r["converged"] and np.max(np.abs(r["gradient"])) <= 4e-4
    assert r["energy"] < r["history"][0]["energy"] - 1e-3
    assert any(h.get("step", ("",))[0] == "gdiis" for h in r["history"])
    h = 1e-3
    for a in range(3):
        for e in range(3):
            xp, xm = r["x"].copy(), r["x"].copy()
            xp[a, e] += h
            xm[a, e] -= h
            assert abs((ev(xp)[0] - ev(xm)[0]) / (2 * h)) < 1e-3
