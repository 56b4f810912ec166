"""Geometry optimisation (OPTIMIZATION, src/poptimization.f90) on the device
SCF and gradient: the same trajectory as the CPU run on the oracle's SCF and
the host gradient entries, and a cc-pVDZ optimisation that ends at a
stationary point of the device energy."""
import numpy as np
import pytest

import polymer_amd as pa
from polymer_amd import optimize as O
from tests.test_optimize_cpu import H2O_START, oracle_evaluate

pytestmark = pytest.mark.gpu


def test_device_trajectory_equals_cpu_trajectory():
    """H2O/STO-3G from the distorted start: the device run takes the same
    number of cycles and steps, every geometry within 1e-7 bohr and energy
    within 1e-9 Eh of the CPU run."""
    print("case: optimisation H2O/STO-3G, device vs CPU", flush=True)
    dev = O.optimize_molecule("STO-3G", H2O_START, schwarz_tol=0.0, scf={"scfconv": 1e-11, "scfcycles": 200})
    cold = O.optimize_molecule("STO-3G", H2O_START, schwarz_tol=0.0, scf={"scfconv": 1e-11, "scfcycles": 200},
                               restart=False)
    print(f"SCF iterations with restart {dev['scf_iterations']} without {cold['scf_iterations']}", flush=True)
    assert sum(dev["scf_iterations"][1:]) < sum(cold["scf_iterations"][1:])
    assert np.max(np.abs(dev["x"] - cold["x"])) < 1e-7
    cpu = O.optimize(oracle_evaluate("STO-3G"), np.array([a[1:] for a in H2O_START]))
    print(f"cycles device {dev['cycles']} cpu {cpu['cycles']}; E {dev['energy']:.10f} {cpu['energy']:.10f}",
          flush=True)
    assert dev["converged"] and dev["cycles"] == cpu["cycles"]
    for a, b in zip(dev["history"], cpu["history"]):
        assert np.max(np.abs(a["x"] - b["x"])) < 1e-7
        assert abs(a["energy"] - b["energy"]) < 1e-9
        assert a.get("step", ("",))[0] == b.get("step", ("",))[0]


def test_ccpvdz_optimisation_reaches_a_stationary_point():
    """H2O/cc-pVDZ to OPTCONV 1e-5: central differences of the device SCF
    energy along every coordinate at the final geometry are below 2e-5
    Eh/bohr, and the energy is below the start's."""
    print("case: optimisation H2O/cc-pVDZ to 1e-5", flush=True)
    r = O.optimize_molecule("cc-pVDZ", H2O_START, optconv=1e-5, scf={"scfconv": 1e-11, "scfcycles": 200})
    print(f"cycles {r['cycles']} E {r['energy']:.10f} max|g| {np.max(np.abs(r['gradient'])):.2e}", flush=True)
    assert r["converged"] and r["energy"] < r["history"][0]["energy"]

    def energy(x):
        geom = [(a[0], *map(float, p)) for a, p in zip(H2O_START, x)]
        ctx = pa.Context(pa.read_basis("cc-pVDZ", geom), schwarz_tol=1e-12)
        try:
            return ctx.self_consistent_field(scfconv=1e-11, diis=2, scfcycles=200)["e_total"]
        finally:
            ctx.close()
    h = 1e-3
    for a in range(3):
        for e in range(3):
            xp, xm = r["x"].copy(), r["x"].copy()
            xp[a, e] += h
            xm[a, e] -= h
            assert abs((energy(xp) - energy(xm)) / (2 * h)) < 2e-5
