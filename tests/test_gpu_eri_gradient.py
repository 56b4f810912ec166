"""Two-electron nuclear gradient on the device (polyeri_eri_gradient,
FOURINDEX_ERI_GRADIENT): equal to the host replay of the same code (itself
pinned by finite differences of the oracle's energy in
tests/test_eri_gradient_cpu.py) on SCF densities, and at the benchmark's
(H2O)16 cc-pVDZ size translationally invariant and unchanged by the
Schwarz screen beyond its threshold."""
import time

import numpy as np
import pytest

import polymer_amd as pa
from polymer_amd import geometry as G
from tests import molecules as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("basis,nw", [("cc-pVDZ", 0), ("6-31Gss", 3)])
def test_eri_gradient_device_equals_host(basis, nw):
    atoms = M.H2O_GOLDEN if nw == 0 else G.water_cluster(nw, 5)
    print(f"case: ERI gradient device vs host, {basis}, {len(atoms)} atoms", flush=True)
    b = pa.read_basis(basis, atoms)
    ctx = pa.Context(b, schwarz_tol=1e-12)
    r = ctx.self_consistent_field(scfconv=1e-8, diis=2, scfcycles=80, solver=1, diis_mode=1)
    g = ctx.eri_gradient(r["P"])
    h = pa.eri_gradient_host(b, r["P"], 1.0, schwarz_tol=1e-12)
    err = np.max(np.abs(g - h))
    print(f"max |device - host| = {err:.2e}, max |g| = {np.max(np.abs(g)):.3e}", flush=True)
    assert err < 1e-10 * max(1.0, np.max(np.abs(h)))
    ctx.close()


def test_eri_gradient_water16_invariants():
    atoms = G.water_cluster(16, 0)
    print("case: ERI gradient (H2O)16 cc-pVDZ invariants", flush=True)
    b = pa.read_basis("cc-pVDZ", atoms)
    ctx = pa.Context(b, schwarz_tol=1e-12)
    r = ctx.self_consistent_field(scfconv=1e-8, diis=2, scfcycles=80, solver=1, diis_mode=1)
    t0 = time.perf_counter()
    g = ctx.eri_gradient(r["P"])
    dt = time.perf_counter() - t0
    drift = np.max(np.abs(g.sum(axis=0)))
    print(f"gradient {dt:.3f} s, max |g| {np.max(np.abs(g)):.3e}, max |sum over atoms| {drift:.2e}", flush=True)
    assert drift < 1e-9 * np.max(np.abs(g))
    # a looser screen changes the result only by what it drops
    c2 = pa.Context(b, schwarz_tol=1e-10)
    g2 = c2.eri_gradient(r["P"])
    print(f"Schwarz 1e-10 vs 1e-12: max diff {np.max(np.abs(g2 - g)):.2e}", flush=True)
    assert np.max(np.abs(g2 - g)) < 1e-6
    ctx.close()
    c2.close()


def test_scf_gradient_matches_energy_differences_gpu():
    """The total RHF gradient of a GPU SCF (Context.scf_gradient: one-electron
    part on the host, two-electron part on the device) against central
    differences of the GPU SCF energy, H2O/cc-pVDZ bent off symmetry."""
    atoms = [(1, 0.1, 1.079252144093028, 1.474611055780858), (8, 0.0, 0.0, 0.0), (1, -0.05, 1.1, -1.4)]
    print("case: SCF gradient vs energy differences, H2O/cc-pVDZ", flush=True)

    def scf(at):
        c = pa.Context(pa.read_basis("cc-pVDZ", at), schwarz_tol=0.0)
        r = c.self_consistent_field(scfconv=1e-11, diis=2, scfcycles=100, solver=1, diis_mode=1)
        return c, r

    ctx, r = scf(atoms)
    g = ctx.scf_gradient(r)
    h = 5e-4
    fd = np.zeros_like(g)
    for a in range(len(atoms)):
        for e in range(3):
            def moved(d):
                at = [list(t) for t in atoms]
                at[a][1 + e] += d
                return [tuple(t) for t in at]
            ep = scf(moved(h))[1]["e_total"]
            em = scf(moved(-h))[1]["e_total"]
            fd[a, e] = (ep - em) / (2 * h)
    err = np.max(np.abs(g - fd))
    print(f"max |g - fd| = {err:.2e}, max |g| = {np.max(np.abs(g)):.3e}", flush=True)
    assert err < 2e-6
    assert np.max(np.abs(g.sum(axis=0))) < 1e-9
