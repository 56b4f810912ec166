"""The allocation-free J/K entry point against the allocating one.

polyeri_jk_build_device_noalloc stages the density and accumulates J/K in
the library's fixed device workspace; its results must equal those of
polyeri_jk_build_device (which uses the context's own buffers) to the J/K
parity tolerance, on two calls in a row on one device, the second with a
smaller system (the workspace keeps the first call's contents), and a
second build on the same context must give the same matrices again.
"""
import numpy as np
import pytest

import polymer_amd as pa
from polymer_amd import geometry as G
from polymer_amd._lib import check, lib

pytestmark = pytest.mark.gpu
TOL = 1e-10   # the J/K parity tolerance of the allocating build against the oracle


def _both(ctx, p):
    n = ctx.ncgs
    nb = n * n * 8
    d_p, d_j, d_k, d_j2, d_k2 = (ctx.device_alloc(nb) for _ in range(5))
    try:
        ctx.to_device(d_p, p)
        check(lib().polyeri_jk_build_device(ctx.h, d_p, d_j, d_k))
        ctx.sync()
        check(lib().polyeri_jk_build_device_noalloc(ctx.h, d_p, d_j2, d_k2))
        ctx.sync()
        out = [ctx.from_device(d, (n, n)) for d in (d_j, d_k, d_j2, d_k2)]
    finally:
        for d in (d_p, d_j, d_k, d_j2, d_k2):
            ctx.device_free(d)
    return out


@pytest.mark.parametrize("nwater", [4, 1])          # the second call is the smaller system
def test_noalloc_matches_allocating_build(nwater):
    print(f"case: no-alloc J/K (H2O){nwater} cc-pVDZ vs polyeri_jk_build_device", flush=True)
    atoms = G.water_cluster(nwater, 3) if nwater > 1 else G.water()
    b = pa.read_basis("cc-pVDZ", atoms)
    ctx = pa.Context(b, schwarz_tol=1e-12)
    rng = np.random.default_rng(nwater)
    p = rng.standard_normal((b.ncgs, b.ncgs))
    p = np.asfortranarray(0.05 * (p + p.T))
    for rep in range(2):
        j, k, j2, k2 = _both(ctx, p)
        dj, dk = float(np.max(np.abs(j2 - j))), float(np.max(np.abs(k2 - k)))
        print(f"rep {rep}: max|dJ| {dj:.2e} max|dK| {dk:.2e} (|J| {np.max(np.abs(j)):.2e})", flush=True)
        assert np.max(np.abs(j)) > 1e-3
        assert dj < TOL and dk < TOL
    ctx.close()
