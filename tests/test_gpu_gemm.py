"""The MP2 transform's matrix-product kernels, one product at a time.

polyeri_gemm_f64 with the FP64 MFMA kernel equals numpy's float64 product to
1e-13 relative, and with the BF16x3 kernel to 1e-6 relative, for ragged
sizes and every operand layout; each output row is located, so a wrong
row order in the result registers names itself in the failure."""
import ctypes

import numpy as np
import pytest

import polymer_amd as pa
from polymer_amd import geometry as Gm
from polymer_amd._lib import check, lib

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return pa.Context(pa.read_basis("STO-3G", Gm.water()))


def _run(ctx, mode, A, B, ta, tb):
    M, K = (A.shape[1], A.shape[0]) if ta else A.shape
    N = B.shape[0] if tb else B.shape[1]
    sam, sak = (1, A.shape[1]) if ta else (A.shape[1], 1)
    sbk, sbn = (1, B.shape[1]) if tb else (B.shape[1], 1)
    ptrs = []
    try:
        def put(x):
            x = np.ascontiguousarray(x, np.float64)
            d = ctx.device_alloc(x.nbytes)
            ctx.to_device_raw(d, x)
            ptrs.append(d)
            return d
        da, db, dc = put(A), put(B), put(np.zeros((M, N)))
        check(lib().polyeri_gemm_f64(mode, M, N, K, da, sam, sak, db, sbk, sbn, dc, N, 1))
        C = np.zeros((M, N))
        ctx.from_device_raw(dc, C)
        return C
    finally:
        for d in ptrs:
            ctx.device_free(d)


@pytest.mark.parametrize("mode,tol", [(1, 1e-13), (2, 1e-6)])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
# shapes that take each tile form of the FP64 kernel: short M (1, 3 and 4 row
# tiles), narrow N, 128-wide and 160-wide column blocks, partial K slices
@pytest.mark.parametrize("M,K,N", [(35, 19, 21), (150, 50, 300), (16, 300, 9), (64, 33, 130), (150, 37, 40),
                                   (100, 70, 150)])
def test_gemm_f64_kernels(ctx, mode, tol, ta, tb, M, K, N):
    rng = np.random.default_rng(7)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    print(f"case: gemm mode {mode} transA {ta} transB {tb} M {M} K {K} N {N}", flush=True)
    C = _run(ctx, mode, A, B, ta, tb)
    ref = (A.T if ta else A) @ (B.T if tb else B)
    where = [int(np.argmin(np.max(np.abs(ref - C[r]), axis=1))) for r in range(M)]
    print("row r of the result equals reference row", where, flush=True)
    assert np.max(np.abs(C - ref)) < tol * np.max(np.abs(ref))
