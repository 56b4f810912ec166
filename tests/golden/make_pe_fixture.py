"""Writes tests/golden/pe_jk_cutoff5_10.npz: the oracle's lattice J/K cells
(or_fourindex_eri_cells, the restatement of src/pfourindexdirect.f90:222-781
with HELIX = 0) for the polyethylene case of pe_case.py at CUTOFF1 5 /
CUTOFF2 10, INTTOL 1e-14, with REDUCED_CEL1X from the oracle's overlap cells
(src/ponee.f90:359-371).  The oracle screens composite primitive pairs
(IJIJ KLKL < INTTOL, src/pfourindexdirect.f90:529); at 1e-12 that screen
drops lattice-sum terms worth 3e-10 in J and 4e-10 in K which the
library's contracted-shell Schwarz screen at 1e-12 keeps, so the fixture is
made with the oracle screen two decades lower.

    OMP_NUM_THREADS=8 python tests/golden/make_pe_fixture.py

About 2 minutes on 8 threads.  The density is regenerated from its seed by
the test, so only the geometry, the sizes and J, K are stored.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import oracle  # noqa: E402
from oracle import periodic  # noqa: E402

import pe_case as C  # noqa: E402

INTTOL = 1e-14


def main():
    atoms = C.cell_atoms()
    path = os.path.join(ROOT, "polymer_amd", "basis", C.BASIS)
    ob = oracle.Basis(path, atoms)
    _, S, _, _, _ = periodic.one_electron_cells(path, atoms, C.PERIOD, C.CEL1, C.CEL1)
    rc1 = periodic.reduced_cel1(S, C.CEL1)
    xyz = np.stack([ob.cgsx, ob.cgsy, ob.cgsz], axis=1)
    real = np.nonzero(np.any(ob.cc != 0.0, axis=1))[0]
    P = C.density_cells(xyz, real)
    t = time.time()
    J, K, info = ob.fourindex_eri_cells(P, C.PERIOD, rc1, C.CEL1, C.CEL2, INTTOL)
    print(f"rc1 {rc1}, {info['composite_quartets']} composite quartets, {time.time() - t:.1f} s")
    np.savez_compressed(os.path.join(HERE, "pe_jk_cutoff5_10.npz"), atoms=np.array(atoms), rc1=rc1,
                        cel1=C.CEL1, cel2=C.CEL2, period=C.PERIOD, inttol=INTTOL, seed=C.SEED, J=J, K=K,
                        p_abs_sum=np.abs(P).sum())


if __name__ == "__main__":
    main()
