"""Host-side checks of libpolyeri that need no GPU: the shared library loads
and exports every symbol include/polyeri.h declares, and its basis reader
(READ3 semantics) reproduces the oracle's reference layout bit for bit."""
import os
import re

import numpy as np
import pytest

import oracle
import polymer_amd as pa
from polymer_amd import _lib, geometry as G
from tests import molecules as M
from tests.conftest import ROOT, basis_path


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "polyeri.h")).read()
    declared = set(re.findall(r"\b(polyeri_[a-z0-9_]+)\s*\(", hdr))
    assert declared == set(_lib.EXPORTS)
    L = _lib.lib()
    for name in declared:
        assert hasattr(L, name), name


@pytest.mark.parametrize("basis,atoms", [
    ("cc-pVDZ", M.H2O_GOLDEN), ("STO-3G", M.BH_GOLDEN), ("STO-3G", M.FH_GOLDEN),
    ("6-31Gss", G.alkane(3, seed=1)), ("cc-pVDZ", G.water_cluster(4, 2)),
])
def test_basis_reader_matches_oracle_layout(basis, atoms):
    b = pa.read_basis(basis, atoms)
    o = oracle.Basis(basis_path(basis), atoms)
    for k in ["pax", "pay", "paz", "pgs_atom", "pgsx", "pgsy", "pgsz", "zt", "cc", "p_shell",
              "p_shell_ang", "cgsx", "cgsy", "cgsz"]:
        assert np.array_equal(getattr(b, k), getattr(o, k)), k


def test_config_sizes():
    """Basis sizes of the benchmark workloads, with ghosts (NPGS/NCGS/NPSHELL)."""
    b = pa.read_basis("cc-pVDZ", G.water_cluster(16, 0))
    assert (b.npgs, b.ncgs, b.npshell, len(b.real_functions)) == (1072, 576, 592, 400)
    b = pa.read_basis("6-31Gss", G.alkane(16, 0))
    assert (b.npgs, b.ncgs, b.npshell, len(b.real_functions)) == (784, 508, 346, 410)
    b = pa.read_basis("6-31Gss", G.water_cluster(8, 0, bond_sigma=0.02, angle_sigma_deg=2.0))
    assert (b.npgs, b.ncgs, b.npshell, len(b.real_functions)) == (384, 248, 168, 200)


def test_water_cluster_constraints():
    for seed in range(5):
        w = G.water_cluster(16, seed)
        o = np.array([a[1:] for a in w if a[0] == 8]) * G.ANG_BOHR
        d = np.linalg.norm(o[:, None] - o[None], axis=-1) + 100 * np.eye(16)
        assert d.min() >= 2.6 and np.linalg.norm(o, axis=1).max() <= 4.5


def test_missing_element_raises():
    with pytest.raises(pa.PolyeriError):
        pa.read_basis("cc-pVDZ", [(17, 0.0, 0.0, 0.0)])


def test_fortran_interface_module_compiles(tmp_path):
    """include/polyeri_mod.f90 (the ISO_C_BINDING module a maintainer adds to
    the reference) compiles with amdflang and links against libpolyeri."""
    import shutil
    import subprocess
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        pytest.skip("amdflang not installed")
    mod = os.path.join(ROOT, "include", "polyeri_mod.f90")
    drv = os.path.join(ROOT, "polymer_amd", "fortran", "pscf_driver.f90")
    subprocess.check_call([fc, "-O0", "-J", str(tmp_path), "-c", mod, "-o", str(tmp_path / "m.o")])
    subprocess.check_call([fc, "-O0", "-I", str(tmp_path), "-J", str(tmp_path), drv, str(tmp_path / "m.o"),
                           "-L" + os.path.join(ROOT, "polymer_amd"), "-lpolyeri", "-o", str(tmp_path / "d")])
