"""Pins the oracle's basis layout (READ3, src/pread.f90:584-1482): the
primitive, contracted and composite-shell counts including ghost functions,
the primitive normalisation and the ghost/composite construction."""
import math

import numpy as np
import pytest

import oracle
from tests import molecules as M
from tests.conftest import basis_path


@pytest.mark.parametrize("case,basis,atoms", [
    ("H2O_ccpVDZ_spherical", "cc-pVDZ", M.H2O_GOLDEN),
    ("BH_STO3G", "STO-3G", M.BH_GOLDEN),
    ("FH_STO3G_zeroT", "STO-3G", M.FH_GOLDEN),
])
def test_counts_match_shipped_outputs(golden, case, basis, atoms):
    b = oracle.Basis(basis_path(basis), atoms)
    g = golden[case]
    assert (b.npgs, b.ncgs, b.npshell) == (g["npgs"], g["ncgs"], g["npshell"])


def test_ghost_layout_h2o():
    b = oracle.Basis(basis_path("cc-pVDZ"), M.H2O_GOLDEN)
    ghosts = [c for c in range(b.ncgs) if not np.any(b.cc[c])]
    assert len(ghosts) == 11            # 5 O p + 4 O d + 2 H p ghost primitives
    assert list(np.bincount(b.p_shell_ang)) == [29, 7, 1]
    # every composite shell carries an s anchor
    assert np.all(b.p_shell[:, 0, 0, 0] >= 0)


@pytest.mark.parametrize("case,basis,atoms", [
    ("H2O_ccpVDZ_spherical", "cc-pVDZ", M.H2O_GOLDEN),
    ("BH_STO3G", "STO-3G", M.BH_GOLDEN),
])
def test_max_contracted_overlap_matches_output(golden, case, basis, atoms):
    """max |S_C| printed by OVERLAP_KINETIC (src/ponee.f90:361-370, F18.13)
    pins the primitive normalisation and the unrenormalised contractions."""
    b = oracle.Basis(basis_path(basis), atoms)
    s_p, s_c, t_c, n_c = b.one_electron()
    assert abs(np.max(np.abs(s_c)) - golden[case]["max_overlap"]) < 6e-14


def test_primitive_self_overlap():
    """An s primitive normalised as (2a/pi)^{3/4} has unit self-overlap."""
    b = oracle.Basis(basis_path("cc-pVDZ"), [(1, 0.0, 0.0, 0.0)])
    s_p, _, _, _ = b.one_electron()
    for p in range(b.npgs):
        if b.pax[p] + b.pay[p] + b.paz[p] == 0:
            a = b.zt[p]
            n = (2 * a / 3.14159265358979) ** 0.75
            assert abs(n * n * s_p[p, p] - 1.0) < 1e-13


@pytest.mark.parametrize("basis,atoms,counts", [
    # the throwaway counter of the survey reproduces these for the configs
    ("6-31Gss", [(6, 0, 0, 0), (1, 2.0, 0, 0), (1, -1.0, 1.7, 0), (1, -1.0, -1.7, 0), (1, 0, 0, 2.0),
                 (6, 0, 0, 5.0), (1, 0, 2, 5), (1, 0, -2, 5)], None),
])
def test_631gss_composite_shells(basis, atoms, counts):
    b = oracle.Basis(basis_path(basis), atoms)
    # C 6-31G**: 6 core s + SP(3)+SP(1) anchors = 10 s-anchored shells; d ghosted
    assert np.sum(b.p_shell_ang == 2) == 2
    assert np.all(b.p_shell[:, 0, 0, 0] >= 0)


def test_water16_counts():
    """(H2O)16 cc-pVDZ exceeds the reference's 1024 cap: NPGS 1072."""
    atoms = []
    for k in range(16):
        atoms += [(a[0], a[1] + 10.0 * k, a[2], a[3]) for a in M.H2O_CONFIG1]
    b = oracle.Basis(basis_path("cc-pVDZ"), atoms)
    assert (b.npgs, b.ncgs, b.npshell) == (1072, 576, 592)
