"""Learned-guess pieces that run without a GPU.

* the CPU reference network (oracle/mlp_ref.py) backpropagates correctly:
  float64 gradients against central finite differences;
* the BF16 rounding of the reference rounds to nearest even;
* the frame transformation of the density blocks: the converged density of
  one water (oracle SCF, 6-31G**) in two orientations gives the same
  frame-local blocks, so the network's targets are rotation invariant; and
  P = T Q T^T inverts Q = T^-1 P T^-T;
* the pair features are invariant under rotation and translation."""
import math

import numpy as np

import oracle
from oracle import mlp_ref
from polymer_amd import geometry as Gm
from polymer_amd import guess as G
from polymer_amd import read_basis
from tests.conftest import basis_path


def test_mlp_reference_gradients_fd():
    rng = np.random.default_rng(0)
    dims = [5, 7, 6, 3]
    npar = sum(dims[l] * dims[l - 1] + dims[l] for l in range(1, len(dims)))
    p = rng.standard_normal(npar) * 0.5
    X = rng.standard_normal((9, 5))
    T = rng.standard_normal((9, 3))
    W = rng.uniform(0.5, 1.5, (9, 3))
    _, g = mlp_ref.loss_and_grads(p, dims, X, T, W, rnd=False)
    h = 1e-6
    for k in rng.choice(npar, 25, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (mlp_ref.loss_and_grads(pp, dims, X, T, W, rnd=False)[0] -
              mlp_ref.loss_and_grads(pm, dims, X, T, W, rnd=False)[0]) / (2 * h)
        assert abs(fd - g[k]) < 1e-7 * max(1.0, abs(fd))


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-3], np.float32)
    r = mlp_ref.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie -> even mantissa
    assert r[2] == 1.0 + 4 * 2 ** -8        # tie -> even (up)
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-3) < 3.0e-3 * 2 ** -8


def _rot(axis, ang):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K


def _water_density(atoms):
    ob = oracle.Basis(basis_path("6-31Gss"), atoms)
    r = ob.scf(inttol=1e-12, scfconv=1e-9, scfrelax=1.0, hfexchange=1.0, lindep=1e-6, temphf=0.0,
               scfcycles=60, diis=2, charge=0, spherical=False, direct=False, guess=None)
    return r["P"]


def test_frame_local_blocks_rotation_invariant():
    w = Gm.water()
    R = _rot([0.3, -0.8, 0.5], 1.1)
    shift = np.array([0.7, -1.2, 2.0])
    w2 = [(z, *(R @ np.array(xyz) + shift)) for z, *xyz in w]
    blocks = []
    for atoms in (w, w2):
        b = read_basis("6-31Gss", atoms)
        P = _water_density(atoms)
        model_free = G.atom_functions(b)
        F = G.frames(atoms)
        T = [G.frame_transform(F[0], model_free[a]) for a in range(3)]
        q = {}
        for i in range(3):
            for j in range(3):
                ii = [c for c, _ in model_free[i]]
                jj = [c for c, _ in model_free[j]]
                blk = P[np.ix_(ii, jj)]
                q[i, j] = np.linalg.solve(T[i], np.linalg.solve(T[j], blk.T).T)
                # inverse map
                assert np.max(np.abs(T[i] @ q[i, j] @ T[j].T - blk)) < 1e-12
        blocks.append(q)
    for k in blocks[0]:
        assert np.max(np.abs(blocks[0][k] - blocks[1][k])) < 1e-7, k


def test_pair_features_invariant():
    atoms = Gm.water_cluster(4, 3, radius=3.0, bond_sigma=0.02, angle_sigma_deg=2.0)
    R = _rot([1.0, 2.0, -0.5], 2.3)
    moved = [(z, *(R @ np.array(xyz) + np.array([3.0, 0.1, -2.0]))) for z, *xyz in atoms]
    p1, p2 = G.pair_list(atoms), G.pair_list(moved)
    assert p1 == p2
    assert np.max(np.abs(G.pair_features(atoms, p1) - G.pair_features(moved, p2))) < 1e-5
