"""The polyethylene lattice J/K case shared by the fixture script and the GPU
test: C2H4 unit cell (zig-zag, C-C along the chain x axis, period 2.55 A,
helix angle 0), 6-31G**, CUTOFF1 5 / CUTOFF2 10, and a seeded density with
the structure of a real one: P(-Q) = P(Q)^T, P(0) symmetric, elements
decaying as exp(-0.35 r) with the distance r (bohr) between the two
functions' centres, zero on ghost functions."""
import numpy as np

ANG = 1.0 / 0.529177249      # src/pmodules.f90:9
PERIOD = 2.55 * ANG
CEL1, CEL2 = 5, 10
BASIS = "6-31Gss"
SEED = 2024


def cell_atoms():
    h = 0.4229
    return [(6, 0.0, 0.0, h * ANG), (6, 1.275 * ANG, 0.0, -h * ANG),
            (1, 0.0, 0.876 * ANG, 1.071 * ANG), (1, 0.0, -0.876 * ANG, 1.071 * ANG),
            (1, 1.275 * ANG, 0.876 * ANG, -1.071 * ANG), (1, 1.275 * ANG, -0.876 * ANG, -1.071 * ANG)]


def density_cells(cgs_xyz, real, seed=SEED, cel1=CEL1, period=PERIOD):
    """(2 cel1 + 1, n, n) cells, index Q + cel1.  cgs_xyz: (n, 3) centres of
    the contracted functions; real: indices of non-ghost functions."""
    n = cgs_xyz.shape[0]
    rng = np.random.default_rng(seed)
    mask = np.zeros((n, n))
    mask[np.ix_(real, real)] = 1.0
    P = np.zeros((2 * cel1 + 1, n, n))
    for q in range(0, cel1 + 1):
        shift = np.array([q * period, 0.0, 0.0])
        r = np.linalg.norm(cgs_xyz[:, None, :] - (cgs_xyz[None, :, :] + shift), axis=2)
        m = rng.uniform(-1.0, 1.0, (n, n)) * np.exp(-0.35 * r) * mask
        if q == 0:
            m = 0.5 * (m + m.T)
        P[cel1 + q] = m
        P[cel1 - q] = m.T
    return P
