"""Seeded synthetic geometries of the benchmark workloads (bohr).

water()            O at the origin, H at (+-0.7572, 0.5865, 0) angstrom
water_cluster()    rigid monomers, uniformly random orientations (unit
                   quaternions), O positions uniform in a ball of the given
                   radius with every O-O distance >= 2.6 angstrom (random
                   sequential placement), optional bond/angle jitter
alkane()           all-trans CnH2n+2 from standard internal coordinates with
                   Gaussian coordinate noise
"""
import math

import numpy as np

ANG_BOHR = 0.529177249  # src/pmodules.f90:9

# monomer in angstrom: O, H, H
_WATER = [(8, 0.0, 0.0, 0.0), (1, 0.7572, 0.5865, 0.0), (1, -0.7572, 0.5865, 0.0)]


def to_bohr(atoms_ang):
    return [(z, x / ANG_BOHR, y / ANG_BOHR, w / ANG_BOHR) for z, x, y, w in atoms_ang]


def water():
    return to_bohr(_WATER)


def _rand_rotation(rng):
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def _monomer(rng, bond_sigma=0.0, angle_sigma_deg=0.0):
    r0 = math.hypot(0.7572, 0.5865)
    th0 = 2.0 * math.atan2(0.7572, 0.5865)
    r1 = r0 + (rng.normal(0, bond_sigma) if bond_sigma else 0.0)
    r2 = r0 + (rng.normal(0, bond_sigma) if bond_sigma else 0.0)
    th = th0 + (math.radians(rng.normal(0, angle_sigma_deg)) if angle_sigma_deg else 0.0)
    h1 = np.array([r1 * math.sin(th / 2), r1 * math.cos(th / 2), 0.0])
    h2 = np.array([-r2 * math.sin(th / 2), r2 * math.cos(th / 2), 0.0])
    return np.array([[0.0, 0.0, 0.0], h1, h2])


def water_cluster(n, seed, radius=4.5, min_oo=2.6, bond_sigma=0.0, angle_sigma_deg=0.0,
                  max_tries=200000):
    """(H2O)n in bohr; positions in angstrom are drawn first."""
    rng = np.random.default_rng(seed)
    centres = []
    tries = 0
    while len(centres) < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place the cluster")
        p = rng.uniform(-radius, radius, 3)
        if np.dot(p, p) > radius * radius:
            continue
        if all(np.linalg.norm(p - c) >= min_oo for c in centres):
            centres.append(p)
    atoms = []
    for c in centres:
        m = _monomer(rng, bond_sigma, angle_sigma_deg) @ _rand_rotation(rng).T + c
        for z, xyz in zip((8, 1, 1), m):
            atoms.append((z, *xyz))
    return to_bohr(atoms)


def alkane(nc=16, seed=0, noise=0.05):
    """All-trans CnH2n+2: C-C 1.53, C-H 1.09 angstrom, tetrahedral angles,
    then N(0, noise) angstrom on every coordinate."""
    rng = np.random.default_rng(seed)
    cc, ch = 1.53, 1.09
    half = math.radians(109.5) / 2
    dx, dy = cc * math.sin(half), cc * math.cos(half)
    carbons = [np.array([i * dx, (dy / 2) * (1 if i % 2 else -1), 0.0]) for i in range(nc)]
    atoms = [(6, *c) for c in carbons]
    hz = ch * math.sin(math.radians(109.5) / 2)
    hy = ch * math.cos(math.radians(109.5) / 2)
    for i, c in enumerate(carbons):
        s = 1 if i % 2 else -1
        atoms.append((1, c[0], c[1] + s * hy, c[2] + hz))
        atoms.append((1, c[0], c[1] + s * hy, c[2] - hz))
    # terminal hydrogens along the chain axis
    atoms.append((1, carbons[0][0] - ch, carbons[0][1], 0.0))
    atoms.append((1, carbons[-1][0] + ch, carbons[-1][1], 0.0))
    out = []
    for z, x, y, w in atoms:
        d = rng.normal(0, noise, 3) if noise else np.zeros(3)
        out.append((z, x + d[0], y + d[1], w + d[2]))
    return to_bohr(out)


def write_xyz(path, atoms_bohr, comment=""):
    sym = {1: "H", 6: "C", 8: "O"}
    with open(path, "w") as f:
        f.write(f"{len(atoms_bohr)}\n{comment}\n")
        for z, x, y, w in atoms_bohr:
            f.write(f"{sym.get(z, z)} {x * ANG_BOHR:.10f} {y * ANG_BOHR:.10f} {w * ANG_BOHR:.10f}\n")
