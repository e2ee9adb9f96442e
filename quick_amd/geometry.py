"""Seeded synthetic geometries for the Fock-build workloads.

The systems are the ones the benchmark configurations name:

* water clusters: O atoms on an ``nx x ny x nz`` cubic grid of 3.1 A spacing,
  each O coordinate jittered by U(-0.3, 0.3) A, every water with O-H 0.957 A and
  H-O-H 104.5 deg in a uniformly random orientation (Shoemake unit quaternion);
* n-alkanes (n-icosane C20H42): C-C 1.53 A, C-H 1.09 A, tetrahedral angles, backbone
  torsions drawn per bond from {180, +60, -60} deg with probabilities
  {0.7, 0.15, 0.15}, redrawn when any non-bonded H...H distance is below 1.8 A.

Coordinates are returned in Angstrom; ``ANGSTROM_TO_BOHR`` is the factor QUICK
itself uses when it reads an input file (quick_constants_module.f90).
"""
from __future__ import annotations

import math
import numpy as np

# QUICK converts input Angstrom with A_TO_BOHRS = 1/0.529177083 (quick_constants_module.f90:71-72).
ANGSTROM_TO_BOHR = 1.0 / 0.529177083

SYMBOL_TO_Z = {"H": 1, "C": 6, "N": 7, "O": 8}


def _random_rotation(rng: np.random.Generator) -> np.ndarray:
    """Uniform random rotation matrix from a Shoemake unit quaternion."""
    u1, u2, u3 = rng.random(3)
    a, b = math.sqrt(1.0 - u1), math.sqrt(u1)
    w = a * math.sin(2.0 * math.pi * u2)
    x = a * math.cos(2.0 * math.pi * u2)
    y = b * math.sin(2.0 * math.pi * u3)
    z = b * math.cos(2.0 * math.pi * u3)
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def water_cluster(nx: int, ny: int, nz: int, seed: int, spacing: float = 3.1,
                  jitter: float = 0.3) -> tuple[list[str], np.ndarray]:
    """(H2O)_{nx*ny*nz} cluster; returns (symbols, xyz[natom,3] in Angstrom).

    Draw order, per water in (i, j, k) row-major order: the 3 jitter values,
    then the 3 quaternion uniforms.
    """
    rng = np.random.default_rng(seed)
    roh = 0.957
    half = math.radians(104.5) / 2.0
    # water in its own frame: O at origin, H's in the xz plane
    frame = np.array([[0.0, 0.0, 0.0],
                      [roh * math.sin(half), 0.0, roh * math.cos(half)],
                      [-roh * math.sin(half), 0.0, roh * math.cos(half)]])
    syms, pos = [], []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                o = np.array([i, j, k], float) * spacing + rng.uniform(-jitter, jitter, 3)
                rot = _random_rotation(rng)
                w = frame @ rot.T + o
                syms += ["O", "H", "H"]
                pos += [w[0], w[1], w[2]]
    return syms, np.array(pos)


def _place(a, b, c, bond, angle, torsion):
    """NeRF: position of atom d bonded to c with angle(b,c,d) and torsion(a,b,c,d)."""
    bc = c - b
    bc /= np.linalg.norm(bc)
    n = np.cross(b - a, bc)
    n /= np.linalg.norm(n)
    m = np.cross(n, bc)
    d2 = np.array([-bond * math.cos(angle),
                   bond * math.sin(angle) * math.cos(torsion),
                   bond * math.sin(angle) * math.sin(torsion)])
    return c + d2[0] * bc + d2[1] * m + d2[2] * n


def n_alkane(ncarbon: int, seed: int, max_tries: int = 1000) -> tuple[list[str], np.ndarray]:
    """n-C_nH_{2n+2} with seeded backbone torsions; returns (symbols, xyz Angstrom).

    Atom order: the carbons along the chain, then for each carbon its hydrogens.
    """
    rng = np.random.default_rng(seed)
    tet = math.acos(-1.0 / 3.0)
    rcc, rch = 1.53, 1.09
    choices = np.radians([180.0, 60.0, -60.0])
    for _ in range(max_tries):
        tors = rng.choice(choices, size=max(ncarbon - 3, 0), p=[0.7, 0.15, 0.15])
        c = [np.array([0.0, 0.0, 0.0]), np.array([rcc, 0.0, 0.0])]
        if ncarbon > 2:
            c.append(c[1] + rcc * np.array([-math.cos(tet), math.sin(tet), 0.0]))
        for n in range(3, ncarbon):
            c.append(_place(c[n - 3], c[n - 2], c[n - 1], rcc, tet, tors[n - 3]))
        hs = []
        for n in range(ncarbon):
            nb = [c[m] for m in (n - 1, n + 1) if 0 <= m < ncarbon]
            if len(nb) == 1:            # terminal CH3: staggered about the C-C bond
                a = nb[0]
                if ncarbon > 2:
                    ref = c[n + 2] if n == 0 else c[n - 2]
                else:                    # ethane: any point off the C-C axis fixes the frame
                    ref = a + np.array([0.0, 1.0, 0.0])
                for t in (math.pi, math.pi / 3, -math.pi / 3):
                    hs.append(_place(ref, a, c[n], rch, tet, t))
            else:                        # CH2: the two remaining tetrahedral directions
                u1 = (nb[0] - c[n]) / np.linalg.norm(nb[0] - c[n])
                u2 = (nb[1] - c[n]) / np.linalg.norm(nb[1] - c[n])
                bis = -(u1 + u2)
                bis /= np.linalg.norm(bis)
                perp = np.cross(u1, u2)
                perp /= np.linalg.norm(perp)
                half = tet / 2.0
                for s in (1.0, -1.0):
                    hs.append(c[n] + rch * (math.cos(half) * bis + s * math.sin(half) * perp))
        hs = np.array(hs)
        d = np.linalg.norm(hs[:, None, :] - hs[None, :, :], axis=-1)
        # geminal H's (same carbon) are bonded-neighbour pairs, skip them
        owner = np.concatenate([[n] * (3 if n in (0, ncarbon - 1) else 2) for n in range(ncarbon)])
        mask = owner[:, None] != owner[None, :]
        if not mask.any() or d[mask].min() >= 1.8:
            syms = ["C"] * ncarbon + ["H"] * len(hs)
            return syms, np.vstack([np.array(c), hs])
    raise RuntimeError("could not draw a clash-free alkane conformation")


# The benchmark configurations (seed per configuration).
CONFIGS = {
    "water4": lambda: water_cluster(2, 2, 1, seed=1),
    "water32": lambda: water_cluster(4, 4, 2, seed=2),
    "icosane": lambda: n_alkane(20, seed=4),
    "water64": lambda: water_cluster(4, 4, 4, seed=5),
}


def small_system(name: str) -> tuple[list[str], np.ndarray]:
    """Tiny systems used by the parity tests (Angstrom)."""
    if name == "h2o":
        half = math.radians(104.5) / 2.0
        return ["O", "H", "H"], np.array([[0.0, 0.0, 0.0],
                                          [0.957 * math.sin(half), 0.0, 0.957 * math.cos(half)],
                                          [-0.957 * math.sin(half), 0.0, 0.957 * math.cos(half)]])
    if name == "water2":
        return water_cluster(2, 1, 1, seed=11)
    if name == "ch4":
        t = 1.09 / math.sqrt(3.0)
        return ["C", "H", "H", "H", "H"], np.array([[0, 0, 0], [t, t, t], [-t, -t, t],
                                                   [-t, t, -t], [t, -t, -t]], float)
    if name == "ethane":
        return n_alkane(2, seed=3)
    if name == "butane":
        return n_alkane(4, seed=3)
    if name in CONFIGS:
        return CONFIGS[name]()
    raise KeyError(name)


def synthetic_density(xyz_bohr, katom, ksumtype, nbasis, seed):
    """Seeded symmetric density whose entries decay as exp(-0.25 R^2) with the
    distance R (bohr) between the two functions' centres, so that QUICK's
    density screening (max |D| over shell blocks) prunes distant quartets the
    way a converged density does."""
    rng = np.random.default_rng(seed)
    centre = np.zeros(nbasis, np.int64)
    for s in range(len(katom)):
        centre[ksumtype[s] - 1:ksumtype[s + 1] - 1] = katom[s] - 1
    r = np.asarray(xyz_bohr)[:, centre].T
    R2 = ((r[:, None, :] - r[None, :, :]) ** 2).sum(-1)
    u = rng.uniform(-1.0, 1.0, (nbasis, nbasis))
    D = 0.6 * np.exp(-0.25 * R2) * (u + u.T) / 2.0
    D[np.diag_indices(nbasis)] += 0.5 + 0.5 * rng.random(nbasis)
    return D


def to_quick_input(symbols, xyz_ang, keywords: str) -> str:
    """A QUICK input deck (keyword card, blank line, atoms in Angstrom)."""
    lines = [keywords, ""]
    for s, (x, y, z) in zip(symbols, xyz_ang):
        lines.append(f" {s:2s} {x:20.10f} {y:20.10f} {z:20.10f}")
    lines.append("")
    return "\n".join(lines) + "\n"
