"""Molecules and basis-set files for the engine's workloads.

Geometries are in bohr (1 angstrom = 1/0.529177210903 bohr, src/masses.rs:28-31).
The synthetic clusters follow the workload definitions: water molecules with
r(OH)=0.9572 A and angle(HOH)=104.52 deg, centred on a cubic lattice, each given
a uniform random rotation (normalised 4-D Gaussian quaternion) and a uniform
+-0.1 A jitter of its centre, all drawn from numpy's PCG64 with a fixed seed.
"""
import math
import os

import numpy as np

ANGSTROM_TO_BOHR = 1.0 / 0.529177210903
Z = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10}
BASIS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "basis")
BASIS_FILES = {
    "sto-3g": "sto-3g.g94", "6-31g*": "6-31g_star.g94", "def2-svp": "def2-svp.g94", "cc-pvdz": "cc-pvdz.g94",
    "cc-pvtz": "cc-pvtz.g94", "def2-universal-jkfit": "def2-universal-jkfit.g94",
    "def2-svp-rifit": "def2-svp-rifit.g94", "def2-svp-ri": "def2-svp-rifit.g94",
}


def basis_text(name):
    key = name.strip().lower()
    if key not in BASIS_FILES:
        raise KeyError("basis %r is not shipped (have: %s)" % (name, ", ".join(sorted(BASIS_FILES))))
    with open(os.path.join(BASIS_DIR, BASIS_FILES[key])) as f:
        return f.read()


class Molecule:
    def __init__(self, symbols, xyz_bohr, charge=0, name=""):
        self.symbols = list(symbols)
        self.xyz = np.asarray(xyz_bohr, dtype=np.float64).reshape(-1, 3)
        self.charge = charge
        self.name = name

    @property
    def zxyz(self):
        return np.column_stack([np.array([Z[s] for s in self.symbols], float), self.xyz])

    def n_electrons(self):
        return sum(Z[s] for s in self.symbols) - self.charge


def water_config1():
    """H2O at the fixed geometry of the STO-3G in-core case (bohr)."""
    return Molecule(["O", "H", "H"], [[0, 0, 0.2217], [0, 1.4309, -0.8867], [0, -1.4309, -0.8867]], name="h2o")


def _water_local():
    r = 0.9572 * ANGSTROM_TO_BOHR
    half = math.radians(104.52) / 2
    return np.array([[0.0, 0.0, 0.0], [r * math.sin(half), 0.0, r * math.cos(half)],
                     [-r * math.sin(half), 0.0, r * math.cos(half)]])


def _rotation(q):
    q = q / np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def water_cluster(grid, spacing_angstrom=2.9, seed=2, name=None):
    """(H2O)_m on an nx*ny*nz lattice, seeded rotations and jitter."""
    nx, ny, nz = grid
    rng = np.random.Generator(np.random.PCG64(seed))
    local = _water_local()
    symbols, xyz = [], []
    for ix in range(nx):
        for iy in range(ny):
            for iz in range(nz):
                centre = np.array([ix, iy, iz], float) * spacing_angstrom * ANGSTROM_TO_BOHR
                rot = _rotation(rng.standard_normal(4))
                jitter = rng.uniform(-0.1, 0.1, 3) * ANGSTROM_TO_BOHR
                for s, p in zip("OHH", local):
                    symbols.append(s)
                    xyz.append(centre + jitter + rot @ p)
    m = nx * ny * nz
    return Molecule(symbols, xyz, name=name or "(H2O)%d" % m)


def lattice_for(m):
    """Smallest near-cubic lattice holding m molecules (first m sites used)."""
    k = 1
    while k ** 3 < m:
        k += 1
    return k


def water_cluster_m(m, seed, spacing_angstrom=2.9):
    k = lattice_for(m)
    full = water_cluster((k, k, k), spacing_angstrom, seed)
    return Molecule(full.symbols[:3 * m], full.xyz[:3 * m], name="(H2O)%d" % m)


def alkane(n):
    """All-trans CnH2n+2 zig-zag: C-C 1.54 A, C-H 1.09 A, tetrahedral angles."""
    cc, ch = 1.54 * ANGSTROM_TO_BOHR, 1.09 * ANGSTROM_TO_BOHR
    theta = math.acos(-1.0 / 3.0)
    dx = cc * math.sin(theta / 2)
    dy = cc * math.cos(theta / 2)
    carbons = [np.array([i * dx, (i % 2) * dy, 0.0]) for i in range(n)]
    symbols, xyz = [], []
    for i, c in enumerate(carbons):
        symbols.append("C")
        xyz.append(c)
    for i, c in enumerate(carbons):
        up = 1.0 if i % 2 == 0 else -1.0
        # two out-of-plane hydrogens per carbon, bisecting the C-C-C plane
        out = ch * math.cos(theta / 2)
        inplane = ch * math.sin(theta / 2)
        base = c + np.array([0.0, -up * out * 0.0, 0.0])
        for s in (1.0, -1.0):
            symbols.append("H")
            xyz.append(base + np.array([0.0, -up * out, s * inplane]))
        if i == 0:
            symbols.append("H")
            xyz.append(c + ch * np.array([-math.sin(theta / 2), up * math.cos(theta / 2), 0.0]) *
                       np.array([1.0, -1.0, 1.0]))
        if i == n - 1 and n > 1:
            d = carbons[i] - carbons[i - 1]
            symbols.append("H")
            xyz.append(c + ch * d / np.linalg.norm(d))
    if n == 1:
        symbols, xyz = ["C"], [carbons[0]]
        t = ch / math.sqrt(3.0)
        for v in ((1, 1, 1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1)):
            symbols.append("H")
            xyz.append(carbons[0] + t * np.array(v, float))
    return Molecule(symbols, xyz, name="C%dH%d" % (n, 2 * n + 2))
