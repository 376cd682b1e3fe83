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
Z = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10,
     "Na": 11, "Mg": 12, "Al": 13, "Si": 14, "P": 15, "S": 16, "Cl": 17, "Ar": 18}
BASIS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "basis")
BASIS_FILES = {
    "sto-3g": "sto-3g.g94", "6-31g": "6-31g.g94", "6-31g*": "6-31g_star.g94", "def2-svp": "def2-svp.g94", "cc-pvdz": "cc-pvdz.g94",
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
    """All-trans CnH2n+2: planar C zig-zag with C-C 1.54 A and C-C-C
    109.471 deg, C-H 1.09 A, every carbon tetrahedral.  A terminal carbon gets
    a hydrogen where the zig-zag would continue."""
    cc, ch = 1.54 * ANGSTROM_TO_BOHR, 1.09 * ANGSTROM_TO_BOHR
    phi = (math.pi - math.acos(-1.0 / 3.0)) / 2  # bond direction against the chain axis
    bond = [np.array([math.cos(phi), math.sin(phi), 0.0]), np.array([math.cos(phi), -math.sin(phi), 0.0])]
    carbons = [np.zeros(3)]
    for i in range(n - 1):
        carbons.append(carbons[-1] + cc * bond[i % 2])
    symbols = ["C"] * n
    xyz = list(carbons)
    s = math.sqrt(2.0 / 3.0)
    for i, c in enumerate(carbons):
        u_prev = -bond[(i - 1) % 2]  # toward C(i-1), or where it would be
        u_next = bond[i % 2]         # toward C(i+1), or where it would be
        if i == 0:
            symbols.append("H")
            xyz.append(c + ch * u_prev)
        if i == n - 1:
            symbols.append("H")
            xyz.append(c + ch * u_next)
        normal = np.cross(u_prev, u_next)
        normal /= np.linalg.norm(normal)
        mid = -(u_prev + u_next) / 2
        for sign in (1.0, -1.0):
            symbols.append("H")
            xyz.append(c + ch * (mid + sign * s * normal))
    return Molecule(symbols, xyz, name="C%dH%d" % (n, 2 * n + 2))


def batch_config3(count=512, seed=3):
    """The batched workload: half all-trans alkanes CnH2n+2 with n ~ U{1..12},
    half water clusters (H2O)m with m ~ U{1..8} on the smallest near-cubic
    lattice, in a seeded order (PCG64)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    mols = []
    for i in range(count):
        if i % 2 == 0:
            mols.append(alkane(int(rng.integers(1, 13))))
        else:
            mols.append(water_cluster_m(int(rng.integers(1, 9)), seed=int(rng.integers(0, 2 ** 31))))
    return mols
