"""Oracle J and K of a seeded density for the (H2O)16/cc-pVDZ workload.

The oracle's direct J/K (contract_jk of src/scf.rs:337-364 over every
screened shell quartet of src/shell_integrals.rs:156-240, Schwarz 1e-12) on
the seeded 4x2x2 water cluster (2.9 A, PCG64 seed 2; 400 Cartesian
functions), for the density D = C C^T with C = 0.2 * standard_normal((400,
40)) from PCG64 seed 11.  Stores the upper triangles of J and K (both are
symmetric), so that the GPU test compares the full-size build against the
oracle without the oracle's minutes of CPU time.

Usage:  python tests/golden/make_h2o16_jk.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "hf-rs_amd"))
import oracle as O  # noqa: E402
import molecules as M  # noqa: E402

SEED = 11


def density(n):
    c = np.random.default_rng(SEED).standard_normal((n, 40)) * 0.2
    return c @ c.T


def main():
    mol = M.water_cluster((4, 2, 2), seed=2)
    b = O.Basis.build(M.basis_text("cc-pvdz"), mol.symbols, mol.xyz)
    d = density(b.n)
    t0 = time.time()
    j, k, _ = b.direct_jk(d)
    iu = np.triu_indices(b.n)
    np.savez_compressed(os.path.join(HERE, "h2o16_ccpvdz_jk.npz"), n=b.n, seed=SEED, j_upper=j[iu], k_upper=k[iu])
    print("n %d, %.1f s, |J| %.3g, |K| %.3g" % (b.n, time.time() - t0, np.abs(j).max(), np.abs(k).max()))


if __name__ == "__main__":
    main()
