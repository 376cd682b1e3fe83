"""Oracle RHF energy and iteration count for the (H2O)16/cc-pVDZ workload.

Runs the CPU oracle (in-core packed tensor and contract_jk, the reference's own
algorithm, src/scf.rs:410-627) on the seeded 4x2x2 water cluster used by
bench.py, with the same options as the engine's SCF there: spherical
orthogonaliser, tol 1e-10, ADIIS->CDIIS.  Needs ~27 GB of host memory.

Usage:  OMP_NUM_THREADS=6 python tests/golden/make_h2o16_reference.py
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "hf-rs_amd"))
import oracle as O  # noqa: E402
import molecules as M  # noqa: E402


def main():
    mol = M.water_cluster((4, 2, 2), seed=2)
    b = O.Basis.build(M.basis_text("cc-pvdz"), mol.symbols, mol.xyz)
    t0 = time.time()
    res, _, _, eps = O.scf_rhf(b, mol.zxyz, tol=1e-10, conv="adiissoscf", spherical=True, direct=False,
                               verbose=True)
    out = {"source": "oracle/hfo.cpp in-core RHF (packed tensor + contract_jk), FP64",
           "workload": "(H2O)16 cc-pVDZ, 4x2x2 lattice 2.9 A, PCG64 seed 2; spherical X; tol 1e-10; adiissoscf",
           "n_cart": b.n, "total_energy": res.total_energy, "iterations": res.iterations,
           "converged": bool(res.converged), "fock_builds": res.fock_builds,
           "seconds_eri": res.seconds_eri, "seconds_total": time.time() - t0,
           "homo": float(eps[res.n_occ - 1]), "lumo": float(eps[res.n_occ])}
    json.dump(out, open(os.path.join(HERE, "h2o16_ccpvdz_scf.json"), "w"), indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
