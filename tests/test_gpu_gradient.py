"""The analytic RHF nuclear gradient on the GPU (hfg_rhf_gradient,
compute_gradient_from_density, src/gradient.rs:287-512) against PySCF's
gradients in the reference's fixture (tests/scf_reference_test.rs:74-94:
1e-6, translational invariance 1e-9) and against the oracle for the same
density (1e-9)."""
import json
import os

import numpy as np
import pytest

import oracle as O
import molecules as M
from conftest import GOLDEN, load_package

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hf():
    mod = load_package()
    if mod.device_count() < 1:
        pytest.fail("no GPU visible to the HIP engine")
    return mod


def molecule(geometry):
    rows = [r.split() for r in geometry.split(";")]
    return M.Molecule([r[0] for r in rows], [[float(x) for x in r[1:]] for r in rows])


@pytest.mark.parametrize("name,basis", [(m, b) for m in ("h2", "lih", "h2o", "nh3") for b in ("sto-3g", "6-31g*")])
def test_rhf_gradient_matches_pyscf(hf, name, basis):
    sysd = json.load(open(os.path.join(GOLDEN, "scf_energies.json")))["systems"][name]
    ref = sysd["results"][basis]
    mol = molecule(sysd["geometry_bohr"])
    b = hf.Basis.from_g94(M.basis_text(basis), mol.symbols, mol.xyz)
    r = hf.scf_rhf(b, mol.zxyz, tol=1e-12, max_iter=200, conv="cdiis", want_matrices=True)
    assert r["converged"] and abs(r["total_energy"] - ref["energy"]) < 1e-8
    g = hf.rhf_gradient(b, mol.zxyz, r["density"], r["fock"])
    dg = np.abs(g - np.array(ref["gradient"])).max()
    print("%s %s dG %.2e  sum %.1e" % (name, basis, dg, np.abs(g.sum(axis=0)).max()))
    assert dg < 1e-6
    assert np.abs(g.sum(axis=0)).max() < 1e-9


@pytest.mark.parametrize("case", ["water_dimer_ccpvdz", "methane_def2svp"])
def test_rhf_gradient_matches_the_oracle(hf, case):
    """The same density and energy-weighted density through both: every
    class of the derivative blocks, with generally contracted cc-pVDZ shells
    and carbon d shells raised to f."""
    if case == "water_dimer_ccpvdz":
        mol, basis = M.water_cluster((2, 1, 1), seed=2), "cc-pvdz"
    else:
        mol, basis = M.alkane(1), "def2-svp"
    text = M.basis_text(basis)
    b = hf.Basis.from_g94(text, mol.symbols, mol.xyz)
    ob = O.Basis.build(text, mol.symbols, mol.xyz)
    n = b.n
    c = np.random.default_rng(8).standard_normal((n, 6)) * 0.3
    p = c @ c.T
    w = c @ np.diag(np.linspace(-2.0, -0.3, 6)) @ c.T
    ctx = hf.Context(b)
    g = ctx.rhf_gradient(mol.zxyz, p, w)
    ctx.close()
    go = O.rhf_gradient(ob, mol.zxyz, p, w)
    print("%s n=%d  max|dG| %.2e  max|G| %.2e" % (case, n, np.abs(g - go).max(), np.abs(go).max()))
    assert np.abs(g - go).max() < 1e-9 * max(1.0, np.abs(go).max())


def test_rhf_gradient_rejects_bad_input(hf):
    mol = M.water_config1()
    b = hf.Basis.from_g94(M.basis_text("sto-3g"), mol.symbols, mol.xyz)
    ctx = hf.Context(b)
    p = np.eye(b.n)
    with pytest.raises(ValueError):
        ctx.rhf_gradient(mol.zxyz, p[:3], p)
    with pytest.raises(hf.HfError) as e:
        ctx.rhf_gradient(mol.zxyz[:1], p, p)  # the basis names atoms beyond the geometry
    assert e.value.code == hf.HFG_INVALID_INPUT
    ctx.close()
