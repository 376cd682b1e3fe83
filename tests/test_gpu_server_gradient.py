"""Gradients of served trajectories (hfg_server_submit with
HFG_SERVE_GRADIENT, hfg_server_gradient; compute_rhf_gradient,
src/gradient.rs:272-285, of every geometry step): the first step of each
fixture molecule against PySCF's gradient (tests/scf_reference_test.rs:89,
1e-6), the last step against the single-molecule hfg_rhf_gradient of the
returned density with W = 1/2 P F P rebuilt from it (1e-9), and forces
summing to zero (1e-9), with more requests than slots."""
import json
import os

import numpy as np
import pytest

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


@pytest.mark.skip(reason="served gradients are switched off: run after the rest of the GPU suite, this test "
                         "stopped making progress (it passed alone); the cause is not established")
def test_served_gradients_match_pyscf_and_the_single_molecule_gradient(hf):
    sysd = json.load(open(os.path.join(GOLDEN, "scf_energies.json")))["systems"]
    rng = np.random.default_rng(11)
    cases = []
    for name, basis in (("h2o", "6-31g*"), ("nh3", "sto-3g"), ("lih", "6-31g*")):
        mol = molecule(sysd[name]["geometry_bohr"])
        geoms = [mol.xyz, mol.xyz + rng.normal(0.0, 0.01, mol.xyz.shape)]
        cases.append((name, basis, mol, geoms, np.array(sysd[name]["results"][basis]["gradient"])))
    dimer = M.water_cluster((2, 1, 1), seed=2)
    cases.append(("water dimer", "cc-pvdz", dimer, [dimer.xyz, dimer.xyz + rng.normal(0.0, 0.01, dimer.xyz.shape)],
                  None))
    srv = hf.Server(capacity=2)
    ids = []
    for name, basis, mol, geoms, _ in cases:
        steps = [M.Molecule(mol.symbols, g) for g in geoms]
        bases = [hf.Basis.from_g94(M.basis_text(basis), s.symbols, s.xyz) for s in steps]
        ids.append(srv.submit(bases, [s.zxyz for s in steps], tol=1e-12, conv="cdiis", max_iter=200,
                              keep_density=True, gradients=True))
    plain = srv.submit([hf.Basis.from_g94(M.basis_text("sto-3g"), cases[0][2].symbols, cases[0][2].xyz)],
                       [cases[0][2].zxyz], tol=1e-10)
    srv.run()
    assert "gradients" not in srv.result(plain) and srv.result(plain)["ok"]
    for rid, (name, basis, mol, geoms, g_ref) in zip(ids, cases):
        r = srv.result(rid, density=True)
        assert r["ok"], r.get("error")
        g = r["gradients"]
        assert g.shape == (2, len(mol.symbols), 3)
        for k in range(2):
            assert np.abs(g[k].sum(axis=0)).max() < 1e-9
        if g_ref is not None:
            print("%s %s step 0: max|G - PySCF| %.2e" % (name, basis, np.abs(g[0] - g_ref).max()))
            assert np.abs(g[0] - g_ref).max() < 1e-6
        # the last step from its returned density, through the single-molecule entry point
        last = M.Molecule(mol.symbols, geoms[-1])
        b = hf.Basis.from_g94(M.basis_text(basis), last.symbols, last.xyz)
        ctx = hf.Context(b)
        p = r["density"]
        _, t, v = ctx.one_electron(last.zxyz)
        j, kx = ctx.contract_jk(p)
        f = t + v + j - 0.5 * kx
        w = 0.5 * p @ f @ p
        g1 = ctx.rhf_gradient(last.zxyz, p, w)
        ctx.close()
        print("%s %s last step: max|G - single| %.2e" % (name, basis, np.abs(g[1] - g1).max()))
        assert np.abs(g[1] - g1).max() < 1e-9


def test_served_gradients_are_refused_while_switched_off(hf):
    mol = M.water_config1()
    srv = hf.Server(capacity=1)
    with pytest.raises(hf.HfError) as e:
        srv.submit([hf.Basis.from_g94(M.basis_text("sto-3g"), mol.symbols, mol.xyz)], [mol.zxyz], gradients=True)
    assert e.value.code == hf.HFG_UNSUPPORTED
    srv.close()
