"""The one-electron matrices and the SAD density on the device
(hfg_one_electron and hfg_sad_density: compute_one_electron_integrals,
src/scf.rs:246-271, and compute_sad_density, src/sad.rs:217-259) against the
oracle (1e-12 of the largest element; the SAD density 1e-8, the convergence
threshold of the atomic solves, at which tests/sad_test.rs pins them too) on
water/STO-3G, a water
dimer in cc-pVDZ (generally contracted shells) and methane/def2-SVP (d
shells), against closed forms for s and p Gaussians, and the reference's own
SAD checks (tests/sad_test.rs:32-54: tr(P S) equals the electron count of a
cation to 1e-12, translation invariance)."""
import numpy as np
import pytest

import oracle as O
import molecules as M
from conftest import load_package
from test_oracle_golden import TWO_CENTRE_SP, sp_closed_forms

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hf():
    mod = load_package()
    if mod.device_count() < 1:
        pytest.fail("no GPU visible to the HIP engine")
    return mod


CASES = {
    "water_sto3g": (lambda: M.water_config1(), "sto-3g"),
    "water_dimer_ccpvdz": (lambda: M.water_cluster((2, 1, 1), seed=2), "cc-pvdz"),
    "methane_def2svp": (lambda: M.alkane(1), "def2-svp"),
    "water4_631gs": (lambda: M.water_cluster((2, 2, 1), seed=4), "6-31g*"),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_one_electron_matrices_match_the_oracle(hf, case):
    make, bname = CASES[case]
    mol = make()
    text = M.basis_text(bname)
    b = hf.Basis.from_g94(text, mol.symbols, mol.xyz)
    ob = O.Basis.build(text, mol.symbols, mol.xyz)
    ctx = hf.Context(b)
    s, t, v = ctx.one_electron(mol.zxyz)
    ctx.close()
    so, to, vo = ob.one_electron(mol.zxyz)
    for name, x, y in (("S", s, so), ("T", t, to), ("V", v, vo)):
        err = np.abs(x - y).max()
        print("%s %s n=%d max|d| %.2e (max %.2e)" % (case, name, b.n, err, np.abs(y).max()))
        assert err < 1e-12 * max(1.0, np.abs(y).max()), (name, err)
        assert np.array_equal(x, x.T)


def test_one_electron_matrices_match_closed_forms(hf):
    xyz = [[0.1, -0.2, 0.3], [1.0, 0.4, -0.5]]
    b = hf.Basis.from_g94(TWO_CENTRE_SP, ["H", "H"], xyz)
    zxyz = np.array([[1.0] + xyz[0], [1.0] + xyz[1]])
    ctx = hf.Context(b)
    s, t, v = ctx.one_electron(zxyz)
    ctx.close()
    ref = sp_closed_forms(xyz)
    assert np.abs(np.diag(s) - 1.0).max() < 1e-14
    for j, (sv, tv, vv) in ((0, ref["s_aa"]), (4, ref["s_ab"])):
        assert abs(s[0, j] - sv) < 1e-14 and abs(t[0, j] - tv) < 1e-13 and abs(v[0, j] - vv) < 1e-13
    assert abs(s[1, 4] - ref["p_x_A_s_B"]) < 1e-14
    assert abs(t[1, 1] - ref["t_pp"]) < 1e-13 and abs(t[3, 3] - ref["t_pp"]) < 1e-13


@pytest.mark.parametrize("case,charge", [("water_sto3g", 0), ("water_dimer_ccpvdz", 1), ("methane_def2svp", 2)])
def test_sad_density_matches_the_oracle_and_the_reference_checks(hf, case, charge):
    make, bname = CASES[case]
    mol = make()
    text = M.basis_text(bname)
    b = hf.Basis.from_g94(text, mol.symbols, mol.xyz)
    ob = O.Basis.build(text, mol.symbols, mol.xyz)
    ctx = hf.Context(b)
    p = ctx.sad_density(mol.zxyz, charge=charge)
    s = ctx.one_electron(mol.zxyz)[0]
    ctx.close()
    so = ob.one_electron(mol.zxyz)[0]
    po = O.sad_density(ob, mol.zxyz, so, charge)
    nel = mol.zxyz[:, 0].sum() - charge
    print("%s charge %d max|dP| %.2e tr(PS)-N %.1e" % (case, charge, np.abs(p - po).max(), (p * s).sum() - nel))
    assert np.abs(p - po).max() < 1e-8
    assert abs((p * s).sum() - nel) < 1e-12 * nel
    # translation invariance (tests/sad_test.rs:44-53)
    shift = np.array([1.3, 2.6, 3.9])
    b2 = hf.Basis.from_g94(text, mol.symbols, mol.xyz + shift)
    z2 = np.column_stack([mol.zxyz[:, 0], mol.xyz + shift])
    ctx = hf.Context(b2)
    p2 = ctx.sad_density(z2, charge=charge)
    ctx.close()
    assert np.abs(p2 - p).max() < 1e-12


def test_one_electron_and_sad_reject_bad_input(hf):
    mol = M.water_config1()
    b = hf.Basis.from_g94(M.basis_text("sto-3g"), mol.symbols, mol.xyz)
    ctx = hf.Context(b)
    with pytest.raises(hf.HfError) as e:
        ctx.sad_density(mol.zxyz[:2])  # the basis names an atom beyond the geometry
    assert e.value.code == hf.HFG_INVALID_INPUT
    with pytest.raises(hf.HfError) as e:
        ctx.sad_density(mol.zxyz, charge=10)  # no electrons left
    assert e.value.code == hf.HFG_INVALID_INPUT
    bad = mol.zxyz.copy()
    bad[0, 1] = np.nan
    with pytest.raises(hf.HfError) as e:
        ctx.one_electron(bad)
    assert e.value.code == hf.HFG_INVALID_INPUT
    ctx.close()
