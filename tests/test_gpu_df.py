"""Density fitting on the GPU vs the oracle and the PySCF fixtures
(tests/df_hf_test.rs:29-163 style): J/K for given occupied coefficients,
DF-RHF energy and every J, K, F, P element at 1e-9."""
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


def water_fixture():
    ref = json.load(open(os.path.join(GOLDEN, "df_hf_water.json")))["water"]
    syms = [a[0] for a in ref["atoms"]]
    xyz = np.array([a[1] for a in ref["atoms"]])
    return ref, syms, xyz


@pytest.mark.parametrize("obas,abas,grid", [("sto-3g", "def2-svp-rifit", None),
                                            ("cc-pvdz", "def2-universal-jkfit", (2, 1, 1))])
def test_df_jk_matches_oracle(hf, obas, abas, grid):
    if grid is None:
        _, syms, xyz = water_fixture()
    else:
        mol = M.water_cluster(grid, seed=2)
        syms, xyz = mol.symbols, mol.xyz
    b = hf.Basis.from_g94(M.basis_text(obas), syms, xyz)
    a = hf.Basis.from_g94(M.basis_text(abas), syms, xyz)
    df = hf.DfTensor(b, a)
    ob = O.Basis.build(M.basis_text(obas), syms, xyz)
    oa = O.Basis.build(M.basis_text(abas), syms, xyz)
    bt = O.df_build(ob, oa)
    assert df.rank == bt.shape[0]
    rng = np.random.default_rng(4)
    c = rng.standard_normal((b.n, 5)) * 0.3
    j, k = df.coulomb_exchange_rhf(c)
    jr, kr = O.df_jk(bt, b.n, c)
    print("%s/%s n=%d rank=%d  J %.2e K %.2e" % (obas, abas, b.n, df.rank, np.abs(j - jr).max(), np.abs(k - kr).max()))
    assert np.abs(j - jr).max() < 1e-10 and np.abs(k - kr).max() < 1e-10


def test_df_rhf_water_matches_pyscf_every_element(hf):
    ref, syms, xyz = water_fixture()
    b = hf.Basis.from_g94(M.basis_text("sto-3g"), syms, xyz)
    a = hf.Basis.from_g94(M.basis_text("def2-svp-rifit"), syms, xyz)
    zxyz = np.column_stack([[M.Z[s] for s in syms], xyz])
    r = hf.df_rhf(b, a, zxyz)
    assert r["converged"] and r["n_aux"] == int(ref["n_aux"])
    assert abs(r["total_energy"] - float(ref["e_scf"])) < 1e-9
    for name, key in (("j", "j"), ("k", "k"), ("fock", "fock"), ("density", "density")):
        assert np.abs(r[key] - np.array(ref[name])).max() < 1e-9, name


def test_df_rhf_water_dimer_matches_oracle(hf):
    mol = M.water_cluster((2, 1, 1), seed=2)
    ob, oa = M.basis_text("cc-pvdz"), M.basis_text("def2-universal-jkfit")
    r = hf.df_rhf(hf.Basis.from_g94(ob, mol.symbols, mol.xyz), hf.Basis.from_g94(oa, mol.symbols, mol.xyz), mol.zxyz)
    o, _, _, _, _ = O.df_rhf(O.Basis.build(ob, mol.symbols, mol.xyz), O.Basis.build(oa, mol.symbols, mol.xyz),
                             mol.zxyz)
    print("DF dimer: E %.12f vs %.12f, iters %d vs %d" % (r["total_energy"], o.total_energy, r["iterations"],
                                                         o.iterations))
    assert r["converged"] and o.converged
    assert abs(r["total_energy"] - o.total_energy) < 1e-8
    assert abs(r["iterations"] - o.iterations) <= 1


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_df_rhf_mixed_precision_reaches_the_fp64_energy(hf, prec):
    """Reduced-precision exchange in the early iterations, FP64 at the end:
    the energy is the FP64 oracle's to 1e-8 Eh (no reference counterpart; the
    reference is FP64 throughout)."""
    mol = M.water_cluster((2, 1, 1), seed=2)
    ob, oa = M.basis_text("cc-pvdz"), M.basis_text("def2-universal-jkfit")
    b, a = hf.Basis.from_g94(ob, mol.symbols, mol.xyz), hf.Basis.from_g94(oa, mol.symbols, mol.xyz)
    r = hf.df_rhf(b, a, mol.zxyz, k_precision=prec)
    o, _, _, _, _ = O.df_rhf(O.Basis.build(ob, mol.symbols, mol.xyz), O.Basis.build(oa, mol.symbols, mol.xyz),
                             mol.zxyz)
    print("DF dimer %s: E %.12f vs %.12f, iters %d vs %d" % (prec, r["total_energy"], o.total_energy,
                                                            r["iterations"], o.iterations))
    assert r["converged"]
    assert abs(r["total_energy"] - o.total_energy) < 1e-8
    assert r["iterations"] <= o.iterations + 4


def test_df_rejects_bad_inputs(hf):
    _, syms, xyz = water_fixture()
    b = hf.Basis.from_g94(M.basis_text("sto-3g"), syms, xyz)
    a = hf.Basis.from_g94(M.basis_text("def2-svp-rifit"), syms, xyz)
    with pytest.raises(hf.HfError):
        hf.DfTensor(b, a, metric_cutoff=-1.0)
    df = hf.DfTensor(b, a)
    with pytest.raises(hf.HfError):
        df.coulomb_exchange_rhf(np.full((b.n, 2), np.nan))
    zxyz = np.column_stack([[M.Z[s] for s in syms], xyz])
    with pytest.raises(hf.HfError):
        hf.df_rhf(b, a, zxyz, charge=1)  # odd electron count
    with pytest.raises(hf.HfError):
        hf.df_rhf(b, a, zxyz, max_iter=1)  # cannot converge in one iteration


def test_df_k_reduced_precision_error_bounds(hf):
    """DF-K on the matrix cores in FP32 and in split BF16 (three products,
    FP32 accumulation) against the FP64 contraction: J unchanged, K within
    the stated relative bounds (FP32 1e-5, BF16x3 1e-3 of max|K|)."""
    mol = M.water_cluster((2, 2, 1), seed=2)
    b = hf.Basis.from_g94(M.basis_text("cc-pvdz"), mol.symbols, mol.xyz)
    a = hf.Basis.from_g94(M.basis_text("def2-universal-jkfit"), mol.symbols, mol.xyz)
    df = hf.DfTensor(b, a)
    rng = np.random.default_rng(12)
    c = rng.standard_normal((b.n, 20)) * 0.2
    j0, k0, _ = df.jk_precision(c, "fp64")
    jr, kr = df.coulomb_exchange_rhf(c)
    assert np.abs(j0 - jr).max() < 1e-13 * max(1, np.abs(jr).max()) and np.abs(k0 - kr).max() < 1e-13 * max(1, np.abs(kr).max())
    scale = np.abs(k0).max()
    for prec, bound in (("fp32", 1e-5), ("bf16x3", 1e-3)):
        j, k, ms = df.jk_precision(c, prec)
        err = np.abs(k - k0).max() / scale
        print("DF-K %s: max|dK|/max|K| = %.2e, %.3f ms" % (prec, err, ms))
        assert np.abs(j - j0).max() < 1e-13 * max(1, np.abs(j0).max())
        assert err < bound
        assert np.allclose(k, k.T, rtol=0, atol=bound * scale)
