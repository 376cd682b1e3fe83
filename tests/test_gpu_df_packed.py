"""Density fitting with orbital f shells (cc-pVTZ) and the packed factor
layout (src/df.rs:60-127: B kept as [rank][n(n+1)/2], formed one column
chunk at a time) against the oracle: DF J/K of the water dimer in cc-pVTZ
with def2-universal-JKFIT to 1e-9 per element in both layouts (the reference's
own DF element bound, tests/df_hf_test.rs; at cc-pVTZ the metric's smallest
kept eigenvalues amplify the integrals' last-digit differences to ~1e-10),
the packed layout equal
to the unpacked one (FP64 1e-12; FP32 and split-BF16 exchange within their
stated bounds of the FP64 result), and the DF-RHF energy of water in cc-pVTZ
within 1e-8 Eh of the oracle's with the iteration count within 1, also
with split-BF16 exchange before the final FP64 builds (water dimer).
HFG_DF_PACKED=1 selects the packed layout for these small systems; the
engine picks it by itself when the unpacked tensors would not fit in HBM."""
import os
from contextlib import contextmanager

import numpy as np
import pytest

import oracle as O
import molecules as M
from conftest import load_package

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hf():
    mod = load_package()
    if mod.device_count() < 1:
        pytest.fail("no GPU visible to the HIP engine")
    return mod


@contextmanager
def layout(packed):
    old = os.environ.get("HFG_DF_PACKED")
    os.environ["HFG_DF_PACKED"] = "1" if packed else "0"
    try:
        yield
    finally:
        if old is None:
            del os.environ["HFG_DF_PACKED"]
        else:
            os.environ["HFG_DF_PACKED"] = old


@pytest.fixture(scope="module")
def dimer_tz():
    mol = M.water_cluster((2, 1, 1), seed=2)
    ob = O.Basis.build(M.basis_text("cc-pvtz"), mol.symbols, mol.xyz)
    oa = O.Basis.build(M.basis_text("def2-universal-jkfit"), mol.symbols, mol.xyz)
    return mol, O.df_build(ob, oa)


@pytest.mark.parametrize("packed", [False, True])
def test_df_jk_with_f_shells_matches_oracle(hf, dimer_tz, packed):
    mol, bt = dimer_tz
    b = hf.Basis.from_g94(M.basis_text("cc-pvtz"), mol.symbols, mol.xyz)
    a = hf.Basis.from_g94(M.basis_text("def2-universal-jkfit"), mol.symbols, mol.xyz)
    with layout(packed):
        df = hf.DfTensor(b, a)
    assert df.rank == bt.shape[0]
    c = np.random.default_rng(6).standard_normal((b.n, 10)) * 0.2
    j, k = df.coulomb_exchange_rhf(c)
    jr, kr = O.df_jk(bt, b.n, c)
    print("cc-pVTZ dimer packed=%d n=%d rank=%d  J %.2e K %.2e" % (packed, b.n, df.rank, np.abs(j - jr).max(),
                                                                  np.abs(k - kr).max()))
    assert np.abs(j - jr).max() < 1e-9 and np.abs(k - kr).max() < 1e-9
    df.close()


def test_packed_layout_equals_unpacked(hf):
    mol = M.water_cluster((2, 2, 1), seed=2)
    b = hf.Basis.from_g94(M.basis_text("cc-pvdz"), mol.symbols, mol.xyz)
    a = hf.Basis.from_g94(M.basis_text("def2-universal-jkfit"), mol.symbols, mol.xyz)
    c = np.random.default_rng(12).standard_normal((b.n, 20)) * 0.2
    with layout(False):
        du = hf.DfTensor(b, a)
    with layout(True):
        dp = hf.DfTensor(b, a)
    j0, k0, _ = du.jk_precision(c, "fp64")
    j1, k1, _ = dp.jk_precision(c, "fp64")
    print("packed vs unpacked FP64: J %.2e K %.2e" % (np.abs(j1 - j0).max(), np.abs(k1 - k0).max()))
    assert np.abs(j1 - j0).max() < 1e-12 * max(1, np.abs(j0).max())
    assert np.abs(k1 - k0).max() < 1e-12 * max(1, np.abs(k0).max())
    scale = np.abs(k0).max()
    for prec, bound in (("fp32", 1e-5), ("bf16x3", 1e-3)):
        j, k, _ = dp.jk_precision(c, prec)
        err = np.abs(k - k0).max() / scale
        print("packed DF-K %s: max|dK|/max|K| = %.2e" % (prec, err))
        assert np.abs(j - j0).max() < 1e-12 * max(1, np.abs(j0).max())
        assert err < bound
    du.close()
    dp.close()


@pytest.mark.parametrize("packed", [False, True])
def test_df_rhf_cc_pvtz_water_matches_oracle(hf, packed):
    mol = M.water_config1()
    ob, oa = M.basis_text("cc-pvtz"), M.basis_text("def2-universal-jkfit")
    with layout(packed):
        r = hf.df_rhf(hf.Basis.from_g94(ob, mol.symbols, mol.xyz), hf.Basis.from_g94(oa, mol.symbols, mol.xyz),
                      mol.zxyz)
    o, _, _, _, _ = O.df_rhf(O.Basis.build(ob, mol.symbols, mol.xyz), O.Basis.build(oa, mol.symbols, mol.xyz),
                             mol.zxyz)
    print("DF-RHF water cc-pVTZ packed=%d: E %.12f vs %.12f, iters %d vs %d" % (
        packed, r["total_energy"], o.total_energy, r["iterations"], o.iterations))
    assert r["converged"] and o.converged
    assert abs(r["total_energy"] - o.total_energy) < 1e-8
    assert abs(r["iterations"] - o.iterations) <= 1


def test_df_rhf_packed_split_bf16_reaches_the_fp64_energy(hf):
    """Packed factor with split-BF16 exchange (the packed split-BF16 A = B C
    kernel) in the early iterations, FP64 builds at the end: the water-dimer
    cc-pVTZ energy equals the oracle's FP64 DF-RHF energy to 1e-8 Eh (no
    reference counterpart; the reference is FP64 throughout)."""
    mol = M.water_cluster((2, 1, 1), seed=2)
    ob, oa = M.basis_text("cc-pvtz"), M.basis_text("def2-universal-jkfit")
    with layout(True):
        r = hf.df_rhf(hf.Basis.from_g94(ob, mol.symbols, mol.xyz), hf.Basis.from_g94(oa, mol.symbols, mol.xyz),
                      mol.zxyz, k_precision="bf16x3")
    o, _, _, _, _ = O.df_rhf(O.Basis.build(ob, mol.symbols, mol.xyz), O.Basis.build(oa, mol.symbols, mol.xyz),
                             mol.zxyz)
    print("packed DF-RHF dimer cc-pVTZ bf16x3: E %.12f vs %.12f, iters %d vs %d" % (
        r["total_energy"], o.total_energy, r["iterations"], o.iterations))
    assert r["converged"] and o.converged
    assert abs(r["total_energy"] - o.total_energy) < 1e-8
    assert r["iterations"] <= o.iterations + 4
