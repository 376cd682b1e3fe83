"""AO -> MO transformation of the ERIs on the GPU (hfg_eri_mo, replacing
transform_eri_ao_to_mo, src/mo_integrals.rs:156-265), against the textbook
four-index contraction (pq|rs) = sum C_mp C_nq C_lr C_ss (mn|ls) done by
numpy on the oracle's packed tensor (brute force on small bases), and two
properties: with C = I the result is the AO tensor itself, and (pq|rs) has
the 8-fold permutation symmetry."""
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


def unpack(packed, n):
    idx = np.zeros((n, n), dtype=np.int64)
    for i in range(n):
        for j in range(n):
            idx[i, j] = i * (i + 1) // 2 + j if i >= j else j * (j + 1) // 2 + i
    np_ = n * (n + 1) // 2
    pij = idx.reshape(-1)
    a = np.minimum.outer(pij, pij)
    b = np.maximum.outer(pij, pij)
    return packed[b * (b + 1) // 2 + a].reshape(n, n, n, n)


def mo_tensor(out, nmo):
    mp = nmo * (nmo + 1) // 2
    full = out.reshape(mp, mp)
    return unpack_pairs(full, nmo)


def unpack_pairs(full, nmo):
    idx = np.zeros((nmo, nmo), dtype=np.int64)
    for p in range(nmo):
        for q in range(nmo):
            idx[p, q] = p * (p + 1) // 2 + q if p >= q else q * (q + 1) // 2 + p
    return full[np.ix_(idx.reshape(-1), idx.reshape(-1))].reshape(nmo, nmo, nmo, nmo).transpose(2, 3, 0, 1)


@pytest.mark.parametrize("basis,grid,nmo_frac", [("sto-3g", (1, 1, 1), 1.0), ("6-31g*", (1, 1, 1), 0.6),
                                                 ("sto-3g", (2, 1, 1), 1.0)])
def test_mo_integrals_match_the_four_index_contraction(hf, basis, grid, nmo_frac):
    mol = M.water_cluster(grid, seed=2)
    text = M.basis_text(basis)
    ctx = hf.Context(hf.Basis.from_g94(text, mol.symbols, mol.xyz))
    n = ctx.n
    nmo = max(1, int(round(n * nmo_frac)))
    rng = np.random.default_rng(7)
    C = rng.standard_normal((n, nmo)) * 0.5
    got = mo_tensor(ctx.eri_mo(C), nmo)
    ao = unpack(O.Basis.build(text, mol.symbols, mol.xyz).eri_packed(), n)
    ref = np.einsum("mp,nq,mnlt,lr,ts->pqrs", C, C, ao, C, C, optimize=True)
    err = np.abs(got - ref).max()
    print("%s %s n=%d nmo=%d: max |d(pq|rs)| %.2e" % (basis, grid, n, nmo, err))
    assert err < 1e-10 * max(1.0, np.abs(ref).max())
    assert np.abs(got - got.transpose(1, 0, 2, 3)).max() == 0.0
    assert np.abs(got - got.transpose(2, 3, 0, 1)).max() < 1e-12


def test_identity_coefficients_give_the_ao_tensor(hf):
    mol = M.water_config1()
    text = M.basis_text("6-31g*")
    ctx = hf.Context(hf.Basis.from_g94(text, mol.symbols, mol.xyz))
    n = ctx.n
    got = mo_tensor(ctx.eri_mo(np.eye(n)), n)
    ao = unpack(ctx.eri_packed(), n)
    assert np.abs(got - ao).max() < 1e-13
