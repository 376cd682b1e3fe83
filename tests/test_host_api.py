"""CPU tests of the Python mirror's argument checks: every wrapper that hands
an array to the C ABI checks its shape first and raises ValueError, so a
wrong-sized array never reaches C (which reads exactly n*n or n*k doubles).
No GPU call is made: the checks run before the library is touched."""
import numpy as np
import pytest

from conftest import load_package


@pytest.fixture(scope="module")
def hf():
    return load_package()


def bare(cls, **attrs):
    obj = cls.__new__(cls)
    obj.h = None
    for k, v in attrs.items():
        setattr(obj, k, v)
    return obj


@pytest.mark.parametrize("shape", [(4, 4), (5,), (5, 4), (25,), (5, 5, 1)])
def test_context_rejects_misshapen_densities(hf, shape):
    ctx = bare(hf.Context, n=5)
    bad = np.zeros(shape)
    with pytest.raises(ValueError):
        ctx.contract_jk(bad)
    with pytest.raises(ValueError):
        ctx.set_density(bad)
    with pytest.raises(ValueError):
        ctx.build_fock(np.zeros((5, 5)), bad)
    with pytest.raises(ValueError):
        ctx.build_fock(bad, np.zeros((5, 5)))


def test_df_tensor_rejects_misshapen_orbitals(hf):
    df = bare(hf.DfTensor, n=6)
    for bad in (np.zeros((5, 2)), np.zeros(6), np.zeros((6, 0))):
        with pytest.raises(ValueError):
            df.coulomb_exchange_rhf(bad)
        with pytest.raises(ValueError):
            df.jk_precision(bad)
    with pytest.raises(ValueError):
        df.jk_precision(np.zeros((6, 2)), "fp16")


def test_batch_rejects_misshapen_density_and_index(hf):
    class B:
        n = 3
    b = bare(hf.Batch, bases=[B(), B()])
    with pytest.raises(ValueError):
        b.set_density(0, np.zeros((4, 4)))
    with pytest.raises(ValueError):
        b.set_density(2, np.zeros((3, 3)))


def test_atoms_are_reshaped_or_rejected(hf):
    z = hf._atoms([8, 0, 0, 0, 1, 0, 1, 0])
    assert z.shape == (2, 4)
    for bad in ([8, 0, 0], np.zeros((2, 3)), np.zeros((0, 4))):
        with pytest.raises(ValueError):
            hf._atoms(bad)


def test_scf_rejects_unknown_convergence_algorithm(hf):
    class B:
        n = 7
        view = None
    with pytest.raises(ValueError):
        hf.scf_rhf(B(), np.zeros((3, 4)), conv="newton")
