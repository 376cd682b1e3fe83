"""The allocation-free J/K entry point (hfg_contract_jk_device_noalloc; its
scratch -- the molecule's table, job counters, per-XCC accumulators -- in a
file-scope device array) against the allocating one (hfg_contract_jk_device)
through the C ABI, on two calls in a row, the second with a smaller molecule,
and the error path for a basis beyond the array's size."""
import ctypes as C

import numpy as np
import pytest

import molecules as M
from conftest import load_package

pytestmark = pytest.mark.gpu


class Hip:
    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def put(self, dst, arr):
        assert self.L.hipMemcpy(dst, arr.ctypes.data_as(C.c_void_p), arr.nbytes, 1) == 0  # host to device

    def get(self, src, shape):
        out = np.zeros(shape)
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), src, out.nbytes, 2) == 0  # device to host
        return out


@pytest.fixture(scope="module")
def hf():
    mod = load_package()
    if mod.device_count() < 1:
        pytest.fail("no GPU visible to the HIP engine")
    L = mod.lib()
    L.hfg_contract_jk_device_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    return mod


def jk_pair(hf, hip, ctx, p):
    n = ctx.n
    nb = n * n * 8
    dp, dj, dk, dj2, dk2 = (hip.alloc(nb) for _ in range(5))
    try:
        hip.put(dp, np.ascontiguousarray(p))
        L = hf.lib()
        assert L.hfg_contract_jk_device(ctx.h, dp, dj, dk) == 0, L.hfg_last_error()
        assert L.hfg_contract_jk_device_noalloc(ctx.h, dp, dj2, dk2) == 0, L.hfg_last_error()
        return [hip.get(x, (n, n)) for x in (dj, dk, dj2, dk2)]
    finally:
        for x in (dp, dj, dk, dj2, dk2):
            hip.L.hipFree(x)


def test_noalloc_jk_matches_the_allocating_entry_point(hf):
    hip = Hip()
    rng = np.random.default_rng(3)
    for mol, basis in ((M.water_cluster((2, 1, 1), seed=2), "cc-pvdz"), (M.alkane(2), "6-31g*")):
        ctx = hf.Context(hf.Basis.from_g94(M.basis_text(basis), mol.symbols, mol.xyz))
        c = rng.standard_normal((ctx.n, 5))
        p = c @ c.T
        j, k, j2, k2 = jk_pair(hf, hip, ctx, p)
        scale = max(1.0, np.abs(j).max())
        print("%s n=%d: |dJ| %.2e |dK| %.2e" % (basis, ctx.n, np.abs(j - j2).max(), np.abs(k - k2).max()))
        assert np.abs(j - j2).max() < 1e-12 * scale and np.abs(k - k2).max() < 1e-12 * scale
        assert np.abs(j).max() > 0


def test_noalloc_rejects_a_basis_beyond_its_scratch(hf):
    mol = M.water_cluster((2, 2, 2), seed=2)  # (H2O)8 cc-pVDZ: n = 200 <= 256; use a larger one
    mol = M.water_cluster((3, 2, 2), seed=2)  # (H2O)12: n = 300
    ctx = hf.Context(hf.Basis.from_g94(M.basis_text("cc-pvdz"), mol.symbols, mol.xyz))
    assert ctx.n > 256
    L = hf.lib()
    st = L.hfg_contract_jk_device_noalloc(ctx.h, C.c_void_p(1), C.c_void_p(1), C.c_void_p(1))
    assert st == hf.HFG_INVALID_INPUT
    assert b"at most 256" in L.hfg_last_error()
