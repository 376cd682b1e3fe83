"""CPU tests of the engine's host side: basis construction, the C ABI's
exported symbols, and the generated recurrences (compiled for the CPU)
against the oracle and the PySCF fixture.  No GPU call is made here."""
import ctypes as C
import json
import os
import re

import numpy as np
import pytest

import oracle as O
import molecules as M
from conftest import GOLDEN, ROOT, PKG, load_package


def header_symbols():
    text = open(os.path.join(ROOT, "include", "hfg.h")).read()
    return sorted(set(re.findall(r"\b(hfg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    hf = load_package()
    lib = hf.lib()
    syms = header_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), s
    assert set(hf.EXPORTED) <= set(syms)


@pytest.mark.parametrize("basis", ["sto-3g", "6-31g*", "cc-pvdz", "def2-svp", "cc-pvtz"])
def test_basis_builder_matches_oracle(basis):
    hf = load_package()
    mol = M.water_cluster((2, 1, 1), seed=7)
    a = hf.Basis.from_g94(M.basis_text(basis), mol.symbols, mol.xyz).a
    b = O.Basis.build(M.basis_text(basis), mol.symbols, mol.xyz).arrays()
    for k in ("atom_index", "lxyz", "prim_offset", "exponent"):
        assert np.array_equal(np.asarray(a[k]).ravel(), np.asarray(b[k]).ravel()), k
    assert np.allclose(a["center"].ravel(), b["center"].ravel(), rtol=0, atol=0)
    assert np.abs(a["coefficient"] - b["coefficient"]).max() <= 1e-13 * np.abs(b["coefficient"]).max()


def test_basis_builder_rejects_missing_element():
    hf = load_package()
    with pytest.raises(hf.HfError):
        hf.Basis.from_g94(M.basis_text("sto-3g"), ["Xx"], [[0, 0, 0]])


class View(C.Structure):
    _fields_ = [("n", C.c_int32)] + [(k, C.c_void_p) for k in ("a", "c", "l", "po", "e", "co")]


def cpucheck():
    path = os.path.join(PKG, "libhfg_cpucheck.so")
    if not os.path.exists(path):
        pytest.skip("CPU check library not built")
    L = C.CDLL(path)
    L.hfgc_boys.argtypes = [C.c_int, C.c_double, C.c_void_p]
    L.hfgc_shell_quartet.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
    L.hfgc_last_error.restype = C.c_char_p
    return L


def test_engine_boys_against_mpmath_fixture():
    L = cpucheck()
    worst = 0.0
    for row in json.load(open(os.path.join(GOLDEN, "boys.json")))["rows"]:
        out = np.zeros(25)
        assert L.hfgc_boys(24, row["t"], out.ctypes.data) == 0
        for n in range(25):
            ref = float(row["f"][n])
            worst = max(worst, abs(out[n] - ref) / ref)
    assert worst < 1e-14, worst


def test_generated_recurrences_match_pyscf_spd_blocks():
    L = cpucheck()
    meta = json.load(open(os.path.join(GOLDEN, "spdf_meta.json")))
    arr = O.Basis.build(meta["basis_gaussian94"], meta["symbols"], meta["geometry_bohr"]).arrays()
    keep = [np.ascontiguousarray(arr[k]) for k in ("atom_index", "center", "lxyz", "prim_offset", "exponent",
                                                   "coefficient")]
    v = View(len(keep[0]), *[k.ctypes.data for k in keep])
    ref = np.load(os.path.join(GOLDEN, "spdf_eri.npz"))["eri"].reshape(20, 20, 20, 20)
    ranges = [(0, 1), (1, 4), (4, 10)]
    worst = 0.0
    for blk in range(81):
        ang = [blk // 27, blk // 9 % 3, blk // 3 % 3, blk % 3]
        dims = [ranges[x][1] - ranges[x][0] for x in ang]
        out = np.zeros(int(np.prod(dims)))
        st = L.hfgc_shell_quartet(C.byref(v), *[4 * a + ang[a] for a in range(4)], out.ctypes.data)
        assert st == 0, L.hfgc_last_error()
        exp = ref[tuple(slice(*ranges[ang[a]]) for a in range(4))]
        worst = max(worst, np.abs(out.reshape(dims) - exp).max())
    assert worst < 1e-13, worst


def test_generated_recurrences_match_oracle_on_contracted_shells():
    L = cpucheck()
    mol = M.water_cluster((2, 1, 1), seed=3)
    ob = O.Basis.build(M.basis_text("cc-pvdz"), mol.symbols, mol.xyz)
    arr = ob.arrays()
    keep = [np.ascontiguousarray(arr[k]) for k in ("atom_index", "center", "lxyz", "prim_offset", "exponent",
                                                   "coefficient")]
    v = View(len(keep[0]), *[k.ctypes.data for k in keep])
    ss = ob.shells()
    ns = len(ss) - 1
    rng = np.random.default_rng(11)
    worst = 0.0
    for _ in range(150):
        q = [int(x) for x in rng.integers(0, ns, 4)]
        exp = ob.shell_quartet(*q)
        out = np.zeros(exp.size)
        assert L.hfgc_shell_quartet(C.byref(v), *q, out.ctypes.data) == 0
        worst = max(worst, np.abs(out.reshape(exp.shape) - exp).max())
    assert worst < 1e-12, worst


@pytest.mark.parametrize("compact", [0, 1])
def test_kernel_boys_path_against_mpmath_fixture(compact):
    """The kernels' per-class evaluation (full or compact LDS slice, T >= 50
    branch) for M = 0..10."""
    L = cpucheck()
    L.hfgc_boys_class.argtypes = [C.c_int, C.c_int, C.c_double, C.c_void_p]
    worst = 0.0
    for row in json.load(open(os.path.join(GOLDEN, "boys.json")))["rows"]:
        for M in range(11):
            out = np.zeros(M + 1)
            assert L.hfgc_boys_class(M, compact, row["t"], out.ctypes.data) == 0
            for m in range(M + 1):
                ref = float(row["f"][m])
                worst = max(worst, abs(out[m] - ref) / ref)
    assert worst < 5e-15, worst


@pytest.mark.parametrize("compact", [0, 1])
def test_kernel_boys_path_between_grid_rows(compact):
    """Both slice layouts on a dense T grid (every eighth between the slice
    rows, where the Taylor step is longest, up to and across T = 50) against
    the oracle's Boys function, pinned to mpmath at 1e-14 elsewhere in this
    suite: 1e-14 relative, F_0..F_M for M = 0..10."""
    L = cpucheck()
    L.hfgc_boys_class.argtypes = [C.c_int, C.c_int, C.c_double, C.c_void_p]
    ts = np.concatenate([np.arange(0.0, 52.0, 1.0 / 64), [49.999, 50.0, 50.001, 75.0, 200.0]])
    worst = 0.0
    for t in ts:
        ref = O.boys(10, float(t))
        for M in (0, 3, 6, 10):
            out = np.zeros(M + 1)
            assert L.hfgc_boys_class(M, compact, float(t), out.ctypes.data) == 0
            worst = max(worst, float(np.max(np.abs(out - ref[:M + 1]) / ref[:M + 1])))
    assert worst < 1e-14, worst


@pytest.mark.parametrize("basis,cells", [("sto-3g", (2, 1, 1)), ("6-31g*", (2, 1, 1)), ("cc-pvdz", (2, 1, 1)),
                                         ("def2-svp", (2, 1, 1)), ("cc-pvdz", (2, 2, 1))])
def test_jk_plan_replay_matches_oracle(basis, cells):
    """The J/K kernels' work plan (merged generally contracted shells, views,
    ket groups, tasks, weights, per-combination Schwarz test, primitive-pair
    screening) and the generated combination blocks, replayed on the CPU,
    against the oracle's direct J/K (src/scf.rs:337-364 on
    src/shell_integrals.rs:156-240)."""
    L = cpucheck()
    L.hfgc_jk.argtypes = [C.c_void_p] * 5
    L.hfgc_jk_last_error.restype = C.c_char_p
    mol = M.water_cluster(cells, seed=5)
    ob = O.Basis.build(M.basis_text(basis), mol.symbols, mol.xyz)
    arr = ob.arrays()
    keep = [np.ascontiguousarray(arr[k]) for k in ("atom_index", "center", "lxyz", "prim_offset", "exponent",
                                                   "coefficient")]
    v = View(len(keep[0]), *[k.ctypes.data for k in keep])
    n = ob.n
    rng = np.random.default_rng(2)
    D = rng.standard_normal((n, n))
    D = D + D.T
    J = np.zeros((n, n))
    K = np.zeros((n, n))
    stats = np.zeros(5, np.int64)
    assert L.hfgc_jk(C.byref(v), D.ctypes.data, J.ctypes.data, K.ctypes.data, stats.ctypes.data) == 0, \
        L.hfgc_jk_last_error()
    Jr, Kr, _ = ob.direct_jk(D)
    assert np.abs(J - Jr).max() < 1e-10 * max(1.0, np.abs(Jr).max())
    assert np.abs(K - Kr).max() < 1e-10 * max(1.0, np.abs(Kr).max())
    if basis == "cc-pvdz":  # O: s (9 exponents, 3 contractions), p, d; H: s (2 contractions), p
        assert stats[0] == 7 * len(mol.symbols) // 3
    # Segmented bases (no two shells of one atom and L share exponents) run
    # the one-member kernel variants of the capped classes; cc-pVDZ's s and p
    # shells are generally contracted, so every capped class keeps its caps.
    if basis == "cc-pvdz":
        assert stats[4] == 0
    else:
        assert stats[4] == int(re.search(r"HFG_UNIT_VARIANT_MASK (0x[0-9a-f]+)u",
                                         open(os.path.join(PKG, "csrc", "gen", "class_list.inc")).read()).group(1), 16)
