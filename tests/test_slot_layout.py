"""CPU test of the server slot layout (bscf_layout.h): every matrix of a slot
lies inside the slot's buffer, 256-byte aligned and disjoint from the others,
for every (n, m) the server can meet -- the size and the carving come from one
function, so a sub-buffer can never run past the allocation."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import PKG

NAMES = ["S", "H", "X", "F", "P", "Pprev", "G", "Pb", "T1", "T2", "E", "Fu", "Fo", "Xs", "X2", "cF", "aF", "aP", "cE"]


def lengths(n, m):
    nn, mm = n * n, m * m
    elen = mm if m < n else nn
    return [nn, nn, n * m] + [nn] * 9 + [mm] * 3 + [8 * nn] * 3 + [8 * elen]


def test_slot_matrices_are_disjoint_aligned_and_inside_the_buffer():
    path = os.path.join(PKG, "libhfg_cpucheck.so")
    if not os.path.exists(path):
        pytest.skip("CPU check library not built")
    L = C.CDLL(path)
    L.hfgc_slot_layout.argtypes = [C.c_size_t, C.c_size_t, C.c_void_p]
    L.hfgc_slot_layout.restype = C.c_size_t
    rng = np.random.default_rng(0)
    cases = [(1, 1), (2, 2), (25, 25), (25, 24), (33, 31), (310, 310), (800, 768)]
    cases += [(int(n), int(n) - int(d)) for n, d in zip(rng.integers(1, 900, 200), rng.integers(0, 40, 200))]
    for n, m in cases:
        m = max(1, min(m, n))
        off = np.zeros(len(NAMES), dtype=np.uint64)
        total = L.hfgc_slot_layout(n, m, off.ctypes.data_as(C.c_void_p))
        ln = lengths(n, m)
        spans = sorted((int(o), int(o) + l, name) for o, l, name in zip(off, ln, NAMES))
        for (a0, a1, an), (b0, b1, bn) in zip(spans, spans[1:]):
            assert a1 <= b0, (n, m, an, bn)
        assert spans[-1][1] <= total, (n, m)
        assert all(int(o) % 32 == 0 for o in off), (n, m)
