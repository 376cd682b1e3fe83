"""CPU pins of the DIIS family that decides the SCF iteration count (checked
+-1 against the oracle): the oracle's restatement (oracle/hfo.cpp, following
src/diis.rs:72-137, 195-266, 289-359) and the engine's host copy
(hf-rs_amd/csrc/diis_host.cpp), each checked against what the mathematics
fixes rather than against itself:
  * Euclidean projection onto the probability simplex: its KKT conditions
    (x = max(v - theta, 0), sum x = 1), which determine it uniquely;
  * the simplex-constrained quadratic minimiser: brute force over every face
    of the simplex (the equality-constrained stationary point of each face,
    feasible ones compared) for m <= 4;
  * CDIIS coefficients: numpy's solve of the bordered Pulay system built here,
    and the optimality condition B c = lambda 1;
  * the ADIIS model (Q, b): with an exactly quadratic energy
    E(P) = tr(P H) + 1/2 tr(P G(P)) and F = H + G(P) (G linear and
    self-adjoint, as for Hartree-Fock), E_n + b^T c + 1/2 c^T Q c must equal
    E(sum_i c_i P_i) for every c with sum c = 1 (the reference's own statement
    of the model, src/diis.rs:269-281)."""
import ctypes as C
import itertools
import os

import numpy as np
import pytest

import oracle as O
from conftest import PKG

dptr = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def olib():
    L = O.lib()
    L.hfo_project_simplex.argtypes = [dptr, C.c_int]
    L.hfo_min_quad_simplex.argtypes = [dptr, dptr, C.c_int, dptr]
    L.hfo_cdiis_coefficients.argtypes = [dptr, C.c_int, C.c_int, dptr]
    L.hfo_adiis_model.argtypes = [dptr, dptr, C.c_int, C.c_int, dptr, dptr]
    return L


def plib():
    path = os.path.join(PKG, "libhfg_cpucheck.so")
    if not os.path.exists(path):
        pytest.skip("CPU check library not built")
    L = C.CDLL(path)
    L.hfgc_simplex_minimise.argtypes = [dptr, dptr, C.c_int, dptr]
    L.hfgc_simplex_project.argtypes = [dptr, C.c_int]
    L.hfgc_solve_small.argtypes = [dptr, dptr, C.c_int, dptr]
    L.hfgc_solve_small.restype = C.c_int
    return L


def projectors():
    o, p = olib(), plib()

    def op(v):
        x = np.ascontiguousarray(v, dtype=np.float64).copy()
        o.hfo_project_simplex(x, len(x))
        return x

    def pp(v):
        x = np.ascontiguousarray(v, dtype=np.float64).copy()
        p.hfgc_simplex_project(x, len(x))
        return x
    return {"oracle": op, "engine": pp}


def minimisers():
    o, p = olib(), plib()

    def om(q, b):
        c = np.zeros(len(b))
        o.hfo_min_quad_simplex(np.ascontiguousarray(q), np.ascontiguousarray(b), len(b), c)
        return c

    def pm(q, b):
        c = np.zeros(len(b))
        p.hfgc_simplex_minimise(np.ascontiguousarray(q), np.ascontiguousarray(b), len(b), c)
        return c
    return {"oracle": om, "engine": pm}


@pytest.mark.parametrize("who", ["oracle", "engine"])
def test_simplex_projection_satisfies_its_kkt_conditions(who):
    proj = projectors()[who]
    rng = np.random.default_rng(1)
    for trial in range(400):
        m = 1 + trial % 8
        v = rng.normal(0, 1 + 3 * rng.random(), m)
        if trial % 5 == 0:
            v[: m // 2] = v[-1]  # ties
        x = proj(v)
        assert (x >= 0).all()
        assert abs(x.sum() - 1.0) < 1e-13
        pos = x > 0
        theta = (v - x)[pos]
        assert np.ptp(theta) < 1e-12          # one threshold for the kept entries
        assert (v[~pos] <= theta[0] + 1e-12).all()  # and the dropped ones lie below it


def test_a_wrong_projection_fails_the_kkt_check():
    """The check above is live: clipping then renormalising is not the projection."""
    v = np.array([0.9, 0.5, -0.2])
    x = np.maximum(v, 0) / np.maximum(v, 0).sum()
    theta = (v - x)[x > 0]
    assert np.ptp(theta) > 1e-3


def brute_force_simplex_qp(q, b):
    m = len(b)
    qs = 0.5 * (q + q.T)
    best, arg = np.inf, None
    for r in range(1, m + 1):
        for S in itertools.combinations(range(m), r):
            S = list(S)
            k = len(S)
            A = np.zeros((k + 1, k + 1))
            A[:k, :k] = qs[np.ix_(S, S)]
            A[:k, k] = 1.0
            A[k, :k] = 1.0
            rhs = np.concatenate([-b[S], [1.0]])
            try:
                sol = np.linalg.solve(A, rhs)
            except np.linalg.LinAlgError:
                continue
            c = np.zeros(m)
            c[S] = sol[:k]
            if (c < -1e-12).any():
                continue
            c = np.maximum(c, 0)
            f = 0.5 * c @ qs @ c + b @ c
            if f < best - 1e-15:
                best, arg = f, c
    return arg, best


@pytest.mark.parametrize("who", ["oracle", "engine"])
def test_simplex_qp_minimiser_matches_brute_force(who):
    solve = minimisers()[who]
    rng = np.random.default_rng(2)
    worst = 0.0
    for trial in range(300):
        m = 2 + trial % 3
        U, _ = np.linalg.qr(rng.normal(size=(m, m)))
        lam = rng.uniform(1.0, 20.0, m)
        q = U @ np.diag(lam) @ U.T
        q = q + rng.normal(0, 0.3, (m, m)) * (trial % 2)  # unsymmetric part: only sym(Q) matters
        b = rng.normal(0, 10.0 if trial % 4 == 0 else 2.0, m)
        c_ref, f_ref = brute_force_simplex_qp(q, b)
        c = solve(q, b)
        qs = 0.5 * (q + q.T)
        f = 0.5 * c @ qs @ c + b @ c
        assert abs(c.sum() - 1) < 1e-12 and (c >= 0).all()
        assert f <= f_ref + 1e-10 * max(1.0, abs(f_ref))
        worst = max(worst, np.abs(c - c_ref).max())
    assert worst < 1e-10, worst


def test_cdiis_coefficients_solve_the_bordered_pulay_system():
    o = olib()
    rng = np.random.default_rng(3)
    for m in range(2, 9):
        E = rng.normal(size=(m, 40))
        c = np.zeros(m)
        assert o.hfo_cdiis_coefficients(np.ascontiguousarray(E), m, 40, c) == 0
        B = E @ E.T
        A = np.zeros((m + 1, m + 1))
        A[:m, :m] = B
        A[:m, m] = A[m, :m] = -1.0
        rhs = np.zeros(m + 1)
        rhs[m] = -1.0
        ref = np.linalg.solve(A, rhs)[:m]
        assert np.abs(c - ref).max() < 1e-10 * max(1.0, np.abs(ref).max())
        assert abs(c.sum() - 1.0) < 1e-12
        g = B @ c  # stationarity of |sum c_i e_i|^2 on sum c = 1: B c is a multiple of 1
        assert np.ptp(g) < 1e-9 * max(1.0, np.abs(g).max())


def test_engine_small_solver_matches_numpy():
    """The engine's bordered-system solver (Gaussian elimination with partial
    pivoting, src/diis.rs:532-589) on the same Pulay systems."""
    p = plib()
    rng = np.random.default_rng(4)
    for m in range(2, 9):
        E = rng.normal(size=(m, 40))
        A = np.zeros((m + 1, m + 1))
        A[:m, :m] = E @ E.T
        A[:m, m] = A[m, :m] = -1.0
        rhs = np.zeros(m + 1)
        rhs[m] = -1.0
        x = np.zeros(m + 1)
        assert p.hfgc_solve_small(np.ascontiguousarray(A), rhs, m + 1, x) == 1
        assert np.abs(x - np.linalg.solve(A, rhs)).max() < 1e-10


def quadratic_energy_model(n, rng):
    """A Hartree-Fock-like energy: H symmetric, G(P) = T : P with T_abcd
    symmetric in a<->b, c<->d and (ab)<->(cd) (self-adjoint), E = tr(PH) + 1/2 tr(P G(P))."""
    H = rng.normal(size=(n, n))
    H = H + H.T
    T = rng.normal(size=(n, n, n, n))
    T = T + T.transpose(1, 0, 2, 3)
    T = T + T.transpose(0, 1, 3, 2)
    T = T + T.transpose(2, 3, 0, 1)

    def G(P):
        return np.einsum("abcd,cd->ab", T, P)

    def E(P):
        return np.sum(P * H) + 0.5 * np.sum(P * G(P))
    return H, G, E


def test_adiis_model_is_the_energy_of_the_interpolated_density():
    o = olib()
    rng = np.random.default_rng(5)
    n = 6
    H, G, E = quadratic_energy_model(n, rng)
    for m in range(2, 7):
        Ps = [0.5 * (lambda a: a + a.T)(rng.normal(size=(n, n))) for _ in range(m)]
        Fs = [H + G(P) for P in Ps]
        dens = np.ascontiguousarray(np.array(Ps).reshape(m, -1))
        focks = np.ascontiguousarray(np.array(Fs).reshape(m, -1))
        q = np.zeros(m * m)
        b = np.zeros(m)
        assert o.hfo_adiis_model(dens, focks, m, n * n, q, b) == 0
        q = q.reshape(m, m)
        e_n = E(Ps[-1])
        for _ in range(5):
            c = rng.normal(size=m)
            c /= c.sum()
            model = e_n + b @ c + 0.5 * c @ q @ c
            exact = E(sum(ci * P for ci, P in zip(c, Ps)))
            assert abs(model - exact) < 1e-10 * max(1.0, abs(exact)), (m, model, exact)


def test_the_adiis_check_catches_a_dropped_half():
    """A model with the quadratic term doubled (a plausible slip) fails the
    check above."""
    o = olib()
    rng = np.random.default_rng(6)
    n, m = 5, 3
    H, G, E = quadratic_energy_model(n, rng)
    Ps = [(lambda a: a + a.T)(rng.normal(size=(n, n))) for _ in range(m)]
    Fs = [H + G(P) for P in Ps]
    q = np.zeros(m * m)
    b = np.zeros(m)
    o.hfo_adiis_model(np.ascontiguousarray(np.array(Ps).reshape(m, -1)),
                      np.ascontiguousarray(np.array(Fs).reshape(m, -1)), m, n * n, q, b)
    q = q.reshape(m, m)
    c = np.array([0.2, 0.3, 0.5])
    wrong = E(Ps[-1]) + b @ c + c @ q @ c
    assert abs(wrong - E(sum(ci * P for ci, P in zip(c, Ps)))) > 1e-6
