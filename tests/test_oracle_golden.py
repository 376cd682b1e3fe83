"""Pins the CPU oracle (the checker) on the reference's own fixtures, closed forms
and invariants.  Each check is chosen so that a plausible mistake (dropped term,
wrong sign or index, transposed operand) fails it."""
import json
import math
import os

import numpy as np
import pytest

import oracle as O
import molecules as M
from conftest import GOLDEN


def golden(name):
    return json.load(open(os.path.join(GOLDEN, name)))


def parse_geometry(text):
    syms, xyz = [], []
    for part in text.split(";"):
        t = part.split()
        syms.append(t[0])
        xyz.append([float(v) for v in t[1:4]])
    return syms, np.array(xyz)


# --- Boys function (src/integrals.rs:20-101; tests/boys_test.rs:12-40) -------
def test_boys_matches_mpmath_fixture():
    worst = 0.0
    for row in golden("boys.json")["rows"]:
        f = O.boys(40, row["t"])
        for n, s in enumerate(row["f"]):
            ref = float(s)
            worst = max(worst, abs(f[n] - ref) / abs(ref))
    assert worst < 1e-14, worst


@pytest.mark.parametrize("t", [0.0, 1e-9, 0.37, 4.2, 17.3, 49.99, 50.01, 88.0, 500.0])
def test_boys_closed_forms(t):
    f = O.boys(12, t)
    if t == 0.0:
        assert np.allclose(f, [1.0 / (2 * n + 1) for n in range(13)], rtol=1e-15, atol=0)
    else:
        f0 = 0.5 * math.sqrt(math.pi / t) * math.erf(math.sqrt(t))
        assert abs(f[0] - f0) / f0 < 2e-15
        # downward recursion identity F_{n-1} = (2T F_n + e^-T)/(2n-1)
        for n in range(1, 13):
            assert abs(f[n - 1] - (2 * t * f[n] + math.exp(-t)) / (2 * n - 1)) <= 1e-14 * f[n - 1]


# --- Shell quartet ERIs (tests/shell_eri_test.rs:215-277) ----------------------
def spdf_basis():
    meta = golden("spdf_meta.json")
    return O.Basis.build(meta["basis_gaussian94"], meta["symbols"], meta["geometry_bohr"]), meta


def test_spdf_all_256_classes_match_pyscf():
    B, meta = spdf_basis()
    assert B.n == 80
    ref = np.load(os.path.join(GOLDEN, "spdf_eri.npz"))["eri"].reshape(20, 20, 20, 20)
    ss = B.shells()  # 16 shells: per atom s, p, d, f
    ranges = [(0, 1), (1, 4), (4, 10), (10, 20)]
    worst = 0.0
    for blk in range(256):
        ang = [blk // 64, blk // 16 % 4, blk // 4 % 4, blk % 4]
        shells = [4 * a + ang[a] for a in range(4)]
        got = B.shell_quartet(*shells)
        exp = ref[tuple(slice(*ranges[ang[a]]) for a in range(4))]
        worst = max(worst, np.abs(got - exp).max())
    assert worst < 2e-10, worst


def test_scalar_pair_engine_matches_shell_engine():
    B, _ = spdf_basis()
    blk = B.shell_quartet(2, 6, 9, 13)   # (d p | p p) on four centres
    ss = B.shells()
    for idx in [(0, 0, 0, 0), (5, 2, 1, 2), (3, 1, 2, 0)]:
        i, j, k, l = (ss[s] + o for s, o in zip((2, 6, 9, 13), idx))
        assert abs(B.contracted_eri(i, j, k, l) - blk[idx]) < 2e-11


def test_ss_ss_closed_form_and_symmetries():
    # (ss|ss) for single s primitives has a closed form through F_0.
    text = "H 0\nS 1 1\n0.8 1\n****\nHe 0\nS 1 1\n1.3 1\n****"
    xyz = np.array([[0.1, -0.2, 0.3], [0.9, 0.4, -0.6]])
    B = O.Basis.build(text, ["H", "He"], xyz)
    a, b = 0.8, 1.3
    na = (2 * a / math.pi) ** 0.75
    nb = (2 * b / math.pi) ** 0.75
    p = q = None
    v = B.shell_quartet(0, 1, 0, 1)[0, 0, 0, 0]
    p = a + b
    mu = a * b / p
    r2 = float(((xyz[0] - xyz[1]) ** 2).sum())
    kab = math.exp(-mu * r2)
    pref = 2 * math.pi ** 2.5 / (p * p * math.sqrt(2 * p))
    expected = na * nb * na * nb * kab * kab * pref * 1.0  # P == Q so F_0(0)=1
    assert abs(v - expected) < 1e-14 * abs(expected)
    B2, _ = spdf_basis()
    x = B2.shell_quartet(1, 6, 11, 12)
    y = B2.shell_quartet(11, 12, 1, 6)
    assert np.abs(x - y.transpose(2, 3, 0, 1)).max() < 1e-14
    z = B2.shell_quartet(6, 1, 12, 11)
    assert np.abs(x - z.transpose(1, 0, 3, 2)).max() < 1e-14


def test_schwarz_inequality_and_screened_packing():
    mol = M.water_config1()
    B = O.Basis.build(M.basis_text("6-31g*"), mol.symbols, mol.xyz)
    ss = B.shells()
    ns = len(ss) - 1
    q = B.schwarz()
    pairs = [(a, b) for a in range(ns) for b in range(a + 1)]
    rng = np.random.default_rng(1)
    for _ in range(40):
        i, k = rng.integers(0, len(pairs), 2)
        blk = B.shell_quartet(*pairs[i], *pairs[k])
        assert np.abs(blk).max() <= q[i] * q[k] * (1 + 1e-12)


# --- J/K digestion (tests/fock_contraction_test.rs:4-44) ----------------------
@pytest.mark.parametrize("n", range(1, 8))
def test_contract_jk_matches_explicit_ordered_contraction(n):
    np_ = n * (n + 1) // 2
    eri = np.sin((np.arange(np_ * (np_ + 1) // 2) + 1) * 0.37)
    pa = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1):
            pa[i, j] = pa[j, i] = math.sin(i * 3 + j + 1)

    def cidx(i, j):
        return i * (i + 1) // 2 + j if i >= j else j * (j + 1) // 2 + i

    full = np.zeros((n, n, n, n))
    for a in range(n):
        for b in range(n):
            for c in range(n):
                for d in range(n):
                    full[a, b, c, d] = eri[cidx(cidx(a, b), cidx(c, d))]
    j, k = O.contract_jk(pa, eri, n)
    assert np.abs(j - np.einsum("abcd,cd->ab", full, pa)).max() < 1e-12
    assert np.abs(k - np.einsum("acbd,cd->ab", full, pa)).max() < 1e-12


def test_direct_jk_equals_packed_tensor_contraction():
    mol = M.water_config1()
    B = O.Basis.build(M.basis_text("6-31g*"), mol.symbols, mol.xyz)
    rng = np.random.default_rng(3)
    p = rng.standard_normal((B.n, B.n))
    p = p + p.T
    j1, k1 = O.contract_jk(p, B.eri_packed(), B.n)
    j2, k2, _ = B.direct_jk(p)
    assert np.abs(j1 - j2).max() < 1e-12 and np.abs(k1 - k2).max() < 1e-12
    # sampled bra pairs partition the work exactly
    parts = [B.direct_jk(p, 3, o) for o in range(3)]
    assert np.abs(sum(x[0] for x in parts) - j2).max() < 1e-12
    assert np.abs(sum(x[1] for x in parts) - k2).max() < 1e-12


# --- SCF energies (tests/scf_reference_test.rs:72-94; spherical_basis_test.rs:136-166)
RHF_CASES = [(name, b) for name in ("h2", "lih", "h2o", "nh3") for b in ("sto-3g", "6-31g*")]


@pytest.mark.parametrize("name,basis", RHF_CASES)
def test_rhf_energy_matches_pyscf(name, basis):
    sysd = golden("scf_energies.json")["systems"][name]
    syms, xyz = parse_geometry(sysd["geometry_bohr"])
    mol = M.Molecule(syms, xyz)
    B = O.Basis.build(M.basis_text(basis), syms, xyz)
    ref = sysd["results"][basis]
    assert B.n == ref["n_ao"]
    res, _, _, _ = O.scf_rhf(B, mol.zxyz, max_iter=200, tol=1e-12, conv="cdiis")
    assert res.converged
    assert abs(res.total_energy - ref["energy"]) < 1e-8


def test_water_ccpvdz_spherical_energy():
    sysd = golden("scf_energies.json")["systems"]["water_ccpvdz_spherical"]
    syms, xyz = parse_geometry(sysd["geometry_bohr"])
    mol = M.Molecule(syms, xyz)
    B = O.Basis.build(M.basis_text("cc-pvdz"), syms, xyz)
    assert B.spherical_transform().shape == (25, 24)
    res, _, _, _ = O.scf_rhf(B, mol.zxyz, max_iter=150, tol=1e-13, conv="cdiis", spherical=True)
    assert res.converged
    assert abs(res.total_energy - sysd["results"]["cc-pvdz"]["energy"]) < 1e-9


def test_default_accelerator_reaches_same_energy():
    mol = M.water_config1()
    B = O.Basis.build(M.basis_text("sto-3g"), mol.symbols, mol.xyz)
    a, _, _, _ = O.scf_rhf(B, mol.zxyz, tol=1e-10, conv="adiissoscf")
    c, _, _, _ = O.scf_rhf(B, mol.zxyz, tol=1e-10, conv="cdiis", direct=True)
    assert abs(a.total_energy - (-74.96302777726257)) < 1e-8
    assert abs(c.total_energy - (-74.96302777726257)) < 1e-8


# --- SAD (tests/sad_test.rs:4-30) ---------------------------------------------
@pytest.mark.parametrize("sym", ["H", "C", "N", "O"])
def test_atomic_fractional_hf_matches_pyscf(sym):
    ref = golden("sad_reference.json")["atoms"][sym]
    B = O.Basis.build(M.basis_text("sto-3g"), [sym], [[0.3, -0.2, 0.7]])
    e, d = O.atomic_hf(B, float(M.Z[sym]))
    assert abs(e - ref["energy"]) < 1e-9
    assert np.abs(d - np.array(ref["density"])).max() < 1e-8


def test_sad_charge_and_translation_invariance():
    mol = M.water_config1()
    B = O.Basis.build(M.basis_text("sto-3g"), mol.symbols, mol.xyz)
    s, _, _ = B.one_electron(mol.zxyz)
    p = O.sad_density(B, mol.zxyz, s)
    assert abs((p * s).sum() - 10.0) < 1e-12
    shifted = mol.xyz + np.array([0.3, -1.1, 2.0])
    B2 = O.Basis.build(M.basis_text("sto-3g"), mol.symbols, shifted)
    z2 = np.column_stack([mol.zxyz[:, 0], shifted])
    s2, _, _ = B2.one_electron(z2)
    assert np.abs(O.sad_density(B2, z2, s2) - p).max() < 1e-12


# --- eigensolver (src/matrix.rs:537-554 style checks) --------------------------
@pytest.mark.parametrize("n", [2, 5, 7, 8, 30, 64])
def test_eigh_decomposes_and_follows_sign_convention(n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n))
    a = a + a.T
    w, v = O.eigh(a)
    assert np.all(np.diff(w) >= 0)
    assert np.abs(v @ np.diag(w) @ v.T - a).max() < 1e-12 * max(1, np.abs(a).max()) * n
    assert np.abs(v.T @ v - np.eye(n)).max() < 1e-12
    idx = np.abs(v).argmax(axis=0)
    assert np.all(v[idx, np.arange(n)] > 0)
    assert np.abs(w - np.linalg.eigvalsh(a)).max() < 1e-12 * n


# --- density fitting (tests/df_test.rs, tests/df_hf_test.rs) -------------------
def df_water():
    syms = ["O", "H", "H"]
    xyz = np.array([[0.0, 0, 0], [1.43, 0.2, 1.1], [-1.2, -0.3, 1.3]])
    B = O.Basis.build(M.basis_text("sto-3g"), syms, xyz)
    A = O.Basis.build(M.basis_text("def2-svp-rifit"), syms, xyz)
    return B, A


def test_df_metric_closed_form():
    text = "H 0\nS 1 1\n0.7 1\nS 1 1\n1.2 1\n****"
    A = O.Basis.build(text, ["H"], [[0.1, -0.2, 0.3]])
    m = O.df_metric(A)
    arr = A.arrays()
    for i in range(2):
        for j in range(2):
            a, b = arr["exponent"][i], arr["exponent"][j]
            exp = 2 * math.pi ** 2.5 / (a * b * math.sqrt(a + b)) * arr["coefficient"][i] * arr["coefficient"][j]
            assert abs(m[i, j] - exp) < 1e-12


def test_df_metric_and_three_center_match_pyscf():
    ref = golden("df_reference.json")
    B, A = df_water()
    assert A.n == ref["n_aux"] == 85
    m = O.df_metric(A)
    assert np.abs(m - np.array(ref["metric"])).max() < 2e-9
    t = O.df_three_center(B, A)
    three = np.array(ref["three"])  # [i][j][P]
    worst = 0.0
    for i in range(B.n):
        for j in range(i + 1):
            worst = max(worst, np.abs(t[:, i * (i + 1) // 2 + j] - three[i, j]).max())
    assert worst < 2e-10


def test_df_rhf_water_matches_pyscf_every_j_k_element():
    ref = golden("df_hf_water.json")["water"]
    syms = [a[0] for a in ref["atoms"]]
    xyz = np.array([a[1] for a in ref["atoms"]])
    B = O.Basis.build(M.basis_text("sto-3g"), syms, xyz)
    A = O.Basis.build(M.basis_text("def2-svp-rifit"), syms, xyz)
    zxyz = np.column_stack([[M.Z[s] for s in syms], xyz])
    res, p, f, j, k = O.df_rhf(B, A, zxyz)
    assert res.converged and res.n_aux == int(ref["n_aux"])
    assert abs(res.total_energy - float(ref["e_scf"])) < 1e-9
    for name, mat in (("j", j), ("k", k), ("fock", f), ("density", p)):
        assert np.abs(mat - np.array(ref[name])).max() < 1e-9, name


def test_flop_model_matches_md_loop_counts():
    expected = {(0, 0, 0, 0): 58, (1, 0, 0, 0): 124, (1, 1, 0, 0): 334, (1, 0, 1, 0): 340, (1, 1, 1, 1): 4309,
                (2, 0, 0, 0): 259, (2, 1, 1, 0): 2896, (2, 2, 0, 0): 1933, (2, 1, 2, 1): 29599,
                (2, 2, 2, 2): 196048, (3, 3, 3, 3): 3497683}
    for k, v in expected.items():
        assert O.md_flops(*k) == v


def _cd_cases():
    data = json.load(open(os.path.join(GOLDEN, "cd_hf_reference.json")))
    names = {"STO_3G_GBS": "sto-3g", "BASIS_6_31G_GBS": "6-31g"}
    out = []
    for c in data["cases"]:
        mol = M.Molecule([a["symbol"] for a in c["atoms"]], [a["position"] for a in c["atoms"]], name=c["name"])
        out.append((c, mol, names[c["basis_constant"]]))
    return out


@pytest.mark.parametrize("case", range(3))
def test_cholesky_pivots_and_cd_hf_energies_match_the_dense_factorization(case):
    """The oracle's pivoted Cholesky (src/cholesky.rs:44-192) against the
    reference's fixture (tests/data/cd_hf_reference.json: an independent dense
    NumPy factorization of the PySCF supermatrix, PySCF RHF on the
    reconstructed integrals): the exact pivot sequence, the vector count, the
    largest residual (1e-12) and the CD-HF energy (1e-9), at three thresholds
    each for LiH/STO-3G, water/6-31G and HCl/6-31G."""
    c, mol, bname = _cd_cases()[case]
    b = O.Basis.build(M.basis_text(bname), mol.symbols, mol.xyz)
    for t in c["thresholds"]:
        fac, mr, piv = O.cholesky_eri(b, t["threshold"])
        assert fac.shape[0] == t["n_vectors"]
        assert piv == t["pivots"]
        assert mr <= t["threshold"] and abs(mr - t["max_residual"]) < 1e-12
        res, _ = O.cd_rhf(b, mol.zxyz, t["threshold"])
        assert res.converged and res.n_aux == t["n_vectors"]
        assert abs(res.total_energy - t["e_cd_hf"]) < 1e-9, (c["name"], t["threshold"], res.total_energy)


def test_cholesky_reconstruction_obeys_the_residual_bound():
    """|(pq|rs) - sum_K L_K(pq) L_K(rs)| <= max residual diagonal for every
    packed integral (the positive-semidefinite residual bounds every element
    by the geometric mean of two diagonals), water/6-31G."""
    c, mol, bname = [x for x in _cd_cases() if x[0]["name"] == "water"][0]
    b = O.Basis.build(M.basis_text(bname), mol.symbols, mol.xyz)
    exact = b.eri_packed()
    npair = b.n * (b.n + 1) // 2
    il = np.tril_indices(npair)  # row-major lower triangle: packed index pq(pq+1)/2 + rs, pq >= rs
    for delta in (1e-2, 1e-4, 1e-7):
        fac, mr, _ = O.cholesky_eri(b, delta)
        approx = (fac.T @ fac)[il]
        worst = np.abs(approx - exact).max()
        assert worst <= mr + 1e-13, (delta, worst, mr)


def test_cholesky_rejects_invalid_requests():
    mol = M.Molecule(["H", "H"], [[0, 0, -0.7], [0, 0, 0.7]])
    b = O.Basis.build(M.basis_text("sto-3g"), mol.symbols, mol.xyz)
    with pytest.raises(RuntimeError):
        O.cholesky_eri(b, -1.0)
    with pytest.raises(RuntimeError):
        O.cholesky_eri(b, 1e-8, budget=16)
    with pytest.raises(RuntimeError):
        O.cholesky_eri(b, 10.0)


RHF_GRADIENT_CASES = [(m, b) for m in ("h2", "lih", "h2o", "nh3") for b in ("sto-3g", "6-31g*")]


@pytest.mark.parametrize("name,basis", RHF_GRADIENT_CASES)
def test_rhf_gradient_matches_pyscf(name, basis):
    """The oracle's analytic gradient (src/gradient.rs:287-512) at the converged
    RHF density (CDIIS, tol 1e-12, as tests/scf_reference_test.rs:74-94) against
    PySCF's gradients in the reference's fixture: 1e-6 (the reference's own
    bound), and the forces sum to zero (translational invariance) to 1e-9.
    W = 1/2 P F P = 2 C_occ eps C_occ^T at convergence."""
    sysd = golden("scf_energies.json")["systems"][name]
    ref = sysd["results"][basis]
    syms, xyz = parse_geometry(sysd["geometry_bohr"])
    mol = M.Molecule(syms, xyz)
    b = O.Basis.build(M.basis_text(basis), syms, xyz)
    res, p, f, _ = O.scf_rhf(b, mol.zxyz, tol=1e-12, max_iter=200, conv="cdiis")
    assert res.converged and abs(res.total_energy - ref["energy"]) < 1e-8
    g = O.rhf_gradient(b, mol.zxyz, p, 0.5 * p @ f @ p)
    assert np.abs(g - np.array(ref["gradient"])).max() < 1e-6
    assert np.abs(g.sum(axis=0)).max() < 1e-9


# --- one-electron matrices: closed forms for s and p Gaussians -----------------
# (Szabo and Ostlund, Modern Quantum Chemistry, appendix A: overlap, kinetic
# and nuclear-attraction integrals of normalised s Gaussians; the p function
# x_A exp(-a r_A^2) has norm 2 sqrt(a) (2a/pi)^(3/4) and kinetic energy 5a/2.)
TWO_CENTRE_SP = """H     0
S   1   1.00
      0.80000000             1.00000000
P   1   1.00
      0.50000000             1.00000000
****
"""


def sp_closed_forms(xyz):
    """S, T, V of the functions (s_A, p_A, s_B, p_B) of two protons at xyz."""
    a, b = 0.8, 0.5
    A, B = np.asarray(xyz[0], float), np.asarray(xyz[1], float)

    def ns(x):
        return (2 * x / math.pi) ** 0.75

    def f0(t):
        return 1.0 if t < 1e-14 else 0.5 * math.sqrt(math.pi / t) * math.erf(math.sqrt(t))

    def ss(x, X, y, Y):  # (S, T, V) of normalised s functions on X and Y
        p, mu = x + y, x * y / (x + y)
        r2 = float(((X - Y) ** 2).sum())
        P = (x * X + y * Y) / p
        s = ns(x) * ns(y) * (math.pi / p) ** 1.5 * math.exp(-mu * r2)
        t = mu * (3 - 2 * mu * r2) * s
        v = sum(-ns(x) * ns(y) * 2 * math.pi / p * math.exp(-mu * r2) * f0(p * float(((P - C) ** 2).sum()))
                for C in (A, B))
        return s, t, v

    s_aa = ss(a, A, a, A)
    s_ab = ss(a, A, a, B)
    # <p_x on A (exponent b) | s on B (exponent a)> = N_p N_s (P - A)_x (pi/p)^(3/2) exp(-mu R^2)
    p, mu = a + b, a * b / (a + b)
    r2 = float(((A - B) ** 2).sum())
    P = (b * A + a * B) / p
    pxs = 2 * math.sqrt(b) * ns(b) * ns(a) * (P - A)[0] * (math.pi / p) ** 1.5 * math.exp(-mu * r2)
    return {"s_aa": s_aa, "s_ab": s_ab, "p_x_A_s_B": pxs, "t_pp": 2.5 * b}


def test_oracle_one_electron_matches_closed_forms():
    xyz = [[0.1, -0.2, 0.3], [1.0, 0.4, -0.5]]
    B = O.Basis.build(TWO_CENTRE_SP, ["H", "H"], xyz)
    zxyz = np.array([[1.0] + xyz[0], [1.0] + xyz[1]])
    s, t, v = B.one_electron(zxyz)
    ref = sp_closed_forms(xyz)
    # functions: s_A, px_A, py_A, pz_A, s_B, px_B, py_B, pz_B
    assert np.abs(np.diag(s) - 1.0).max() < 1e-14
    for k, (sv, tv, vv) in (("s_aa", ref["s_aa"]), ("s_ab", ref["s_ab"])):
        j = 0 if k == "s_aa" else 4
        assert abs(s[0, j] - sv) < 1e-14 and abs(t[0, j] - tv) < 1e-13 and abs(v[0, j] - vv) < 1e-13
    assert abs(s[1, 4] - ref["p_x_A_s_B"]) < 1e-14
    assert abs(t[1, 1] - ref["t_pp"]) < 1e-13 and abs(t[2, 2] - ref["t_pp"]) < 1e-13
    for m in (s, t, v):
        assert np.abs(m - m.T).max() < 1e-15
