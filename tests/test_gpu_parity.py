"""Parity of the HIP engine with the oracle and the PySCF fixtures, through the
C ABI (include/hfg.h).  Tolerances: ERIs 2e-10 abs (tests/shell_eri_test.rs),
J/K max-abs 1e-10, total RHF energy 1e-8 Eh (1e-9 for the spherical water
case), SCF iteration count within +-1 of the oracle."""
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


def sym_density(n, seed):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal((n, n)) * 0.3
    return p + p.T


def test_spd_quartets_match_pyscf(hf):
    meta = json.load(open(os.path.join(GOLDEN, "spdf_meta.json")))
    ob = O.Basis.build(meta["basis_gaussian94"], meta["symbols"], meta["geometry_bohr"])
    arr = ob.arrays()
    # keep only the s, p, d shells of each centre (the engine compiles L <= 2)
    keep = [i for i in range(80) if (i % 20) < 10]
    po = arr["prim_offset"]
    sub = dict(atom_index=arr["atom_index"][keep], center=arr["center"][keep], lxyz=arr["lxyz"][keep],
               prim_offset=np.concatenate([[0], np.cumsum([po[i + 1] - po[i] for i in keep])]).astype(np.int32),
               exponent=np.concatenate([arr["exponent"][po[i]:po[i + 1]] for i in keep]),
               coefficient=np.concatenate([arr["coefficient"][po[i]:po[i + 1]] for i in keep]))
    ctx = hf.Context(hf.Basis(sub))
    ref = np.load(os.path.join(GOLDEN, "spdf_eri.npz"))["eri"].reshape(20, 20, 20, 20)
    ranges = [(0, 1), (1, 4), (4, 10)]
    worst = 0.0
    for blk in range(81):
        ang = [blk // 27, blk // 9 % 3, blk // 3 % 3, blk % 3]
        dims = [ranges[x][1] - ranges[x][0] for x in ang]
        got = ctx.shell_quartet(*[3 * a + ang[a] for a in range(4)], dims)
        exp = ref[tuple(slice(*ranges[ang[a]]) for a in range(4))]
        worst = max(worst, np.abs(got - exp).max())
    print("spd quartets vs PySCF: max |dev| %.2e" % worst)
    assert worst < 2e-10


def test_schwarz_bounds_match_oracle(hf):
    mol = M.water_cluster((2, 1, 1), seed=5)
    text = M.basis_text("cc-pvdz")
    ctx = hf.Context(hf.Basis.from_g94(text, mol.symbols, mol.xyz))
    q = ctx.schwarz()
    qr = O.Basis.build(text, mol.symbols, mol.xyz).schwarz()
    assert np.abs(q - qr).max() <= 1e-12 * qr.max()


@pytest.mark.parametrize("basis", ["sto-3g", "6-31g*"])
def test_packed_tensor_matches_oracle(hf, basis):
    mol = M.water_config1()
    text = M.basis_text(basis)
    ctx = hf.Context(hf.Basis.from_g94(text, mol.symbols, mol.xyz))
    got = ctx.eri_packed()
    exp = O.Basis.build(text, mol.symbols, mol.xyz).eri_packed()
    assert got.shape == exp.shape
    assert np.abs(got - exp).max() < 1e-12
    # entries left at zero by one engine (screened quartets, or values that
    # vanish by symmetry) are zero up to rounding in the other
    assert np.abs(got[exp == 0.0]).max(initial=0.0) < 1e-15
    assert np.abs(exp[got == 0.0]).max(initial=0.0) < 1e-15


@pytest.mark.parametrize("basis,grid", [("sto-3g", (1, 1, 1)), ("6-31g*", (2, 1, 1)), ("cc-pvdz", (2, 1, 1)),
                                        ("def2-svp", (2, 2, 1)), ("cc-pvdz", (2, 2, 1))])
def test_direct_jk_matches_oracle(hf, basis, grid):
    mol = M.water_cluster(grid, seed=2)
    text = M.basis_text(basis)
    ctx = hf.Context(hf.Basis.from_g94(text, mol.symbols, mol.xyz))
    p = sym_density(ctx.n, 42)
    j, k = ctx.contract_jk(p)
    jr, kr, _ = O.Basis.build(text, mol.symbols, mol.xyz).direct_jk(p)
    ej, ek = np.abs(j - jr).max(), np.abs(k - kr).max()
    print("%s %s n=%d  J %.2e  K %.2e" % (basis, grid, ctx.n, ej, ek))
    assert ej <= 1e-10 and ek <= 1e-10
    assert np.abs(j - j.T).max() == 0.0 and np.abs(k - k.T).max() == 0.0


@pytest.mark.parametrize("n_carbon,basis", [(3, "def2-svp"), (2, "cc-pvdz"), (4, "6-31g*")])
def test_direct_jk_alkane_matches_oracle(hf, n_carbon, basis):
    """All-trans alkanes (the config-3 molecule family): carbon d shells with
    other contraction patterns than water's, through every sliced class."""
    mol = M.alkane(n_carbon)
    text = M.basis_text(basis)
    ctx = hf.Context(hf.Basis.from_g94(text, mol.symbols, mol.xyz))
    p = sym_density(ctx.n, 43)
    j, k = ctx.contract_jk(p)
    jr, kr, _ = O.Basis.build(text, mol.symbols, mol.xyz).direct_jk(p)
    ej, ek = np.abs(j - jr).max(), np.abs(k - kr).max()
    print("C%dH%d %s n=%d  J %.2e  K %.2e" % (n_carbon, 2 * n_carbon + 2, basis, ctx.n, ej, ek))
    assert ej <= 1e-10 and ek <= 1e-10


def test_engine_rejects_what_it_does_not_support(hf):
    """Error paths of the C ABI: f shells (cc-pVTZ) are refused with the
    unsupported status, an odd electron count and a non-positive tolerance
    with the invalid-input status (src/error.rs:5-15 semantics), and the
    engine stays usable afterwards."""
    mol = M.water_config1()
    with pytest.raises(hf.HfError) as e:
        hf.Context(hf.Basis.from_g94(M.basis_text("cc-pvtz"), mol.symbols, mol.xyz))
    assert e.value.code == hf.HFG_UNSUPPORTED
    b = hf.Basis.from_g94(M.basis_text("sto-3g"), mol.symbols, mol.xyz)
    with pytest.raises(hf.HfError) as e:
        hf.scf_rhf(b, mol.zxyz, charge=1)
    assert e.value.code == hf.HFG_INVALID_INPUT
    with pytest.raises(hf.HfError) as e:
        hf.scf_rhf(b, mol.zxyz, tol=0.0)
    assert e.value.code == hf.HFG_INVALID_INPUT
    r = hf.scf_rhf(b, mol.zxyz, tol=1e-10, conv="cdiis")
    assert abs(r["total_energy"] - (-74.96302777726257)) < 1e-8


def test_build_fock_is_h_plus_j_minus_half_k(hf):
    mol = M.water_config1()
    ctx = hf.Context(hf.Basis.from_g94(M.basis_text("6-31g*"), mol.symbols, mol.xyz))
    p = sym_density(ctx.n, 1)
    h = sym_density(ctx.n, 2)
    j, k = ctx.contract_jk(p)
    f = ctx.build_fock(h, p)
    assert np.abs(f - (h + j - 0.5 * k)).max() < 1e-13


def test_jk_linear_in_density_at_full_size(hf):
    """(H2O)16/cc-pVDZ: J and K are linear in P and symmetric (size-independent checks)."""
    mol = M.water_cluster((4, 2, 2), seed=2)
    ctx = hf.Context(hf.Basis.from_g94(M.basis_text("cc-pvdz"), mol.symbols, mol.xyz))
    assert ctx.n == 400
    p1, p2 = sym_density(400, 3), sym_density(400, 4)
    j1, k1 = ctx.contract_jk(p1)
    j2, k2 = ctx.contract_jk(p2)
    j3, k3 = ctx.contract_jk(p1 + 2.0 * p2)
    assert np.abs(j3 - (j1 + 2 * j2)).max() < 1e-10 * max(1.0, np.abs(j3).max())
    assert np.abs(k3 - (k1 + 2 * k2)).max() < 1e-10 * max(1.0, np.abs(k3).max())
    # Coulomb energy is positive for a positive semidefinite density
    c = np.random.default_rng(9).standard_normal((400, 20))
    pp = c @ c.T
    jp, kp = ctx.contract_jk(pp)
    assert (pp * jp).sum() > 0 and (pp * kp).sum() > 0


def test_jk_at_full_size_matches_oracle_fixture(hf):
    """(H2O)16/cc-pVDZ, n = 400: J and K of the seeded density against the
    oracle's direct J/K (tests/golden/make_h2o16_jk.py, upper triangles), max
    abs 1e-10 -- the whole quartet set of the full-size workload, every class
    and every molecule-sized table, not only a small cluster."""
    fx = np.load(os.path.join(GOLDEN, "h2o16_ccpvdz_jk.npz"))
    mol = M.water_cluster((4, 2, 2), seed=2)
    ctx = hf.Context(hf.Basis.from_g94(M.basis_text("cc-pvdz"), mol.symbols, mol.xyz))
    n = int(fx["n"])
    assert ctx.n == n == 400
    c = np.random.default_rng(int(fx["seed"])).standard_normal((n, 40)) * 0.2
    j, k = ctx.contract_jk(c @ c.T)
    iu = np.triu_indices(n)
    ej, ek = np.abs(j[iu] - fx["j_upper"]).max(), np.abs(k[iu] - fx["k_upper"]).max()
    print("n=%d  J %.2e  K %.2e  (|J| %.2f, |K| %.2f)" % (n, ej, ek, np.abs(j).max(), np.abs(k).max()))
    assert ej <= 1e-10 and ek <= 1e-10


def test_rhf_h2o_sto3g_matches_pyscf_and_oracle(hf):
    mol = M.water_config1()
    text = M.basis_text("sto-3g")
    r = hf.scf_rhf(hf.Basis.from_g94(text, mol.symbols, mol.xyz), mol.zxyz, tol=1e-10, conv="cdiis")
    assert r["converged"]
    assert abs(r["total_energy"] - (-74.96302777726257)) < 1e-8
    o, _, _, _ = O.scf_rhf(O.Basis.build(text, mol.symbols, mol.xyz), mol.zxyz, tol=1e-10, conv="cdiis")
    assert abs(r["iterations"] - o.iterations) <= 1


def parse_geometry(text):
    syms, xyz = [], []
    for part in text.split(";"):
        t = part.split()
        syms.append(t[0])
        xyz.append([float(v) for v in t[1:4]])
    return syms, np.array(xyz)


@pytest.mark.parametrize("name,basis", [(n, b) for n in ("h2", "lih", "h2o", "nh3") for b in ("sto-3g", "6-31g*")])
def test_rhf_energies_match_pyscf(hf, name, basis):
    sysd = json.load(open(os.path.join(GOLDEN, "scf_energies.json")))["systems"][name]
    syms, xyz = parse_geometry(sysd["geometry_bohr"])
    mol = M.Molecule(syms, xyz)
    r = hf.scf_rhf(hf.Basis.from_g94(M.basis_text(basis), syms, xyz), mol.zxyz, max_iter=200, tol=1e-12,
                   conv="cdiis")
    assert r["converged"]
    assert abs(r["total_energy"] - sysd["results"][basis]["energy"]) < 1e-8


def test_rhf_spherical_water_ccpvdz(hf):
    sysd = json.load(open(os.path.join(GOLDEN, "scf_energies.json")))["systems"]["water_ccpvdz_spherical"]
    syms, xyz = parse_geometry(sysd["geometry_bohr"])
    mol = M.Molecule(syms, xyz)
    r = hf.scf_rhf(hf.Basis.from_g94(M.basis_text("cc-pvdz"), syms, xyz), mol.zxyz, max_iter=150, tol=1e-13,
                   conv="cdiis", spherical=True)
    assert r["converged"]
    assert abs(r["total_energy"] - sysd["results"]["cc-pvdz"]["energy"]) < 1e-9


@pytest.mark.parametrize("conv,spherical", [("adiissoscf", True), ("cdiis", False)])
def test_rhf_water_dimer_matches_oracle(hf, conv, spherical):
    mol = M.water_cluster((2, 1, 1), seed=2)
    text = M.basis_text("cc-pvdz")
    r = hf.scf_rhf(hf.Basis.from_g94(text, mol.symbols, mol.xyz), mol.zxyz, tol=1e-10, conv=conv,
                   spherical=spherical)
    o, _, _, _ = O.scf_rhf(O.Basis.build(text, mol.symbols, mol.xyz), mol.zxyz, tol=1e-10, conv=conv,
                           spherical=spherical, direct=True)
    print("dimer %s sph=%s: E %.12f vs %.12f, iters %d vs %d" % (conv, spherical, r["total_energy"],
                                                                o.total_energy, r["iterations"], o.iterations))
    assert r["converged"] and o.converged
    assert abs(r["total_energy"] - o.total_energy) < 1e-8
    assert abs(r["iterations"] - o.iterations) <= 1


def test_rhf_h2o16_ccpvdz_matches_oracle_at_full_size(hf):
    """The bench workload itself: energy within 1e-8 Eh and the same iteration
    count (+-1) as the oracle's in-core run (tests/golden/make_h2o16_reference.py)."""
    ref = json.load(open(os.path.join(GOLDEN, "h2o16_ccpvdz_scf.json")))
    mol = M.water_cluster((4, 2, 2), seed=2)
    r = hf.scf_rhf(hf.Basis.from_g94(M.basis_text("cc-pvdz"), mol.symbols, mol.xyz), mol.zxyz, tol=1e-10,
                   conv="adiissoscf", spherical=True)
    print("(H2O)16: E %.12f vs oracle %.12f, iterations %d vs %d" % (r["total_energy"], ref["total_energy"],
                                                                     r["iterations"], ref["iterations"]))
    assert r["converged"]
    assert abs(r["total_energy"] - ref["total_energy"]) < 1e-8
    assert abs(r["iterations"] - ref["iterations"]) <= 1


@pytest.mark.parametrize("inc", [False, True])
def test_rhf_water_ccpvdz_spherical_incremental_builds(hf, inc):
    """With incremental Fock builds (J/K of the density change, threshold
    1e-12 / max|dP|, full first build and final rebuild) and without, the
    PySCF energy of spherical water/cc-pVDZ (tests/data/spherical_reference.json:
    2283-2285) within 1e-9."""
    sysd = json.load(open(os.path.join(GOLDEN, "scf_energies.json")))["systems"]["water_ccpvdz_spherical"]
    syms, xyz = parse_geometry(sysd["geometry_bohr"])
    mol = M.Molecule(syms, xyz)
    r = hf.scf_rhf(hf.Basis.from_g94(M.basis_text("cc-pvdz"), syms, xyz), mol.zxyz, max_iter=150, tol=1e-13,
                   conv="cdiis", spherical=True, incremental=inc)
    assert r["converged"]
    assert abs(r["total_energy"] - sysd["results"]["cc-pvdz"]["energy"]) < 1e-9


def test_rhf_h2o16_incremental_matches_oracle(hf):
    """The config-2 solve with incremental builds: energy within 1e-8 Eh and
    iterations within 1 of the oracle's in-core run."""
    ref = json.load(open(os.path.join(GOLDEN, "h2o16_ccpvdz_scf.json")))
    mol = M.water_cluster((4, 2, 2), seed=2)
    r = hf.scf_rhf(hf.Basis.from_g94(M.basis_text("cc-pvdz"), mol.symbols, mol.xyz), mol.zxyz, tol=1e-10,
                   conv="adiissoscf", spherical=True, incremental=True)
    print("(H2O)16 incremental: E %.12f vs %.12f, iterations %d vs %d" % (r["total_energy"], ref["total_energy"],
                                                                           r["iterations"], ref["iterations"]))
    assert r["converged"]
    assert abs(r["total_energy"] - ref["total_energy"]) < 1e-8
    assert abs(r["iterations"] - ref["iterations"]) <= 1


def test_rhf_densities_by_purification_match_the_eigensolver(hf, monkeypatch):
    """The iteration densities from SP2 purification (default) and from the
    eigensolver (HFG_SCF_EIGH=1) give the same converged energy to 1e-10 Eh
    and iteration counts within 1, on the spherical (H2O)4/cc-pVDZ solve."""
    mol = M.water_cluster((2, 2, 1), seed=2)
    b = hf.Basis.from_g94(M.basis_text("cc-pvdz"), mol.symbols, mol.xyz)
    monkeypatch.delenv("HFG_SCF_EIGH", raising=False)
    r = hf.scf_rhf(b, mol.zxyz, tol=1e-10, conv="adiissoscf", spherical=True)
    monkeypatch.setenv("HFG_SCF_EIGH", "1")
    e = hf.scf_rhf(b, mol.zxyz, tol=1e-10, conv="adiissoscf", spherical=True)
    print("(H2O)4 SP2 vs eigh: E %.12f vs %.12f, iterations %d vs %d" % (r["total_energy"], e["total_energy"],
                                                                        r["iterations"], e["iterations"]))
    assert r["converged"] and e["converged"]
    assert abs(r["total_energy"] - e["total_energy"]) < 1e-10
    assert abs(r["iterations"] - e["iterations"]) <= 1


def test_batch_jk_matches_single_molecule_builds(hf):
    """Fused multi-molecule J/K (mixed sizes, bases and shell classes) equals
    each molecule's own build and the oracle's J/K."""
    mols = [(M.alkane(1), "def2-svp"), (M.water_cluster((2, 1, 1), seed=4), "cc-pvdz"), (M.alkane(3), "sto-3g"),
            (M.water_cluster_m(3, seed=9), "6-31g*"), (M.alkane(4), "def2-svp")]
    bases = [hf.Basis.from_g94(M.basis_text(b), m.symbols, m.xyz) for m, b in mols]
    batch = hf.Batch(bases)
    info = batch.info()
    assert info["count"] == len(mols) and info["n_max"] == max(b.n for b in bases)
    rng = np.random.default_rng(8)
    dens = []
    for i, b in enumerate(bases):
        c = rng.standard_normal((b.n, 6))
        dens.append(c @ c.T)
        batch.set_density(i, dens[-1])
    batch.jk()
    for i, (b, (m, bn)) in enumerate(zip(bases, mols)):
        j, k = batch.get_jk(i)
        j1, k1 = hf.Context(b).contract_jk(dens[i])
        scale = max(1.0, np.abs(j1).max())
        assert np.abs(j - j1).max() < 1e-11 * scale, (i, np.abs(j - j1).max())
        assert np.abs(k - k1).max() < 1e-11 * scale, (i, np.abs(k - k1).max())
        if i in (0, 2):
            jo, ko, _ = O.Basis.build(M.basis_text(bn), m.symbols, m.xyz).direct_jk(dens[i])
            assert np.abs(j - jo).max() < 1e-10 * scale and np.abs(k - ko).max() < 1e-10 * scale
    batch.jk_each()
    for i in range(len(bases)):
        j, k = batch.get_jk(i)
        j1, k1 = hf.Context(bases[i]).contract_jk(dens[i])
        assert np.abs(j - j1).max() < 1e-11 * max(1.0, np.abs(j1).max())


@pytest.mark.timeout(900)
def test_batch_config3_subset_matches_oracle(hf):
    """The first 16 molecules of the batched def2-SVP workload (alkanes and
    water clusters, seed 3; n = 25..285): one fused J/K build of all of them,
    each molecule's J and K against the oracle's direct J/K at 1e-10."""
    mols = M.batch_config3(512, seed=3)[:16]
    text = M.basis_text("def2-svp")
    bases = [hf.Basis.from_g94(text, m.symbols, m.xyz) for m in mols]
    batch = hf.Batch(bases)
    rng = np.random.default_rng(16)
    dens = []
    for i, b in enumerate(bases):
        c = rng.standard_normal((b.n, max(1, b.n // 5))) * 0.3
        dens.append(c @ c.T)
        batch.set_density(i, dens[-1])
    batch.jk()
    worst = 0.0
    for i, m in enumerate(mols):
        j, k = batch.get_jk(i)
        jo, ko, _ = O.Basis.build(text, m.symbols, m.xyz).direct_jk(dens[i])
        ej, ek = np.abs(j - jo).max(), np.abs(k - ko).max()
        print("%s n=%d  J %.2e  K %.2e" % (m.name, bases[i].n, ej, ek), flush=True)
        worst = max(worst, ej, ek)
        assert ej <= 1e-10 and ek <= 1e-10, (i, m.name)
    print("config-3 subset: worst |dJ|,|dK| vs oracle %.2e" % worst)


def test_global_atomic_single_stream_path_matches_oracle(hf, monkeypatch):
    """The J/K kernels without LDS rows (k_jk, global per-XCC atomics; the
    path classes take when the rows do not fit the LDS) on one stream
    (HFG_JK_ROWS=0, HFG_SIDE_BLOCKS=0), against the oracle at 1e-10."""
    monkeypatch.setenv("HFG_JK_ROWS", "0")
    monkeypatch.setenv("HFG_SIDE_BLOCKS", "0")
    for mol, basis in ((M.water_cluster((2, 1, 1), seed=2), "cc-pvdz"), (M.alkane(3), "def2-svp")):
        text = M.basis_text(basis)
        ctx = hf.Context(hf.Basis.from_g94(text, mol.symbols, mol.xyz))
        p = sym_density(ctx.n, 44)
        j, k = ctx.contract_jk(p)
        jr, kr, _ = O.Basis.build(text, mol.symbols, mol.xyz).direct_jk(p)
        ej, ek = np.abs(j - jr).max(), np.abs(k - kr).max()
        print("global-atomic %s %s n=%d  J %.2e  K %.2e" % (mol.name, basis, ctx.n, ej, ek))
        assert ej <= 1e-10 and ek <= 1e-10


def test_fock_build_graph_replays_match_eager_builds(hf, monkeypatch):
    """The one-molecule Fock build captured as a HIP graph on its second call
    and replayed after: every replay (with eager work and synchronisations of
    the program in between) against the eager build of the same density, for
    five densities; (H2O)2/cc-pVDZ with both streams."""
    mol = M.water_cluster((2, 1, 1), seed=2)
    b = hf.Basis.from_g94(M.basis_text("cc-pvdz"), mol.symbols, mol.xyz)
    monkeypatch.setenv("HFG_GRAPH", "0")
    eager = hf.Context(b)
    monkeypatch.setenv("HFG_GRAPH", "1")
    graph = hf.Context(b)
    monkeypatch.delenv("HFG_GRAPH")
    for i in range(5):
        p = sym_density(b.n, 100 + i)
        j0, k0 = eager.contract_jk(p)
        j1, k1 = graph.contract_jk(p)
        jm, km = graph.contract_jk_multi(p, [p])  # the multi-density entry, through the same build
        assert np.abs(j1 - j0).max() < 1e-12 and np.abs(k1 - k0).max() < 1e-12, i
        assert np.abs(jm - j0).max() < 1e-12 and np.abs(km[0] - k0).max() < 1e-12, i
    assert eager.graph_launches() == 0
    assert graph.graph_launches() >= 4
    eager.close()
    graph.close()
