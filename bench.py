"""Throughput of the Fock-build hot path on one MI355X (one GPU).

Headline (value): SCF solves/s of the batched form on the batched workload,
512 RHF solves of def2-SVP molecules (half all-trans alkanes CnH2n+2,
n ~ U{1..12}; half water clusters (H2O)m, m ~ U{1..8}; PCG64 seed 3;
Cartesian, ADIIS->CDIIS, tol 1e-10, FP64 throughout), solved together by
the server (hfg_server_*: continuous batching over resident slots, slot
set-ups beside the iterations, every iteration's Fock builds fused into one
launch per shell-quartet class, the DIIS and purification linear algebra
batched) in its serving form (inside the iterations J/K of the density
change with a raised Schwarz threshold; first, every 12th and final builds
full), against the plain single-molecule GPU driver (hfg_scf_rhf, full
builds, the reference's algorithm) on a sample of the same molecules, with
the energies compared (the serving form is the headline only when every
sampled energy is within 1e-8 Eh of the plain path's).  The same server
with full builds is reported beside it.  A step (--steps) is one
fused Fock build of the whole 512-molecule batch (J and K of every
molecule's resident density); as many builds again with every class on one
stream give each class kernel's own time, for the roofline of the dominant
kernel.

Also reported: the fused against the molecule-by-molecule Fock builds of the
batch; the served stream (Poisson arrivals of geometry trajectories, p50/p99
latency against the plain path); the single-molecule direct Fock build and
SCF solve on (H2O)16/cc-pVDZ (400 Cartesian functions, seed 2); density
fitting on (H2O)32/cc-pVDZ; and the CPU baseline: the oracle's restatement
of the reference algorithm (in-core packed ERI tensor, contract_jk, SAD,
ADIIS/CDIIS) on a random sample of the batched workload's molecules.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "hf-rs_amd")
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 vector peak


def progress(msg):
    """One line per finished phase on stderr (the JSON result is stdout's only line)."""
    print("bench: %s" % msg, file=sys.stderr, flush=True)


def load_pkg():
    spec = importlib.util.spec_from_file_location("hf_rs_amd", os.path.join(PKG, "__init__.py"),
                                                  submodule_search_locations=[PKG])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["hf_rs_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def cpu_baseline(M, count, threads, budget_s, seed=11):
    """The reference algorithm on the host: the oracle's in-core RHF
    (screened packed ERI tensor built once, contract_jk every iteration, SAD,
    ADIIS->CDIIS; src/scf.rs:410-627), FP64 with OpenMP, on a sample of the
    batched workload's molecules (a seeded random order, as many as fit a
    third of budget_s, at least two), after one warm-up solve; the sample is
    solved three times and the median rate reported, as solves/s and as Fock
    builds/s (the builds of the solves over their time, ERIs included)."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    mols = M.batch_config3(count, seed=3)
    text = M.basis_text("def2-svp")
    order = np.random.Generator(np.random.PCG64(seed)).permutation(count)
    warm = mols[int(order[-1])]
    O.scf_rhf(O.Basis.build(text, warm.symbols, warm.xyz), warm.zxyz, tol=1e-10, conv="adiissoscf")

    def solve(i):
        m = mols[int(i)]
        b = O.Basis.build(text, m.symbols, m.xyz)
        t0 = time.perf_counter()
        r, _, _, _ = O.scf_rhf(b, m.zxyz, tol=1e-10, conv="adiissoscf")
        return time.perf_counter() - t0, r.fock_builds, "%s(n=%d)" % (m.name, b.n)

    sample, names, rates, brates = [], [], [], []
    spent, builds = 0.0, 0
    for i in order[:-1]:
        dt, nb, name = solve(i)
        sample.append(i)
        names.append(name)
        spent += dt
        builds += nb
        if spent > budget_s / 3 and len(sample) >= 2:
            break
    rates.append(len(sample) / spent)
    brates.append(builds / spent)
    for _ in range(2):
        spent, builds = 0.0, 0
        for i in sample:
            dt, nb, _ = solve(i)
            spent += dt
            builds += nb
        rates.append(len(sample) / spent)
        brates.append(builds / spent)
    cpu = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": float(np.median(rates)), "unit": "SCF solves/s", "cores": threads, "kind": "port",
            "cpu_model": cpu, "fock_builds_per_s": float(np.median(brates)),
            "runs": [float(x) for x in rates],
            "sample": "oracle in-core RHF (reference algorithm) on %d of the %d batched-workload molecules in a "
                      "seeded random order (%s), median of 3 runs after a warm-up solve" % (
                          len(sample), count, ", ".join(names))}


def df_report(hf, M, steps=5, solves=True, basis="cc-pvdz"):
    """Density-fitted J/K on (H2O)32 (4x4x2 lattice, seed 4) in `basis`
    (cc-pVDZ; cc-pVTZ is the survey's config 4, n = 2080) with the
    def2-universal-JKFIT auxiliary basis (Cartesian), 160 occupied orbitals
    (random, fixed seed): the exchange contraction in FP64, FP32 and split
    BF16 on the matrix cores, each with its error against FP64.  At cc-pVTZ
    the factor is kept packed ([rank][n(n+1)/2], 74 GB) and formed one
    column chunk at a time."""
    mol = M.water_cluster((4, 4, 2), seed=4)
    b = hf.Basis.from_g94(M.basis_text(basis), mol.symbols, mol.xyz)
    a = hf.Basis.from_g94(M.basis_text("def2-universal-jkfit"), mol.symbols, mol.xyz)
    t0 = time.perf_counter()
    df = hf.DfTensor(b, a)
    build = time.perf_counter() - t0
    progress("DF %s: factor built in %.1f s" % (basis, build))
    o = 5 * 32
    c = np.random.default_rng(7).standard_normal((b.n, o)) * 0.1
    n, rank = b.n, df.rank
    k_flops = 2.0 * rank * n * n * o + 2.0 * n * n * rank * o  # A = B C and K = 2 A A^T as computed
    # algorithmic count of DF-K: A = B C plus the symmetric A A^T counted once
    k_alg = 2.0 * rank * n * n * o + 1.0 * rank * n * (n + 1) * o
    out = {"workload": "DF-J/K (H2O)32 %s / def2-universal-JKFIT, n %d, n_aux %d, rank %d, %d occupied"
                       % ({"cc-pvdz": "cc-pVDZ", "cc-pvtz": "cc-pVTZ"}.get(basis, basis), n, df.n_aux, rank, o),
           "build_s": build, "k_flops_per_build": k_flops,
           "k_algorithmic_flops_per_build": k_alg}
    j0, k0, _ = df.jk_precision(c, "fp64")
    scale = float(np.abs(k0).max())
    for prec, peak in (("fp64", 78.6), ("fp32", 157.3), ("bf16x3", 2500.0)):
        ms = []
        for _ in range(steps):
            j, k, t = df.jk_precision(c, prec)
            ms.append(t)
        t = float(np.median(ms))
        mult = 3 if prec == "bf16x3" else 1
        out[prec] = {"ms_per_build": t, "fock_builds_per_s": 1e3 / t,
                     "k_tflops_upper": k_flops / (t * 1e-3) / 1e12,
                     "k_algorithmic_tflops": k_alg / (t * 1e-3) / 1e12,
                     "mfma_peak_tflops": peak, "bf16_products": mult,
                     "max_abs_dK_rel": float(np.abs(k - k0).max()) / scale}
    df.close()
    if not solves:
        return out
    # One DF-RHF solve (run_df_hf_with_aux: core-H guess, CDIIS, FP64), DF
    # build included.
    r = hf.df_rhf(b, a, mol.zxyz)
    out["df_rhf_solve"] = {"seconds": r["seconds_total"], "seconds_fock": r["seconds_fock"],
                           "iterations": r["iterations"], "converged": bool(r["converged"]),
                           "energy": r["total_energy"], "solves_per_s": 1.0 / r["seconds_total"]}
    # The same with split-BF16 exchange until the residual is below 1e-4, FP64 after.
    m = hf.df_rhf(b, a, mol.zxyz, k_precision="bf16x3", switch_residual=1e-4)
    out["df_rhf_solve_bf16x3_then_fp64"] = {
        "seconds": m["seconds_total"], "seconds_fock": m["seconds_fock"], "iterations": m["iterations"],
        "converged": bool(m["converged"]), "energy": m["total_energy"], "solves_per_s": 1.0 / m["seconds_total"],
        "dE_vs_fp64_solve": m["total_energy"] - r["total_energy"]}
    return out


PROFILE = os.path.join(ROOT, "profiles", "fused_jk_config3_def2svp.json")


def fused_report(hf, M, count, steps, warmup, unbatched=True):
    """The hot path of the batched workload: one step = J and K of all
    `count` molecules' resident densities in fused launches (one per
    shell-quartet class over the whole batch).  Timed over `steps` builds;
    then `steps` builds with every class on one stream and per-class HIP
    events give each class kernel's own time; next to it the same builds
    molecule after molecule (each molecule's own launches).  The roofline is
    that of the class kernel with the largest time per launch: its executed
    FP64 work per launch (rocprofv3 SQ instruction counters, profiles/) over
    that time, against the FP64 peak."""
    mols = M.batch_config3(count, seed=3)
    text = M.basis_text("def2-svp")
    bases = [hf.Basis.from_g94(text, m.symbols, m.xyz) for m in mols]
    t0 = time.perf_counter()
    batch = hf.Batch(bases)
    setup = time.perf_counter() - t0
    progress("batch of %d molecules set up in %.1f s" % (count, setup))
    info = batch.info()
    rng = np.random.default_rng(6)
    for i, b in enumerate(bases):
        c = rng.standard_normal((b.n, max(1, b.n // 4))) * 0.3
        batch.set_density(i, c @ c.T)
    for _ in range(warmup):
        batch.jk()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        s0 = time.perf_counter()
        batch.jk()  # returns after the build is complete on the device
        lat.append(time.perf_counter() - s0)
    elapsed = time.perf_counter() - t0
    # Per-class times from as many builds again with every class on one
    # stream, so that each class kernel's time is its own (in the timed builds
    # the high-L classes overlap the others on a second stream).
    batch.profile(True, exclusive=True)
    for _ in range(steps):
        batch.jk()
    batch.profile(False)
    stats = batch.class_stats()
    each = float("nan")
    if unbatched:
        batch.jk_each()
        t1 = time.perf_counter()
        for _ in range(max(1, steps // 4)):
            batch.jk_each()
        each = (time.perf_counter() - t1) / max(1, steps // 4)
    fused = elapsed / steps
    dom = max(stats, key=lambda x: x["ms"] / max(1, x["launches"]))
    dom_ms = dom["ms"] / max(1, dom["launches"])
    prof = json.load(open(PROFILE))["per_kernel"] if os.path.exists(PROFILE) else {}
    # The class's kernel is k_jk_rows<la,lb,lc,ld,variant>: the profile entry of
    # that class with the longest dispatches (a batch of one basis runs one variant).
    prefix = "k_jk_rows<%d,%d,%d,%d," % tuple(dom["l"])
    cand = [k for k in prof if k.startswith(prefix)]
    kname = max(cand, key=lambda k: prof[k].get("avg_us", 0)) if cand else prefix + "?>"
    roof = {"bound": "fp64-valu", "kernel": kname, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "ms_per_launch": dom_ms, "achieved": None, "frac": None, "traffic": None,
            "md_equivalent": {"tflops": dom["flops"] / (dom_ms * 1e-3) / 1e12 if dom_ms > 0 else None,
                              "flops_per_launch": dom["flops"],
                              "model": "McMurchie-Davidson loop count per primitive quartet of every surviving "
                                       "shell quartet (the reference's work; general contraction and HGP do less, "
                                       "so this rate can exceed the peak)"}}
    if prof:
        t = prof.get(kname)
        if t and dom_ms > 0:
            ef = t["executed"]["f64_flops"]
            ach = ef / (dom_ms * 1e-3) / 1e12
            roof.update(achieved=ach, frac=ach / PEAK_FP64_TFLOPS, flops_per_launch=ef,
                        traffic=t["fetch_bytes"] + t["write_bytes"],
                        source="profiles/fused_jk_config3_def2svp.json: executed FP64 FLOPs per launch from "
                               "SQ_INSTS_VALU_{FMA,MUL,ADD,TRANS}_F64 x 64 lanes (FMA = 2); traffic = FETCH_SIZE "
                               "+ WRITE_SIZE, copy-calibrated",
                        profile_avg_us=t.get("avg_us"))
            assert roof["frac"] <= 1.0, roof
    out = {"workload": "%d molecules def2-SVP (alkanes C1-C12, water clusters 1-8; seed 3), Cartesian n %d..%d"
                       % (count, min(b.n for b in bases), max(b.n for b in bases)),
           "fock_builds_per_s": count / fused, "ms_per_batch_build": 1e3 * fused,
           "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)), "p99": 1e3 * float(np.percentile(lat, 99))},
           "shell_quartets": info["quartets"], "md_algorithmic_tflops": info["flops"] / fused / 1e12,
           "unbatched_fock_builds_per_s": count / each if unbatched else None,
           "unbatched_ms_per_batch": 1e3 * each if unbatched else None,
           "speedup_vs_unbatched": each / fused if unbatched else None, "setup_s": setup,
           "classes": [{"l": x["l"], "ms": x["ms"] / max(1, x["launches"]), "quartets": x["quartets"]}
                       for x in stats if x["quartets"]]}
    batch.close()
    return out, roof


def fock_report(hf, M, steps, warmup, solve_repeats, scf):
    """Direct RHF/cc-pVDZ on the seeded (H2O)16 cluster (4x2x2 lattice, 2.9 A,
    seed 2; 400 Cartesian / 384 spherical functions): Fock builds of one
    molecule (density and results in HBM), and full SCF solves (SAD,
    ADIIS->CDIIS, spherical orthogonaliser, tol 1e-10) by the plain driver."""
    mol = M.water_cluster((4, 2, 2), seed=2)
    text = M.basis_text("cc-pvdz")
    basis = hf.Basis.from_g94(text, mol.symbols, mol.xyz)
    t0 = time.perf_counter()
    ctx = hf.Context(basis)
    setup_s = time.perf_counter() - t0
    info = ctx.info()
    rng = np.random.default_rng(5)
    c = rng.standard_normal((basis.n, 40)) * 0.2
    ctx.set_density(c @ c.T)
    for _ in range(warmup):
        ctx.jk_resident()
    ctx.profile(True)
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        s0 = time.perf_counter()
        ctx.jk_resident()
        lat.append(time.perf_counter() - s0)
    elapsed = time.perf_counter() - t0
    ctx.profile(False)
    stats = ctx.class_stats()
    total_flops, prim_q = ctx.fock_flops()
    out = {"workload": "direct RHF/cc-pVDZ (H2O)16, 4x2x2 lattice 2.9 A, seed 2, Schwarz 1e-12",
           "n_cart": basis.n, "n_sph": 384, "shell_quartets": info["n_quartets"], "primitive_quartets": prim_q,
           "fock_builds_per_s": steps / elapsed, "ms_per_build": 1e3 * elapsed / steps,
           "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)), "p99": 1e3 * float(np.percentile(lat, 99))},
           "md_algorithmic_tflops": total_flops / (elapsed / steps) / 1e12, "setup_s": setup_s,
           "classes": [{"l": x["l"], "ms": x["ms"] / max(1, x["launches"]), "quartets": x["quartets"]}
                       for x in stats if x["quartets"]]}
    ctx.close()
    if scf:
        cold = hf.scf_rhf(basis, mol.zxyz, tol=1e-10, conv="adiissoscf", spherical=True)
        full = hf.scf_rhf(basis, mol.zxyz, tol=1e-10, conv="adiissoscf", spherical=True)
        sl = []
        for _ in range(solve_repeats):
            r = hf.scf_rhf(basis, mol.zxyz, tol=1e-10, conv="adiissoscf", spherical=True, incremental=True)
            sl.append(r["seconds_total"])
        out["scf"] = {"solves_per_s": 1.0 / float(np.median(sl)), "iterations": r["iterations"],
                      "converged": r["converged"], "energy": r["total_energy"],
                      "solve_latency_ms": {"p50": 1e3 * float(np.percentile(sl, 50)),
                                           "p99": 1e3 * float(np.percentile(sl, 99)), "solves": len(sl)},
                      "phases_s": {k: r["seconds_" + k] for k in ("setup", "one_electron", "guess", "fock",
                                                                  "linalg")},
                      "cold_seconds": cold["seconds_total"],
                      "fock": "incremental (J/K of the density change inside the iterations; full first and final)",
                      "full_build_solve": {"seconds": full["seconds_total"], "iterations": full["iterations"],
                                           "energy": full["total_energy"]}}
        ref = os.path.join(ROOT, "tests", "golden", "h2o16_ccpvdz_scf.json")
        if os.path.exists(ref):
            o = json.load(open(ref))
            out["scf"]["dE_vs_oracle"] = r["total_energy"] - o["total_energy"]
            out["scf"]["full_build_solve"]["dE_vs_oracle"] = full["total_energy"] - o["total_energy"]
            out["scf"]["oracle_iterations"] = o["iterations"]
            # the reference algorithm's in-core timing of this solve, from the fixture's own run
            out["scf"]["oracle_incore_record"] = {k: o[k] for k in o if "second" in k or "threads" in k}
    return out


def scf_batch_report(hf, M, count=512, capacity=0, plain_every=16):
    """Config 3 as RHF solves: `count` def2-SVP molecules (alkanes and water
    clusters, seed 3, Cartesian) solved together by the server (continuous
    batching over `capacity` slots, all by default), against the plain
    single-molecule driver on a sample of the same molecules (indices
    i % plain_every in {0, 1}: alkanes and water clusters alike).  The plain
    path runs the reference's defaults (ADIIS->CDIIS, tol 1e-10, FP64, full
    Fock builds).  The server runs twice: with the same full builds, and in its
    serving form, where the builds inside the iterations are J/K of the density
    change with the Schwarz threshold raised by 1/max|dP| (every 12th and the
    final build full, FP64 throughout); both forms' energies and iteration
    counts are compared with the plain path's."""
    mols = M.batch_config3(count, seed=3)
    text = M.basis_text("def2-svp")
    bases = [hf.Basis.from_g94(text, m.symbols, m.xyz) for m in mols]
    kw = dict(tol=1e-10, conv="adiissoscf")

    def served(incremental):
        srv = hf.Server(capacity=capacity or count)
        for b, m in zip(bases, mols):
            srv.submit([b], [m.zxyz], incremental=incremental, **kw)
        t0 = time.perf_counter()
        srv.run()
        dt = time.perf_counter() - t0
        res = [srv.result(i) for i in range(count)]
        st = srv.stats()
        srv.close()
        progress("batched RHF (%s builds): %d solves in %.2f s" % ("incremental" if incremental else "full", count,
                                                                   dt))
        return dt, res, st

    full = served(False)
    inc = served(True)
    sample = [i for i in range(count) if i % plain_every in (0, 1)]
    t0 = time.perf_counter()
    plain = {i: hf.scf_rhf(bases[i], mols[i].zxyz, **kw) for i in sample}
    tp = time.perf_counter() - t0
    progress("plain RHF: %d solves in %.2f s" % (len(sample), tp))
    plain_solves = len(sample) / tp

    def form(run):
        dt, res, st = run
        de = [res[i]["steps_result"][0]["energy"] - plain[i]["total_energy"] for i in sample]
        dit = [res[i]["steps_result"][0]["iterations"] - plain[i]["iterations"] for i in sample]
        lat = [r["finished_s"] - r["arrival_s"] for r in res]
        return {"solves_per_s": count / dt, "seconds": dt, "all_converged": all(r["converged"] for r in res),
                "solve_latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)),
                                     "p99": 1e3 * float(np.percentile(lat, 99))},
                "fused_fock_builds": st["fused_builds"],
                "seconds_waiting_for_setup": st["seconds_setup"],  # slot set-ups run beside the iterations
                "jk_config": "%d of %d (autotuned)" % (st["jk_config"], st["jk_candidates"]),
                "speedup_vs_plain": (count / dt) / plain_solves,
                "max_abs_dE_vs_plain": float(np.abs(de).max()), "max_abs_diter_vs_plain": int(np.abs(dit).max())}

    out = {"workload": "%d RHF solves def2-SVP (alkanes C1-C12, water clusters 1-8; seed 3), Cartesian n %d..%d, "
                       "ADIIS->CDIIS tol 1e-10, FP64" % (count, min(b.n for b in bases), max(b.n for b in bases)),
           "slots": capacity or count,
           "plain_gpu": {"solves_per_s": plain_solves, "sample": len(sample), "fock": "full builds",
                         "sample_rule": "molecules with index %% %d in {0, 1}" % plain_every},
           "full_builds": form(full), "serving_form": form(inc)}
    out["serving_form"]["fock"] = "J/K of the density change, threshold 1e-12/max|dP|; full first, every 12th, final"
    # the headline: the serving form when every sampled energy is within
    # 1e-8 Eh of the plain path's, else the full-build form
    head = out["serving_form"] if out["serving_form"]["max_abs_dE_vs_plain"] <= 1e-8 else out["full_builds"]
    out["headline"] = "serving_form" if head is out["serving_form"] else "full_builds"
    out.update(solves_per_s=head["solves_per_s"], speedup_vs_plain=head["speedup_vs_plain"])
    return out


def config5_stream(hf, M, n_req, lam, steps, seed=5):
    """The served workload: requests with Poisson arrivals (rate lam), each a
    `steps`-step geometry trajectory (seeded Gaussian displacements, sigma
    0.01 bohr per coordinate and step) of an alkane (n ~ U{1..20}) or a water
    cluster (m ~ U{1..32}), half and half, in STO-3G, 6-31G* or cc-pVDZ."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out, t = [], 0.0
    for i in range(n_req):
        t += rng.exponential(1.0 / lam)
        if rng.random() < 0.5:
            mol = M.alkane(int(rng.integers(1, 21)))
        else:
            mol = M.water_cluster_m(int(rng.integers(1, 33)), seed=int(rng.integers(0, 2 ** 31)))
        basis = ["sto-3g", "6-31g*", "cc-pvdz"][int(rng.integers(0, 3))]
        geoms = [mol.xyz]
        for _ in range(steps - 1):
            geoms.append(geoms[-1] + rng.normal(0.0, 0.01, mol.xyz.shape))
        out.append((t, mol, basis, geoms))
    return out


def serve_report(hf, M, n_req=16, lam=20.0, steps=10, capacity=16):
    stream = config5_stream(hf, M, n_req, lam, steps)
    kw = dict(tol=1e-10, conv="adiissoscf")
    reqs = []
    for t, mol, basis, geoms in stream:
        text = M.basis_text(basis)
        mols = [M.Molecule(mol.symbols, g) for g in geoms]
        reqs.append((t, [hf.Basis.from_g94(text, m.symbols, m.xyz) for m in mols], [m.zxyz for m in mols], mol, basis))
    srv = hf.Server(capacity=capacity)
    for t, bases, zs, _, _ in reqs:
        srv.submit(bases, zs, arrival=t, incremental=True, **kw)  # the serving form (see scf_batch_report)
    t0 = time.perf_counter()
    srv.run()
    dt = time.perf_counter() - t0
    res = [srv.result(i) for i in range(len(reqs))]
    srv.close()
    lat = [r["finished_s"] - r["arrival_s"] for r in res]
    solves = sum(r["steps"] for r in res)
    progress("served: %d requests (%d solves) in %.2f s" % (len(reqs), solves, dt))
    # The plain path: one request after another, each step solved alone from
    # the previous step's density (hfg_scf_rhf_guess); latency from the
    # queue of a single worker fed by the same arrivals.
    t_free, plat, e_err, ptime = 0.0, [], 0.0, 0.0
    for (t, bases, zs, _, _), r in zip(reqs, res):
        s0 = time.perf_counter()
        guess = None
        for k, (b, z) in enumerate(zip(bases, zs)):
            o = hf.scf_rhf(b, z, guess=guess, want_matrices=True, **kw)
            guess = o["density"]
            e_err = max(e_err, abs(o["total_energy"] - r["steps_result"][k]["energy"]))
        svc = time.perf_counter() - s0
        ptime += svc
        start = max(t, t_free)
        t_free = start + svc
        plat.append(t_free - t)
    progress("plain served: %.2f s of solves" % ptime)
    return {"workload": "%d requests x %d geometry steps, Poisson %.0f req/s (seed 5): alkanes C1-C20 / water clusters "
                        "1-32 in STO-3G/6-31G*/cc-pVDZ, n %d..%d" % (len(reqs), steps, lam,
                                                                   min(r["n"] for r in res), max(r["n"] for r in res)),
            "slots": capacity, "solves_per_s": solves / dt, "seconds": dt,
            "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)), "p99": 1e3 * float(np.percentile(lat, 99))},
            "all_converged": all(r["converged"] for r in res),
            "plain_gpu": {"solves_per_s": solves / ptime,
                          "latency_ms": {"p50": 1e3 * float(np.percentile(plat, 50)),
                                         "p99": 1e3 * float(np.percentile(plat, 99))},
                          "rule": "requests in arrival order, one at a time, each step warm-started"},
            "fock": "server: J/K of the density change inside the iterations (full first, every 12th, final); "
                    "plain: full builds",
            "speedup_vs_plain": (solves / dt) / (solves / ptime), "max_abs_dE_vs_plain": e_err}


def gradient_report(hf, M):
    """Analytic RHF gradients (hfg_rhf_gradient) at converged densities of
    molecules of the served stream's families, timed per gradient; the forces
    of each sum to zero (translational invariance)."""
    cases = [("H2O 6-31G*", M.water_config1(), "6-31g*"), ("(H2O)4 cc-pVDZ", M.water_cluster((2, 2, 1), seed=5),
                                                            "cc-pvdz"),
             ("C6H14 def2-SVP", M.alkane(6), "def2-svp")]
    out = []
    for name, mol, basis in cases:
        b = hf.Basis.from_g94(M.basis_text(basis), mol.symbols, mol.xyz)
        r = hf.scf_rhf(b, mol.zxyz, tol=1e-10, conv="adiissoscf", want_matrices=True)
        ctx = hf.Context(b)
        p, f = r["density"], r["fock"]
        w = 0.5 * p @ f @ p
        ctx.rhf_gradient(mol.zxyz, p, w)  # warm-up
        t0 = time.perf_counter()
        g = ctx.rhf_gradient(mol.zxyz, p, w)
        dt = time.perf_counter() - t0
        ctx.close()
        out.append({"molecule": name, "n": b.n, "atoms": len(mol.symbols), "seconds": dt,
                    "max_abs_force": float(np.abs(g).max()), "max_abs_force_sum": float(np.abs(g.sum(0)).max())})
    return out


def summarize_profiles(trace_csv, pmc_dirs, copy_dirs, out_json, workload, commands, last=3):
    """Turns rocprofv3 CSV output into the per-kernel summary bench.py reads:
    average duration of each kernel's last `last` dispatches (kernel trace:
    the timed builds, after the warm-up and tuning ones), HBM bytes per dispatch (FETCH_SIZE and
    WRITE_SIZE in KiB, scaled by the same counters on a copy kernel of known
    size) and executed FP64 FLOPs per dispatch (SQ instruction counters x 64
    lanes, FMA = 2 FLOPs)."""
    import csv
    import glob

    def rows(d, pat):
        out = []
        for f in glob.glob(os.path.join(d, "**", pat), recursive=True):
            out += list(csv.DictReader(open(f)))
        return out

    def short(name):
        k = name.replace(" ", "").split("(")[0].replace("hfg::", "")
        return k[4:] if k.startswith("void") else k
    per = {}
    disp = {}
    for r in rows(trace_csv, "*kernel_trace.csv"):
        disp.setdefault(short(r["Kernel_Name"]), []).append((int(r["Start_Timestamp"]), int(r["End_Timestamp"])))
    for k, v in disp.items():
        v.sort()
        d = [(e - s0) / 1e3 for s0, e in v[-last:]]
        per.setdefault(k, {})["avg_us"] = sum(d) / len(d)
        per[k]["dispatches"] = len(v)
    cal = {}
    for ctr, d in copy_dirs.items():
        v = [float(r["Counter_Value"]) for r in rows(d, "*counter_collection.csv")
             if "copy_calibration" in r["Kernel_Name"] and r["Counter_Name"] == ctr]
        cal[ctr] = (2 ** 30 / 1024.0) / max(v) if v else 1.0  # a 1 GiB copy reads and writes 1 GiB
    sums = {}
    for d in pmc_dirs:
        for r in rows(d, "*counter_collection.csv"):
            k = short(r["Kernel_Name"])
            sums.setdefault(k, {}).setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    for k, ctrs in sums.items():
        e = per.setdefault(k, {})
        avg = {c: sum(v) / len(v) for c, v in ctrs.items()}
        if "FETCH_SIZE" in avg:
            e["fetch_bytes"] = avg["FETCH_SIZE"] * 1024 * cal.get("FETCH_SIZE", 1.0)
        if "WRITE_SIZE" in avg:
            e["write_bytes"] = avg["WRITE_SIZE"] * 1024 * cal.get("WRITE_SIZE", 1.0)
        f64 = {c: avg[c] for c in avg if c.startswith("SQ_")}
        if "SQ_INSTS_VALU_FMA_F64" in f64:
            f64["f64_flops"] = 64.0 * (2 * f64["SQ_INSTS_VALU_FMA_F64"] + f64.get("SQ_INSTS_VALU_MUL_F64", 0) +
                                       f64.get("SQ_INSTS_VALU_ADD_F64", 0) + f64.get("SQ_INSTS_VALU_TRANS_F64", 0))
            e["executed"] = f64
    json.dump({"workload": workload, "copy_calibration": cal, "commands": commands, "per_kernel": per},
              open(out_json, "w"), indent=1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed fused Fock builds of the batch")
    ap.add_argument("--warmup", type=int, default=6,
                    help="untimed fused builds (the first five also tune the class-launch configuration)")
    ap.add_argument("--batch-count", type=int, default=512)
    ap.add_argument("--only-fused", action="store_true", help="only the fused Fock builds (profiling runs)")
    ap.add_argument("--only-fock", action="store_true",
                    help="only the (H2O)16/cc-pVDZ Fock builds, without solves (profiling runs)")
    ap.add_argument("--only-df", action="store_true",
                    help="only the DF build and DF J/K builds, without the DF-RHF solves (profiling runs)")
    ap.add_argument("--no-solves", action="store_true", help="skip the batched SCF solves (the headline)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=16)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--no-fock", action="store_true", help="skip the (H2O)16 single-molecule measurements")
    ap.add_argument("--fock-steps", type=int, default=20)
    ap.add_argument("--solve-repeats", type=int, default=5)
    ap.add_argument("--no-df", action="store_true")
    ap.add_argument("--no-df4", action="store_true", help="skip the config-4 DF (cc-pVTZ, packed factor)")
    ap.add_argument("--only-df4", action="store_true",
                    help="only the config-4 DF build and DF J/K builds, without solves (profiling runs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="run only the counter-calibration copy kernel over BYTES bytes and exit")
    ap.add_argument("--summarize-profiles", nargs=3, metavar=("TRACE_DIR", "PMC_DIRS", "COPY_DIRS"),
                    help="write the per-kernel profile summary from rocprofv3 CSV directories "
                         "(PMC_DIRS comma-separated; COPY_DIRS as FETCH_SIZE=dir,WRITE_SIZE=dir) and exit")
    ap.add_argument("--profile-out", default=PROFILE)
    ap.add_argument("--profile-last", type=int, default=3, help="timed dispatches per kernel in the trace")
    args = ap.parse_args()
    if args.summarize_profiles:
        tdir, pdirs, cdirs = args.summarize_profiles
        copy = dict(x.split("=", 1) for x in cdirs.split(",") if x)
        summarize_profiles(tdir, pdirs.split(","), copy, args.profile_out,
                           "fused J/K of the %d-molecule def2-SVP batch (seed 3), per launch" % args.batch_count,
                           {"trace": tdir, "pmc": pdirs, "copy": cdirs}, last=args.profile_last)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this engine's workloads are single-GPU; run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    hf = load_pkg()
    sys.path.insert(0, PKG)
    if args.calibrate_copy:
        ms = hf.calibration_copy(args.calibrate_copy)
        print(json.dumps({"copy_bytes": args.calibrate_copy, "ms": ms,
                          "GB_per_s": 2 * args.calibrate_copy / ms / 1e6}))
        return
    import molecules as M
    if hf.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.only_fock or args.only_df or args.only_df4:
        sub = (fock_report(hf, M, args.fock_steps, 3, 0, scf=False) if args.only_fock
               else df_report(hf, M, steps=args.steps, solves=False, basis="cc-pvtz" if args.only_df4 else "cc-pvdz"))
        print(json.dumps(sub), flush=True)
        return
    fused, roof = fused_report(hf, M, args.batch_count, args.steps, args.warmup, unbatched=not args.only_fused)
    progress("fused builds: %.1f ms per batch build" % fused["ms_per_batch_build"])
    line = {"metric": "Fock builds/s & SCF solves/s at |dE|<=1e-8 Eh vs plain GPU; p50/p99 solve latency",
            "value": None, "unit": "SCF solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": fused["ms_per_batch_build"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded geometries: all-trans alkanes, water clusters)",
            "config": {"workload": "batched RHF: %d def2-SVP solves (alkanes C1-C12 / water clusters 1-8, seed 3), "
                                   "FP64, served form (incremental Fock builds, final build full); a step = one "
                                   "fused Fock build of the whole batch" % args.batch_count,
                       "molecules": args.batch_count, "basis": "def2-SVP (Cartesian)", "parallelism": "single GPU"},
            "roofline": roof, "fused_fock_builds": fused}
    if args.only_fused:
        print(json.dumps(line), flush=True)
        return
    if not args.no_solves:
        sb = scf_batch_report(hf, M, args.batch_count)
        line["value"] = sb["solves_per_s"]
        line["batched_solves"] = sb
        progress("batched solves: %.2f solves/s (%s; full builds %.2f; plain %.2f)" % (
            sb["solves_per_s"], sb["headline"], sb["full_builds"]["solves_per_s"], sb["plain_gpu"]["solves_per_s"]))
    if not args.no_serve:
        line["served"] = serve_report(hf, M, args.serve_requests, args.serve_rate)
        progress("served stream done")
    if not args.no_serve:
        line["gradient"] = gradient_report(hf, M)
        progress("gradients done")
    if not args.no_fock:
        line["h2o16_ccpvdz"] = fock_report(hf, M, args.fock_steps, 3, args.solve_repeats, scf=True)
        progress("(H2O)16 done")
    if not args.no_df:
        line["df"] = df_report(hf, M)
        progress("DF done")
    if not args.no_df4:
        line["df_config4"] = df_report(hf, M, steps=3, basis="cc-pvtz")
        progress("DF config 4 done")
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(M, args.batch_count, args.cpu_threads, args.cpu_seconds)
        progress("CPU baseline done")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
