"""Throughput of the Fock-build hot path on one MI355X.

Workload: direct RHF/cc-pVDZ on a seeded (H2O)16 cluster (4x2x2 lattice, 2.9 A
spacing, PCG64 seed 2, uniform rotations, +-0.1 A jitter), 400 Cartesian /
384 spherical functions, Schwarz threshold 1e-12, FP64 throughout.  A step is
one Fock build: J and K of one density over every surviving shell quartet,
with the density and the results resident in HBM.  value = Fock builds/s.
Also reported: a full SCF solve (SAD guess, ADIIS->CDIIS, spherical
orthogonaliser, tol 1e-10, incremental Fock builds inside the iterations) as
SCF solves/s with its iteration count and energy error against the oracle's
in-core solve (a later solve in the process; the first, with one-time costs,
and a solve with full builds only, beside it), the
per-step latency percentiles, the roofline of the dominant shell-quartet class
kernel and of the whole build, and the oracle's CPU time on a bounded sample.
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


def load_pkg():
    spec = importlib.util.spec_from_file_location("hf_rs_amd", os.path.join(PKG, "__init__.py"),
                                                  submodule_search_locations=[PKG])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["hf_rs_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def cpu_baseline(mol, text, stride, threads, density):
    """Oracle direct J/K (the reference's McMurchie-Davidson algorithm, FP64,
    OpenMP) on the bra pairs ij % stride == 0, scaled by algorithmic FLOPs."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    ob = O.Basis.build(text, mol.symbols, mol.xyz)
    total, _, _ = ob.fock_flops()
    sample, _, _ = ob.fock_flops(stride, 0)
    _, _, sec = ob.direct_jk(density, stride, 0)
    est = sec * total / sample
    return {"value": 1.0 / est, "unit": "Fock builds/s", "cores": threads, "kind": "port",
            "sample": "oracle direct J/K over bra shell pairs ij %% %d == 0 (%.1f%% of the algorithmic FLOPs), "
                      "%.1f s on %d threads, scaled to the full build" % (stride, 100 * sample / total, sec,
                                                                          threads)}


def df_report(hf, M, steps=5):
    """Density-fitted J/K on (H2O)32/cc-pVDZ (4x4x2 lattice, seed 4) with the
    def2-universal-JKFIT auxiliary basis (Cartesian), 160 occupied orbitals
    (random, fixed seed): the exchange contraction in FP64, FP32 and split
    BF16 on the matrix cores, each with its error against FP64."""
    mol = M.water_cluster((4, 4, 2), seed=4)
    b = hf.Basis.from_g94(M.basis_text("cc-pvdz"), mol.symbols, mol.xyz)
    a = hf.Basis.from_g94(M.basis_text("def2-universal-jkfit"), mol.symbols, mol.xyz)
    t0 = time.perf_counter()
    df = hf.DfTensor(b, a)
    build = time.perf_counter() - t0
    o = 5 * 32
    c = np.random.default_rng(7).standard_normal((b.n, o)) * 0.1
    n, rank = b.n, df.rank
    k_flops = 2.0 * rank * n * n * o + 2.0 * n * n * rank * o  # A = B C and K = 2 A A^T as computed
    # algorithmic count of DF-K: A = B C plus the symmetric A A^T counted once
    k_alg = 2.0 * rank * n * n * o + 1.0 * rank * n * (n + 1) * o
    out = {"workload": "DF-J/K (H2O)32 cc-pVDZ / def2-universal-JKFIT, n %d, n_aux %d, rank %d, %d occupied"
                       % (n, df.n_aux, rank, o), "build_s": build, "k_flops_per_build": k_flops,
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


def batch_report(hf, M, count, steps=5):
    """The batched workload: `count` molecules (half all-trans alkanes
    CnH2n+2, n ~ U{1..12}, half water clusters (H2O)m, m ~ U{1..8}; PCG64 seed
    3) in def2-SVP; one step = J/K of every molecule in fused launches, next to
    the same builds molecule after molecule."""
    mols = M.batch_config3(count, seed=3)
    text = M.basis_text("def2-svp")
    bases = [hf.Basis.from_g94(text, m.symbols, m.xyz) for m in mols]
    t0 = time.perf_counter()
    batch = hf.Batch(bases)
    setup = time.perf_counter() - t0
    info = batch.info()
    rng = np.random.default_rng(6)
    for i, b in enumerate(bases):
        c = rng.standard_normal((b.n, max(1, b.n // 4))) * 0.3
        batch.set_density(i, c @ c.T)
    batch.jk()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.jk()
    fused = (time.perf_counter() - t0) / steps
    batch.jk_each()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.jk_each()
    each = (time.perf_counter() - t0) / steps
    return {"workload": "%d molecules def2-SVP (alkanes C1-C12, water clusters 1-8; seed 3), Cartesian n %d..%d"
                        % (count, min(b.n for b in bases), max(b.n for b in bases)),
            "fock_builds_per_s": count / fused, "ms_per_batch": 1e3 * fused,
            "algorithmic_tflops": info["flops"] / fused / 1e12, "shell_quartets": info["quartets"],
            "unbatched_fock_builds_per_s": count / each, "unbatched_ms_per_batch": 1e3 * each,
            "speedup_vs_unbatched": each / fused, "setup_s": setup}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-stride", type=int, default=24)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-scf", action="store_true")
    ap.add_argument("--solve-repeats", type=int, default=8, help="warm SCF solves for the solve-latency percentiles")
    ap.add_argument("--batch", action="store_true",
                    help="also time the batched multi-molecule workload (off: see DESIGN.md, batches)")
    ap.add_argument("--batch-count", type=int, default=512)
    ap.add_argument("--no-df", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="run only the counter-calibration copy kernel over BYTES bytes and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this engine's workloads are single-GPU (see DESIGN.md, sharding); "
              "run with --gpus 1", file=sys.stderr)
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
    mol = M.water_cluster((4, 2, 2), seed=2)
    text = M.basis_text("cc-pvdz")
    basis = hf.Basis.from_g94(text, mol.symbols, mol.xyz)
    t0 = time.perf_counter()
    ctx = hf.Context(basis)
    setup_s = time.perf_counter() - t0
    info = ctx.info()
    rng = np.random.default_rng(5)
    c = rng.standard_normal((basis.n, 40)) * 0.2
    dens = c @ c.T  # symmetric positive semidefinite; work does not depend on values
    ctx.set_density(dens)
    for _ in range(args.warmup):
        ctx.jk_resident()
    ctx.profile(True)
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        ctx.jk_resident()  # returns after the build completes on the device
        lat.append(time.perf_counter() - s0)
    elapsed = time.perf_counter() - t0
    ctx.profile(False)  # keeps the accumulated statistics until the next enable
    stats = ctx.class_stats()
    total_flops, prim_q = ctx.fock_flops()
    value = args.steps / elapsed
    ms_classes = sum(s["ms"] for s in stats) / args.steps
    # Dominant kernel: the class carrying the most algorithmic work (the
    # high-L classes run on a second stream, where event times include the
    # overlap with the main stream's classes).
    dom = max(stats, key=lambda s: s["flops"])
    dom_ms = dom["ms"] / max(1, dom["launches"])
    dom_tflops = dom["flops"] / (dom_ms * 1e-3) / 1e12 if dom_ms > 0 else 0.0
    ms_step = 1e3 * elapsed / args.steps
    whole_tflops = total_flops / (ms_step * 1e-3) / 1e12
    # The J/K kernels keep K rows in LDS (k_jk_rows) unless HFG_JK_ROWS=0.
    kname = ("k_jk_rows<%d, %d, %d, %d>" if os.environ.get("HFG_JK_ROWS", "1") != "0"
             else "k_jk<%d, %d, %d, %d>") % tuple(dom["l"])
    traffic, traffic_src, executed = None, None, None
    tpath = os.path.join(ROOT, "profiles", "traffic_jk_h2o16_ccpvdz.json")
    if os.path.exists(tpath):
        t = json.load(open(tpath))["per_kernel"].get(kname.replace(" ", ""))
        if t and "fetch_bytes" in t:
            traffic = t["fetch_bytes"] + t["write_bytes"]
            traffic_src = "profiles/traffic_jk_h2o16_ccpvdz.json (FETCH_SIZE + WRITE_SIZE, copy-calibrated)"
        if t and "executed" in t and dom_ms > 0:
            # FP64 work the kernel executes (SQ instruction counters x 64 lanes)
            # over this run's launch time: the hardware utilisation, where the
            # algorithmic figure above counts the reference's per-primitive
            # work, which general contraction avoids in part.
            ef = t["executed"]["f64_flops"]
            executed = {"flops_per_launch": ef, "achieved": ef / (dom_ms * 1e-3) / 1e12,
                        "frac": ef / (dom_ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
                        "source": "profiles/traffic_jk_h2o16_ccpvdz.json (SQ_INSTS_VALU_{FMA,MUL,ADD,TRANS}_F64 x 64)"}
    line = {
        "metric": "Fock builds/s (direct FP64 J+K, |dE|<=1e-8 Eh vs FP64 oracle)",
        "value": value, "unit": "Fock builds/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded water cluster geometry)",
        "config": {"workload": "direct RHF/cc-pVDZ (H2O)16, 4x2x2 lattice 2.9 A, seed 2, Schwarz 1e-12",
                   "n_cart": basis.n, "n_sph": 384, "shell_quartets": info["n_quartets"],
                   "primitive_quartets": prim_q, "parallelism": "single GPU"},
        "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)), "p99": 1e3 * float(np.percentile(lat, 99))},
        "setup_s": setup_s,
        "roofline": {"bound": "fp64-valu", "kernel": kname,
                     "achieved": dom_tflops, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": dom_tflops / PEAK_FP64_TFLOPS, "traffic": traffic, "traffic_source": traffic_src,
                     "flops_per_launch": dom["flops"], "ms_per_launch": dom_ms,
                     "flop_model": "MD per primitive quartet of each contracted quartet (the reference's loops)",
                     "executed": executed},
        "fock_build_roofline": {"achieved": whole_tflops, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                                "frac": whole_tflops / PEAK_FP64_TFLOPS, "algorithmic_flops": total_flops,
                                "wall_ms": ms_step, "class_kernel_ms_summed": ms_classes},
        "classes": [{"l": s["l"], "ms": s["ms"] / max(1, s["launches"]), "gflop": s["flops"] / 1e9,
                     "quartets": s["quartets"]} for s in stats if s["quartets"]],
    }
    if not args.no_scf:
        # The first solve in a process also pays one-time costs (library
        # initialisation, the per-element SAD cache); the second is what a
        # serving process sees.  Both are reported, solves/s from the second.
        cold = hf.scf_rhf(basis, mol.zxyz, tol=1e-10, conv="adiissoscf", spherical=True)
        full = hf.scf_rhf(basis, mol.zxyz, tol=1e-10, conv="adiissoscf", spherical=True)
        r = hf.scf_rhf(basis, mol.zxyz, tol=1e-10, conv="adiissoscf", spherical=True, incremental=True)
        # Solve latency over repeated warm solves (the same molecule, each a
        # complete solve from the geometry: set-up, guess, iterations).
        sl = [r["seconds_total"]]
        for _ in range(args.solve_repeats - 1):
            sl.append(hf.scf_rhf(basis, mol.zxyz, tol=1e-10, conv="adiissoscf", spherical=True,
                                 incremental=True)["seconds_total"])
        line["scf"] = {"solves_per_s": 1.0 / r["seconds_total"], "seconds": r["seconds_total"],
                       "solve_latency_ms": {"p50": 1e3 * float(np.percentile(sl, 50)),
                                            "p99": 1e3 * float(np.percentile(sl, 99)), "solves": len(sl)},
                       "cold_seconds": cold["seconds_total"],
                       "iterations": r["iterations"], "converged": r["converged"], "fock_builds": r["fock_builds"],
                       "energy": r["total_energy"], "cold_energy": cold["total_energy"],
                       "phases_s": {k: r["seconds_" + k] for k in ("setup", "one_electron", "guess", "fock",
                                                                   "linalg")},
                       "cold_phases_s": {k: cold["seconds_" + k] for k in ("setup", "one_electron", "guess",
                                                                           "fock", "linalg")},
                       "fock": "incremental (J/K of the density change inside the iterations; full first and "
                               "final builds)",
                       "full_build_solve": {"seconds": full["seconds_total"], "iterations": full["iterations"],
                                            "energy": full["total_energy"]}}
        ref = os.path.join(ROOT, "tests", "golden", "h2o16_ccpvdz_scf.json")
        if os.path.exists(ref):
            o = json.load(open(ref))
            line["scf"]["dE_vs_oracle"] = r["total_energy"] - o["total_energy"]
            line["scf"]["full_build_solve"]["dE_vs_oracle"] = full["total_energy"] - o["total_energy"]
            line["scf"]["oracle_iterations"] = o["iterations"]
    if args.batch:
        line["batch"] = batch_report(hf, M, args.batch_count)
    if not args.no_df:
        line["df"] = df_report(hf, M)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(mol, text, args.cpu_stride, args.cpu_threads, dens)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
