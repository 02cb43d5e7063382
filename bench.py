"""Throughput of the LOCO compat Wald scan on one MI355X: Wald tests/s (SNP x phenotype x permutation).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 2] [--precision P] [--no-compare] [--no-cpu]
    python bench.py --config 3      (uneven NA groups: grouped launch vs per-group loop, see run_grouped)
    python bench.py --config 5      (large-cohort low-rank LOCO, see run_lowrank)

Workload (one "step"): the fastlmm-compat leave-one-chromosome-out scan that perm_threshold drives, over one
synthetic panel resident in HBM -- per chromosome the LOCO kinship, the projected FP64 eigendecomposition, the
rotation of every scan column, the per-column delta search, the rotation of the chromosome's SNPs and the fused
Wald scan, with full F/beta/SE/SFVE/nullh2 detail for the real phenotypes and the genome-wide max F of every
column.  Default workload: N=1,000 x M=100,000 (22 chromosomes) x 6,484 RINT phenotypes x (1 + 100
permutations), C=4 -- 6.5488e10 Wald tests per step.

The timed path is split-FP16 MFMA (power-of-two-scaled fp16 hi + lo, 3 products, FP32 accumulation, FP64 epilogue);
the same step on the split-BF16 and FP32 MFMA paths and the FP64 CPU oracle (on a bounded sample) are reported
beside it, followed (unless --no-others) by one step of the uneven-NA-group workload (--config 3) and of the
large-cohort low-rank workload (--config 5) as extra keys.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "fast-er-lmm_amd"))
sys.path.insert(0, str(ROOT))

MI355X_PEAK_BF16_TFLOPS = 2500.0  # dense bf16 MFMA
MI355X_PEAK_FP32_TFLOPS = 157.3   # dense fp32 MFMA


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--n-perm", type=int, default=None)
    ap.add_argument("--pheno", type=int, default=None, help="override the number of real phenotypes")
    ap.add_argument("--snps", type=int, default=None, help="override M")
    ap.add_argument("--n-samples", type=int, default=None, help="override N (config 5)")
    ap.add_argument("--rank", type=int, default=None, help="kinship rank (config 5)")
    ap.add_argument("--precision", default="split_fp16", choices=["split_fp16", "split_bf16", "fp32"])
    ap.add_argument("--no-compare", action="store_true", help="skip the split-BF16 / FP32 comparison steps")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-others", action="store_true",
                    help="config 2: skip the uneven-NA-group and low-rank cohort summaries that follow the headline")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--pipeline", action="store_true",
                    help="LOCO: prepare chromosome c + 1 on a second stream while chromosome c scans")
    ap.add_argument("--screen", default="auto", choices=["auto", "off"],
                    help="permutation columns: the certified screen (auto) or the full split scan (off)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run one plain device copy of this many MiB first (reference kernel for HBM counters)")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(cfg_id, n_perm, pheno, snps):
    import numpy as np
    import torch
    from fasterlmm_amd import sim
    cfg = dict(sim.CONFIGS[cfg_id])
    if pheno:
        cfg["P"] = pheno
    if snps:
        cfg["M"] = snps
    n_perm = cfg["n_perm"] if n_perm is None else n_perm
    t0 = time.time()
    panel = sim.make_panel(cfg, seed=19930909)
    log(f"[bench] panel N={cfg['N']} M={cfg['M']} P={cfg['P']} C={cfg['C']} in {time.time() - t0:.1f}s")
    dev = torch.device("cuda")
    from fasterlmm_amd.io import standardise_columns
    Z = standardise_columns(torch.from_numpy(panel.G).to(dev))  # int8 -> FP64 standardised on the GPU
    X = torch.from_numpy(panel.X).to(dev)
    y_real = torch.from_numpy(panel.Y).to(dev)
    N, B = y_real.shape
    # permutation columns: per-phenotype row shuffles drawn on the GPU (synthetic data; the numpy-seeded
    # reference scheme lives in fasterlmm_amd.perms and is exercised by the tests)
    g = torch.Generator(device=dev)
    g.manual_seed(19930909)
    Y_all = torch.empty(N, B * (1 + n_perm), dtype=torch.float64, device=dev)
    Y_all[:, :B] = y_real
    for b0 in range(0, B, 256):
        b1 = min(B, b0 + 256)
        orders = torch.rand(b1 - b0, n_perm, N, generator=g, device=dev).argsort(dim=2)  # (b, j, N)
        yp = torch.gather(y_real[:, b0:b1].T.unsqueeze(1).expand(b1 - b0, n_perm, N), 2, orders)
        Y_all[:, B + b0 * n_perm:B + b1 * n_perm] = yp.reshape((b1 - b0) * n_perm, N).T
    torch.cuda.synchronize()
    return cfg, panel, Z, X, Y_all, B, n_perm


STEP_SETTINGS: dict = {}  # engine settings of the timed steps (--pipeline, --screen)


def run_step(Z, X, Y_all, chrom, B, precision):
    from fasterlmm_amd import core, engine
    with engine.configure(precision=precision, **STEP_SETTINGS):
        return core.loco_scan_compat(Z, X, Y_all, chrom, n_real=B, n_grid=100)


def compat_cpu_check(panel, res, B, others=None, *, Z=None, X=None, Y_all=None, n_top=50, frac=0.01, n_perm_cols=2000):
    """Full-size parity of the GPU results against the FP64 oracle (numpy, host cores): chromosome 1's LOCO
    kinship, projected eigh, rotation and per-column delta (100-point grid + 50 golden-section steps) for the B
    real phenotypes, then the Wald tests of a 1% SNP sample of chromosome 1 plus its n_top SNPs of largest GPU F,
    against every real phenotype.  Reports max |d log10 p| (tests whose p is below the double range in scipy's
    tail are compared by F), max |d beta| / max(|beta|, se), the relative beta error on tests with |beta| >= se,
    and the GPU's log delta against the oracle's, for the headline result and for each of `others` ({name: result},
    e.g. the split-BF16 run).  With Z / X / Y_all (the GPU inputs) it also checks the permutation columns: a random
    sample of n_perm_cols of them is scanned on the GPU over chromosome 1 with its LOCO kinship (the same kernels
    as the headline) and their maximum F over the chromosome's SNPs is compared with the oracle's."""
    import numpy as np
    import torch
    from scipy import stats
    from fasterlmm_amd import core
    from oracle import lmm
    t0 = time.perf_counter()
    Zs = lmm.standardise_columns(panel.G.astype(np.float64))
    chrom = panel.chrom
    c = sorted(np.unique(chrom))[0]
    test = np.flatnonzero(chrom == c)
    K = lmm.grm(Zs[:, chrom != c])
    spec = lmm.compat_rotate(K, panel.X, panel.Y[:, :B])
    ld = lmm.fit_delta_grid_compat(spec, n_grid=100)
    f_gpu = res.f.index_select(0, torch.from_numpy(test).to(res.f.device)).cpu().numpy()
    rng = np.random.default_rng(7)
    top = np.argsort(-f_gpu.max(axis=1))[:n_top]
    pick = np.unique(np.concatenate([rng.choice(test.size, max(1, int(frac * test.size)), replace=False), top]))
    rows = test[pick]
    ref = lmm.snp_wald_scan_compat(spec, ld, Zs[:, rows], n_real=B)
    df = spec["s"].shape[0] - 1
    ri = torch.from_numpy(rows).to(res.f.device)
    lp = lambda F: -stats.f.logsf(F, 1, df) / np.log(10.0)  # noqa: E731
    b_ref = lp(ref["f"])
    big = np.abs(ref["beta"]) >= ref["se"]

    def compare(r):
        got = {k: getattr(r, k).index_select(0, ri).cpu().numpy() for k in ("f", "beta", "se", "nullh2")}
        a = lp(got["f"])
        fin = np.isfinite(a) & np.isfinite(b_ref)
        dlp = float(np.abs(a - b_ref)[fin].max())
        db = np.abs(got["beta"] - ref["beta"])
        ld_gpu = np.log(1.0 / got["nullh2"][0] - 1.0)
        dld = np.abs(ld_gpu - ld)
        worst = int(np.argmax(np.abs(a - b_ref).max(axis=0) * fin.all(axis=0)))
        return {"max_abs_dlog10p": dlp, "max_dbeta_rel": float((db / np.maximum(np.abs(ref["beta"]), ref["se"])).max()),
                "max_dbeta_over_abs_beta_where_abs_beta_ge_se": float((db[big] / np.abs(ref["beta"][big])).max()),
                "tests_with_abs_beta_ge_se": int(big.sum()),
                "max_abs_dlog_delta": float(dld.max()), "columns_dlog_delta_gt_1e-3": int((dld > 1e-3).sum()),
                "worst_column": {"col": worst, "log_delta_gpu": float(ld_gpu[worst]),
                                 "log_delta_oracle": float(ld[worst]), "dlog10p": float(np.abs(a - b_ref)[:, worst].max())},
                "not_finite_in_scipy": int((~fin).sum())}

    out = {"sample": f"chromosome {c}: {rows.size} SNPs ({frac:.0%} random + {n_top} of largest GPU F) x {B} real "
                     f"phenotypes = {rows.size * B} tests; FP64 oracle LOCO pipeline",
           "max_F_oracle": float(ref["f"].max()), **compare(res)}
    for name, r in (others or {}).items():
        out[name] = compare(r)
    if Z is not None and Y_all is not None and Y_all.shape[1] > B:
        P = Y_all.shape[1]
        sel = np.sort(rng.choice(np.arange(B, P), size=min(n_perm_cols, P - B), replace=False))
        Ysel = Y_all[:, torch.from_numpy(sel).to(Y_all.device)]
        on_t = torch.from_numpy(test).to(Z.device)
        rg = core.loco_scan_compat(Z.index_select(1, on_t), X, Ysel, chrom[test], n_real=0, Z_kin=Z, chrom_kin=chrom,
                                   n_grid=100)
        pspec = lmm.compat_rotate(K, panel.X, Ysel.cpu().numpy())
        pld = lmm.fit_delta_grid_compat(pspec, n_grid=100)
        pref = lmm.snp_wald_scan_compat(pspec, pld, Zs[:, test], n_real=0)
        rel = np.abs(rg.max_F.cpu().numpy() - pref["max_F"]) / pref["max_F"]
        out["perm_max_F"] = {"columns": int(sel.size), "snps": int(test.size), "max_rel_err": float(rel.max()),
                             "median_rel_err": float(np.median(rel)),
                             "min_log_delta_oracle": float(pld.min()), "max_F_oracle": float(pref["max_F"].max())}
    out["oracle_s"] = round(time.perf_counter() - t0, 1)
    return out


def cpu_baseline(panel, cfg, n_perm, budget_s):
    """FP64 oracle (numpy, all host cores) on a bounded sample of the same workload: the first chromosome's
    SNPs against a slice of the scan columns, full compat pipeline (LOCO kinship, projected eigh, rotation,
    grid + 50 golden-section steps, scan).  Rate = Wald tests / wall seconds."""
    import numpy as np
    from oracle import lmm
    threads = len(os.sched_getaffinity(0))
    Zs = lmm.standardise_columns(panel.G.astype(np.float64))
    chrom = panel.chrom
    X = panel.X
    n_ph = 160  # real phenotypes in the sample: ~10-15 s of host work at N = 1,000
    n_cols = n_ph * (1 + n_perm)
    rng = np.random.default_rng(1)
    Yr = panel.Y[:, :n_ph]
    Yp = np.concatenate([Yr] + [Yr[rng.permutation(Yr.shape[0])] for _ in range(n_perm)], axis=1)[:, :n_cols]
    c = sorted(np.unique(chrom))[0]
    test = chrom == c
    t0 = time.perf_counter()
    K = lmm.grm(Zs[:, ~test])
    spec = lmm.compat_rotate(K, X, Yp)
    ld = lmm.fit_delta_grid_compat(spec, n_grid=100)
    res = lmm.snp_wald_scan_compat(spec, ld, Zs[:, test], n_real=n_ph)
    dt = time.perf_counter() - t0
    tests = int(test.sum()) * n_cols
    assert np.isfinite(res["max_F"]).all()
    return {"value": tests / dt, "unit": "Wald tests/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP64 compat LOCO pipeline for chromosome {c} ({int(test.sum())} SNPs) x {n_cols} columns "
                      f"({n_ph} phenotypes x {1 + n_perm}), N={Zs.shape[0]}, {dt:.1f}s wall, numpy/BLAS on {threads} threads; "
                      f"this is smaller than the planned slice (all 654,884 columns over 1-2 chromosomes): one batch, "
                      f"so the kinship, projected eigh and rotation are computed once, not once per 256-phenotype "
                      f"batch as the reference's CLI does, which makes the rate an upper bound for the reference"}


def traffic_per_launch(prec, scan):
    """HBM bytes per scan launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload,
    when one exists for this kernel (prec "screen": the screen kernel) and launch size (else None)."""
    name = ("wald_screen_traffic.json" if prec == "screen" else
            "wald_scan_factored_traffic.json" if scan["factored"] else "wald_scan_wide_traffic.json")
    f = ROOT / "profiles" / name
    if not f.exists():
        return None, None
    d = json.loads(f.read_text())
    e = d.get(prec)
    if not e or e.get("tests_per_launch") != round(scan["tests"] / max(scan["launches"], 1)):
        return None, None
    return e["bytes_per_launch"], {"fetch_bytes": e["fetch_bytes"], "write_bytes": e["write_bytes"],
                                   "source": f"profiles/{name}"}


def timed_run(Z, X, Y_all, chrom, B, precision, steps):
    """(results of the last step, seconds for `steps` steps, scan-kernel summary)."""
    import torch
    from fasterlmm_amd import engine
    engine.scan_timer.reset(True)
    engine.stage_timer.reset(True)
    engine.screen_stats.reset(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        res = run_step(Z, X, Y_all, chrom, B, precision)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    scan = engine.scan_timer.summary()
    scan["by_kind"] = {k: engine.scan_timer.summary(k) for k in ("full", "screen", "resolve")}
    scan["screen_evaluations"] = engine.screen_stats.summary()
    scan["stages_ms_per_step"] = {k: round(v / steps, 2) for k, v in engine.stage_timer.summary().items()}
    engine.scan_timer.reset(False)
    engine.stage_timer.reset(False)
    engine.screen_stats.reset(False)
    return res, dt, scan


def run_grouped(args, emit=True):
    """Uneven batching (N=5,000 x M=200,000 x 2,000 RINT phenotypes in 64 seeded missing-at-random NA groups with
    retained N ~ U(3,000, 5,000), C=6, 100 permutations each): every group's full kinship, then the first
    chromosome's per-group LOCO bases, delta fits and SNP planes, and the scans of all 64 groups for that
    chromosome timed as ONE grouped launch (autotuned, replayed from a captured HIP graph) against the per-group
    launch loop on the same planes.  A step is one grouped scan of the chromosome's SNP block."""
    import numpy as np
    import torch
    from fasterlmm_amd import engine, grouped, sim
    from fasterlmm_amd.io import group_by_na
    cfg = dict(sim.CONFIGS[3])
    if args.pheno:
        cfg["P"] = args.pheno
    if args.snps:
        cfg["M"] = args.snps
    n_perm = cfg["n_perm"] if args.n_perm is None else args.n_perm
    t0 = time.perf_counter()
    panel = sim.make_panel(cfg, seed=19930909)
    log(f"[bench3] panel N={cfg['N']} M={cfg['M']} P={cfg['P']} C={cfg['C']} in {time.perf_counter() - t0:.1f}s")
    dev = torch.device("cuda")
    groups = group_by_na(torch.from_numpy(panel.Y), list(range(cfg["P"])))
    gen = torch.Generator(device=dev)
    gen.manual_seed(19930909)
    Ys, keeps, n_reals = [], [], []
    for keep, cols in groups:
        y = torch.from_numpy(panel.Y[np.asarray(keep)][:, cols]).to(dev)
        n, b = y.shape
        orders = torch.rand(b, n_perm, n, generator=gen, device=dev).argsort(dim=2)
        yp = torch.gather(y.T.unsqueeze(1).expand(b, n_perm, n), 2, orders).reshape(b * n_perm, n).T
        Ys.append(torch.cat([y, yp], dim=1))
        keeps.append(keep)
        n_reals.append(b)
    Ns = [len(k) for k in keeps]
    log(f"[bench3] {len(groups)} NA groups, retained N {min(Ns)}..{max(Ns)}, columns per group "
        f"{min(y.shape[1] for y in Ys)}..{max(y.shape[1] for y in Ys)}")
    Z = torch.from_numpy(panel.G).to(dev).double()
    X = torch.from_numpy(panel.X).to(dev)
    prec = engine.prec_code(args.precision)
    engine.stage_timer.reset(True)
    with engine.configure(precision=args.precision):
        t0 = time.perf_counter()
        gl = grouped.GroupedLoco(Z, X, Ys, keeps, panel.chrom, n_reals=n_reals, n_grid=100,
                                 progress=lambda c, g: log(f"[bench3] chromosome {c} group {g} basis ready")
                                 if g % 8 == 7 else None)
        torch.cuda.synchronize()
        t_setup = time.perf_counter() - t0
        log(f"[bench3] per-group kinship and column planes in {t_setup:.1f}s")
        c = gl.chroms[0]
        t0 = time.perf_counter()
        jobs = next(gl.chromosome_jobs(c))
        torch.cuda.synchronize()
        t_prep = time.perf_counter() - t0
    stages = engine.stage_timer.summary()
    engine.stage_timer.reset(False)
    log(f"[bench3] chromosome {c}: bases, delta fits and SNP planes of {len(jobs)} groups in {t_prep:.1f}s {stages}")
    tests = sum(j.mc * j.P for j in jobs)
    # the launched algorithm: sWy over Neff, sWs over Neff (exact weights) or the R-term exponential sum (factored)
    alg_flops = sum(2.0 * (j.d.neff + (j.d.fac.R if j.factored else j.d.neff)) * j.mc * j.P for j in jobs)
    ref_flops = sum(4.0 * j.d.neff * j.mc * j.P for j in jobs)
    runner = grouped.GroupedScan(prec, len(jobs), max(j.P for j in jobs))
    t0 = time.perf_counter()
    runner.run(jobs)  # autotune + graph capture
    torch.cuda.synchronize()
    t_tune = time.perf_counter() - t0
    key = runner.tuner.key(prec, jobs)
    log(f"[bench3] autotune {t_tune:.2f}s: choice {runner.last_choice}, trials {runner.tuner.trials.get(key)}")
    want = [{k: getattr(j.out, k).clone() for k in ("f", "beta", "max_F")} for j in jobs]

    def timed(fn, reps):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        a = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - a) / reps

    t_grouped = timed(lambda: runner.run(jobs), args.steps)
    same = all(torch.equal(getattr(j.out, k), w[k]) for j, w in zip(jobs, want) for k in w)
    t_loop = timed(lambda: grouped.per_group_loop(jobs), args.steps)
    runner_eager = grouped.GroupedScan(prec, len(jobs), max(j.P for j in jobs), graph=False, tuner=runner.tuner)
    t_eager = timed(lambda: runner_eager.run(jobs), args.steps)
    value = tests / t_grouped
    split = args.precision != "fp32"
    products = 3 if split else 1
    published = MI355X_PEAK_BF16_TFLOPS if split else MI355X_PEAK_FP32_TFLOPS
    alg = alg_flops / t_grouped / 1e12
    ref = ref_flops / t_grouped / 1e12
    log(f"[bench3] grouped {1000 * t_grouped:.2f} ms ({value:.4e} tests/s), eager grouped {1000 * t_eager:.2f} ms, "
        f"per-group loop {1000 * t_loop:.2f} ms -> {t_loop / t_grouped:.3f}x")
    variant, blocks = runner.last_choice
    line = {
        "metric": "Wald tests/s, grouped scan of uneven NA groups (one chromosome's SNP block, all groups)",
        "value": value, "unit": "Wald tests/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1000 * t_grouped, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"split_fp16": "fp16", "split_bf16": "bf16", "fp32": "fp32"}[args.precision], "data": "synthetic",
        "config": {"workload": f"uneven NA groups: N={cfg['N']} M={cfg['M']} ({cfg['n_chrom']} chromosomes) "
                               f"P={cfg['P']} phenotypes in {len(jobs)} groups (retained N {min(Ns)}..{max(Ns)}) "
                               f"x (1+{n_perm} perms), C={cfg['C']}; step = chromosome {c} ({jobs[0].mc} SNPs) "
                               f"against every group's columns", "tests_per_step": tests,
                   "parallelism": "single GPU"},
        "grouped": {"variant": {0: "square 128x128", 1: "wide 128x256",
                                2: "wide 128x256, factored weights"}[variant], "blocks": blocks,
                    "graph_replay_ms": 1000 * t_grouped, "eager_ms": 1000 * t_eager,
                    "per_group_loop_ms": 1000 * t_loop, "speedup_vs_loop": t_loop / t_grouped,
                    "identical_to_first_run": same, "autotune_s": t_tune,
                    "autotune_trials_ms": [(round(ms, 3), v, b) for ms, v, b in runner.tuner.trials.get(key, [])]},
        "roofline": {"bound": "mfma", "achieved": round(ref, 2), "peak": round(published / products, 1),
                     "unit": "TFLOP/s", "frac": round(ref * products / published, 4), "traffic": None,
                     "kernel": "wald_scan_grouped_kernel", "mfma_issued_tflops": round(products * alg, 2),
                     "launched_algorithm_tflops": round(alg, 2),
                     "note": "achieved = 4 (N_g - C) flop per Wald test (the compat scan's algorithmic work) x tests "
                             "/ time; mfma_issued_tflops = 3 x the launched algorithm's flops (factored weights: "
                             "2 (N_g - C) + 2 R per test) / time"},
        "pipeline_s": {"kinship_and_columns_all_groups": t_setup, "chromosome_bases_deltas_snp_planes": t_prep,
                       "stages_ms": stages},
        "cpu_baseline": None,
    }
    if emit:
        print(json.dumps(line), flush=True)
    return line


def run_lowrank(args, emit=True):
    """Large-cohort low-rank path: N=50,000 x M=500,000 int8 dosages resident in HBM, scanned in 16,384-SNP blocks
    against a rank-2,000 kinship of 50,000 evenly strided markers (randomised range finder per LOCO kinship),
    P=1,000 phenotypes x (1 + 20 permutations), C=10, LOCO over 22 chromosomes.  One step = the whole LOCO scan.
    The CPU leg runs the FP64 low-rank oracle on a 1% SNP subsample of the first and last chromosomes (real
    phenotypes, with the GPU's deterministic basis and deltas) and reports its rate and the GPU's error against it."""
    import numpy as np
    import torch
    from fasterlmm_amd import engine, extreme_scan, sim
    from fasterlmm_amd.io import standardise_columns
    cfg = dict(sim.CONFIGS[5])
    for k, v in (("P", args.pheno), ("M", args.snps), ("N", args.n_samples)):
        if v:
            cfg[k] = v
    if args.snps:
        cfg["n_kin"] = min(cfg["n_kin"], cfg["M"] // 10)
    n_perm = cfg["n_perm"] if args.n_perm is None else args.n_perm
    rank = cfg["rank"] if args.rank is None else args.rank
    t0 = time.perf_counter()
    panel = sim.make_panel_gpu(cfg, seed=19930909, block=cfg["block"])
    dev = panel.G.device
    B = cfg["P"]
    N = cfg["N"]
    gen = torch.Generator(device=dev)
    gen.manual_seed(7)
    Y_all = torch.empty(N, B * (1 + n_perm), dtype=torch.float64, device=dev)
    Y_all[:, :B] = panel.Y
    for b0 in range(0, B, 64):
        b1 = min(B, b0 + 64)
        orders = torch.rand(b1 - b0, n_perm, N, generator=gen, device=dev).argsort(dim=2)
        yp = torch.gather(panel.Y[:, b0:b1].T.unsqueeze(1).expand(b1 - b0, n_perm, N), 2, orders)
        Y_all[:, B + b0 * n_perm:B + b1 * n_perm] = yp.reshape((b1 - b0) * n_perm, N).T
        del orders, yp
    kin = torch.from_numpy(panel.kin_idx).to(dev)
    Gk = standardise_columns(panel.G.index_select(1, kin))
    g_chrom = panel.chrom[panel.kin_idx]
    torch.cuda.synchronize()
    log(f"[bench5] panel N={N} M={cfg['M']} P={B}x(1+{n_perm}) C={cfg['C']} kinship markers {Gk.shape[1]} "
        f"rank {rank} in {time.perf_counter() - t0:.1f}s")
    prec = args.precision
    M = cfg["M"]
    P = Y_all.shape[1]
    tests = M * P

    def step():
        with engine.configure(precision=prec):
            return extreme_scan.loco_scan_resident(
                panel.G, panel.chrom, Gk, g_chrom, panel.X, Y_all, n_real=B, block_size=cfg["block"], rank=rank,
                on_chrom=lambda k, n: log(f"[bench5] chromosome {k}/{n} done") if k % 4 == 0 or k == n else None)

    for w in range(args.warmup):
        a = time.perf_counter()
        step()
        torch.cuda.synchronize()
        log(f"[bench5] warmup {w} {time.perf_counter() - a:.1f}s")
    engine.scan_timer.reset(True)
    engine.stage_timer.reset(True)
    torch.cuda.synchronize()
    a = time.perf_counter()
    for _ in range(args.steps):
        res = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - a) / args.steps
    scan = engine.scan_timer.summary()
    stages = {k: round(v / args.steps, 1) for k, v in engine.stage_timer.summary().items()}
    engine.scan_timer.reset(False)
    engine.stage_timer.reset(False)
    value = tests / dt
    split = prec != "fp32"
    products = 3 if split else 1
    published = MI355X_PEAK_BF16_TFLOPS if split else MI355X_PEAK_FP32_TFLOPS
    alg = scan["alg_flops"] / (scan["total_ms"] * 1e-3) / 1e12 if scan["total_ms"] else 0.0
    log(f"[bench5] {prec}: {dt:.2f} s/step, {value:.4e} tests/s, scan kernels {scan['launches']} launches "
        f"{scan['total_ms'] / args.steps:.0f} ms/step ({alg:.1f} TF/s alg), stages {stages}")
    cpu = None
    if not args.no_cpu:
        cpu = lowrank_cpu_check(panel, Gk, g_chrom, Y_all, B, res, rank, prec)
        log(f"[bench5] cpu leg: {cpu}")
    line = {
        "metric": "Wald tests/s (SNP x pheno x perm), low-rank LOCO scan, int8 genotypes resident in HBM",
        "value": value, "unit": "Wald tests/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1000 * dt, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"split_fp16": "fp16", "split_bf16": "bf16", "fp32": "fp32"}[prec], "data": "synthetic",
        "config": {"workload": f"low-rank LOCO: N={N} M={M} int8 in {cfg['block']}-SNP blocks "
                               f"({cfg['n_chrom']} chromosomes), rank-{rank} kinship from {Gk.shape[1]} strided "
                               f"markers, P={B} x (1+{n_perm} perms) = {P} columns, C={cfg['C']}",
                   "tests_per_step": tests, "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": round(alg, 2), "peak": round(published / products, 1),
                     "unit": "TFLOP/s", "frac": round(alg * products / published, 4), "traffic": None,
                     "kernel": "wald_scan_lowrank_kernel", "mean_launch_ms": round(scan["mean_ms"], 3),
                     "launches_per_step": scan["launches"] // args.steps,
                     "note": "algorithmic flops per test = 4 k (on-rank products) + 2 N (off-rank cross term)"},
        "stages_ms_per_step": stages, "cpu_baseline": cpu,
    }
    if emit:
        print(json.dumps(line), flush=True)
    return line


def lowrank_cpu_check(panel, Gk, g_chrom, Y_all, B, res, rank, prec):
    """The FP64 low-rank oracle (numpy, host cores) on a 1% SNP subsample of the first and last chromosomes and the
    real phenotypes, given the GPU's basis and deltas for those chromosomes (recomputed; the GPU path is
    deterministic); returns the rate and the GPU's error against it."""
    import numpy as np
    import torch
    from scipy import stats
    from fasterlmm_amd import engine, lowrank
    from oracle import lmm, lowrank as olr
    threads = len(os.sched_getaffinity(0))
    chroms = sorted(np.unique(panel.chrom).tolist())
    rng = np.random.default_rng(5)
    t_cpu, n_tests, dlp, dbeta, drelf, n_inf = 0.0, 0, 0.0, 0.0, 0.0, 0
    X = panel.X.cpu().numpy()
    Nn = X.shape[0]
    df = Nn - X.shape[1] - 1
    factor = None
    for c in (chroms[0], chroms[-1]):
        keep = torch.from_numpy((g_chrom != c).astype(np.float64)).to(Gk.device)
        with engine.configure(precision=prec):
            if factor is None:
                factor = lowrank.SplitFactor(Gk, engine.prec_code())  # deterministic: the LOCO engine's planes
            b = lowrank._basis(Gk, panel.X, keep, rank=rank, factor=factor)  # the LOCO engine's call, same bits
            UY, UUY = lowrank.rotate_phenos(b, Y_all)
            spec = lowrank.LowRankSpectrum(s=b.s, U_eff=b.U_eff, X=b.X, Xpinv=b.Xpinv, Neff=b.Neff, UY=UY, UUY=UUY)
            ld = lowrank.fit_delta_grid_lowrank(spec)
        cols = np.flatnonzero(panel.chrom == c)
        sub = np.sort(rng.choice(cols, size=max(1, len(cols) // 100), replace=False))
        S = lmm.standardise_columns(panel.G[:, torch.from_numpy(sub).to(Gk.device)].cpu().numpy().astype(np.float64))
        ospec = {"s": b.s.cpu().numpy(), "U_eff": b.U_eff.cpu().numpy(), "X": X, "Xpinv": b.Xpinv.cpu().numpy(),
                 "Neff": b.Neff, "UY": UY[:, :B].cpu().numpy(), "UUY": UUY[:, :B].cpu().numpy()}
        ldr = ld[:B].cpu().numpy()
        del b, UY, UUY, spec
        a = time.perf_counter()
        ref = olr.snp_wald_scan_lowrank(ospec, ldr, S, n_real=B)
        t_cpu += time.perf_counter() - a
        n_tests += S.shape[1] * B
        f = res.f[torch.from_numpy(sub).to(res.f.device)].cpu().numpy()
        beta = res.beta[torch.from_numpy(sub).to(res.f.device)].cpu().numpy()
        lp = lambda F: -stats.f.logsf(F, 1, df) / np.log(10.0)  # noqa: E731
        la, lb = lp(f), lp(ref["f"])
        fin = np.isfinite(la) & np.isfinite(lb)  # scipy's tail underflows to -inf for the very largest F
        dlp = max(dlp, float(np.abs(la[fin] - lb[fin]).max()) if fin.any() else 0.0)
        n_inf += int((~fin).sum())
        drelf = max(drelf, float((np.abs(f - ref["f"]) / np.maximum(ref["f"], 1.0)).max()))
        dbeta = max(dbeta, float((np.abs(beta - ref["beta"]) / np.maximum(np.abs(ref["beta"]), ref["se"])).max()))
    return {"value": n_tests / t_cpu, "unit": "Wald tests/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle low-rank Wald scan (numpy/BLAS, {threads} threads) of a 1% SNP subsample of "
                      f"chromosomes {chroms[0]} and {chroms[-1]} x {B} real phenotypes, given the GPU basis and "
                      f"deltas of those chromosomes: {n_tests} tests in {t_cpu:.1f}s",
            "gpu_vs_oracle": {"max_abs_dlog10p": dlp, "max_dbeta_over_max_beta_se": dbeta, "max_dF_over_max_F_1": drelf,
                              "tests_with_p_below_double_range": n_inf}}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; use --gpus 1"}))
        log("bench.py: multi-GPU runs are not part of this build (one MI355X); run with --gpus 1")
        sys.exit(2)
    if args.config not in (1, 2, 3, 5):
        log(f"bench.py: --config {args.config} is not a bench workload (1 and 2: the LOCO compat scan, 3: uneven NA "
            "groups, 5: the low-rank cohort); the learned delta proposal is evaluated by "
            "`python -m fasterlmm_amd.learned --eval`")
        sys.exit(2)
    if args.config == 3:
        run_grouped(args)
        return
    if args.config == 5:
        run_lowrank(args)
        return
    import numpy as np
    import torch

    if args.copy_probe_mib > 0:
        # a copy of known size: what FETCH_SIZE / WRITE_SIZE read for it calibrates those counters
        n = args.copy_probe_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copy probe: {args.copy_probe_mib} MiB read + {args.copy_probe_mib} MiB written")
        del src, dst

    STEP_SETTINGS.update(pipeline=args.pipeline, screen=args.screen)
    cfg, panel, Z, X, Y_all, B, n_perm = build_workload(args.config, args.n_perm, args.pheno, args.snps)
    N, M = Z.shape
    P = Y_all.shape[1]
    tests_per_step = M * P
    chrom = panel.chrom
    df = N - X.shape[1] - 1
    prec = args.precision

    for w in range(args.warmup):
        t0 = time.perf_counter()
        run_step(Z, X, Y_all, chrom, B, prec)
        torch.cuda.synchronize()
        log(f"[bench] warmup {w} {time.perf_counter() - t0:.3f}s")

    res, dt, scan = timed_run(Z, X, Y_all, chrom, B, prec, args.steps)
    ms_per_step = 1000.0 * dt / args.steps
    value = tests_per_step * args.steps / dt
    log(f"[bench] {prec}: {ms_per_step:.1f} ms/step, {value:.4e} tests/s, scan kernels {scan}")

    def tflops(sc, key="alg_flops"):
        return sc[key] / (sc["total_ms"] * 1e-3) / 1e12 if sc["total_ms"] else 0.0

    def screen_kernel_flops(sc):
        """The screen's MFMA flop: the recorded 2 (kpad + kpad2) per test, less 2 kpad2 for each test on the
        bucketed tile (sWy only; its sWs comes from FP64 node sums)."""
        ev = scan["screen_evaluations"]
        return sc["kernel_flops"] - 2.0 * ev.get("kpad2", 0) * ev.get("swy_only_tests", 0)

    split = prec != "fp32"
    alg = tflops(scan, "ref_flops")  # the compat scan's 4 (N - C) flop per Wald test, over every scan kernel's time
    products = 3 if split else 1
    published = MI355X_PEAK_BF16_TFLOPS if split else MI355X_PEAK_FP32_TFLOPS
    full, scr, rsv = (scan["by_kind"][k] for k in ("full", "screen", "resolve"))
    per_step = lambda sc: round(sc["total_ms"] / args.steps, 2)  # noqa: E731
    kernels = {"full_split_scan": {"ms_per_step": per_step(full), "launches_per_step": full["launches"] // args.steps,
                                   "issued_tflops": round(products * tflops(full, "kernel_flops"), 2)},
               "screen": {"ms_per_step": per_step(scr), "launches_per_step": scr["launches"] // args.steps,
                          "issued_tflops": round(screen_kernel_flops(scr) / max(scr["total_ms"], 1e-9) / 1e9, 2),
                          "swy_only_fraction": round(scan["screen_evaluations"].get("swy_only_fraction", 0.0), 4)},
               "resolve": {"ms_per_step": per_step(rsv), "launches_per_step": rsv["launches"] // args.steps},
               "fp64_evaluations": scan["screen_evaluations"]}
    if scr["launches"]:
        # the dominant kernel: the hi-plane screen of the permutation columns, one fp16 product per K element
        dom, issued = scr, screen_kernel_flops(scr) / max(scr["total_ms"], 1e-9) / 1e9
        kernel = ("the certified screen of the permutation columns (one flmm_wald_screen call per chromosome): "
                  "screen_bucketed_kernel (sWy on hi fp16 planes; sWs interpolated per log-delta bucket from FP64 "
                  "node sums, screen_nodes_kernel)")
        traffic, traffic_detail = traffic_per_launch("screen", scr)
        how = ("achieved = MFMA flop the screen executes per call (one fp16 product over the padded K: kpad per "
               "test on the bucketed tile, kpad + kpad2 on the general one where it runs) / the call's mean time "
               "(HIP events on the launch stream, around the node-sum and screen kernels), against the published "
               "dense fp16 peak: frac is its issued MFMA utilisation.")
    else:
        dom, issued = full, products * tflops(full, "kernel_flops")
        kernel = (f"wald_scan_lowrank_kernel<{prec}> (factored weights)" if full["factored"]
                  else f"wald_scan_wide_kernel<{prec}>")
        traffic, traffic_detail = traffic_per_launch(prec, full)
        how = ("achieved = MFMA flop the split scan executes per launch (3 products over the padded K per test) / its "
               "mean launch time (HIP events on the launch stream), against the published dense MFMA peak: frac is "
               "its issued MFMA utilisation.")
    roofline = {"bound": "mfma", "achieved": round(issued, 2), "peak": published, "unit": "TFLOP/s",
                "frac": round(issued / published, 4), "traffic": traffic, "traffic_detail": traffic_detail,
                "kernel": kernel, "mean_launch_ms": round(dom["mean_ms"], 4),
                "launches_per_step": dom["launches"] // args.steps,
                "algorithmic_tflops": round(alg, 2), "algorithmic_frac_of_peak": round(alg / published, 4),
                "launched_algorithm_tflops": round(tflops(scan), 2), "factored_terms": scan["R"],
                "scan_kernels": kernels,
                "note": how + " algorithmic_tflops counts the compat scan's 4 (N - C) flop per Wald test (the "
                        "reference algorithm's work) over the time of every scan kernel (screen, FP64 resolution "
                        "and the full split scan of the real-column tiles); launched_algorithm_tflops counts "
                        "2 (N - C) + 2 R per test (the factored weights' algorithm)."}

    extra = {}
    kept = {}
    from fasterlmm_amd.pvalues import neglog10p
    lp_head = neglog10p(res.f, df)  # on the GPU: the host would need minutes for 1e9 F values
    for other in [] if args.no_compare else [p for p in ("split_bf16", "fp32") if p != prec]:
        r2, dt2, s2 = timed_run(Z, X, Y_all, chrom, B, other, 1)
        rate = tests_per_step / dt2
        dlp = float((neglog10p(r2.f, df) - lp_head).abs().max())
        a2 = tflops(s2)
        pk = MI355X_PEAK_BF16_TFLOPS / 3.0 if other != "fp32" else MI355X_PEAK_FP32_TFLOPS
        # factored or exact as the headline's weights setting chooses for that precision (fp32: exact)
        extra[f"{other}_path"] = {"value": rate, "ms_per_step": 1000 * dt2, "scan_kernel_ms": s2["total_ms"],
                                  "scan_tflops": round(a2, 2), "scan_frac_of_peak": round(a2 / pk, 4),
                                  "max_abs_dlog10p_vs_headline": dlp,
                                  "headline_speedup": value / rate,
                                  "headline_scan_kernel_speedup": (s2["total_ms"] / (scan["total_ms"] / args.steps))
                                  if scan["total_ms"] else None}
        log(f"[bench] {other}: {1000 * dt2:.1f} ms/step, {rate:.4e} tests/s; {prec}/{other} = {value / rate:.2f}x")
        if other == "split_bf16":
            kept[other] = r2  # checked against the oracle below
        del r2

    cpu = None
    if not args.no_cpu:
        t2 = time.perf_counter()
        cpu = cpu_baseline(panel, cfg, n_perm, args.cpu_seconds)
        log(f"[bench] cpu baseline {cpu['value']:.4e} tests/s in {time.perf_counter() - t2:.1f}s")
        if args.config == 2:
            cpu["gpu_vs_oracle"] = compat_cpu_check(panel, res, B, kept, Z=Z, X=X, Y_all=Y_all)
            log(f"[bench] full-size parity vs the FP64 oracle: {cpu['gpu_vs_oracle']}")
    kept.clear()

    desc = {"split_fp16": "split-FP16 (power-of-two-scaled fp16 hi+lo, 3 MFMA products, FP32 accumulate)",
            "split_bf16": "split-BF16 (bf16 hi+lo, 3 MFMA products, FP32 accumulate)",
            "fp32": "FP32 MFMA"}[prec]
    line = {
        "metric": "Wald tests/s (SNP x pheno x perm), split-precision MFMA LOCO compat scan",
        "value": value, "unit": "Wald tests/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"split_fp16": "fp16", "split_bf16": "bf16", "fp32": "fp32"}[prec], "data": "synthetic",
        "config": {"workload": f"LOCO compat scan, N={N} M={M} ({len(np.unique(chrom))} chromosomes) "
                               f"P={B} phenotypes x (1+{n_perm} perms) = {P} columns, C={X.shape[1]}, 100-point grid",
                   "precision": desc + "; FP64 epilogue, delta search, projections and eigendecompositions",
                   "tests_per_step": tests_per_step, "parallelism": "single GPU"},
        "roofline": roofline, "cpu_baseline": cpu, "stages_ms_per_step": scan["stages_ms_per_step"], **extra,
    }
    if args.config == 2 and not args.no_others:
        # the uneven-NA-group and large-cohort workloads, measured in the same run (one step each)
        del res, Z, X, Y_all, lp_head
        torch.cuda.empty_cache()
        sub = argparse.Namespace(**{**vars(args), "steps": 1, "warmup": 1, "n_perm": None, "pheno": None, "snps": None})
        line["uneven_na_groups"] = run_grouped(sub, emit=False)
        torch.cuda.empty_cache()
        line["low_rank_cohort"] = run_lowrank(argparse.Namespace(**{**vars(sub), "warmup": 0}), emit=False)
        torch.cuda.empty_cache()
        line["learned_delta"] = run_learned()
    print(json.dumps(line), flush=True)


def run_learned():
    """The learned delta proposal at its evaluation workload (N = 2,000 Balding-Nichols panel, 10,000 held-out
    phenotypes x (1 + 100 permutations), the shipped network): fallback rate, log-likelihood gap and |d log10 p|
    against the plain grid + Newton search, and both delta searches' times (learned.evaluate)."""
    from fasterlmm_amd import learned
    t0 = time.perf_counter()
    info = learned.evaluate(learned.model(), n=2000, n_pheno=10_000, n_perm=100)
    info["workload"] = ("learned delta proposal: N=2000, 20000 kinship + 1000 test SNPs, 10000 phenotypes "
                        "(h2 ~ U(0, 0.95)) x (1+100 perms) = 1010000 columns, 100-point grid, shipped network")
    info["delta_search_speedup_learned_vs_grid_newton"] = (info["seconds_delta_grid_newton"]
                                                           / info["seconds_delta_learned"])
    info["seconds_total"] = time.perf_counter() - t0
    log(f"learned delta: fallback {info['fallback_rate']:.3f}, grid+Newton {info['seconds_delta_grid_newton']:.3f}s, "
        f"learned {info['seconds_delta_learned']:.3f}s ({info['seconds_total']:.1f}s total)")
    return info


if __name__ == "__main__":
    main()
