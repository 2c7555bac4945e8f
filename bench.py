"""Benchmark: SwiftLink's MCMC linkage hot path on one MI355X.

Workload ped150_snp1000: 150 people over 6 generations with 3 consanguineous
loops, 1000 SNPs at 0.25 cM, 10% missing genotypes, 5 trait positions per
interval; one chain, L-sampler:M-sampler 1:1, a scored graph every 10th sweep.

A step is one sweep of the chain on the device (an L-sweep or an M-sweep as
the schedule draw says), with the LOD scoring of the chain's graph on every
10th sweep, so `value` is whole-job MCMC sweeps/s (FP32-rescaled samplers and
peel, FP64 LOD accumulation).  Reported beside it: the FP64 chain's sweeps/s,
the sweeps/s of 64 replicate chains run together in one context, LOD evals/s
of the trait peel over a batch of 64 device-resident graphs in
FP32-rescaled and FP64 with the largest |dLOD| between the two, the roofline
of the dominant kernel (the M-sampler's one-launch meiosis step, timed with
HIP events on the library's stream) and of the peel kernel, and a CPU
baseline: the reference's own CPU path (oracle/_ref/refharness) on this
host's cores, or the oracle restatement when that binary is absent.  The
`adjoint` object times the trait-peel adjoint (LOD and its gradient with
respect to penetrances, disease allele frequency and flanking thetas at every
position) over 64 resident graphs of ped200_snp2000 (200 people, 7
generations, 4 loops, 2000 SNPs), beside the oracle's forward-mode gradient
on one core.  `sequential_imputation` times 1000 replicates of the chains'
starting-graph search on the bench pedigree, and `elod` 262140 ELOD
replicates of it under the default ELOD model, each beside the reference's
own code on one core.

`python bench.py --large` instead runs the large combined workload
ped600_snp20000 (600 people over 8 generations with 10 consanguineous loops,
20000 SNPs at 0.1 cM, 20% missing genotypes, 5 positions per interval): 256
replicate chains from the gene-drop graph, L-sampler:M-sampler 1:1, a scored
graph every 100th sweep.  It times sweeps of all 256 chains together and one
scoring pass over 256 resident graphs, and reports sweeps/s over the 256
chains with the scoring's share of 1/100, LOD evals/s in FP32-rescaled and
FP64, and over every position of 4 graphs the non-finite LODs of each (the
unscaled FP64 peel underflows at this size) and the largest |dLOD| where both
are finite.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MCMC sweeps/s + LOD evals/s, FP32-rescaled vs FP64 at |ΔLOD|≤1e-4"
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
LPROB, PERIOD, SEED = 0.5, 10, 20250101
ADJ_CONFIG = "ped200_snp2000"  # the differentiation workload: 200 people, 7 generations, 4 loops, 2000 SNPs
SI_REPS = 1000          # sequential-imputation replicates (the reference's -s default)
ELOD_REPS = 4 * 65535   # ELOD replicates (four launches' worth)
LARGE_CONFIG, LARGE_CHAINS, LARGE_PERIOD = "ped600_snp20000", 256, 100  # the large combined workload


def profile_traffic(kernel: str, workload: str):
    """HBM bytes per launch from the committed rocprofv3 counter summary, if any."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_pmc.txt")
    if not os.path.exists(path):
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


def cpu_threads() -> int:
    """The host cores this process may use: OMP_NUM_THREADS when set (the
    share of host cores a GPU is granted), else the CPUs in its affinity
    mask."""
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    if env > 0:
        return env
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(tmpdir: str, base: str, case, threads: int):
    """Times the reference CPU path on a bounded sample of the same schedule:
    L-sweeps (LocusSampler::step over both parities), M-sweeps (MeiosisSampler
    reset + steps, OpenMP over loci) and Peeler::process over every interval,
    combined as sweeps/s = 1 / (p tL + (1-p) tM + tLOD / period)."""
    harness = os.path.join(ROOT, "oracle", "_ref", "refharness")
    p, m, d = base + ".ped", base + ".map", base + ".dat"
    if os.path.exists(harness):
        out = os.path.join(tmpdir, "cpu")
        os.makedirs(out, exist_ok=True)
        subprocess.run([harness, "graphs", p, m, d, out, "5", "1", "1"], check=True, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL)
        # five samples; each component is the median of its five timings
        nl, nm, ng, reps = 100, 60, 4 * threads, 5
        runs, wall = [], 0.0
        for _ in range(reps):
            subprocess.run([harness, "bench", p, m, d, out, os.path.join(out, "graphs.i32"), str(threads), str(nl),
                            str(nm), str(ng)], check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
            b = json.load(open(os.path.join(out, "bench.json")))
            runs.append((b["l_seconds"] / nl, b["m_seconds"] / nm, b["lod_seconds"] / ng,
                         b["lod_evals"] / b["lod_seconds"]))
            wall += b["l_seconds"] + b["m_seconds"] + b["lod_seconds"]
        tl, tm, tg, lev = (float(np.median([r[i] for r in runs])) for i in range(4))
        sweep = LPROB * tl + (1 - LPROB) * tm + tg / PERIOD
        return {"value": 1.0 / sweep, "unit": "sweeps/s", "cores": threads, "kind": "reference",
                "lod_evals_per_s": lev, "l_sweep_ms": tl * 1e3, "m_sweep_ms": tm * 1e3, "lod_graph_ms": tg * 1e3,
                "samples": reps, "os_cpu_count": os.cpu_count(),
                "sample": f"median of {reps} samples, each {nl} L-sweeps, {nm} M-sweeps and {ng} graphs x "
                          f"{case.M - 1} intervals x {case.nlod} positions with the reference's LocusSampler, "
                          f"MeiosisSampler and Peeler, FP64, {threads} OpenMP threads ({wall:.1f} s wall in all)"}
    # the reference binary is absent: the oracle restatement, one core
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    o = Oracle(case)
    g = np.load(os.path.join(tmpdir, "g.npy"))
    sweeps = 10
    t0 = time.perf_counter()
    o.chain(g, SEED, 0, sweeps, 0, PERIOD, LPROB)
    dt = time.perf_counter() - t0
    return {"value": sweeps / dt, "unit": "sweeps/s", "cores": 1, "kind": "port",
            "sample": f"{sweeps} sweeps of the oracle chain (same schedule, scored every {PERIOD}th), FP64, "
                      f"1 core, {dt:.1f} s"}


def cpu_impute_elod_baseline(tmpdir: str, base: str):
    """The reference's LocusSampler::sequential_imputation (20 replicates)
    and Elod::run's replicate loop (3000 replicates, default model) on one
    core, timed inside the harness (set-up excluded)."""
    harness = os.path.join(ROOT, "oracle", "_ref", "refharness")
    if not os.path.exists(harness):
        return None, None
    p, m, d = base + ".ped", base + ".map", base + ".dat"
    si_out, el_out = os.path.join(tmpdir, "cpu_si"), os.path.join(tmpdir, "cpu_elod")
    reps, ereps = 20, 3000
    subprocess.run([harness, "impute", p, m, d, si_out, str(SEED), str(reps)], check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    subprocess.run([harness, "elod", p, m, d, el_out, "0.0001", "0", "0", "1", "0.05", str(ereps), str(SEED), "0"],
                   check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    si = json.load(open(os.path.join(si_out, "impute_meta.json")))
    el = json.load(open(os.path.join(el_out, "elod_meta.json")))
    return ({"value": si["replicates"] * si["loci"] / si["seconds"], "unit": "replicate-loci/s", "cores": 1,
             "kind": "reference",
             "sample": f"{reps} replicates of the reference's sequential_imputation, FP64, 1 core "
                       f"({si['seconds']:.1f} s; parallel_run spreads replicates over OpenMP threads)"},
            {"value": ereps / el["seconds"], "unit": "replicates/s", "cores": 1, "kind": "reference",
             "sample": f"{ereps} replicates of Elod::run's loop (start_from + Peeler::process), FP64, 1 core "
                       f"({el['seconds']:.1f} s)"})


MOD_MODEL = ((0.05, 0.3, 0.85), 0.12)  # trait model of the adjoint workload


def cpu_adjoint_baseline(case, graph, jac, positions: int):
    """The oracle's forward-mode FP64 gradient (the adjoint's reference) on one
    core over the first `positions` (interval, position) pairs of one graph."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    o = Oracle(case)
    t0 = time.perf_counter()
    done = 0
    for l in range(case.M - 1):
        for p in range(case.nlod):
            o.trait_grad(graph, l, p + 1, jac)
            done += 1
            if done >= positions:
                break
        if done >= positions:
            break
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "gradient evals/s", "cores": 1, "kind": "port",
            "sample": f"{done} (graph, position) gradients of one graph with the oracle's forward-mode FP64 "
                      f"gradient, 1 core, {dt:.1f} s"}


def large_workload(args):
    """The large combined workload: LARGE_CHAINS chains of LARGE_CONFIG
    through the sweep schedule on one device, one sweep per call so that each
    prints its time, then the trait peel over resident graphs."""
    import torch
    from swiftlink_amd import native, synth
    from swiftlink_amd.case import pack_graphs

    def note(msg):
        print(msg, file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    case, data = synth.make_case(LARGE_CONFIG, tempfile.mkdtemp(prefix="sllarge_"))
    g0 = synth.true_graph(data, case)
    note(f"{LARGE_CONFIG}: {case.N} people ({case.F} founders), {case.M} SNPs, made in "
         f"{time.perf_counter() - t0:.1f} s")
    C = LARGE_CHAINS
    sched = (0, LARGE_PERIOD, LPROB, SEED)
    with native.Context(case, native.SL_FP32_RESCALED) as ctx:
        init = np.repeat(g0[None], C, axis=0)  # the reference layout: C x 96 MB at this size
        ctx.mcmc_init(init)
        del init
        # sweep 0 is a scored one (0 mod the period): the warm-up sweeps start there
        for s in range(args.warmup):
            t0 = time.perf_counter()
            ctx.mcmc_run(1, *sched, first_sweep=s)
            ctx.sync()
            note(f"warm-up sweep {s}: {time.perf_counter() - t0:.2f} s")
        kinds = {"l_sweeps": 0, "m_sweeps": 0, "scored": 0}
        times = []
        for s in range(args.warmup, args.warmup + args.steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st = ctx.mcmc_run(1, *sched, first_sweep=s)
            ctx.sync()
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
            for k in kinds:
                kinds[k] += st[k]
            note(f"sweep {s}: {times[-1]:.2f} s, L-sweep chains {st['l_sweeps']}, M-sweep chains {st['m_sweeps']}")
        near = int(ctx.mcmc_near_edge().sum())
        ok = bool(np.all(ctx.mcmc_status() == 0))

    # the scoring of a graph of every chain: the trait peel over C resident graphs
    packed = np.repeat(pack_graphs(case, g0[None]), C, axis=0)  # packed once, then tiled
    packed = torch.from_numpy(packed.reshape(-1).view(np.int32)).cuda()
    torch.cuda.synchronize()
    evals = C * (case.M - 1) * case.nlod

    def lod(precision, graphs, launches):
        with native.Context(case, precision) as lctx:
            lctx.lod_accumulate_packed(packed.data_ptr(), graphs)
            lctx.sync()
            lctx.lod_reset()
            t0 = time.perf_counter()
            for _ in range(launches):
                lctx.lod_accumulate_packed(packed.data_ptr(), graphs)
            lctx.sync()
            dt = (time.perf_counter() - t0) / launches
            _, _, lodv = lctx.lod_results()
        return dt, lodv

    score_s, _ = lod(native.SL_FP32_RESCALED, C, 2)
    note(f"scoring {C} graphs (FP32): {score_s:.2f} s")
    DG = 4
    d32, lod32 = lod(native.SL_FP32_RESCALED, DG, 1)
    d64, lod64 = lod(native.SL_FP64, DG, 1)
    note(f"{DG} graphs: FP32 {d32:.2f} s, FP64 {d64:.2f} s")
    ok32, ok64 = np.isfinite(lod32), np.isfinite(lod64)
    both = ok32 & ok64
    sweep_s = float(np.mean(times))
    per_sweep = sweep_s + score_s / LARGE_PERIOD
    res = {
        "metric": METRIC,
        "value": C / per_sweep,
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": per_sweep * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded pedigree and gene-drop genotypes; every chain starts from the gene-drop graph)",
        "config": {"workload": f"{LARGE_CONFIG}: {case.N} people ({case.F} founders), {case.M} SNPs, "
                               f"{case.nlod} positions/interval; {C} chains, L:M 1:1, scored every "
                               f"{LARGE_PERIOD}th sweep",
                   "chains": C, "lprob": LPROB, "period": LARGE_PERIOD},
        "sweeps": dict(kinds, near_edge_draws=near, chains_legal=ok),
        "sweep_s": sweep_s,
        "score_s": score_s,
        "lod_evals_per_s": evals / score_s,
        "fp64_lod_evals_per_s": DG * (case.M - 1) * case.nlod / d64,
        "lod_check": {
            "graphs": DG, "positions": int(lod32.size),
            "fp32_nonfinite": int(np.count_nonzero(~ok32)), "fp64_nonfinite": int(np.count_nonzero(~ok64)),
            "max_abs_dlod_where_both_finite": float(np.abs(lod32[both] - lod64[both]).max()) if both.any() else None,
            "note": "the unscaled FP64 peel underflows at this size (ln L below about -745) as the reference's "
                    "does, and loses bits in subnormals just above it; the FP32-rescaled peel stays finite"},
        "note": f"value = {C} / (mean sweep time of all {C} chains + one scoring pass over {C} graphs / "
                f"{LARGE_PERIOD})",
    }
    print(json.dumps(res))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed sweeps (default 200; 2 with --large)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed sweeps (default 20; 1 with --large)")
    ap.add_argument("--batch", type=int, default=64, help="descent graphs per LOD-throughput launch")
    ap.add_argument("--config", default="ped150_snp1000")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the 16-pedigree batch and the adjoint workloads (profiling passes)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 5 times (reference for HBM byte counters)")
    ap.add_argument("--large", action="store_true",
                    help=f"only the large combined workload ({LARGE_CONFIG}, {LARGE_CHAINS} chains)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 2 if args.large else 200
    if args.warmup is None:
        args.warmup = 1 if args.large else 20
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import torch
    if args.calibrate_copy:
        n = args.calibrate_copy * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": 4 * n, "launches": 5}))
        return
    if args.large:
        large_workload(args)
        return
    from swiftlink_amd import native, synth
    from swiftlink_amd.case import pack_graphs

    tmp = tempfile.mkdtemp(prefix="slbench_")
    case, data = synth.make_case(args.config, tmp)
    g0 = synth.true_graph(data, case)
    np.save(os.path.join(tmp, "g.npy"), g0)

    # ---- MCMC: chains through the schedule, scoring on the device ----
    def chain(precision, chains, steps, warmup):
        ctx = native.Context(case, precision)
        ctx.mcmc_init(np.repeat(g0[None], chains, axis=0))
        ctx.lod_reset()
        ctx.mcmc_run(warmup, 0, PERIOD, LPROB, SEED)
        ctx.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = ctx.mcmc_run(steps, 0, PERIOD, LPROB, SEED, first_sweep=warmup)
        ctx.sync()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        near = int(ctx.mcmc_near_edge().sum())
        # kernel averages from a separate, shorter pass with events around every launch
        ctx.profile(True)
        ctx.mcmc_run(max(10, steps // 10), 0, PERIOD, LPROB, SEED, first_sweep=warmup + steps)
        prof = {k: ctx.profile_read(k) for k in (native.SL_KERNEL_LOD, native.SL_KERNEL_LSAMPLER,
                                                 native.SL_KERNEL_MSAMPLER, native.SL_KERNEL_MSAMPLER_FB)}
        ctx.close()
        return dt, st, prof, near

    dt32, st32, prof32, near32 = chain(native.SL_FP32_RESCALED, 1, args.steps, args.warmup)
    n64 = max(20, args.steps // 4)
    dt64, st64, _, _ = chain(native.SL_FP64, 1, n64, args.warmup)
    value = args.steps / dt32
    fp64_sweeps = n64 / dt64
    CB, nb = 64, max(10, args.steps // 10)
    dtb, stb, _, nearb = chain(native.SL_FP32_RESCALED, CB, nb, 2)
    batch_sweeps = CB * nb / dtb

    # ---- LOD throughput on a batch of resident graphs ----
    B = args.batch
    packed = torch.from_numpy(pack_graphs(case, np.repeat(g0[None], B, axis=0)).reshape(-1).view(np.int32)).cuda()
    torch.cuda.synchronize()
    evals_per_launch = B * (case.M - 1) * case.nlod

    def lod(precision, launches):
        ctx = native.Context(case, precision)
        for _ in range(3):
            ctx.lod_accumulate_packed(packed.data_ptr(), B)
        ctx.sync()
        ctx.lod_reset()
        ctx.profile(True)
        t0 = time.perf_counter()
        for _ in range(launches):
            ctx.lod_accumulate_packed(packed.data_ptr(), B)
        ctx.sync()
        dt = time.perf_counter() - t0
        kms, nl = ctx.profile_read(native.SL_KERNEL_LOD)
        _, _, lodv = ctx.lod_results()
        flops, _ = ctx.lod_work()
        kern = ctx.lod_kernel_for(B)
        ctx.close()
        return launches * evals_per_launch / dt, kms / max(nl, 1), lodv, flops, kern

    lod32, k32, lodv32, flops, kern32 = lod(native.SL_FP32_RESCALED, 20)
    lod64, _, lodv64, _, _ = lod(native.SL_FP64, 8)

    # ---- 16 pedigrees of uneven size, 4 chains each, one captured sweep per
    # context replayed side by side (sl_batch_*), shapes from the autotuner ----
    extras = not args.no_extras
    from swiftlink_amd.tune import autotune
    if extras:
        bcases = synth.make_batch(tempfile.mkdtemp(prefix="slbatch_"))
        bctx, tunes = [], []
        BCH, BK, BW = 4, 10, 2
        for bcase, bdata in bcases:
            c = native.Context(bcase, native.SL_FP32_RESCALED)
            c.mcmc_init(np.repeat(synth.true_graph(bdata, bcase)[None], BCH, axis=0))
            tu = autotune(c)
            tunes.append({"people": bcase.N, "snps": bcase.M, "block": tu["block"], "gpw": tu["gpw"],
                          "lod_kernel": tu["lod_kernel"],
                          "msampler_mode": tu["msampler_mode"]})
            bctx.append(c)
        bsched = (0, PERIOD, LPROB, SEED)
        t0 = time.perf_counter()
        for c in bctx:
            c.mcmc_run(BK, *bsched)
        for c in bctx:
            c.sync()
        bseq = time.perf_counter() - t0
        with native.Batch(bctx) as bb:
            bb.run(BW, *bsched, first_sweep=BK)
            t0 = time.perf_counter()
            bstats = bb.run(BK, *bsched, first_sweep=BK + BW)
            bdt = time.perf_counter() - t0
        bchains = BCH * len(bctx)
        for c in bctx:
            c.close()

    # ---- adjoint of the trait peel: LOD and its gradient over resident
    # graphs of the differentiation workload (ped200_snp2000) ----
    if extras:
        acase, adata = synth.make_case(ADJ_CONFIG, tempfile.mkdtemp(prefix="sladj_"))
        ag0 = synth.true_graph(adata, acase)
        AG = 64
        apacked = torch.from_numpy(pack_graphs(acase, np.repeat(ag0[None], AG, axis=0)).reshape(-1).view(np.int32)).cuda()
        torch.cuda.synchronize()
        with native.Context(acase, native.SL_FP64) as actx:
            actx.set_trait_model(*MOD_MODEL)
            actx.lod_grad_packed(apacked.data_ptr(), 2)
            t0 = time.perf_counter()
            alod, adlod = actx.lod_grad_packed(apacked.data_ptr(), AG)
            adt = time.perf_counter() - t0
        apos = (acase.M - 1) * acase.nlod

    # ---- sequential imputation (the chains' starting graph, 1000 replicates
    # as the reference's -s default) and ELOD (replicates of start_from at the
    # trait locus, scored on the two-marker map) on the bench pedigree ----
    if extras:
        from swiftlink_amd import elod as elodmod
        with native.Context(case, native.SL_FP32_RESCALED) as ictx:
            ictx.sequential_imputation(SEED, 16)
            t0 = time.perf_counter()
            _, iw, ibest, _ = ictx.sequential_imputation(SEED, SI_REPS)
            idt = time.perf_counter() - t0
        e3, e2 = elodmod.elod_cases(case)
        with native.Context(e3, native.SL_FP32_RESCALED) as es, native.Context(e2, native.SL_FP64) as esc:
            native.run_elod(es, esc, SEED, 2048)
            t0 = time.perf_counter()
            elod_value = native.run_elod(es, esc, SEED, ELOD_REPS)
            edt = time.perf_counter() - t0

    # ---- roofline of the meiosis-step kernel (one chain: mspec_kernel, the
    # forward-backward pass of step k beside the founder-allele likelihoods of
    # step k+1 for both values of meiosis k's bit): algorithmic bytes per
    # launch = per (chain, meiosis, locus) ceil(2N/8) descent-graph bytes +
    # ceil(2 N_typed/8) genotype bytes, times 2M likelihood jobs ----
    ms_ms, ms_launches = prof32[native.SL_KERNEL_MSAMPLER]
    avg = lambda k: prof32[k][0] / max(prof32[k][1], 1)
    n_typed = int(np.count_nonzero(case.typed))
    bytes_per_unit = -(-2 * case.N // 8) + -(-2 * n_typed // 8)
    bytes_per_launch = 2 * case.M * bytes_per_unit
    ms_avg = ms_ms / max(ms_launches, 1)
    achieved = bytes_per_launch / (ms_avg * 1e-3) / 1e9 if ms_launches else 0.0
    lod_achieved = flops * evals_per_launch / (k32 * 1e-3) / 1e12
    lod_name = ("lane_lod_kernel (one (graph, position) evaluation per lane)" if kern32 == native.SL_LOD_LANES
                else "lod_kernel<float,true>")

    res = {
        "metric": METRIC,
        "value": value,
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt32 * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded pedigree and gene-drop genotypes; the chain starts from the gene-drop graph)",
        "config": {"workload": f"{args.config}: {case.N} people ({case.F} founders), {case.M} SNPs, "
                               f"{case.nlod} positions/interval; 1 chain, L:M 1:1, scored every {PERIOD}th sweep",
                   "chains": 1, "lprob": LPROB, "period": PERIOD},
        "sweeps": {"l_sweeps": st32["l_sweeps"], "m_sweeps": st32["m_sweeps"], "scored": st32["scored"],
                   "near_edge_draws": near32},
        "fp64_sweeps_per_s": fp64_sweeps,
        "batched": {"chains": CB, "sweeps_per_chain": nb, "sweeps_per_s": batch_sweeps,
                    "l_sweeps": stb["l_sweeps"], "m_sweeps": stb["m_sweeps"], "scored": stb["scored"],
                    "near_edge_draws": nearb,
                    "note": "64 replicate chains of the same pedigree in one context, same schedule"},
        "lod_evals_per_s": lod32,
        "fp64_lod_evals_per_s": lod64,
        "max_abs_dlod_fp32_vs_fp64": float(np.abs(lodv32 - lodv64).max()),
        "kernel_ms": {"msampler_likelihoods": ms_avg,
                      "msampler_forward_backward": avg(native.SL_KERNEL_MSAMPLER_FB),
                      "lsampler_parity": avg(native.SL_KERNEL_LSAMPLER),
                      "lod_scoring": avg(native.SL_KERNEL_LOD)},
        "roofline": {"bound": "hbm", "kernel": "mspec_kernel (one meiosis step: forward-backward pass of step k, "
                                               "likelihoods of step k+1 for both values of meiosis k's bit)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": profile_traffic("mspec_kernel", args.config),
                     "bytes_per_launch": bytes_per_launch, "launch_ms": ms_avg},
        "lod_roofline": {"bound": "mfma", "compute": "fp32 VALU (no MFMA: tables of 4^k cells)",
                         "kernel": lod_name, "achieved": lod_achieved, "peak": FP32_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": lod_achieved / FP32_PEAK_TFLOPS,
                         "traffic": profile_traffic(lod_name.split("<")[0].split(" ")[0], args.config),
                         "flops_per_eval": flops, "launch_ms": k32},
    }
    if extras:
        res["batch_pedigrees"] = {
            "workload": f"{len(bcases)} pedigrees of 40-400 people (1-6 loops) and 500-5000 SNPs at 0.1-0.5 cM, "
                        f"{BCH} chains each, FP32-rescaled samplers and peel, FP64 LOD accumulation, L:M 1:1, "
                        f"scored every {PERIOD}th sweep",
            "chains": bchains, "sweeps_per_chain": BK, "sweeps_per_s": bchains * BK / bdt,
            "sequential_sweeps_per_s": bchains * BK / bseq,
            "l_sweeps": int(sum(x["l_sweeps"] for x in bstats)), "m_sweeps": int(sum(x["m_sweeps"] for x in bstats)),
            "note": "sweeps_per_s: one hipGraph-captured sweep per pedigree, replayed on the pedigrees' streams side by "
                    "side (sl_batch_run); sequential_sweeps_per_s: sl_mcmc_run on each pedigree in turn, same shapes",
            "autotuned": tunes}
        res["adjoint"] = {
            "workload": f"{ADJ_CONFIG}: {acase.N} people, {acase.M} SNPs, {apos} positions; {AG} resident graphs "
                        f"(copies of the gene-drop graph), model penetrance {MOD_MODEL[0]}, frequency {MOD_MODEL[1]}",
            "graphs": AG, "positions": apos, "ms": adt * 1e3, "gradient_evals_per_s": AG * apos / adt,
            "partials_per_eval": 6, "dtype": "f64", "max_lod": float(alod.max()),
            "note": "sl_lod_grad: ln L and d ln L/d(3 penetrances, frequency, theta_L, theta_R) per (graph, position), "
                    "reduced on the device to the LOD and its gradient at every position"}
    if extras:
        res["sequential_imputation"] = {
            "workload": f"{args.config}: {SI_REPS} replicates x {case.M} loci (LocusSampler::sequential_imputation), "
                        f"FP32-rescaled peel, FP64 weights",
            "seconds": idt, "replicate_loci_per_s": SI_REPS * case.M / idt, "best_weight": float(iw[ibest])}
        res["elod"] = {
            "workload": f"{args.config} pedigree, default ELOD model (frequency 1e-4, penetrance 0:0:1, markers "
                        f"0.05 M apart), {ELOD_REPS} replicates; FP32-rescaled sampler, FP64 scoring",
            "seconds": edt, "replicates_per_s": ELOD_REPS / edt, "elod": elod_value}
    if not args.no_cpu:
        threads = cpu_threads()
        res["cpu_baseline"] = cpu_baseline(tmp, os.path.join(tmp, args.config), case, threads)
    if not args.no_cpu and extras:
        _, ajac = native.trait_model_tables(MOD_MODEL[0], MOD_MODEL[1], acase.affection, acase.F)
        from swiftlink_amd.linkage import disease_probs
        import dataclasses
        dp = np.array([disease_probs(int(acase.affection[i]), i < acase.F, MOD_MODEL[1], MOD_MODEL[0])
                       for i in range(acase.N)])
        res["adjoint"]["cpu_baseline"] = cpu_adjoint_baseline(
            dataclasses.replace(acase, disease_prob=dp.reshape(acase.disease_prob.shape)), ag0, ajac, 15000)
    if extras:
        # the reference's peel-sequence search (greedy, then random-restart
        # downhill) as host C++, beside the reference's own code
        iters = 5000
        from swiftlink_amd import linkage
        rows = linkage.read_ped(os.path.join(tmp, args.config) + ".ped")
        perm = native.person_order([1 if (r[2] == "0" and r[3] == "0") else 0 for r in rows])
        pos = {rows[i][1]: k for k, i in enumerate(perm)}
        pm = np.array([-1 if rows[i][3] == "0" else pos[rows[i][3]] for i in perm], np.int32)
        pf = np.array([-1 if rows[i][2] == "0" else pos[rows[i][2]] for i in perm], np.int32)
        t0 = time.perf_counter()
        _, pcost = native.peel_sequence(pm, pf, iters, 7, 7, threads=1)
        pdt = time.perf_counter() - t0
        res["peel_sequence"] = {"workload": f"{args.config} pedigree, {iters} downhill swaps per restart, seeds 7",
                                "seconds": pdt, "cost": pcost, "cores": 1}
        harness = os.path.join(ROOT, "oracle", "_ref", "refharness")
        if not args.no_cpu and os.path.exists(harness):
            base = os.path.join(tmp, args.config)
            out = os.path.join(tmp, "cpu_peel")
            subprocess.run([harness, f"peeliters={iters}", "peelseed=7", "peelseq", base + ".ped", base + ".map",
                            base + ".dat", out], check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
            pj = json.load(open(os.path.join(out, "peelseq.json")))
            res["peel_sequence"]["cpu_baseline"] = {
                "value": pj["seconds"], "unit": "s", "cores": 1, "kind": "reference",
                "sample": f"PeelSequenceGenerator::build_peel_sequence({iters}) on the same pedigree file"}
    if not args.no_cpu and extras and "elod" in res:
        si_cpu, elod_cpu = cpu_impute_elod_baseline(tmp, os.path.join(tmp, args.config))
        if si_cpu:
            res["sequential_imputation"]["cpu_baseline"] = si_cpu
            res["elod"]["cpu_baseline"] = elod_cpu
    print(json.dumps(res))


if __name__ == "__main__":
    main()
