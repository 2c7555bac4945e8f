"""Benchmark: LOD trait peel throughput on one MI355X.

Workload (150 people over 6 generations with 3 consanguineous loops, 1000
SNPs at 0.25 cM, 10% missing genotypes, 5 trait positions per interval):
each step accumulates the LOD scores of a batch of descent graphs already
resident in HBM, i.e. the trait peel at every (graph, interval, position).
The metric's LOD evals/s is reported for the FP32-rescaled kernel; the plain
FP64 kernel is timed on the same graphs, and |dLOD| between the two is stated.

Prints one JSON line (see the measurement contract in the design notes):
value = LOD evals/s of the FP32-rescaled path, roofline of the peel kernel
(HIP events on the library's stream), and a CPU baseline measured on this
host with the reference's own CPU path (oracle/_ref/refharness) when present.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MCMC sweeps/s + LOD evals/s, FP32-rescaled vs FP64 at |ΔLOD|≤1e-4"
FP32_PEAK_TFLOPS = 157.3


def cpu_baseline(tmpdir: str, base: str, case, threads: int):
    """Times the reference CPU path (Peeler::process over every interval,
    one Peeler per OpenMP thread) on a bounded sample of the same workload."""
    harness = os.path.join(ROOT, "oracle", "_ref", "refharness")
    p, m, d = base + ".ped", base + ".map", base + ".dat"
    if os.path.exists(harness):
        out = os.path.join(tmpdir, "cpu")
        os.makedirs(out, exist_ok=True)
        subprocess.run([harness, "graphs", p, m, d, out, "5", "1", "1"], check=True, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL)
        n_graphs = 4 * threads
        subprocess.run([harness, "bench", p, m, d, out, os.path.join(out, "graphs.i32"), str(threads), "0", "0",
                        str(n_graphs)], check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        b = json.load(open(os.path.join(out, "bench.json")))
        return {"value": b["lod_evals"] / b["lod_seconds"], "unit": "LOD evals/s", "cores": threads,
                "kind": "reference",
                "sample": f"{n_graphs} graphs x {case.M - 1} intervals x {case.nlod} positions "
                          f"({b['lod_evals']} evals, {b['lod_seconds']:.1f} s wall) with the reference's "
                          f"Peeler::process, FP64, {threads} OpenMP threads"}
    # the reference binary is absent: time the oracle restatement, one core
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    from swiftlink_amd import synth  # noqa
    o = Oracle(case)
    g = np.load(os.path.join(tmpdir, "g.npy"))
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < 10.0:
        for l in range(0, case.M - 1, 7):
            o.lod_terms(g, l)
            n += case.nlod
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "LOD evals/s", "cores": 1, "kind": "port",
            "sample": f"{n} evals over every 7th interval of one graph, FP64 oracle, 1 core, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=64, help="descent graphs per step")
    ap.add_argument("--config", default="ped150_snp1000")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--gpw", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp32-only", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 5 times (reference for HBM byte counters)")
    args = ap.parse_args()
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
    from swiftlink_amd import native, synth
    from swiftlink_amd.case import pack_graphs

    tmp = tempfile.mkdtemp(prefix="slbench_")
    case, data = synth.make_case(args.config, tmp)
    g0 = synth.true_graph(data, case)
    np.save(os.path.join(tmp, "g.npy"), g0)
    B = args.batch
    graphs = np.repeat(g0[None], B, axis=0)
    packed = torch.from_numpy(pack_graphs(case, graphs).reshape(-1).view(np.int32)).cuda()
    torch.cuda.synchronize()
    evals_per_step = B * (case.M - 1) * case.nlod

    def run(precision, steps, warmup):
        ctx = native.Context(case, precision, block=args.block, gpw=args.gpw)
        for _ in range(warmup):
            ctx.lod_accumulate_packed(packed.data_ptr(), B)
        ctx.sync()
        ctx.profile(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.lod_accumulate_packed(packed.data_ptr(), B)
        ctx.sync()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        kms, nl = ctx.profile_read(native.SL_KERNEL_LOD)
        acc, count, lod = ctx.lod_results()
        flops, _ = ctx.lod_work()
        ctx.close()
        return dt, kms / max(nl, 1), lod, flops

    dt32, k32, lod32, flops = run(native.SL_FP32_RESCALED, args.steps, args.warmup)
    if args.fp32_only:
        dt64, k64, lod64 = float("nan"), float("nan"), lod32
    else:
        dt64, k64, lod64, _ = run(native.SL_FP64, max(2, args.steps // 4), args.warmup)
    # same graphs, same count of accumulations -> comparable LOD
    value = args.steps * evals_per_step / dt32
    fp64_value = max(2, args.steps // 4) * evals_per_step / dt64
    achieved = flops * evals_per_step / (k32 * 1e-3) / 1e12
    res = {
        "metric": METRIC,
        "value": value,
        "unit": "LOD evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt32 * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded pedigree, gene-drop genotypes; the gene-drop descent graph replicated per batch)",
        "config": {"workload": f"{args.config}: {case.N} people ({case.F} founders), {case.M} SNPs, "
                               f"{case.nlod} positions/interval, {B} graphs per step",
                   "graphs_per_step": B, "evals_per_step": evals_per_step},
        "fp64_lod_evals_per_s": fp64_value,
        "fp32_vs_fp64_speedup": value / fp64_value,
        "max_abs_dlod_fp32_vs_fp64": None if args.fp32_only else float(np.abs(lod32 - lod64).max()),
        "roofline": {"bound": "mfma", "compute": "fp32 VALU (no MFMA: tables of 4^k cells)",
                     "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                     "kernel_ms": k32, "flops_per_eval": flops, "kernel": "lod_kernel<float,true>"},
    }
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
        res["cpu_baseline"] = cpu_baseline(tmp, os.path.join(tmp, args.config), case, threads)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
