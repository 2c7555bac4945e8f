"""Writes the golden vectors under tests/golden/ from the compiled reference.

Run in a container where the reference sources exist (REFSRC, default
/root/reference/src): it builds oracle/_ref/refharness (oracle/Makefile,
target `ref`), runs it on the reference's example pedigrees and on a seeded
synthetic pedigree, and stores inputs and outputs as compressed .npz files.
Each fixture holds:
  case_*      the pedigree/map/peel tables exactly as the reference built them
  graphs      legal descent graphs (reference int32 layout [G][M][2N])
  lod_terms   per-graph ln L - R - T for every interval/position (Peeler::process)
  lod_acc     LODscores accumulators over those graphs, lod_final their LOD
  fag         FounderAlleleGraph4 likelihood per graph and locus
  lsweep_*    one L-sweep (LocusSampler::step over a locus order) under the
              Philox stream (seed, chain 0, sweep 5)
  msweep_*    one M-sweep (MeiosisSampler reset + steps) under (seed, 0, sweep 6)
  chain_*     a 40-sweep chain in the shared schedule with scoring, and its CODA
              trace: DescentGraph::get_likelihood at every scored sweep
  impute_*    4 sequential-imputation replicates (LocusSampler::sequential_imputation)
              under (seed 55, chain r, 0, step 0xFFFFFFFC), their weights and the pick
  impute0_*   LocusSampler::locus_by_locus under (seed 56, 0, 0, step 0xFFFFFFFC)
  writer_*    LinkageWriter's LOD table for the graphs as one family and split
              into two, the LOD values it printed, marker names and positions
  (and peelseq_ped150.npz: the reference's peel table of the synthetic
   150-person pedigree, see make_peelseq; linkage/: the example pedigrees'
   LINKAGE input files the loop and east fixtures were made from)
  elod_*      Elod::run's set-up and replicate loop (400 replicates, default
              model, replicate j under (seed 31, j, 0, step 0xFFFFFFFC)): the
              ELOD, accumulator, P(T), the first 6 replicates' graphs, both
              maps' tables and the peel sequence it built
  elod2_*     the same under penetrance 0.01:0.5:0.9, frequency 0.05, markers 0.1 M
              apart (300 replicates, seed 41)
"""
from __future__ import annotations

import dataclasses
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from swiftlink_amd.case import Case, PeelOps  # noqa: E402
from swiftlink_amd import synth  # noqa: E402

REF = os.environ.get("REFSRC", "/root/reference/src")
EXAMPLES = os.path.join(os.path.dirname(REF), "examples")
HARNESS = os.path.join(ROOT, "oracle", "_ref", "refharness")
OUT = os.path.dirname(os.path.abspath(__file__))


def run(*args):
    subprocess.run([HARNESS, *map(str, args)], check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)


def make(name, base, work):
    p, m, d = base + ".ped", base + ".map", base + ".dat"
    run("dump", p, m, d, work)
    run("graphs", p, m, d, work, 7, 4, 3)
    G = 4
    run("lod", p, m, d, work, os.path.join(work, "graphs.i32"), G)
    run("fag", p, m, d, work, os.path.join(work, "graphs.i32"), G)
    case = Case.from_harness_dir(work)
    M, N = case.M, case.N
    order = np.array(list(range(0, M, 2)) + list(range(1, M, 2)), np.int32)
    order.tofile(os.path.join(work, "lorder.i32"))
    ml = case.meioses()
    morder = np.random.default_rng(3).permutation(ml).astype(np.int32)
    morder.tofile(os.path.join(work, "morder.i32"))
    gi = os.path.join(work, "graph_init.i32")
    for sub in ("ls", "ms", "ch", "si", "si0", "el", "el2"):
        os.makedirs(os.path.join(work, sub), exist_ok=True)
    run("lsweep", p, m, d, os.path.join(work, "ls"), gi, os.path.join(work, "lorder.i32"), 99, 0, 5)
    run("msweep", p, m, d, os.path.join(work, "ms"), gi, os.path.join(work, "morder.i32"), 99, 0, 6)
    run("chain", p, m, d, os.path.join(work, "ch"), gi, 77, 0, 40, 10, 3, 0.5)
    run("writer", p, m, d, work, os.path.join(work, "graphs.i32"), G)
    run("elod", p, m, d, os.path.join(work, "el"), 0.0001, 0, 0, 1, 0.05, 400, 31, 6)
    run("elod", p, m, d, os.path.join(work, "el2"), 0.05, 0.01, 0.5, 0.9, 0.1, 300, 41, 0)
    run("impute", p, m, d, os.path.join(work, "si"), 55, 4)
    run("impute", p, m, d, os.path.join(work, "si0"), 56, 0)
    rd = lambda f, dt: np.fromfile(os.path.join(work, f), dt)
    fx = {"case_" + k: v for k, v in case.to_npz_dict().items()}
    fx.update(
        graphs=rd("graphs.i32", np.int32).reshape(-1, M, 2 * N)[:G],
        graph_init=rd("graph_init.i32", np.int32).reshape(M, 2 * N),
        lod_terms=rd("lod_terms.f64", np.float64).reshape(G, M - 1, case.nlod),
        lod_acc=rd("lod_acc.f64", np.float64), lod_final=rd("lod_final.f64", np.float64),
        lod_count=np.array([json.load(open(os.path.join(work, "lod_meta.json")))["count"]]),
        fag=rd("fag.f64", np.float64).reshape(G, M),
        lsweep_order=order, lsweep_seed=np.array([99, 0, 5]),
        lsweep_out=rd("ls/graph_out.i32", np.int32).reshape(M, 2 * N), lsweep_draws=rd("ls/draws.i32", np.int32),
        msweep_order=morder, msweep_seed=np.array([99, 0, 6]),
        msweep_out=rd("ms/graph_out.i32", np.int32).reshape(M, 2 * N),
        chain_params=np.array([77, 0, 40, 10, 3]), chain_lprob=np.array([0.5]),
        chain_out=rd("ch/graph_out.i32", np.int32).reshape(M, 2 * N),
        chain_kinds=rd("ch/sweep_kinds.i32", np.int32), chain_acc=rd("ch/lod_acc.f64", np.float64),
        chain_count=np.array([json.load(open(os.path.join(work, "ch/chain_meta.json")))["count"]]),
        chain_coda=rd("ch/coda.f64", np.float64), chain_coda_iter=rd("ch/coda_iter.i32", np.int32),
        impute_seed=np.array([55, 4]),
        impute_graphs=rd("si/impute_graphs.i32", np.int32).reshape(4, M, 2 * N),
        impute_weights=rd("si/impute_weights.f64", np.float64), impute_best=rd("si/impute_best.i32", np.int32),
        impute0_seed=np.array([56]),
        impute0_graph=rd("si0/impute_graphs.i32", np.int32).reshape(M, 2 * N),
        impute0_weight=rd("si0/impute_weights.f64", np.float64),
        writer_text1=np.array([open(os.path.join(work, "linkage1.txt")).read()]),
        writer_text2=np.array([open(os.path.join(work, "linkage2.txt")).read()]),
        writer_lod=rd("writer_lod.f64", np.float64).reshape(3, M - 1, case.nlod),
        writer_names=np.array(open(os.path.join(work, "marker_names.txt")).read().split("\n")[:M]),
        writer_gpos=rd("gpos.f64", np.float64),
    )
    em = json.load(open(os.path.join(work, "el", "elod_meta.json")))
    eops = PeelOps.from_lists([o["type"] for o in em["ops"]], [o["peel"] for o in em["ops"]],
                              [o["cutset"] for o in em["ops"]], [o["prev"] for o in em["ops"]],
                              [o["children"] for o in em["ops"]])
    fx.update(
        elod_params=np.array([0.0001, 0.0, 0.0, 1.0, 0.05]), elod_run=np.array([400, 31]),
        elod_value=np.array([em["elod"]]), elod_acc=np.array([em["acc"]]), elod_count=np.array([em["count"]]),
        elod_trait_prob=np.array([em["trait_prob"]]),
        elod_graphs=rd("el/elod_graphs.i32", np.int32).reshape(6, 3, 2 * N),
        elod_disease_prob=rd("el/elod_disease_prob.f64", np.float64).reshape(N, 4),
        elod_geno_prob=rd("el/elod_geno_prob.f64", np.float64).reshape(N, 3, 4),
        elod_theta3=np.array(em["theta3"]), elod_theta2=np.array([em["theta2"]]),
        elod_partial2=np.array([em["partial2"]]), elod_ids=np.array(em["ids"]),
        **{"elod_op_" + f.name: getattr(eops, f.name) for f in dataclasses.fields(PeelOps)},
    )
    em2 = json.load(open(os.path.join(work, "el2", "elod_meta.json")))
    fx.update(
        elod2_params=np.array([0.05, 0.01, 0.5, 0.9, 0.1]), elod2_run=np.array([300, 41]),
        elod2_value=np.array([em2["elod"]]), elod2_acc=np.array([em2["acc"]]),
        elod2_trait_prob=np.array([em2["trait_prob"]]),
        elod2_disease_prob=rd("el2/elod_disease_prob.f64", np.float64).reshape(N, 4),
        elod2_theta3=np.array(em2["theta3"]), elod2_partial2=np.array([em2["partial2"]]),
    )
    path = os.path.join(OUT, name + ".npz")
    np.savez_compressed(path, **fx)
    print(f"{name}: N={N} M={M} ops={case.ops.n} -> {os.path.getsize(path)} bytes")


def make_peelseq(tmp):
    """peelseq_ped150.npz: the reference's peel table of the 150-person
    synthetic pedigree (random downhill search, 2000 swaps per restart,
    seeds 7) with the parent indices in its internal order."""
    d = os.path.join(tmp, "ps150")
    synth.write_config("ped150_snp1000", d)
    b = os.path.join(d, "ped150_snp1000")
    run("peeliters=2000", "peelseed=7", "dump", b + ".ped", b + ".map", b + ".dat", os.path.join(d, "dump"))
    case = Case.from_harness_dir(os.path.join(d, "dump"))
    fx = {"mother": case.mother, "father": case.father, "params": np.array([2000, 7])}
    fx.update({"op_" + f.name: getattr(case.ops, f.name) for f in dataclasses.fields(PeelOps)})
    path = os.path.join(OUT, "peelseq_ped150.npz")
    np.savez_compressed(path, **fx)
    print(f"peelseq_ped150: ops={case.ops.n} -> {os.path.getsize(path)} bytes")


def copy_linkage_inputs():
    """The example pedigrees' LINKAGE input files (data: .ped/.map/.dat),
    kept beside the fixtures made from them so that the device pipeline can
    start from the same files where the reference is not mounted."""
    import shutil
    dst = os.path.join(OUT, "linkage")
    os.makedirs(dst, exist_ok=True)
    for ex in ("loop", "east"):
        for ext in (".ped", ".map", ".dat"):
            shutil.copyfile(os.path.join(EXAMPLES, ex + ext), os.path.join(dst, ex + ext))


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", f"REFSRC={REF}"], check=True)
    copy_linkage_inputs()
    with tempfile.TemporaryDirectory() as tmp:
        make_peelseq(tmp)
        for ex in ("loop", "east"):
            make(ex, os.path.join(EXAMPLES, ex), os.path.join(tmp, ex))
        sdir = os.path.join(tmp, "synth")
        synth.write_config("ped30_snp100", sdir)
        make("ped30_snp100", os.path.join(sdir, "ped30_snp100"), os.path.join(tmp, "ped30"))


if __name__ == "__main__":
    main()
