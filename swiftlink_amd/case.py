"""Host-side description of one linkage problem: the pedigree in internal
order, the genetic map, the trait model tables and the peel-operation table.

These are exactly the data the reference hands across its GPU boundary
(gpu_lodscores.cc:163-372 builds the same arrays from Pedigree, GeneticMap and
PeelSequenceGenerator): persons founders-first with parent indices, per-person
disease probabilities in phased-trait order UU, AA, AU, UA (trait.h:21-27),
per-(person, locus) sampler trait probabilities, genotype-elimination masks,
recombination fractions, partial thetas and the ordered peel operations with
their cutsets, consumed tables and children.

A Case can be loaded from a directory written by the reference harness
(oracle/_ref/refharness dump), from a compressed fixture (.npz), or built from
LINKAGE files by swiftlink_amd.linkage.
"""
from __future__ import annotations

import dataclasses
import json
import os
from typing import List, Optional

import numpy as np

OP_CHILD, OP_PARENT, OP_PARTNER, OP_LAST = 1, 2, 3, 4


def _csr(lists: List[List[int]]):
    off = np.zeros(len(lists) + 1, dtype=np.int32)
    for i, l in enumerate(lists):
        off[i + 1] = off[i] + len(l)
    flat = np.array([x for l in lists for x in l], dtype=np.int32)
    return off, flat


@dataclasses.dataclass
class PeelOps:
    type: np.ndarray        # int32 [n]
    peel: np.ndarray        # int32 [n]
    cut_off: np.ndarray     # int32 [n+1]
    cut: np.ndarray         # int32
    prev_off: np.ndarray
    prev: np.ndarray
    child_off: np.ndarray
    child: np.ndarray

    @property
    def n(self) -> int:
        return int(self.type.shape[0])

    def cutset(self, o: int) -> List[int]:
        return [int(x) for x in self.cut[self.cut_off[o]:self.cut_off[o + 1]]]

    def prevs(self, o: int) -> List[int]:
        return [int(x) for x in self.prev[self.prev_off[o]:self.prev_off[o + 1]]]

    def children(self, o: int) -> List[int]:
        return [int(x) for x in self.child[self.child_off[o]:self.child_off[o + 1]]]

    @staticmethod
    def from_lists(types, peels, cutsets, prevs, children) -> "PeelOps":
        co, c = _csr(cutsets)
        po, p = _csr(prevs)
        ho, h = _csr(children)
        return PeelOps(np.asarray(types, np.int32), np.asarray(peels, np.int32), co, c, po, p, ho, h)


@dataclasses.dataclass
class Case:
    N: int
    F: int
    M: int
    nlod: int
    mother: np.ndarray        # int32 [N], -1 founder
    father: np.ndarray
    sex: np.ndarray           # int32 [N] 0 unsexed, 1 male, 2 female
    affection: np.ndarray     # int32 [N] 0 unknown, 1 unaffected, 2 affected
    typed: np.ndarray         # int32 [N]
    geno: np.ndarray          # uint8 [N][M] 0 untyped, 1 het, 2 homA, 3 homB
    disease_prob: np.ndarray  # f64 [N][4]
    geno_prob: np.ndarray     # f64 [N][M][4]
    legal: np.ndarray         # uint8 [N][M] bit t: phased trait t legal
    theta: np.ndarray         # f64 [M-1]
    partial_theta: np.ndarray # f64 [M-1]
    minor: np.ndarray         # f64 [M]
    ops: PeelOps
    ids: Optional[List[str]] = None
    trait_prob: Optional[float] = None
    sex_linked: bool = False

    # ------------------------------------------------------------------
    @staticmethod
    def from_harness_dir(d: str) -> "Case":
        meta = json.load(open(os.path.join(d, "meta.json")))
        N, M = meta["N"], meta["M"]
        rd = lambda name, dt: np.fromfile(os.path.join(d, name), dtype=dt)
        ops = meta["ops"]
        po = PeelOps.from_lists([o["type"] for o in ops], [o["peel"] for o in ops],
                                [o["cutset"] for o in ops], [o["prev"] for o in ops],
                                [o["children"] for o in ops])
        return Case(
            N=N, F=meta["F"], M=M, nlod=meta["nlod"],
            mother=rd("mother.i32", np.int32), father=rd("father.i32", np.int32),
            sex=rd("sex.i32", np.int32), affection=rd("affection.i32", np.int32),
            typed=rd("typed.i32", np.int32),
            geno=rd("genotypes.u8", np.uint8).reshape(N, M),
            disease_prob=rd("disease_prob.f64", np.float64).reshape(N, 4),
            geno_prob=rd("geno_prob.f64", np.float64).reshape(N, M, 4),
            legal=rd("legal.u8", np.uint8).reshape(N, M),
            theta=rd("theta.f64", np.float64), partial_theta=rd("partial_theta.f64", np.float64),
            minor=rd("minor.f64", np.float64), ops=po, ids=meta["ids"],
            trait_prob=meta["trait_prob"], sex_linked=bool(meta["sex_linked"]))

    def to_npz_dict(self) -> dict:
        d = {k: getattr(self, k) for k in ("mother", "father", "sex", "affection", "typed", "geno",
                                            "disease_prob", "legal", "theta", "partial_theta", "minor")}
        d["dims"] = np.array([self.N, self.F, self.M, self.nlod, int(self.sex_linked)], np.int64)
        for f in dataclasses.fields(PeelOps):
            d["op_" + f.name] = getattr(self.ops, f.name)
        d["trait_prob"] = np.array([np.nan if self.trait_prob is None else self.trait_prob])
        if self.ids is not None:
            d["ids"] = np.array(self.ids)
        # geno_prob is derived data (mask x marker prior); store it compactly
        d["geno_prob"] = self.geno_prob.astype(np.float64)
        return d

    @staticmethod
    def from_npz(path_or_dict) -> "Case":
        z = np.load(path_or_dict, allow_pickle=False) if isinstance(path_or_dict, str) else path_or_dict
        N, F, M, nlod, sl = (int(x) for x in z["dims"])
        ops = PeelOps(**{f.name: z["op_" + f.name] for f in dataclasses.fields(PeelOps)})
        tp = float(z["trait_prob"][0])
        return Case(N=N, F=F, M=M, nlod=nlod, mother=z["mother"], father=z["father"], sex=z["sex"],
                    affection=z["affection"], typed=z["typed"], geno=z["geno"],
                    disease_prob=z["disease_prob"], geno_prob=z["geno_prob"], legal=z["legal"],
                    theta=z["theta"], partial_theta=z["partial_theta"], minor=z["minor"], ops=ops,
                    ids=[str(x) for x in z["ids"]] if "ids" in z else None,
                    trait_prob=None if np.isnan(tp) else tp, sex_linked=bool(sl))

    # ------------------------------------------------------------------
    def meioses(self) -> np.ndarray:
        """Non-ignorable meiosis ids 2*(person-F)+parent (markov_chain.cc:65-80,
        person.cc:208-222: a meiosis from a founder with one child is skipped)."""
        nkids = np.zeros(self.N, np.int64)
        for i in range(self.F, self.N):
            nkids[self.mother[i]] += 1
            nkids[self.father[i]] += 1
        out = []
        for m in range(2 * (self.N - self.F)):
            person = self.F + m // 2
            par = self.father[person] if m % 2 else self.mother[person]
            if par < self.F and nkids[par] == 1:
                continue
            out.append(m)
        return np.array(out, np.int32)

    def check_supported(self) -> None:
        """The GPU library covers autosomal pedigrees where every child in a
        PARENT peel's cutset has its other parent in that cutset (true when
        nobody has children with two different mates)."""
        if self.sex_linked:
            raise ValueError("sex-linked traits are not supported by the GPU path "
                             "(the reference GPU path rejects them too, main.cc:534-537)")
        for o in range(self.ops.n):
            if self.ops.type[o] != OP_PARENT:
                continue
            cut = set(self.ops.cutset(o))
            p = int(self.ops.peel[o])
            for ch in self.ops.children(o):
                other = self.father[ch] if self.mother[ch] == p else self.mother[ch]
                if other not in cut:
                    raise ValueError(f"op {o}: child {ch}'s other parent is not in the cutset")


def pack_graphs(case: "Case", graphs: np.ndarray) -> np.ndarray:
    """Reference-layout graphs int32 [G][M][2N] -> the device layout uint32
    [G][M][W], bit b = 2*(person-F) + parent, W = ceil(2(N-F)/32)."""
    g = np.asarray(graphs).reshape(-1, case.M, 2 * case.N)
    bits = (g[:, :, 2 * case.F:] & 1).astype(np.uint8)
    nm = 2 * (case.N - case.F)
    W = (nm + 31) // 32
    pad = np.zeros((g.shape[0], case.M, 32 * W), np.uint8)
    pad[:, :, :nm] = bits
    packed = np.packbits(pad, axis=2, bitorder="little")
    return np.ascontiguousarray(packed).view(np.uint32).reshape(g.shape[0], case.M, W)
