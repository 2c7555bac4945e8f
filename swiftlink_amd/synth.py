"""Seeded synthetic linkage data: pedigrees with consanguineous loops, a SNP
map, gene-dropped genotypes and a recessive trait.

The reference ships no generator; this one writes LINKAGE ped/map/dat files
in the format the reference parses (pedigree_parser.cc:80-161 for ped lines,
map_parser.cc:30-68 for "chr cM name" map lines, linkage_parser.cc for the
DAT layout of examples/east.dat), so the same inputs feed the reference CPU
path and this library.

Model
-----
* Pedigree: one founder couple, then generation by generation each couple has
  children; children that continue the line marry either a married-in founder
  or, to close a loop, a cousin (child of a sibling of one of their parents).
  Everyone has at most one mate, so every couple's children share both parents.
* Map: SNPs at fixed spacing in cM, Haldane recombination.
* Genotypes: founder haplotypes drawn from per-SNP minor-allele frequencies,
  dropped down the pedigree with Haldane crossovers, then a fraction set
  missing.
* Trait: one founder haplotype of the loop ancestors carries the disease
  allele at the map midpoint; affected persons are homozygous for it; the
  model written to the DAT file is recessive, penetrance (0, 0, 1).
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import List, Optional, Tuple

import numpy as np


@dataclasses.dataclass
class SynthSpec:
    n_people: int = 30
    n_generations: int = 4
    n_loops: int = 1
    n_markers: int = 100
    spacing_cm: float = 0.5
    maf_low: float = 0.1
    maf_high: float = 0.5
    missing: float = 0.05
    disease_freq: float = 1e-4
    penetrance: Tuple[float, float, float] = (0.0, 0.0, 1.0)
    seed: int = 1001
    min_affected: int = 2


# The configurations the benchmarks and parity tests are quoted on.
CONFIGS = {
    # 30 people / 4 generations / 1 first-cousin loop, 100 SNPs at 0.5 cM, 5% missing
    "ped30_snp100": SynthSpec(30, 4, 1, 100, 0.5, missing=0.05, seed=1001),
    # 150 people / 6 generations / 3 loops, 1000 SNPs at 0.25 cM, 10% missing
    "ped150_snp1000": SynthSpec(150, 6, 3, 1000, 0.25, missing=0.10, seed=1002),
    # 200 people / 7 generations / 4 loops, 2000 SNPs
    "ped200_snp2000": SynthSpec(200, 7, 4, 2000, 0.25, missing=0.10, seed=1004),
    # 600 people / 8 generations / 10 loops, 20000 SNPs at 0.1 cM, 20% missing
    "ped600_snp20000": SynthSpec(600, 8, 10, 20000, 0.1, missing=0.20, seed=1005),
}


@dataclasses.dataclass
class SynthPedigree:
    ids: List[str]
    father: List[int]      # -1 for founders
    mother: List[int]
    sex: List[int]         # 1 male, 2 female
    generation: List[int]
    loops: int


class _Builder:
    def __init__(self, rng: np.random.Generator):
        self.rng = rng
        self.father: List[int] = []
        self.mother: List[int] = []
        self.sex: List[int] = []
        self.gen: List[int] = []
        self.mate: List[int] = []

    def add(self, father: int, mother: int, sex: int, gen: int) -> int:
        self.father.append(father)
        self.mother.append(mother)
        self.sex.append(sex)
        self.gen.append(gen)
        self.mate.append(-1)
        return len(self.sex) - 1

    def grandparents(self, i: int) -> set:
        out = set()
        for p in (self.father[i], self.mother[i]):
            if p >= 0:
                out.update(x for x in (self.father[p], self.mother[p]) if x >= 0)
        return out

    def siblings(self, a: int, b: int) -> bool:
        return self.father[a] >= 0 and self.father[a] == self.father[b] and self.mother[a] == self.mother[b]


def _children_counts(rng, n_couples: int, budget: int) -> List[int]:
    """Split `budget` children over couples, each couple at least one."""
    if n_couples == 0:
        return []
    budget = max(budget, n_couples)
    counts = [1] * n_couples
    for _ in range(budget - n_couples):
        counts[int(rng.integers(n_couples))] += 1
    return counts


def make_pedigree(spec: SynthSpec, rng: np.random.Generator) -> SynthPedigree:
    """Grow a pedigree of exactly `spec.n_people` members over
    `spec.n_generations` generations with `spec.n_loops` cousin marriages."""
    G = spec.n_generations
    N = spec.n_people
    if G < 2 or N < 4:
        raise ValueError("need at least 2 generations and 4 people")
    # Descendant counts per generation grow geometrically; married-in spouses
    # add roughly one person per continuing descendant.  Solve for the growth
    # factor r so that the expected total is N.
    def total(r):
        t, d = 2.0, 2.0
        for g in range(1, G):
            d = d * r if g > 1 else max(2.0, r * 1.0)
            cont = 0.0 if g == G - 1 else d * 0.75
            t += d + cont
        return t
    lo, hi = 0.5, 8.0
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if total(mid) < N:
            lo = mid
        else:
            hi = mid
    r = 0.5 * (lo + hi)

    b = _Builder(rng)
    f0 = b.add(-1, -1, 1, 0)
    m0 = b.add(-1, -1, 2, 0)
    b.mate[f0], b.mate[m0] = m0, f0
    couples = [(f0, m0)]
    loops_left = spec.n_loops
    d_expect = 2.0
    for g in range(1, G):
        d_expect = d_expect * r if g > 1 else max(2.0, r)
        last = g == G - 1
        if last:
            budget = N - len(b.sex)
        else:
            budget = int(round(d_expect))
        budget = max(budget, len(couples))
        counts = _children_counts(rng, len(couples), budget)
        kids = []
        for (fa, mo), k in zip(couples, counts):
            for _ in range(k):
                kids.append(b.add(fa, mo, int(rng.integers(1, 3)), g))
        if last:
            break
        # Decide who continues the line.  Loops first: pair cousins of
        # opposite sex (different parents, a shared grandparent).
        order = list(kids)
        rng.shuffle(order)
        married = set()
        new_couples = []
        if loops_left > 0 and g >= 2:
            for a in order:
                if loops_left == 0:
                    break
                if a in married:
                    continue
                ga = b.grandparents(a)
                for c in order:
                    if c == a or c in married or b.sex[c] == b.sex[a]:
                        continue
                    if b.siblings(a, c) or not (ga & b.grandparents(c)):
                        continue
                    fa, mo = (a, c) if b.sex[a] == 1 else (c, a)
                    b.mate[a], b.mate[c] = c, a
                    married.update((a, c))
                    new_couples.append((fa, mo))
                    loops_left -= 1
                    break
        # remaining continuing descendants marry in a founder spouse; the
        # fraction is chosen so the next generation has enough parents.
        remaining = N - len(b.sex)
        gens_left = G - 1 - g
        want_cont = max(1, int(round(len(kids) * 0.75)))
        # never create more couples than people we can still afford
        # (each couple: one spouse + at least one child per later generation)
        for a in order:
            if len(new_couples) >= want_cont:
                break
            if a in married:
                continue
            cost = 1 + gens_left
            if remaining - cost < 0:
                break
            spouse = b.add(-1, -1, 3 - b.sex[a], g)
            b.mate[a], b.mate[spouse] = spouse, a
            married.update((a, spouse))
            fa, mo = (a, spouse) if b.sex[a] == 1 else (spouse, a)
            new_couples.append((fa, mo))
            remaining -= cost
        if not new_couples:
            raise RuntimeError("pedigree growth stalled")
        couples = new_couples
    n = len(b.sex)
    if n != N:
        raise RuntimeError(f"generated {n} people, wanted {N}")
    ids = [str(i + 1) for i in range(n)]
    return SynthPedigree(ids, b.father, b.mother, b.sex, b.gen, spec.n_loops - loops_left)


def _haldane(d_morgan: np.ndarray) -> np.ndarray:
    return 0.5 * (1.0 - np.exp(-2.0 * d_morgan))


def gene_drop(ped: SynthPedigree, pos_cm: np.ndarray, maf: np.ndarray, disease_pos: int,
              disease_founder_hap: int, rng: np.random.Generator):
    """Drop haplotypes down the pedigree.  Returns (alleles [N][2][M] in {1,2},
    disease_ibd [N][2] bool) where disease_ibd marks haplotypes carrying the
    founder disease haplotype at the disease position (between markers
    disease_pos-1 and disease_pos)."""
    N = len(ped.ids)
    M = len(pos_cm)
    # the disease locus sits at the midpoint of the interval before disease_pos
    loc_cm = np.concatenate([pos_cm[:disease_pos], [0.5 * (pos_cm[disease_pos - 1] + pos_cm[disease_pos])],
                             pos_cm[disease_pos:]])
    th = _haldane(np.diff(loc_cm) / 100.0)
    L = M + 1
    hap = np.zeros((N, 2, L), dtype=np.int8)
    inherit = np.zeros((N, 2, L), dtype=np.int8)   # meiosis indicators of the drop
    origin = np.full((N, 2), -1, dtype=np.int64)   # founder haplotype id at the disease locus
    order = sorted(range(N), key=lambda i: ped.generation[i])
    for i in order:
        if ped.father[i] < 0:
            for s in range(2):
                h = (rng.random(M) < maf).astype(np.int8) + 1
                hap[i, s] = np.concatenate([h[:disease_pos], [1], h[disease_pos:]])
                origin[i, s] = 2 * i + s
            continue
        for s, par in ((0, ped.mother[i]), (1, ped.father[i])):
            sw = rng.random(L - 1) < th
            which = np.concatenate([[int(rng.integers(2))], sw.astype(np.int64)])
            which = np.cumsum(which) % 2
            hap[i, s] = hap[par, which, np.arange(L)]
            inherit[i, s] = which
            origin[i, s] = origin[par, which[disease_pos]] if origin[par, which[disease_pos]] >= 0 else -1
            # origin is tracked through the disease column only
    alleles = np.delete(hap, disease_pos, axis=2)
    carry = origin == disease_founder_hap
    return alleles, carry, np.delete(inherit, disease_pos, axis=2)


def generate(spec: SynthSpec, max_drops: int = 400):
    """Return a dict with pedigree, map and genotype arrays for `spec`."""
    rng = np.random.default_rng(spec.seed)
    ped = make_pedigree(spec, rng)
    N, M = len(ped.ids), spec.n_markers
    pos = 10.0 + spec.spacing_cm * np.arange(M)
    maf = rng.uniform(spec.maf_low, spec.maf_high, size=M)
    disease_pos = M // 2
    # candidate disease haplotypes: those of the first founder couple, which
    # are ancestral to every loop
    best = None
    for attempt in range(max_drops):
        drng = np.random.default_rng([spec.seed, attempt])
        fh = int(drng.integers(4))
        alleles, carry, inherit = gene_drop(ped, pos, maf, disease_pos, fh, drng)
        affected = carry[:, 0] & carry[:, 1]
        naff = int(affected.sum())
        if best is None or naff > best[0]:
            best = (naff, alleles, affected, drng, inherit)
        if naff >= spec.min_affected:
            break
    naff, alleles, affected, drng, inherit = best
    missing = drng.random((N, M)) < spec.missing
    return dict(spec=spec, ped=ped, pos_cm=pos, maf=maf, alleles=alleles, missing=missing,
                affected=affected, disease_pos=disease_pos, inherit=inherit)


def true_graph(data: dict, case) -> np.ndarray:
    """The gene drop's meiosis indicators as a descent graph in the case's
    internal order, reference layout int32 [M][2N] (founder slots 0).  It is
    legal by construction: the genotypes were dropped along it."""
    ped: SynthPedigree = data["ped"]
    inh = data["inherit"]
    pos = {pid: i for i, pid in enumerate(ped.ids)}
    M = inh.shape[2]
    g = np.zeros((M, 2 * case.N), np.int32)
    for j, pid in enumerate(case.ids):
        i = pos[pid]
        if ped.father[i] < 0:
            continue
        g[:, 2 * j] = inh[i, 0]
        g[:, 2 * j + 1] = inh[i, 1]
    return g


def batch_specs(n: int = 16, seed: int = 1003) -> List[SynthSpec]:
    """The uneven batch: n pedigrees of 40-400 people (4 + people/100
    generations, 1-6 loops), 500-5000 SNPs at 0.1-0.5 cM, 10% missing."""
    rng = np.random.default_rng(seed)
    people = np.linspace(40, 400, n).round().astype(int)
    rng.shuffle(people)
    specs = []
    for i, npeople in enumerate(people):
        markers = int(round(500 + 4500 * ((i * 7) % n) / max(n - 1, 1)))
        specs.append(SynthSpec(int(npeople), 4 + int(npeople) // 100, 1 + i % 6, markers, [0.1, 0.2, 0.25, 0.5][i % 4],
                               missing=0.10, seed=2000 + i))
    return specs


def make_batch(outdir: str, n: int = 16, nlod: int = 5):
    """Cases and data of the uneven batch (batch_specs)."""
    from . import linkage
    out = []
    for i, spec in enumerate(batch_specs(n)):
        data = generate(spec)
        name = f"batch{i:02d}_ped{spec.n_people}_snp{spec.n_markers}"
        p, m, d = write_linkage(data, outdir, name)
        out.append((linkage.load(p, m, d, nlod=nlod), data))
    return out


def make_case(name: str, outdir: str, nlod: int = 5, **overrides):
    """Generate config `name`, write its LINKAGE files and load them.
    Returns (case, data)."""
    from . import linkage
    spec = dataclasses.replace(CONFIGS[name], **overrides) if overrides else CONFIGS[name]
    data = generate(spec)
    p, m, d = write_linkage(data, outdir, name)
    return linkage.load(p, m, d, nlod=nlod), data


def write_linkage(data: dict, outdir: str, name: str = "case") -> Tuple[str, str, str]:
    """Write <name>.ped/.map/.dat; returns their paths."""
    os.makedirs(outdir, exist_ok=True)
    spec: SynthSpec = data["spec"]
    ped: SynthPedigree = data["ped"]
    al = data["alleles"]
    miss = data["missing"]
    M = spec.n_markers
    pth = os.path.join(outdir, name + ".ped")
    with open(pth, "w") as f:
        for i in range(len(ped.ids)):
            fa = ped.ids[ped.father[i]] if ped.father[i] >= 0 else "0"
            mo = ped.ids[ped.mother[i]] if ped.mother[i] >= 0 else "0"
            aff = 2 if data["affected"][i] else 1
            g = np.where(miss[i][None, :], 0, al[i]).T  # [M][2]
            gs = " ".join(f"{a} {b}" for a, b in g)
            f.write(f"1 {ped.ids[i]} {fa} {mo} {ped.sex[i]} {aff} {gs}\n")
    mth = os.path.join(outdir, name + ".map")
    with open(mth, "w") as f:
        f.write("#Chr Genpos Marker\n")
        for l in range(M):
            f.write(f"1 {data['pos_cm'][l]:.6f} m{l + 1}\n")
    dth = os.path.join(outdir, name + ".dat")
    pen = spec.penetrance
    with open(dth, "w") as f:
        f.write(f"{M + 1} 0 0 5\n0 0.0 0.0 0\n")
        f.write(" ".join(str(i + 1) for i in range(M + 1)) + "\n")
        f.write("1 2 # TRAIT\n")
        f.write(f"{1.0 - spec.disease_freq:.10g} {spec.disease_freq:.10g}\n1\n")
        f.write(f"{pen[0]:.6f} {pen[1]:.6f} {pen[2]:.6f}\n")
        for l in range(M):
            q = data["maf"][l]
            f.write(f"3 2 # m{l + 1}\n{1.0 - q:.10f} {q:.10f}\n")
        f.write("0 0\n")
        th = _haldane(np.diff(data["pos_cm"]) / 100.0)
        f.write("0.1 " + " ".join(f"{t:.12g}" for t in th) + "\n")
        f.write("1 2.0 1.0\n")
    return pth, mth, dth


def write_config(name: str, outdir: str, **overrides) -> Tuple[str, str, str]:
    spec = dataclasses.replace(CONFIGS[name], **overrides) if overrides else CONFIGS[name]
    return write_linkage(generate(spec), outdir, name)
