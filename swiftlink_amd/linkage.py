"""LINKAGE ped/map/dat input -> Case (host pre-computation).

Restates what the reference's parsers and domain classes compute before the
MCMC starts, so the GPU library can be driven without the reference:

* map: cM -> Morgans (map_parser.cc:62), Haldane theta between adjacent
  markers and the linear partial theta (genetic_map.cc:14-64);
* trait model from the DAT file (linkage_parser.cc: trait line, liability
  class, penetrances) and per-person disease probabilities
  (disease_model.cc:146-187, person.cc:85-119);
* the sampler's per-(person, locus) trait-probability cache
  (person.cc:224-299).  The reference fills it while parsing, before parent
  ids are resolved, so every person takes the founder branch: the genotype
  mask times the marker prior, normalised (pedigree_parser.cc:141-157,
  person.h:168).  This is reproduced as is;
* founders-first internal order (pedigree.cc:154-162): by default a stable
  one; with peel="reference" the reference's own (libstdc++ std::sort, not
  stable) through sl_person_order;
* genotype elimination masks (elimination.cc:87-125, 393-395);
* a peel sequence: this package's greedy order (swiftlink_amd.peelseq) or
  the reference's search (sl_peel_sequence).

And the output side: the LOD table the reference writes
(LinkageWriter::write, linkage_writer.cc:14-91), `write_lod_table`.
"""
from __future__ import annotations

import math
from typing import List, Optional, Tuple

import numpy as np

from .case import Case
from . import peelseq

# phased genotype bits (genotype.h:51-57) and the trait -> genotype map
# (genotype.cc:88-103): UU->AA, AA->BB, AU->BA, UA->AB
G_AA, G_AB, G_BA, G_BB = 8, 4, 2, 1
TRAIT_TO_GENO = (G_AA, G_BB, G_BA, G_AB)
UNTYPED, HETERO, HOMOZ_A, HOMOZ_B = 0, 1, 2, 3


def _haldane(m: float) -> float:
    return 0.5 * (1.0 - math.exp(-2.0 * m))


def _inverse_haldane(r: float) -> float:
    return -0.5 * math.log(1 - (2 * r))


def read_map(path: str) -> Tuple[List[str], np.ndarray]:
    names, pos = [], []
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t or t[0].startswith("#"):
                continue
            names.append(t[2])
            pos.append(float(t[1]) / 100.0)
    return names, np.array(pos)


def _fmt(x: float) -> str:
    """A double as std::ostream writes it by default (6 significant
    digits, %g rules)."""
    if math.isnan(x):
        return "nan" if math.copysign(1.0, x) > 0 else "-nan"
    return format(x, "g")


def format_lod_table(names, gpos, partial_theta, family_lods) -> str:
    """The text of LinkageWriter::write (linkage_writer.cc:52-84).

    names: M marker names; gpos: M genetic positions (Morgans);
    partial_theta: M-1 partial thetas (genetic_map.cc:58-62);
    family_lods: one array [M-1][nlod] of LOD scores per pedigree
    (LODscores::get).  Each interval lists its nlod positions at
    g_i + inverse_haldane(partial_theta_i (j+1)) cM with the LOD summed over
    pedigrees in order, then each pedigree's own LOD when there is more than
    one."""
    fams = [np.asarray(f, np.float64) for f in family_lods]
    M = len(names)
    nlod = fams[0].shape[1]
    out = ["marker\tposition\tlod\n"]
    for i in range(M - 1):
        out.append(f"{names[i]}\t{_fmt(100.0 * float(gpos[i]))}\n")
        for j in range(nlod):
            tmp = float(fams[0][i, j])
            for f in fams[1:]:
                tmp += float(f[i, j])
            pos = float(gpos[i]) + _inverse_haldane(float(partial_theta[i]) * (j + 1))
            line = f"-\t{_fmt(100 * pos)}\t{_fmt(tmp)}"
            if len(fams) > 1:
                line += "".join(f"\t{_fmt(float(f[i, j]))}" for f in fams)
            out.append(line + "\n")
    out.append(f"{names[M - 1]}\t{_fmt(100.0 * float(gpos[M - 1]))}\n")
    return "".join(out)


def write_lod_table(path: str, names, gpos, partial_theta, family_lods) -> None:
    """Writes format_lod_table(...) to `path` (truncating it)."""
    with open(path, "w") as f:
        f.write(format_lod_table(names, gpos, partial_theta, family_lods))


def read_dat(path: str):
    """Returns (disease_freq, penetrance[3], minor_freqs[list])."""
    toks = [l.split("#")[0].split() for l in open(path)]
    toks = [t for t in toks if t]
    nloci = int(toks[0][0])
    i = 3
    freq = None
    pen = None
    minors = []
    while i < len(toks) and len(minors) + (1 if freq is not None else 0) < nloci:
        code = int(toks[i][0])
        if code == 1:  # affection status
            af = [float(x) for x in toks[i + 1]]
            freq = af[1]
            pen = [float(x) for x in toks[i + 3]]
            i += 4
        elif code == 3:  # numbered alleles
            af = [float(x) for x in toks[i + 1]]
            minors.append(af[1])
            i += 2
        else:
            raise ValueError(f"unsupported locus type {code} in {path}")
    if freq is None:
        raise ValueError("no affection-status locus in the DAT file")
    return freq, pen, minors


def read_ped(path: str):
    rows = []
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t:
                continue
            rows.append(t)
    return rows


def _genotype(a: str, b: str) -> int:
    if a == "0" or b == "0":
        return UNTYPED
    if a == b:
        return HOMOZ_A if a == "1" else HOMOZ_B
    return HETERO


def disease_probs(aff: int, founder: bool, freq: float, pen) -> List[float]:
    """[UU, AA, AU, UA] (person.cc:85-119 with get_penetrance_prob2 /
    get_apriori_prob2, disease_model.cc:146-187), autosomal."""
    def one(t):  # t: 0 homo U, 1 hetero, 2 homo A
        if not founder:
            if aff == 0:
                return 0.25
            return pen[t] if aff == 2 else 1.0 - pen[t]
        tmp = {2: freq * freq, 1: freq * (1.0 - freq), 0: (1.0 - freq) * (1.0 - freq)}[t]
        if aff == 2:
            return tmp * pen[t]
        if aff == 1:
            return tmp * (1.0 - pen[t])
        return (tmp * pen[t]) + (tmp * (1.0 - pen[t]))
    return [one(0), one(2), one(1), one(1)]


def marker_prior(minor: float) -> List[float]:
    """Snp::init_probs (genetic_map.h:68-87): [UU, AA, AU, UA] normalised."""
    major = 1.0 - minor
    p = [major * major, minor * minor, minor * major, minor * major]
    total = p[0] + p[2] + p[3] + p[1]
    return [x / total for x in p]


def sampler_cache(geno: np.ndarray, minors) -> np.ndarray:
    """[N][M][4] trait probabilities of the L-sampler (person.cc:224-299,
    founder branch for everyone)."""
    N, M = geno.shape
    out = np.zeros((N, M, 4))
    for l in range(M):
        mp = marker_prior(minors[l])
        for g in range(4):
            mask = {UNTYPED: (1, 1, 1, 1), HETERO: (0, 0, 1, 1), HOMOZ_A: (1, 0, 0, 0), HOMOZ_B: (0, 1, 0, 0)}[g]
            probs = [mp[t] if mask[t] else 0.0 for t in range(4)]
            total = probs[0] + probs[1] + probs[2] + probs[3]
            row = [x / total for x in probs]
            sel = geno[:, l] == g
            if sel.any():
                out[sel, l, :] = row
    return out


def eliminate(geno: np.ndarray, mother: np.ndarray, father: np.ndarray, F: int) -> np.ndarray:
    """Genotype elimination to a fixpoint (elimination.cc:41-125), vectorised
    over loci.  Returns legal[N][M], bit t set when phased trait t is
    possible."""
    N, M = geno.shape
    poss = np.full((N, M), G_AA | G_AB | G_BA | G_BB, np.uint8)
    poss[geno == HOMOZ_A] = G_AA
    poss[geno == HETERO] = G_AB | G_BA
    poss[geno == HOMOZ_B] = G_BB

    def parent_homoz(par, other, child, homoz, maternal):
        only = poss[par] == homoz
        if not only.any():
            return 0
        if (maternal and homoz == G_AA) or ((not maternal) and homoz == G_BB):
            het, other_het = G_AB, G_BA
        else:
            het, other_het = G_BA, G_AB
        other_homoz = G_BB if homoz == G_AA else G_AA
        before = poss[child].copy(), poss[other].copy()
        poss[child] = np.where(only, poss[child] & ~np.uint8(other_homoz | other_het), poss[child])
        cond = only & ((poss[child] & het) != 0) & ((poss[child] & homoz) == 0)
        poss[other] = np.where(cond, poss[other] & ~np.uint8(homoz), poss[other])
        return int((poss[child] != before[0]).sum() + (poss[other] != before[1]).sum())

    def child_homoz(mo, fa, child, homoz):
        only = poss[child] == homoz
        if not only.any():
            return 0
        other_homoz = G_BB if homoz == G_AA else G_AA
        b0, b1 = poss[mo].copy(), poss[fa].copy()
        poss[mo] = np.where(only, poss[mo] & ~np.uint8(other_homoz), poss[mo])
        poss[fa] = np.where(only, poss[fa] & ~np.uint8(other_homoz), poss[fa])
        return int((poss[mo] != b0).sum() + (poss[fa] != b1).sum())

    while True:
        changes = 0
        for i in range(F, N):
            mo, fa = int(mother[i]), int(father[i])
            changes += parent_homoz(mo, fa, i, G_AA, True)
            changes += parent_homoz(mo, fa, i, G_BB, True)
            changes += parent_homoz(fa, mo, i, G_AA, False)
            changes += parent_homoz(fa, mo, i, G_BB, False)
            changes += child_homoz(mo, fa, i, G_AA)
            changes += child_homoz(mo, fa, i, G_BB)
        if (poss == 0).any():
            bad = np.argwhere(poss == 0)[0]
            raise ValueError(f"genotypes inconsistent at person {bad[0]}, locus {bad[1]}")
        if changes == 0:
            break
    legal = np.zeros((N, M), np.uint8)
    for t, g in enumerate(TRAIT_TO_GENO):
        legal |= (((poss & g) != 0).astype(np.uint8) << t)
    return legal


def load(ped_path: str, map_path: str, dat_path: str, nlod: int = 5, peel_seed: int = 1,
         peel_tries: int = 64, peel: str = "greedy", peel_iterations: int = 1000000) -> Case:
    """The Case of the first pedigree in the files.  peel="greedy": this
    package's fast elimination order (peelseq.build, peel_tries random
    tie-breaks) over a stable founders-first order; peel="reference": the
    reference's person order (sl_person_order) and its own search
    (PeelSequenceGenerator::build_peel_sequence, through sl_peel_sequence)
    with peel_iterations swaps per restart and peel_seed for both its random
    sources, which gives the reference's tables."""
    names, gpos = read_map(map_path)
    freq, pen, minors = read_dat(dat_path)
    M = len(names)
    if len(minors) != M:
        raise ValueError(f"DAT has {len(minors)} markers, MAP has {M}")
    rows = read_ped(ped_path)
    fam = rows[0][0]
    rows = [r for r in rows if r[0] == fam]  # first pedigree only
    ids = [r[1] for r in rows]
    if peel == "reference":
        # the reference's own (unstable) founders-first order
        from . import native
        perm = native.person_order([1 if (r[2] == "0" and r[3] == "0") else 0 for r in rows])
        order = [rows[i] for i in perm]
        founders = [r for r in order if r[2] == "0" and r[3] == "0"]
    else:
        founders = [r for r in rows if r[2] == "0" and r[3] == "0"]
        others = [r for r in rows if not (r[2] == "0" and r[3] == "0")]
        order = founders + others
    idx = {r[1]: i for i, r in enumerate(order)}
    N, F = len(order), len(founders)
    mother = np.array([-1 if r[3] == "0" else idx[r[3]] for r in order], np.int32)
    father = np.array([-1 if r[2] == "0" else idx[r[2]] for r in order], np.int32)
    sex = np.array([int(r[4]) for r in order], np.int32)
    aff = np.array([int(r[5]) for r in order], np.int32)
    geno = np.zeros((N, M), np.uint8)
    for i, r in enumerate(order):
        al = r[6:]
        if len(al) != 2 * M:
            raise ValueError(f"person {r[1]} has {len(al) // 2} genotypes, map has {M}")
        geno[i] = [_genotype(al[2 * l], al[2 * l + 1]) for l in range(M)]
    typed = (geno != UNTYPED).any(axis=1).astype(np.int32)
    dprob = np.array([disease_probs(int(aff[i]), i < F, freq, pen) for i in range(N)])
    gprob = sampler_cache(geno, minors)
    legal = eliminate(geno, mother, father, F)
    theta = np.array([_haldane(gpos[l + 1] - gpos[l]) for l in range(M - 1)])
    if (theta == 0).any():
        raise ValueError("zero recombination fraction between adjacent markers (genetic_map.cc:48-53)")
    ptheta = np.array([_haldane(_inverse_haldane(t) / float(nlod + 1)) for t in theta])
    if peel == "reference":
        from . import native
        ops, _ = native.peel_sequence(mother, father, peel_iterations, seed=peel_seed, rand_seed=peel_seed)
    elif peel == "greedy":
        ops = peelseq.build(N, F, mother, father, seed=peel_seed, tries=peel_tries)
    else:
        raise ValueError(f"unknown peel search {peel!r}")
    case = Case(N=N, F=F, M=M, nlod=nlod, mother=mother, father=father, sex=sex, affection=aff, typed=typed,
                geno=geno, disease_prob=dprob, geno_prob=gprob, legal=legal, theta=theta, partial_theta=ptheta,
                minor=np.array(minors, np.float64), ops=ops, ids=[r[1] for r in order])
    return case
