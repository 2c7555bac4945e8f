"""Expected LOD score of pedigrees before genotyping (Elod, elod.h, elod.cc).

For each pedigree the reference builds two maps (elod.h:35-66):
  * a three-locus map marker (0) - trait (sep/2) - marker (sep), markers at
    minor frequency 1/2, on which the L-sampler runs with no genotypes and the
    trait locus' sampler probabilities replaced by the disease probabilities
    (elod.h:70-88, Person::copy_disease_probs, person.h:204-208);
  * a two-marker map (0, sep) with one trait position per interval, on which
    the trait peel scores the markers' descent graph.
Each replicate samples the three-locus graph by LocusSampler::start_from at
the trait locus (elod.cc:58) and scores its marker columns
(elod.cc:60-63); the ELOD is the LOD of all replicates at that position, and
the total is the sum over pedigrees (elod.cc:72-83).

Here the replicates run side by side on the device (sl_elod): replicate j
draws its uniforms in sequence from Philox chain j at step 0xFFFFFFFC, so the
result does not depend on how replicates are spread over workgroups.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Sequence, Tuple

import numpy as np

from .case import Case
from . import linkage, native

# defaults.h:4-7
DEFAULT_FREQUENCY = 0.0001
DEFAULT_PENETRANCE = (0.0, 0.0, 1.0)
DEFAULT_SEPARATION = 0.05
DEFAULT_REPLICATES = 1000000


def _haldane(m: float) -> float:
    return 0.5 * (1.0 - math.exp(-2.0 * m))


def _inverse_haldane(r: float) -> float:
    return -0.5 * math.log(1 - (2 * r))


def elod_cases(base: Case, frequency: float = DEFAULT_FREQUENCY, penetrance=DEFAULT_PENETRANCE,
               separation: float = DEFAULT_SEPARATION) -> Tuple[Case, Case]:
    """(three-locus sampler case, two-marker scorer case) of one pedigree.

    `base` supplies the pedigree (order, parents, sex, affection) and its peel
    sequence; its genotypes, map and disease model are not used."""
    if separation <= 0.0:
        raise ValueError("ELOD marker separation must be greater than zero (main.cc:442-444)")
    if len(penetrance) != 3:
        raise ValueError("ELOD penetrance needs three values")
    N, F = base.N, base.F
    dprob = np.array([linkage.disease_probs(int(base.affection[i]), i < F, frequency, list(penetrance))
                      for i in range(N)])
    # Markers have minor frequency 1/2 and no genotypes: founders take the
    # prior (1/4 each), others a flat 1/4 (person.cc:224-299, normalised).
    marker = np.full((N, 4), 0.25)
    half = separation / 2
    # GeneticMap::sanity_check recomputes thetas from the positions and the
    # partial thetas with one position per interval (genetic_map.cc:41-61)
    theta3 = np.array([_haldane(half - 0.0), _haldane(separation - half)])
    theta2 = np.array([_haldane(separation - 0.0)])
    part = lambda th: np.array([_haldane(_inverse_haldane(t) / 2.0) for t in th])
    common = dict(N=N, F=F, nlod=1, mother=base.mother, father=base.father, sex=base.sex,
                  affection=base.affection, disease_prob=dprob, ops=base.ops, ids=base.ids, trait_prob=None,
                  sex_linked=False)
    case3 = Case(M=3, typed=np.zeros(N, np.int32), geno=np.zeros((N, 3), np.uint8),
                 geno_prob=np.stack([marker, dprob, marker], axis=1), legal=np.full((N, 3), 15, np.uint8),
                 theta=theta3, partial_theta=part(theta3), minor=np.array([0.5, frequency, 0.5]), **common)
    case2 = Case(M=2, typed=np.zeros(N, np.int32), geno=np.zeros((N, 2), np.uint8),
                 geno_prob=np.stack([marker, marker], axis=1), legal=np.full((N, 2), 15, np.uint8),
                 theta=theta2, partial_theta=part(theta2), minor=np.array([0.5, 0.5]), **common)
    return case3, case2


@dataclasses.dataclass
class ElodResult:
    total: float
    per_pedigree: List[float]


def elod(pedigrees: Sequence[Case], frequency: float = DEFAULT_FREQUENCY, penetrance=DEFAULT_PENETRANCE,
         separation: float = DEFAULT_SEPARATION, replicates: int = DEFAULT_REPLICATES, seed: int = 1,
         precision: int = native.SL_FP64, device: int = 0) -> ElodResult:
    """Elod::run on the device: the ELOD of every pedigree and their sum."""
    if replicates <= 0:
        raise ValueError("replicates must be positive")
    per = []
    for base in pedigrees:
        c3, c2 = elod_cases(base, frequency, penetrance, separation)
        with native.Context(c3, precision, device=device) as sampler, \
                native.Context(c2, native.SL_FP64, device=device) as scorer:
            per.append(native.run_elod(sampler, scorer, seed, replicates))
    return ElodResult(total=float(sum(per)), per_pedigree=per)
