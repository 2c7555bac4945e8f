"""From LINKAGE files to device results: the reference's example pedigrees
(loop, east: their .ped/.map/.dat kept with the fixtures) and the seeded
30-person pedigree (written by synth) are read by linkage.load with the
reference's person order and peel search, which must reproduce the tables
the reference built (persons, peel table, disease and genotype
probabilities, legality, map); the HIP path then scores the fixture graphs
(Peeler::process terms within 1e-12) and runs the 40-sweep golden chain
(markov_chain.cc:314-404), which must end on the reference chain's graph bit
for bit with its LOD accumulators within 1e-9."""
import dataclasses as dc
import os

import numpy as np
import pytest

from conftest import load_golden, GOLDEN_CASES
from swiftlink_amd import linkage, native, synth
from swiftlink_amd.case import PeelOps

pytestmark = pytest.mark.gpu

EXAMPLES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "linkage")


def _load(name, tmp_path):
    if name == "ped30_snp100":
        p, m, d = synth.write_config(name, str(tmp_path))
    else:
        base = os.path.join(EXAMPLES, name)
        p, m, d = base + ".ped", base + ".map", base + ".dat"
    return linkage.load(p, m, d, peel="reference", peel_iterations=20000, peel_seed=12345)


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_linkage_files_through_the_device_match_the_reference(name, tmp_path):
    ref, fx = load_golden(name)
    case = _load(name, tmp_path)
    assert case.ids == ref.ids
    for f in dc.fields(PeelOps):
        assert np.array_equal(getattr(case.ops, f.name), getattr(ref.ops, f.name)), f.name
    for k in ("mother", "father", "typed", "geno", "disease_prob", "geno_prob", "legal", "theta", "partial_theta",
              "minor"):
        assert np.array_equal(getattr(case, k), getattr(ref, k)), k
    with native.Context(case, native.SL_FP64) as ctx:
        terms = ctx.lod_terms(fx["graphs"])
        np.testing.assert_allclose(terms, fx["lod_terms"], rtol=1e-12, atol=1e-10)
        seed, chain, sweeps, burnin, period = (int(x) for x in fx["chain_params"])
        ctx.mcmc_init(fx["graph_init"][None])
        ctx.lod_reset()
        ctx.mcmc_run(sweeps, burnin, period, float(fx["chain_lprob"][0]), seed)
        out = ctx.mcmc_graphs()[0]
        acc, cnt, _ = ctx.lod_results()
    assert np.array_equal(out, fx["chain_out"])
    assert cnt == int(fx["chain_count"][0])
    np.testing.assert_allclose(acc, fx["chain_acc"], rtol=0, atol=1e-9)
