"""The C oracle against golden vectors written by the compiled reference
(tests/golden/make_golden.py).  Everything is compared bit for bit."""
import numpy as np

from oracle_lib import Oracle


def test_trait_prob(golden):
    case, fx = golden
    assert Oracle(case).trait_prob() == case.trait_prob


def test_lod_terms(golden):
    case, fx = golden
    o = Oracle(case)
    for g, graph in enumerate(fx["graphs"]):
        mine = np.array([o.lod_terms(graph, l) for l in range(case.M - 1)])
        assert np.array_equal(mine, fx["lod_terms"][g])


def test_lod_accumulate(golden):
    case, fx = golden
    acc = Oracle(case).lod_accumulate(fx["graphs"])
    assert np.array_equal(acc, fx["lod_acc"])
    lod = (acc - np.log(len(fx["graphs"])) - case.trait_prob) / np.log(10.0)
    np.testing.assert_allclose(lod, fx["lod_final"], rtol=0, atol=1e-12)


def test_fag(golden):
    case, fx = golden
    o = Oracle(case)
    mine = np.array([[o.fag(g, l) for l in range(case.M)] for g in fx["graphs"]])
    assert np.array_equal(mine, fx["fag"])


def test_lsweep(golden):
    case, fx = golden
    o = Oracle(case)
    seed, chain, sweep = (int(x) for x in fx["lsweep_seed"])
    g = fx["graph_init"].copy()
    draws = []
    for l in fx["lsweep_order"]:
        g, n = o.lstep(g, int(l), seed, chain, sweep)
        draws.append(n)
    assert np.array_equal(g, fx["lsweep_out"])
    assert np.array_equal(np.array(draws), fx["lsweep_draws"])


def test_msweep(golden):
    case, fx = golden
    o = Oracle(case)
    seed, chain, sweep = (int(x) for x in fx["msweep_seed"])
    g = o.msweep(fx["graph_init"], fx["msweep_order"], seed, chain, sweep)
    assert np.array_equal(g, fx["msweep_out"])


def test_chain(golden):
    case, fx = golden
    seed, chain, sweeps, burnin, period = (int(x) for x in fx["chain_params"])
    g, acc, cnt, kinds = Oracle(case).chain(fx["graph_init"], seed, chain, sweeps, burnin, period,
                                            float(fx["chain_lprob"][0]))
    assert np.array_equal(kinds, fx["chain_kinds"])
    assert np.array_equal(g, fx["chain_out"])
    assert cnt == int(fx["chain_count"][0])
    assert np.array_equal(acc, fx["chain_acc"])


def test_meiosis_list(golden):
    case, fx = golden
    assert np.array_equal(Oracle(case).meioses(), case.meioses())


def test_impute(golden):
    """Sequential-imputation replicates (LocusSampler::sequential_imputation)
    and the replicate parallel_run keeps."""
    case, fx = golden
    o = Oracle(case)
    seed, R = (int(x) for x in fx["impute_seed"])
    w = []
    for r in range(R):
        g, wr = o.impute(seed, r)
        assert np.array_equal(g, fx["impute_graphs"][r]), r
        w.append(wr)
    assert np.array_equal(np.array(w), fx["impute_weights"])
    assert int(np.argmax(w)) == int(fx["impute_best"][0])


def test_impute_by_locus(golden):
    case, fx = golden
    g, w = Oracle(case).impute(int(fx["impute0_seed"][0]), 0, by_locus=True)
    assert np.array_equal(g, fx["impute0_graph"])
    assert w == float(fx["impute0_weight"][0])


def test_impute_weight_recomputed(golden):
    """A replicate's weight from its final graph: the locus likelihoods along
    its walk (each locus conditions only on the neighbour sampled before it,
    which the later steps never rewrite)."""
    from oracle_lib import uniform
    case, fx = golden
    o = Oracle(case)
    seed, R = (int(x) for x in fx["impute_seed"])
    for r in range(R):
        start = min(int(uniform(seed, r, 0, 0xFFFFFFFC, 0) * case.M), case.M - 1)
        assert o.impute_weight(fx["impute_graphs"][r], start) == fx["impute_weights"][r]
