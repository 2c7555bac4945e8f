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
