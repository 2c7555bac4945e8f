"""Whole chains on the GPU (sl_mcmc_run) against the reference and the oracle.

FP64: after the schedule, every chain's graph is identical to the reference's
MarkovChain run fed the same Philox stream (golden fixture for chain 0) and to
the oracle's (other chains), the scored-graph count matches, and the LOD
accumulators agree within 1e-9 (the device merges graphs with a max-shifted
log-sum-exp instead of the serial log_sum, so only rounding differs).  Several
chains merge as the reference's -R replicates do: log-sum over chains.
"""
import numpy as np
import pytest

from conftest import load_golden, GOLDEN_CASES
from oracle_lib import Oracle
from swiftlink_amd import native

pytestmark = pytest.mark.gpu

ACC_TOL = 1e-9


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_chain_fp64_matches_reference(name):
    case, fx = load_golden(name)
    seed, chain, sweeps, burnin, period = (int(x) for x in fx["chain_params"])
    lprob = float(fx["chain_lprob"][0])
    assert chain == 0
    with native.Context(case, native.SL_FP64) as ctx:
        ctx.mcmc_init(fx["graph_init"][None])
        ctx.lod_reset()
        st = ctx.mcmc_run(sweeps, burnin, period, lprob, seed)
        out = ctx.mcmc_graphs()[0]
        acc, cnt, _ = ctx.lod_results()
        assert np.all(ctx.mcmc_near_edge() == 0)
    kinds = fx["chain_kinds"]
    assert st["l_sweeps"] == int((kinds == 0).sum()) and st["m_sweeps"] == int((kinds == 1).sum())
    assert np.array_equal(out, fx["chain_out"])
    assert cnt == int(fx["chain_count"][0]) == st["scored"]
    np.testing.assert_allclose(acc, fx["chain_acc"], rtol=0, atol=ACC_TOL)


@pytest.mark.parametrize("mode", [native.SL_MSAMPLER_SPECULATIVE, native.SL_MSAMPLER_TWO_LAUNCH])
@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_chains_merge_like_replicates(name, mode):
    case, fx = load_golden(name)
    o = Oracle(case)
    starts = np.stack([fx["graph_init"], fx["graphs"][0], fx["graphs"][1], fx["graph_init"]])
    sweeps, burnin, period, lprob, seed = 24, 6, 3, 0.5, 31
    with native.Context(case, native.SL_FP64, msampler_mode=mode) as ctx:
        ctx.mcmc_init(starts)
        ctx.lod_reset()
        st = ctx.mcmc_run(sweeps, burnin, period, lprob, seed)
        out = ctx.mcmc_graphs()
        acc, cnt, _ = ctx.lod_results()
        assert np.all(ctx.mcmc_near_edge() == 0)
    accs, total, kinds_all = [], 0, []
    for c in range(len(starts)):
        g, a, n, kinds = o.chain(starts[c], seed, c, sweeps, burnin, period, lprob)
        assert np.array_equal(out[c], g), c
        accs.append(a)
        total += n
        kinds_all.append(kinds)
    kinds_all = np.array(kinds_all)
    assert st["l_sweeps"] == int((kinds_all == 0).sum()) and st["m_sweeps"] == int((kinds_all == 1).sum())
    assert cnt == total == st["scored"]
    np.testing.assert_allclose(acc, np.logaddexp.reduce(np.array(accs), axis=0), rtol=0, atol=ACC_TOL)


@pytest.mark.parametrize("mode", [native.SL_MSAMPLER_SPECULATIVE, native.SL_MSAMPLER_TWO_LAUNCH])
@pytest.mark.parametrize("chains", [2, 4])
@pytest.mark.parametrize("precision", [native.SL_FP64, native.SL_FP32_RESCALED])
def test_captured_sweeps_match_launched_sweeps(precision, chains, mode):
    """With several chains sl_mcmc_run replays captured graphs of its sweeps
    (the sweep index on the device) unless profiling is on, when every kernel
    is launched from the stream with events around it.  Over two runs, with
    LOD scoring launched between the replays, both give the same graphs, LOD
    accumulators, sweep counts and near-edge counts; the chains mix L- and
    M-sweeps within sweeps."""
    case, fx = load_golden("ped30_snp100")
    starts = np.stack([fx["graph_init"], fx["graphs"][0], fx["graphs"][1], fx["graph_init"]])[:chains]
    res = []
    for prof in (False, True):
        with native.Context(case, precision, msampler_mode=mode) as ctx:
            ctx.profile(prof)
            ctx.mcmc_init(starts)
            ctx.lod_reset()
            st = [ctx.mcmc_run(23, 4, 3, 0.5, 17), ctx.mcmc_run(9, 4, 3, 0.5, 17, first_sweep=23)]
            res.append((st, ctx.mcmc_graphs(), ctx.lod_results(), ctx.mcmc_near_edge()))
    (st_a, g_a, (acc_a, n_a, _), ne_a), (st_b, g_b, (acc_b, n_b, _), ne_b) = res
    assert st_a == st_b
    assert np.array_equal(g_a, g_b)
    assert n_a == n_b > 0
    assert np.array_equal(acc_a, acc_b)
    assert np.array_equal(ne_a, ne_b)
    assert st_a[0]["l_sweeps"] > 0 and st_a[0]["m_sweeps"] > 0


def test_run_continues_from_first_sweep():
    case, fx = load_golden("ped30_snp100")
    starts = np.stack([fx["graph_init"], fx["graphs"][0]])
    with native.Context(case, native.SL_FP64) as ctx:
        ctx.mcmc_init(starts)
        ctx.lod_reset()
        ctx.mcmc_run(30, 5, 5, 0.5, 8)
        whole = ctx.mcmc_graphs()
        acc_whole, n_whole, _ = ctx.lod_results()
        ctx.mcmc_init(starts)
        ctx.lod_reset()
        ctx.mcmc_run(13, 5, 5, 0.5, 8)
        ctx.mcmc_run(17, 5, 5, 0.5, 8, first_sweep=13)
        parts = ctx.mcmc_graphs()
        acc_parts, n_parts, _ = ctx.lod_results()
    assert np.array_equal(whole, parts)
    assert n_whole == n_parts
    np.testing.assert_allclose(acc_parts, acc_whole, rtol=0, atol=ACC_TOL)


def test_pure_l_and_pure_m_schedules():
    case, fx = load_golden("east")
    o = Oracle(case)
    with native.Context(case, native.SL_FP64) as ctx:
        for lprob, key in ((1.0, "l_sweeps"), (0.0, "m_sweeps")):
            ctx.mcmc_init(fx["graph_init"][None])
            st = ctx.mcmc_run(5, 100, 1, lprob, 3)
            assert st[key] == 5 and st["scored"] == 0
            g, _, _, kinds = o.chain(fx["graph_init"], 3, 0, 5, 100, 1, lprob)
            assert np.all(kinds == (0 if lprob == 1.0 else 1))
            assert np.array_equal(ctx.mcmc_graphs()[0], g)


def test_chain_fp32_stays_legal_and_scores():
    case, fx = load_golden("ped30_snp100")
    o = Oracle(case)
    starts = np.stack([fx["graph_init"]] * 3)
    with native.Context(case, native.SL_FP32_RESCALED) as ctx:
        ctx.mcmc_init(starts)
        ctx.lod_reset()
        st = ctx.mcmc_run(40, 10, 5, 0.5, 12)
        out = ctx.mcmc_graphs()
        acc, cnt, lod = ctx.lod_results()
        lik = ctx.mcmc_likelihood()
        assert np.all(ctx.mcmc_status() == 0)
    assert cnt == st["scored"] == 3 * 6  # sweeps 10, 15, ..., 35
    assert np.all(np.isfinite(acc)) and np.all(np.isfinite(lod))
    for c in range(3):
        ref = np.array([o.fag(out[c], l) for l in range(case.M)])
        assert np.all(ref > 0)
        assert np.all(np.abs(lik[c] / ref - 1.0) <= 1e-12)


def _coda_value(o, case, dg):
    """DescentGraph::get_likelihood (descent_graph.cc:184-215) of one graph
    from its parts: 2(N-F) ln 1/2 + the recombination log-probability
    between adjacent loci + the sum of ln FAG over loci (the oracle's
    founder-allele-graph likelihood)."""
    nm = 2 * (case.N - case.F)
    m = dg[:, 2 * case.F:2 * case.N]
    cross = (m[1:] != m[:-1]).sum(axis=1)
    rec = (cross * np.log(case.theta) + (nm - cross) * np.log1p(-case.theta)).sum()
    return nm * np.log(0.5) + rec + sum(np.log(o.fag(dg, l)) for l in range(case.M))


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_coda_trace_matches_reference(name):
    """The trace holds the reference chain's CODA likelihoods
    (markov_chain.cc:361-369 writes get_likelihood at each scored sweep): the
    same iterations, and values within 1e-9 relative of the golden ones; the
    last row (the final sweep is scored) equals the final graph's likelihood
    computed from its parts."""
    case, fx = load_golden(name)
    seed, chain, sweeps, burnin, period = (int(x) for x in fx["chain_params"])
    lprob = float(fx["chain_lprob"][0])
    with native.Context(case, native.SL_FP64) as ctx:
        ctx.mcmc_init(fx["graph_init"][None])
        ctx.mcmc_trace(True)
        ctx.mcmc_run(sweeps, burnin, period, lprob, seed)
        it, ll = ctx.mcmc_trace_read()
        out = ctx.mcmc_graphs()[0]
    assert np.array_equal(it, fx["chain_coda_iter"])
    np.testing.assert_allclose(ll[:, 0], fx["chain_coda"], rtol=1e-9, atol=0)
    assert it[-1] == sweeps
    o = Oracle(case)
    assert abs(ll[-1, 0] - _coda_value(o, case, out)) <= 1e-9 * abs(ll[-1, 0])


def test_coda_trace_spans_runs_and_chains():
    """Several chains and a schedule split over two runs record the same trace
    as one run; chain 0 (same start, same Philox stream) repeats the golden
    values; re-initialising clears the trace; turning it off records
    nothing."""
    case, fx = load_golden("ped30_snp100")
    seed, _, sweeps, burnin, period = (int(x) for x in fx["chain_params"])
    lprob = float(fx["chain_lprob"][0])
    starts = np.stack([fx["graph_init"], fx["graphs"][0], fx["graphs"][1]])
    with native.Context(case, native.SL_FP64) as ctx:
        ctx.mcmc_init(starts)
        ctx.mcmc_trace(True)
        ctx.mcmc_run(sweeps, burnin, period, lprob, seed)
        it1, ll1 = ctx.mcmc_trace_read()
        ctx.mcmc_init(starts)
        assert ctx.mcmc_trace_read()[0].size == 0
        ctx.mcmc_run(17, burnin, period, lprob, seed)
        ctx.mcmc_run(sweeps - 17, burnin, period, lprob, seed, first_sweep=17)
        it2, ll2 = ctx.mcmc_trace_read()
        ctx.mcmc_init(starts)
        ctx.mcmc_trace(False)
        ctx.mcmc_run(sweeps, burnin, period, lprob, seed)
        assert ctx.mcmc_trace_read()[0].size == 0
    assert ll1.shape == (len(fx["chain_coda_iter"]), 3)
    assert np.array_equal(it1, it2) and np.array_equal(ll1, ll2)
    np.testing.assert_allclose(ll1[:, 0], fx["chain_coda"], rtol=1e-9, atol=0)
    assert np.all(np.isfinite(ll1))
