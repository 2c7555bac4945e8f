"""Batches: several pedigrees' chains driven by one captured hipGraph per
sweep (sl_batch_*) give exactly what sl_mcmc_run gives on each context.  The
graph is replayed once per sweep and the device sweep counter advances
between replays, so every test compares after many replays, with eager work
and host synchronisation between batch calls."""
import numpy as np
import pytest

from conftest import load_golden, GOLDEN_CASES
from swiftlink_amd import native

pytestmark = pytest.mark.gpu

SCHED = dict(burnin=4, period=3, lprob=0.5, seed=77)


def _starts(fx, chains):
    pool = [fx["graph_init"]] + list(fx["graphs"][:3])
    return np.stack([pool[i % len(pool)] for i in range(chains)])


def _state(ctx):
    acc, cnt, _ = ctx.lod_results()
    return ctx.mcmc_graphs(), acc, cnt, ctx.mcmc_status(), ctx.mcmc_near_edge()


def _eager(name, chains, precision, sweeps, splits):
    case, fx = load_golden(name)
    with native.Context(case, precision) as ctx:
        ctx.mcmc_init(_starts(fx, chains))
        ctx.lod_reset()
        stats = []
        first = 0
        for n in splits(sweeps):
            stats.append(ctx.mcmc_run(n, SCHED["burnin"], SCHED["period"], SCHED["lprob"], SCHED["seed"], first))
            first += n
        return _state(ctx), stats


@pytest.mark.parametrize("precision", [native.SL_FP64, native.SL_FP32_RESCALED])
def test_batch_matches_per_context_runs(precision):
    spec = [("loop", 1), ("east", 3), ("ped30_snp100", 4)]
    sweeps = 21
    parts = lambda n: [7, 14]
    ctxs = []
    for name, chains in spec:
        case, fx = load_golden(name)
        c = native.Context(case, precision)
        c.mcmc_init(_starts(fx, chains))
        c.lod_reset()
        ctxs.append(c)
    try:
        with native.Batch(ctxs) as b:
            first, stats = 0, []
            for n in parts(sweeps):
                stats.append(b.run(n, SCHED["burnin"], SCHED["period"], SCHED["lprob"], SCHED["seed"], first))
                first += n
                for c in ctxs:  # eager work between batch calls
                    c.sync()
                    c.mcmc_graphs()
        for i, (name, chains) in enumerate(spec):
            (g, acc, cnt, st, near), estats = _eager(name, chains, precision, sweeps, parts)
            bg, bacc, bcnt, bst, bnear = _state(ctxs[i])
            assert np.array_equal(bg, g), name
            assert np.array_equal(bacc, acc), name
            assert bcnt == cnt and np.array_equal(bst, st) and np.array_equal(bnear, near)
            for bs, es in zip(stats, estats):
                assert bs[i] == es
    finally:
        for c in ctxs:
            c.close()


def test_batch_rejects_shared_context_and_survives_destroyed_member():
    case, fx = load_golden("loop")
    a = native.Context(case, native.SL_FP64)
    b = native.Context(case, native.SL_FP64)
    for c in (a, b):
        c.mcmc_init(_starts(fx, 2))
        c.lod_reset()
    batch = native.Batch([a, b])
    with pytest.raises(native.SlError):
        native.Batch([a])
    b.close()  # leaves the batch
    batch.run(5, 0, 1, 0.5, 3)
    batch.close()
    a.close()


def test_autotune_keeps_chains_and_results():
    """The autotuner restores the chains it timed, and any launch shape gives
    the same chains (and accumulators up to the log-sum order of graph
    chunks) as the default shapes."""
    from swiftlink_amd.tune import autotune
    case, fx = load_golden("ped30_snp100")
    starts = _starts(fx, 4)
    with native.Context(case, native.SL_FP64) as a, native.Context(case, native.SL_FP64) as b:
        a.mcmc_init(starts)
        tu = autotune(a, blocks=(64, 256), gpws=(1, 4))
        assert np.array_equal(a.mcmc_graphs(), starts)
        assert a.options == {"block": tu["block"], "gpw": tu["gpw"], "msampler_mode": tu["msampler_mode"],
                             "lod_kernel": tu["lod_kernel"]}
        b.mcmc_init(starts)
        for c in (a, b):
            c.lod_reset()
            c.mcmc_run(12, *SCHED.values())
        assert np.array_equal(a.mcmc_graphs(), b.mcmc_graphs())
        acc_a, cnt_a, _ = a.lod_results()
        acc_b, cnt_b, _ = b.lod_results()
        assert cnt_a == cnt_b
        np.testing.assert_allclose(acc_a, acc_b, rtol=1e-12, atol=1e-9)
