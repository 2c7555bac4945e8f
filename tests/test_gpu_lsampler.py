"""L-sampler on the GPU against the oracle and the compiled reference.

FP64: the sampled graph after an L-sweep is bit-identical to the reference's
LocusSampler::step over the same locus order under the same Philox stream
(golden fixtures) and to the oracle, for one and several chains.
FP32 (per-table exponents): each locus is an independent draw given its
neighbours, so a locus can disagree with the FP64 sweep only where one of its
op draws fell within float rounding of a CDF edge, which the kernel counts and
marks per locus (sl_mcmc_near_edge_loci), or where a neighbour it conditioned
on disagreed.
"""
import numpy as np
import pytest

from conftest import load_golden, GOLDEN_CASES
from oracle_lib import Oracle
from swiftlink_amd import native, synth

pytestmark = pytest.mark.gpu


def oracle_lsweep(o, g, M, seed, chain, sweep, p0=0):
    order = list(range(p0, M, 2)) + list(range(1 - p0, M, 2))
    for l in order:
        g, _ = o.lstep(g, l, seed, chain, sweep)
    return g


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_lsweep_fp64_matches_reference(name):
    case, fx = load_golden(name)
    seed, chain, sweep = (int(x) for x in fx["lsweep_seed"])
    assert chain == 0
    with native.Context(case, native.SL_FP64) as ctx:
        ctx.mcmc_init(fx["graph_init"][None])
        ctx.lsweep(seed, sweep, 0)
        out = ctx.mcmc_graphs()[0]
        assert np.all(ctx.mcmc_status() == 0)
    assert np.array_equal(out, fx["lsweep_out"])


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_lsweep_fp64_chains_and_parity(name):
    case, fx = load_golden(name)
    o = Oracle(case)
    starts = np.stack([fx["graph_init"], fx["graphs"][0], fx["graphs"][1]])
    with native.Context(case, native.SL_FP64) as ctx:
        ctx.mcmc_init(starts)
        ctx.lsweep(1234, 3, 1)
        ctx.lsweep(1234, 4, 0)
        out = ctx.mcmc_graphs()
    for c in range(3):
        g = oracle_lsweep(o, starts[c], case.M, 1234, c, 3, 1)
        g = oracle_lsweep(o, g, case.M, 1234, c, 4, 0)
        assert np.array_equal(out[c], g), c


@pytest.fixture(scope="module")
def ped150(tmp_path_factory):
    d = tmp_path_factory.mktemp("ped150s")
    case, data = synth.make_case("ped150_snp1000", str(d))
    return case, synth.true_graph(data, case), Oracle(case)


def test_lsweep_fp64_ped150(ped150):
    case, g0, o = ped150
    with native.Context(case, native.SL_FP64) as ctx:
        ctx.mcmc_init(g0[None])
        ctx.lsweep(77, 0, 0)
        out = ctx.mcmc_graphs()[0]
        assert np.all(ctx.mcmc_status() == 0)
    ref = oracle_lsweep(o, g0, case.M, 77, 0, 0, 0)
    assert np.array_equal(out, ref)


def _check_fp32_sweep(case, out, ref, near, p0):
    """An FP32 L-sweep against the FP64 one from the same graph and stream.
    A first-parity locus depends on the input graph only, so it may differ
    only where its op draws fell near a CDF edge; a second-parity locus may
    also differ because a neighbour did."""
    diff = np.any(out != ref, axis=1)
    for l in np.nonzero(diff)[0]:
        if l % 2 == p0:
            assert near[l], ("first parity", int(l))
        else:
            nb = (l > 0 and diff[l - 1]) or (l + 1 < case.M and diff[l + 1])
            assert near[l] or nb, ("second parity", int(l))
    return int(diff.sum())


def test_lsweep_fp32_ped150(ped150):
    """FP32 tables: every locus that differs from the FP64 oracle's sweep is
    explained by a counted near-edge draw (at it, or at a first-parity
    neighbour it conditioned on); every column stays legal."""
    case, g0, o = ped150
    with native.Context(case, native.SL_FP32_RESCALED) as ctx:
        ctx.mcmc_init(g0[None])
        ctx.lsweep(77, 0, 0)
        out = ctx.mcmc_graphs()[0]
        near = ctx.mcmc_near_edge_loci()[0]
        count = int(ctx.mcmc_near_edge()[0])
        assert np.all(ctx.mcmc_status() == 0)
    ref = oracle_lsweep(o, g0, case.M, 77, 0, 0, 0)
    assert count >= int(near.sum())
    differing = _check_fp32_sweep(case, out, ref, near, 0)
    print(f"ped150 FP32 L-sweep: {differing} loci differ, {count} near-edge draws")
    assert min(o.fag(out, l) for l in range(case.M)) > 0


def test_lsweep_fp64_counts_no_near_edges(ped150):
    case, g0, o = ped150
    with native.Context(case, native.SL_FP64) as ctx:
        ctx.mcmc_init(np.stack([g0, g0]))
        ctx.lsweep(78, 1, 1)
        assert np.all(ctx.mcmc_near_edge() == 0)
        assert not ctx.mcmc_near_edge_loci().any()


def test_lsweep_fp32_narrow_shape_matches_wide(ped150):
    """Two FP32 chains of ped150 make 1000 workgroups per parity launch, past
    the 768 below which the L-sampler takes 512-thread workgroups: they run
    the 256-thread shape and its four-wave level schedule (about 3 cells per
    op at ped150).  One chain makes 500 workgroups and runs the wide shape.
    Each op is evaluated by one wavefront in both shapes and the table
    exponents are integer sums, so chain 0 of the two-chain run, which draws
    the same Philox stream (chain index 0) as the one-chain run, must equal it
    bit for bit.  Chain 1 draws the stream of chain index 1: it must be legal
    at every locus and differ from the FP64 oracle's sweep on that stream only
    where a counted near-edge draw explains it."""
    case, g0, o = ped150
    with native.Context(case, native.SL_FP32_RESCALED) as ctx:
        ctx.mcmc_init(g0[None])
        ctx.lsweep(91, 2, 1)
        one = ctx.mcmc_graphs()[0]
        ctx.mcmc_init(np.stack([g0, g0]))
        ctx.lsweep(91, 2, 1)
        two = ctx.mcmc_graphs()
        assert np.all(ctx.mcmc_status() == 0)
        near = ctx.mcmc_near_edge_loci()
    assert np.array_equal(two[0], one)
    ref1 = oracle_lsweep(o, g0, case.M, 91, 1, 2, 1)
    assert np.any(two[1] != two[0])
    _check_fp32_sweep(case, two[1], ref1, near[1], 1)
    assert min(o.fag(two[1], l) for l in range(case.M)) > 0
