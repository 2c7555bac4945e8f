"""M-sampler on the GPU against the oracle and the compiled reference.

The founder-allele-graph legality (likelihood zero or not) is the
reference's exactly, and the likelihood values agree with its products within
1e-12 relative (the kernel sums logs over components).  With that rounding
and the forward-backward pass reassociated into block scans, a draw can
decide differently from the reference only if its uniform lies within
rounding of its threshold; the kernel counts draws within 1e-12 of their
threshold, and these tests require that count to be zero and the sampled
graphs to be identical to the reference's (golden fixtures) and the oracle's
under the same Philox stream.
"""
import numpy as np
import pytest

from conftest import load_golden, GOLDEN_CASES
from oracle_lib import Oracle, uniform
from swiftlink_amd import native, synth

pytestmark = pytest.mark.gpu

STEP_SHUFFLE = 0xFFFFFFFE


def shuffled_order(o, seed, chain, sweep):
    """Fisher-Yates over the non-ignorable meioses with the chain's shuffle draws."""
    mo = [int(x) for x in o.meioses()]
    d = 0
    for i in range(len(mo) - 1, 0, -1):
        j = min(int(uniform(seed, chain, sweep, STEP_SHUFFLE, d) * (i + 1)), i)
        d += 1
        mo[i], mo[j] = mo[j], mo[i]
    return np.array(mo, np.int32)


# launch modes, each with the founder-allele graphs in registers (the default
# up to 1024 people) and in LDS (the path of larger pedigrees)
MODES = [m | g for m in (native.SL_MSAMPLER_SPECULATIVE, native.SL_MSAMPLER_TWO_LAUNCH)
         for g in (0, native.SL_MSAMPLER_LDS_GRAPHS)]
GRAPHS = [native.SL_MSAMPLER_AUTO, native.SL_MSAMPLER_LDS_GRAPHS]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_msweep_fp64_matches_reference(name, mode):
    case, fx = load_golden(name)
    seed, chain, sweep = (int(x) for x in fx["msweep_seed"])
    assert chain == 0
    with native.Context(case, native.SL_FP64, msampler_mode=mode) as ctx:
        ctx.mcmc_init(fx["graph_init"][None])
        ctx.msweep(seed, sweep, fx["msweep_order"])
        out = ctx.mcmc_graphs()[0]
        assert np.all(ctx.mcmc_status() == 0)
        assert np.all(ctx.mcmc_near_edge() == 0)
    assert np.array_equal(out, fx["msweep_out"])


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_msweep_shuffled_chains_match_oracle(name, mode):
    case, fx = load_golden(name)
    o = Oracle(case)
    starts = np.stack([fx["graph_init"], fx["graphs"][0], fx["graphs"][1]])
    with native.Context(case, native.SL_FP64, msampler_mode=mode) as ctx:
        ctx.mcmc_init(starts)
        ctx.msweep(4321, 7)
        ctx.msweep(4321, 8)
        out = ctx.mcmc_graphs()
        assert np.all(ctx.mcmc_status() == 0)
        assert np.all(ctx.mcmc_near_edge() == 0)
    for c in range(3):
        g = o.msweep(starts[c], shuffled_order(o, 4321, c, 7), 4321, c, 7)
        g = o.msweep(g, shuffled_order(o, 4321, c, 8), 4321, c, 8)
        assert np.array_equal(out[c], g), c


LIK_RTOL = 1e-12


def assert_likelihoods(got, ref):
    assert np.array_equal(got == 0.0, ref == 0.0)
    nz = ref != 0.0
    assert np.all(np.abs(got[nz] / ref[nz] - 1.0) <= LIK_RTOL)


@pytest.mark.parametrize("graphs", GRAPHS)
@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_locus_likelihoods_match_reference(name, graphs):
    case, fx = load_golden(name)
    o = Oracle(case)
    starts = np.stack([fx["graph_init"], fx["graphs"][0], fx["lsweep_out"], fx["msweep_out"]])
    with native.Context(case, native.SL_FP64, msampler_mode=graphs) as ctx:
        ctx.mcmc_init(starts)
        lik = ctx.mcmc_likelihood()
        assert np.all(ctx.mcmc_status() == 0)
    for c in range(len(starts)):
        ref = np.array([o.fag(starts[c], l) for l in range(case.M)])
        assert np.all(ref > 0)
        assert_likelihoods(lik[c], ref)


@pytest.mark.parametrize("graphs", GRAPHS)
@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_flipped_columns_legality_and_likelihood(name, graphs):
    """Graphs with random single flips: many columns become illegal."""
    case, fx = load_golden(name)
    o = Oracle(case)
    rng = np.random.default_rng(11)
    gs = []
    for _ in range(6):
        g = fx["graph_init"].copy()
        for l in range(case.M):
            g[l, rng.integers(2 * case.F, 2 * case.N)] ^= 1
        gs.append(g)
    gs = np.stack(gs)
    with native.Context(case, native.SL_FP64, msampler_mode=graphs) as ctx:
        ctx.mcmc_init(gs)
        lik = ctx.mcmc_likelihood()
    ref = np.array([[o.fag(g, l) for l in range(case.M)] for g in gs])
    assert (ref == 0.0).any() and (ref > 0.0).any()
    assert_likelihoods(lik, ref)


def test_register_and_lds_graphs_agree_bitwise(ped150):
    """Both founder-allele graph builds sum the same per-component terms in the same order."""
    case, g0, _ = ped150
    rng = np.random.default_rng(5)
    gs = [g0]
    for _ in range(3):
        g = g0.copy()
        for l in range(case.M):
            g[l, rng.integers(2 * case.F, 2 * case.N)] ^= 1
        gs.append(g)
    gs = np.stack(gs)
    lik = []
    for graphs in GRAPHS:
        with native.Context(case, native.SL_FP64, msampler_mode=graphs) as ctx:
            ctx.mcmc_init(gs)
            lik.append(ctx.mcmc_likelihood())
    assert (lik[0] == 0.0).any() and (lik[0] > 0.0).any()
    assert np.array_equal(lik[0], lik[1])


def _illegal_graph(case, o, g):
    """A copy of g with one meiosis flipped at one locus so that the column is illegal."""
    for l in range(case.M):
        for b in range(2 * case.F, 2 * case.N):
            h = g.copy()
            h[l, b] ^= 1
            if o.fag(h, l) == 0.0:
                return h, l
    return None, None


def test_illegal_graph_is_reported_and_left_unchanged():
    case, fx = load_golden("loop")
    o = Oracle(case)
    bad, locus = _illegal_graph(case, o, fx["graph_init"])
    assert bad is not None
    starts = np.stack([fx["graph_init"], bad])
    with native.Context(case, native.SL_FP64) as ctx:
        ctx.mcmc_init(starts)
        lik = ctx.mcmc_likelihood()
        assert lik[1, locus] == 0.0 and np.all(lik[0] > 0)
        st = ctx.mcmc_status()
        assert st[0] == 0 and st[1] == native.SL_E_ILLEGAL_GRAPH
        ctx.mcmc_init(starts)
        ctx.msweep(5, 0)
        out = ctx.mcmc_graphs()
        st = ctx.mcmc_status()
    assert st[1] == native.SL_E_ILLEGAL_GRAPH and st[0] == 0
    assert np.array_equal(out[1], bad)
    assert np.array_equal(out[0], o.msweep(starts[0], shuffled_order(o, 5, 0, 0), 5, 0, 0))


def test_lsweep_then_msweep_keeps_graphs_legal():
    case, fx = load_golden("ped30_snp100")
    o = Oracle(case)
    starts = np.stack([fx["graph_init"]] * 4)
    with native.Context(case, native.SL_FP64) as ctx:
        ctx.mcmc_init(starts)
        for s in range(6):
            if s % 2:
                ctx.msweep(99, s)
            else:
                ctx.lsweep(99, s, s % 4 // 2)
        out = ctx.mcmc_graphs()
        lik = ctx.mcmc_likelihood()
        assert np.all(ctx.mcmc_status() == 0)
    for c in range(4):
        ref = np.array([o.fag(out[c], l) for l in range(case.M)])
        assert np.all(ref > 0)
        assert_likelihoods(lik[c], ref)


@pytest.fixture(scope="module")
def ped150(tmp_path_factory):
    d = tmp_path_factory.mktemp("ped150m")
    case, data = synth.make_case("ped150_snp1000", str(d))
    return case, synth.true_graph(data, case), Oracle(case)


@pytest.mark.parametrize("mode", MODES)
def test_msweep_fp64_ped150(ped150, mode):
    case, g0, o = ped150
    with native.Context(case, native.SL_FP64, msampler_mode=mode) as ctx:
        ctx.mcmc_init(g0[None])
        lik = ctx.mcmc_likelihood()[0]
        ctx.msweep(2024, 1)
        out = ctx.mcmc_graphs()[0]
        assert np.all(ctx.mcmc_status() == 0)
        near = int(ctx.mcmc_near_edge()[0])
    ref = np.array([o.fag(g0, l) for l in range(case.M)])
    assert_likelihoods(lik, ref)
    print(f"ped150 M-sweep: {near} draws within 1e-12 of their threshold")
    assert near == 0
    assert np.array_equal(out, o.msweep(g0, shuffled_order(o, 2024, 0, 1), 2024, 0, 1))


@pytest.fixture(scope="module")
def long_map(tmp_path_factory):
    """A small pedigree over a long marker map: the forward-backward pass then
    runs its segments from global memory (M > 8 x 512 in one launch per step)
    or with eight loci per thread (two launches per step)."""
    d = tmp_path_factory.mktemp("ped30m5000")
    case, data = synth.make_case("ped30_snp100", str(d), n_markers=5000, spacing_cm=0.1)
    return case, synth.true_graph(data, case), Oracle(case)


@pytest.mark.parametrize("mode", [native.SL_MSAMPLER_SPECULATIVE, native.SL_MSAMPLER_TWO_LAUNCH])
def test_msweep_long_map_matches_oracle(long_map, mode):
    case, g0, o = long_map
    assert case.M == 5000
    starts = np.stack([g0, g0])
    with native.Context(case, native.SL_FP64, msampler_mode=mode) as ctx:
        ctx.mcmc_init(starts)
        ctx.msweep(606, 2)
        out = ctx.mcmc_graphs()
        assert np.all(ctx.mcmc_status() == 0)
        near = ctx.mcmc_near_edge()
    assert np.all(near == 0)
    for c in range(2):
        assert np.array_equal(out[c], o.msweep(starts[c], shuffled_order(o, 606, c, 2), 606, c, 2)), c


@pytest.fixture(scope="module")
def ped1100(tmp_path_factory):
    """1100 people (301 founders) over 6 SNPs: 1541 non-ignorable meioses, more
    than one pass of the shuffle's swap partners (1024 per pass), and
    founder-allele graphs in LDS (over 1024 people)."""
    d = tmp_path_factory.mktemp("ped1100")
    case, data = synth.make_case("ped600_snp20000", str(d), n_people=1100, n_markers=6, n_generations=10,
                                 n_loops=6)
    return case, synth.true_graph(data, case), Oracle(case)


def test_msweep_shuffle_over_1024_meioses(ped1100):
    case, g0, o = ped1100
    assert len(o.meioses()) > 1024
    starts = np.stack([g0, g0])
    print(f"ped1100: N={case.N} F={case.F} meioses={len(o.meioses())}", flush=True)
    with native.Context(case, native.SL_FP64) as ctx:
        ctx.mcmc_init(starts)
        ctx.msweep(99, 3)
        out = ctx.mcmc_graphs()
        assert np.all(ctx.mcmc_status() == 0)
        near = ctx.mcmc_near_edge()
    assert np.all(near == 0)
    for c in range(2):
        assert np.array_equal(out[c], o.msweep(starts[c], shuffled_order(o, 99, c, 3), 99, c, 3)), c


def _fp32_vs_fp64_steps(case, g0, o, mode, steps, seed):
    """One M-step at a time (sl_msweep with a one-meiosis order) in an FP64 and
    an FP32-rescaled context from the same graph: the FP32 forward-backward
    pass consumes the same draws and must give the FP64 pass's graph whenever
    it counted no draw within 1e-5 of its threshold.  Returns the number of
    steps with such a draw."""
    meioses = o.meioses()
    g = g0
    flagged = 0
    with native.Context(case, native.SL_FP64, msampler_mode=mode) as c64, \
            native.Context(case, native.SL_FP32_RESCALED, msampler_mode=mode) as c32:
        for j in range(steps):
            m = meioses[(7 * j) % len(meioses)]
            outs = []
            for ctx in (c64, c32):
                ctx.mcmc_init(g[None])
                ctx.msweep(seed, j, [m])
                assert np.all(ctx.mcmc_status() == 0)
                outs.append((ctx.mcmc_graphs()[0], int(ctx.mcmc_near_edge()[0])))
            (g64, n64), (g32, n32) = outs
            assert n64 == 0
            if n32 == 0:
                assert np.array_equal(g32, g64), (j, m)
            else:
                flagged += 1
            g = g64
    return flagged


@pytest.mark.parametrize("mode", [native.SL_MSAMPLER_SPECULATIVE, native.SL_MSAMPLER_TWO_LAUNCH])
def test_msweep_fp32_equals_fp64_but_counted_edges_ped150(ped150, mode):
    case, g0, o = ped150
    flagged = _fp32_vs_fp64_steps(case, g0, o, mode, 60, 3141)
    print(f"ped150 FP32 M-steps with a draw near an edge: {flagged} of 60")
    assert flagged <= 6


def test_msweep_fp32_long_map(long_map):
    """5000 loci: the FP32 pass through the global-memory segments (one
    launch per step) against FP64."""
    case, g0, o = long_map
    flagged = _fp32_vs_fp64_steps(case, g0, o, native.SL_MSAMPLER_SPECULATIVE, 12, 2718)
    assert flagged <= 2


def test_msweep_fp32_whole_sweep_keeps_graph_legal(ped150):
    case, g0, o = ped150
    with native.Context(case, native.SL_FP32_RESCALED) as ctx:
        ctx.mcmc_init(np.stack([g0, g0]))
        for s in range(3):
            ctx.msweep(77, s)
        out = ctx.mcmc_graphs()
        assert np.all(ctx.mcmc_status() == 0)
    for c in range(2):
        assert min(o.fag(out[c], l) for l in range(case.M)) > 0
