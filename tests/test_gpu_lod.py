"""LOD trait peel on the GPU against the oracle (parity tests proper).

FP64: per-graph terms within 1e-12 relative of the oracle (same operation
order; only log and the recombination sum differ in rounding), accumulated
LOD within 1e-6 absolute.  FP32 with per-table exponents: |dLOD| <= 1e-4.
"""
import numpy as np
import pytest

from conftest import load_golden, GOLDEN_CASES
from oracle_lib import Oracle
from swiftlink_amd import native, synth

pytestmark = pytest.mark.gpu

FP32_LOD_TOL = 1e-4   # |LOD_fp32 - LOD_fp64 oracle|
FP64_LOD_TOL = 1e-6


def _lod(acc, count, trait_prob):
    return (acc - np.log(count) - trait_prob) / np.log(10.0)


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_trait_prob_fp64(name):
    case, fx = load_golden(name)
    with native.Context(case, native.SL_FP64) as ctx:
        np.testing.assert_allclose(ctx.trait_prob(), case.trait_prob, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_trait_prob_fp32(name):
    case, fx = load_golden(name)
    with native.Context(case, native.SL_FP32_RESCALED) as ctx:
        assert abs(ctx.trait_prob() - case.trait_prob) < 1e-4


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_terms_fp64_match_reference(name):
    case, fx = load_golden(name)
    with native.Context(case, native.SL_FP64) as ctx:
        terms = ctx.lod_terms(fx["graphs"])
    np.testing.assert_allclose(terms, fx["lod_terms"], rtol=1e-12, atol=1e-10)


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_terms_fp32(name):
    case, fx = load_golden(name)
    with native.Context(case, native.SL_FP32_RESCALED) as ctx:
        terms = ctx.lod_terms(fx["graphs"])
    assert np.abs(terms - fx["lod_terms"]).max() / np.log(10.0) < FP32_LOD_TOL


@pytest.mark.parametrize("precision,tol", [(native.SL_FP64, FP64_LOD_TOL), (native.SL_FP32_RESCALED, FP32_LOD_TOL)])
@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_accumulate_matches_reference(name, precision, tol):
    case, fx = load_golden(name)
    with native.Context(case, precision, gpw=3) as ctx:
        ctx.lod_accumulate(fx["graphs"][:3])
        ctx.lod_accumulate(fx["graphs"][3:])
        acc, count, lod = ctx.lod_results()
    assert count == len(fx["graphs"])
    assert np.abs(lod - fx["lod_final"]).max() < tol
    assert np.abs(acc - fx["lod_acc"]).max() < tol * np.log(10.0)


def test_reset_and_empty():
    case, fx = load_golden("loop")
    with native.Context(case, native.SL_FP64) as ctx:
        ctx.lod_accumulate(fx["graphs"])
        ctx.lod_reset()
        acc, count, _ = ctx.lod_results()
        assert count == 0 and np.all(np.isneginf(acc))
        ctx.lod_accumulate(fx["graphs"][:0])
        ctx.lod_accumulate(fx["graphs"][:1])
        acc, count, lod = ctx.lod_results()
    np.testing.assert_allclose(acc, fx["lod_terms"][0].reshape(-1), rtol=1e-12)


def test_merge_of_many_graphs_matches_small_calls():
    """600 graphs in one call fold through the workgroup-per-accumulator
    merge (256 chunks or more); the same graphs in calls of 100 through the
    thread-per-accumulator merge.  Both are log-sums of the same terms, so
    they agree to rounding, and with the sum of the oracle's per-graph terms."""
    case, fx = load_golden("east")
    reps = np.concatenate([fx["graphs"]] * (600 // len(fx["graphs"]) + 1))[:600]
    with native.Context(case, native.SL_FP64) as big, native.Context(case, native.SL_FP64) as small:
        big.lod_accumulate(reps)
        for k in range(0, 600, 100):
            small.lod_accumulate(reps[k:k + 100])
        acc_b, n_b, _ = big.lod_results()
        acc_s, n_s, _ = small.lod_results()
        terms = big.lod_terms(fx["graphs"])
    assert n_b == n_s == 600
    np.testing.assert_allclose(acc_b, acc_s, rtol=1e-12, atol=1e-12)
    per = np.asarray(terms).reshape(len(fx["graphs"]), -1)
    idx = np.arange(600) % len(fx["graphs"])
    m = per.max(axis=0)
    want = m + np.log(np.exp(per[idx] - m).sum(axis=0))
    np.testing.assert_allclose(acc_b, want, rtol=1e-12, atol=1e-12)


@pytest.fixture(scope="module")
def ped150(tmp_path_factory):
    d = tmp_path_factory.mktemp("ped150")
    case, data = synth.make_case("ped150_snp1000", str(d))
    g0 = synth.true_graph(data, case)
    o = Oracle(case)
    graphs = [g0]
    g = g0
    for s in range(3):  # a few oracle L-sweeps to get distinct legal graphs
        for l in list(range(0, case.M, 2)) + list(range(1, case.M, 2)):
            g, _ = o.lstep(g, l, 11, 0, s)
        graphs.append(g)
    return case, np.stack(graphs), o


def test_ped150_fp64_terms(ped150):
    """150 people, 3 loops, 1000 SNPs: spot intervals against the oracle."""
    case, graphs, o = ped150
    with native.Context(case, native.SL_FP64) as ctx:
        terms = ctx.lod_terms(graphs)
    for gi in range(len(graphs)):
        for l in (0, 1, 499, 500, case.M - 2):
            np.testing.assert_allclose(terms[gi, l], o.lod_terms(graphs[gi], l), rtol=1e-12)


def test_ped150_fp32_lod_within_tolerance(ped150):
    case, graphs, o = ped150
    ref = o.lod_accumulate(graphs)
    tp = o.trait_prob()
    lod_ref = _lod(ref, len(graphs), tp)
    with native.Context(case, native.SL_FP32_RESCALED) as ctx:
        ctx.lod_accumulate(graphs)
        acc, count, lod = ctx.lod_results()
    assert np.abs(lod - lod_ref).max() < FP32_LOD_TOL
    with native.Context(case, native.SL_FP64) as ctx:
        ctx.lod_accumulate(graphs)
        acc, count, lod = ctx.lod_results()
    assert np.abs(lod - lod_ref).max() < FP64_LOD_TOL


def _table_fields(text):
    return [line.split("\t") for line in text.splitlines()]


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_lod_table_from_device_scores(name):
    """Two pedigrees' LOD scores from the device (one context each, the golden
    graphs split in two) written as the reference's LinkageWriter table: the
    same rows and positions, LODs within the FP64 tolerance."""
    from swiftlink_amd import linkage
    case, fx = load_golden(name)
    G = len(fx["graphs"])
    lods = []
    for part in (fx["graphs"][:G // 2], fx["graphs"][G // 2:]):
        with native.Context(case, native.SL_FP64) as ctx:
            ctx.lod_accumulate(part)
            lods.append(ctx.lod_results()[2].reshape(case.M - 1, case.nlod))
    text = linkage.format_lod_table(list(fx["writer_names"]), fx["writer_gpos"], case.partial_theta, lods)
    mine, ref = _table_fields(text), _table_fields(str(fx["writer_text2"][0]))
    assert len(mine) == len(ref) and mine[0] == ref[0]  # the header row
    for a, b in zip(mine[1:], ref[1:]):
        assert len(a) == len(b) and a[:2] == b[:2], (a, b)
        for x, y in zip(a[2:], b[2:]):
            assert abs(float(x) - float(y)) <= 1e-5 * max(1.0, abs(float(y))), (a, b)


def _many(fx, n):
    reps = -(-n // len(fx["graphs"]))
    return np.concatenate([fx["graphs"]] * reps)[:n]


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_lane_kernel_fp32_terms_near_workgroup_kernel(name):
    """The lane kernel (one (graph, position) per lane, 70 graphs: one full
    and one partial block of 64) is an FP32 kernel: its terms agree with the
    FP32 workgroup kernel's to FP32 rounding (the two renormalise the tables
    at different points), and both within the LOD bar of the reference's.
    Asked for FP64, the lane option runs the workgroup kernel: the terms
    equal it bit for bit and match the reference to 1e-12."""
    case, fx = load_golden(name)
    graphs = _many(fx, 70)
    with native.Context(case, native.SL_FP32_RESCALED, lod_kernel=native.SL_LOD_LANES) as a, \
            native.Context(case, native.SL_FP32_RESCALED, lod_kernel=native.SL_LOD_WORKGROUP) as b:
        ta, tb = a.lod_terms(graphs), b.lod_terms(graphs)
    assert np.abs(ta - tb).max() <= 1e-4 * max(1.0, np.abs(tb).max())
    assert np.abs(ta[:len(fx["graphs"])] - fx["lod_terms"]).max() / np.log(10.0) < FP32_LOD_TOL
    with native.Context(case, native.SL_FP64, lod_kernel=native.SL_LOD_LANES) as a, \
            native.Context(case, native.SL_FP64, lod_kernel=native.SL_LOD_WORKGROUP) as b:
        ta, tb = a.lod_terms(graphs), b.lod_terms(graphs)
    assert np.array_equal(ta, tb)
    np.testing.assert_allclose(ta[:len(fx["graphs"])], fx["lod_terms"], rtol=1e-12, atol=1e-10)


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_lane_kernel_fp32_and_accumulators(name):
    """FP32 lane kernel: terms within the LOD bar of the reference's, and the
    accumulated LOD of 70 graphs within the bar of the FP64 workgroup
    kernel's; FP64 accumulators under either option agree to rounding of the
    log-sum order."""
    case, fx = load_golden(name)
    graphs = _many(fx, 70)
    with native.Context(case, native.SL_FP32_RESCALED, lod_kernel=native.SL_LOD_LANES) as ctx:
        t32 = ctx.lod_terms(fx["graphs"])
        ctx.lod_accumulate(graphs)
        _, c32, lod32 = ctx.lod_results()
    assert np.abs(t32 - fx["lod_terms"]).max() / np.log(10.0) < FP32_LOD_TOL
    accs = []
    for kernel in (native.SL_LOD_LANES, native.SL_LOD_WORKGROUP):
        with native.Context(case, native.SL_FP64, lod_kernel=kernel) as ctx:
            ctx.lod_accumulate(graphs)
            acc, c64, lod64 = ctx.lod_results()
            accs.append((acc, lod64))
    assert c32 == c64 == 70
    np.testing.assert_allclose(accs[0][0], accs[1][0], rtol=1e-13, atol=1e-12)
    assert np.abs(lod32 - accs[1][1]).max() < FP32_LOD_TOL


def test_lane_kernel_ped150_fp32_and_fp64(ped150):
    """150 people, 96 graphs (one and a half blocks): the FP64 lane option
    (the workgroup kernel) equals the workgroup kernel; FP32 lane terms within
    the bar of the oracle's at every 50th interval."""
    case, legal, oracle = ped150
    rng = np.random.default_rng(150)
    graphs = np.concatenate([legal] * 24)
    nm = 2 * (case.N - case.F)
    for g in graphs[len(legal):]:
        g[rng.integers(0, case.M, 30), 2 * case.F + rng.integers(0, nm, 30)] ^= 1
    with native.Context(case, native.SL_FP64, lod_kernel=native.SL_LOD_LANES) as a, \
            native.Context(case, native.SL_FP64, lod_kernel=native.SL_LOD_WORKGROUP) as b:
        assert np.array_equal(a.lod_terms(graphs), b.lod_terms(graphs))
    with native.Context(case, native.SL_FP32_RESCALED, lod_kernel=native.SL_LOD_LANES) as ctx:
        t32 = ctx.lod_terms(graphs[:8])
    for gi in range(8):
        for l in range(0, case.M - 1, 50):
            assert np.abs(t32[gi, l] - oracle.lod_terms(graphs[gi], l)).max() / np.log(10.0) < FP32_LOD_TOL
