"""ELOD on the device (sl_elod: start_from replicates through the L-sampler
kernel, their marker columns scored by the trait peel) against the compiled
reference's Elod::run and the oracle.

FP64: the ELOD, accumulator and replicate count match the reference's to a
relative 1e-12 (the device log-sums the replicates' terms in another order);
past one launch's worth of replicates (65535) every replicate still draws
from its own stream, checked against the oracle.  FP32 sampling: a replicate
can differ from FP64 only where a uniform falls within float rounding of a
CDF edge, so the ELOD agrees to 1e-3.
"""
import numpy as np
import pytest

from conftest import load_golden, GOLDEN_CASES
from oracle_lib import Oracle
from swiftlink_amd import elod, native

pytestmark = pytest.mark.gpu


def _params(fx):
    freq, p0, p1, p2, sep = (float(x) for x in fx["elod_params"])
    R, seed = (int(x) for x in fx["elod_run"])
    return freq, (p0, p1, p2), sep, R, seed


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_elod_matches_reference(name):
    case, fx = load_golden(name)
    freq, pen, sep, R, seed = _params(fx)
    c3, c2 = elod.elod_cases(case, freq, pen, sep)
    with native.Context(c3, native.SL_FP64) as s, native.Context(c2, native.SL_FP64) as sc:
        value = native.run_elod(s, sc, seed, R)
        acc, count, lod = sc.lod_results()
    assert count == R == int(fx["elod_count"][0])
    assert abs(acc[0] - fx["elod_acc"][0]) <= 1e-12 * max(1.0, abs(fx["elod_acc"][0]))
    assert abs(value - fx["elod_value"][0]) <= 1e-12 * max(1.0, abs(fx["elod_value"][0]))
    assert value == lod[0]


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_elod_other_disease_model_matches_reference(name):
    """Penetrance 0.01:0.5:0.9, frequency 0.05, markers 0.1 M apart."""
    case, fx = load_golden(name)
    freq, p0, p1, p2, sep = (float(x) for x in fx["elod2_params"])
    R, seed = (int(x) for x in fx["elod2_run"])
    c3, c2 = elod.elod_cases(case, freq, (p0, p1, p2), sep)
    with native.Context(c3, native.SL_FP64) as s, native.Context(c2, native.SL_FP64) as sc:
        value = native.run_elod(s, sc, seed, R)
        acc, count, _ = sc.lod_results()
    assert count == R
    assert abs(acc[0] - fx["elod2_acc"][0]) <= 1e-12 * max(1.0, abs(fx["elod2_acc"][0]))
    assert abs(value - fx["elod2_value"][0]) <= 1e-12 * max(1.0, abs(fx["elod2_value"][0]))


def test_elod_total_over_pedigrees():
    """elod.elod sums the pedigrees' ELODs (elod.cc:72-83)."""
    cases = [load_golden(n) for n in ("loop", "east")]
    freq, pen, sep, R, seed = _params(cases[0][1])
    res = elod.elod([c for c, _ in cases], freq, pen, sep, replicates=R, seed=seed)
    want = [float(fx["elod_value"][0]) for _, fx in cases]
    np.testing.assert_allclose(res.per_pedigree, want, rtol=1e-12, atol=1e-12)
    assert abs(res.total - sum(want)) <= 1e-11


def test_elod_past_one_launch_matches_oracle():
    """70000 replicates: two launches of the sampler, replicate j drawing as
    chain j in both; the oracle runs every replicate."""
    case, fx = load_golden("loop")
    R, seed = 70000, 8
    c3, c2 = elod.elod_cases(case)
    o3, o2 = Oracle(c3), Oracle(c2)
    graphs = np.stack([o3.impute_from(seed, j, 1)[0] for j in range(R)])
    acc = o2.lod_accumulate(graphs[:, [0, 2]])[0]
    want = (acc - np.log(R) - o2.trait_prob()) / np.log(10.0)
    with native.Context(c3, native.SL_FP64) as s, native.Context(c2, native.SL_FP64) as sc:
        value = native.run_elod(s, sc, seed, R)
        dacc, count, _ = sc.lod_results()
    assert count == R
    assert abs(dacc[0] - acc) <= 1e-12 * max(1.0, abs(acc))
    assert abs(value - want) <= 1e-11 * max(1.0, abs(want))


def test_elod_fp32_sampler():
    case, fx = load_golden("east")
    freq, pen, sep, _, seed = _params(fx)
    r64 = elod.elod([case], freq, pen, sep, replicates=20000, seed=seed)
    r32 = elod.elod([case], freq, pen, sep, replicates=20000, seed=seed, precision=native.SL_FP32_RESCALED)
    assert abs(r32.total - r64.total) <= 1e-3 * max(1.0, abs(r64.total))


def test_elod_rejects_mismatched_contexts():
    case, fx = load_golden("loop")
    c3, c2 = elod.elod_cases(case)
    with native.Context(c3, native.SL_FP64) as s, native.Context(c2, native.SL_FP64) as sc:
        with pytest.raises(native.SlError):
            native.run_elod(sc, s, 1, 10)  # sampler and scorer swapped
        with pytest.raises(native.SlError):
            native.run_elod(s, sc, 1, 0)


def test_elod_ped150_fp64_matches_oracle_and_fp32_close(tmp_path):
    """ped150 (150 people, 3 loops) under the default ELOD model: 3000
    replicates in FP64 against the oracle's start_from replicates scored by
    its trait peel, and the FP32-sampled ELOD against that FP64 value.  An
    FP32 replicate differs from its FP64 twin only where a uniform falls
    within float rounding of a CDF edge, so the bound is 1e-3 absolute on an
    ELOD of order 1."""
    from swiftlink_amd import synth
    case, _ = synth.make_case("ped150_snp1000", str(tmp_path))
    R, seed = 3000, 21
    c3, c2 = elod.elod_cases(case)
    o3, o2 = Oracle(c3), Oracle(c2)
    graphs = np.stack([o3.impute_from(seed, j, 1)[0] for j in range(R)])
    acc = o2.lod_accumulate(graphs[:, [0, 2]])[0]
    want = (acc - np.log(R) - o2.trait_prob()) / np.log(10.0)
    with native.Context(c3, native.SL_FP64) as s, native.Context(c2, native.SL_FP64) as sc:
        v64 = native.run_elod(s, sc, seed, R)
        dacc, count, _ = sc.lod_results()
    assert count == R
    assert abs(dacc[0] - acc) <= 1e-12 * max(1.0, abs(acc))
    assert abs(v64 - want) <= 1e-11 * max(1.0, abs(want))
    with native.Context(c3, native.SL_FP32_RESCALED) as s, native.Context(c2, native.SL_FP64) as sc:
        v32 = native.run_elod(s, sc, seed, R)
    assert abs(v32 - v64) <= 1e-3, (v32, v64)
