"""The allocation-free LOD entry point (sl_lod_accumulate_packed_noalloc)
against sl_lod_accumulate_packed, through the C ABI: two calls in a row, the
second with fewer graphs, in FP64 and FP32-rescaled, on a case whose graphs
fit one group of the fixed scratch array and on one that takes several
groups; and the capacity error for a map whose partials exceed the array."""
import tempfile

import numpy as np
import pytest

from conftest import load_golden
from swiftlink_amd import native, synth
from swiftlink_amd.case import Case, pack_graphs

pytestmark = pytest.mark.gpu


def _device(case, graphs):
    import torch
    return torch.from_numpy(pack_graphs(case, graphs).reshape(-1).view(np.int32)).cuda()


def _scored(ctx, dev, n, noalloc):
    ctx.lod_reset()
    if noalloc:
        ctx.lod_accumulate_packed_noalloc(dev.data_ptr(), n)
    else:
        ctx.lod_accumulate_packed(dev.data_ptr(), n)
    acc, cnt, _ = ctx.lod_results()
    return acc, cnt


def _check_pair(case, graphs, sizes, precision, exact):
    dev = _device(case, graphs)
    with native.Context(case, precision) as ctx:
        for n in sizes:  # the second call is smaller than the first
            print(f"noalloc: N={case.N} M={case.M} precision={precision} graphs={n}", flush=True)
            ref, cref = _scored(ctx, dev, n, False)
            got, cgot = _scored(ctx, dev, n, True)
            assert cgot == cref == n
            assert np.all(np.isfinite(ref) == np.isfinite(got))
            f = np.isfinite(ref)
            if exact:
                assert np.array_equal(ref[f], got[f])
            else:  # the one-call path merges >= 256 chunks by a tree, the groups in sequence
                np.testing.assert_allclose(got[f], ref[f], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("precision", [native.SL_FP64, native.SL_FP32_RESCALED])
def test_noalloc_matches_accumulate_one_group(precision):
    case, fx = load_golden("ped30_snp100")
    reps = -(-40 // len(fx["graphs"]))
    graphs = np.concatenate([fx["graphs"]] * reps)[:40]
    _check_pair(case, graphs, (40, 7), precision, exact=True)


def test_noalloc_matches_accumulate_several_groups():
    # ped150: (M-1) x 5 positions x 8 B = 39960 B per chunk, 209 chunks per
    # group of the 8 MiB array, so 300 graphs take two groups
    case, data = synth.make_case("ped150_snp1000", tempfile.mkdtemp(prefix="noalloc_"))
    g0 = synth.true_graph(data, case)
    rng = np.random.default_rng(5)
    graphs = np.repeat(g0[None], 300, axis=0)
    # vary the graphs: flip random meiosis bits of non-founders
    nf = 2 * (case.N - case.F)
    for g in graphs[1:]:
        idx = rng.integers(0, nf, size=40)
        loci = rng.integers(0, case.M, size=40)
        g[loci, 2 * case.F + idx] ^= 1
    _check_pair(case, graphs, (300, 150), native.SL_FP64, exact=False)


def test_noalloc_rejects_partials_beyond_the_array():
    # a nuclear family on 70000 markers with 16 positions per interval:
    # (M-1) x 16 x 8 B = 8.96 MB exceed the 8 MiB array
    M, N, F = 70000, 3, 2
    mother = np.array([-1, -1, 0], np.int32)
    father = np.array([-1, -1, 1], np.int32)
    ops, _ = native.peel_sequence(mother, father, 100, 1, 1, threads=1)
    tiny = Case(
        N=N, F=F, M=M, nlod=16, mother=mother, father=father, sex=np.array([2, 1, 1], np.int32),
        affection=np.array([0, 0, 2], np.int32), typed=np.zeros(N, np.int32), geno=np.zeros((N, M), np.uint8),
        disease_prob=np.array([[0.98, 1e-4, 0.01, 0.01]] * 2 + [[0.0, 1.0, 0.0, 0.0]]),
        geno_prob=np.full((N, M, 4), 0.25), legal=np.full((N, M), 15, np.uint8),
        theta=np.full(M - 1, 0.01), partial_theta=np.full(M - 1, 0.01 / 17), minor=np.full(M, 0.3), ops=ops)
    dev = _device(tiny, np.zeros((1, M, 2 * N), np.int32))
    print("noalloc: capacity error on M=70000, 16 positions", flush=True)
    with native.Context(tiny, native.SL_FP64) as ctx:
        with pytest.raises(native.SlError, match="exceed the 8388608-byte scratch array"):
            ctx.lod_accumulate_packed_noalloc(dev.data_ptr(), 1)
        # the allocating entry point still scores it
        ctx.lod_accumulate_packed(dev.data_ptr(), 1)
        ctx.sync()
