"""Autotuner for the launch shapes of one context.

For a context holding its chains, times on the device:
  * the LOD scoring of the chains' graphs for every LOD workgroup shape
    (block_threads x graphs_per_block): the block size sets how many
    wavefronts share a peel level, graphs_per_block how many graphs one
    workgroup folds (its LDS arena is reused across them); and the lane
    kernel, one (graph, position) evaluation per lane;
  * one M-sweep of every chain in each M-sampler launch mode (one launch per
    step with speculative likelihoods, or two launches per step).
The chains' graphs are restored afterwards (and the LOD accumulators
cleared), so the chains continue as they were.  Timing uses the library's
HIP-event profiling and the host clock.  The shapes
change through sl_set_options; results are identical for every shape.
"""
from __future__ import annotations

import itertools
from typing import Dict, Sequence

import numpy as np

from . import native

LOD_BLOCKS = (64, 128, 256, 512)
LOD_GPW = (1, 2, 4, 8)
M_MODES = (native.SL_MSAMPLER_SPECULATIVE, native.SL_MSAMPLER_TWO_LAUNCH)


def _lod_ms(ctx, graphs: np.ndarray, reps: int) -> float:
    # the events bracket the peel launch only, not the graphs' upload
    ctx.lod_accumulate(graphs)  # warm
    ctx.sync()
    ctx.profile(True)
    for _ in range(reps):
        ctx.lod_accumulate(graphs)
    ms, n = ctx.profile_read(native.SL_KERNEL_LOD)
    ctx.profile(False)
    return ms / max(n, 1)


def _msweep_ms(ctx, seed: int, sweeps: int) -> float:
    import time
    ctx.msweep(seed, 0)  # warm
    ctx.sync()
    t0 = time.perf_counter()
    for s in range(1, sweeps + 1):
        ctx.msweep(seed, s)
    ctx.sync()
    return (time.perf_counter() - t0) * 1e3 / sweeps


def autotune(ctx: "native.Context", blocks: Sequence[int] = LOD_BLOCKS, gpws: Sequence[int] = LOD_GPW,
             m_modes: Sequence[int] = M_MODES, lod_reps: int = 3, m_sweeps: int = 2) -> Dict:
    """Picks and applies the fastest shapes for ctx's chains; returns the
    choice and every timing (ms).  The LOD candidates are the workgroup
    kernel at every (block, graphs per block) and the lane kernel (one
    evaluation per lane), which serves any chain count at which it fits."""
    graphs = ctx.mcmc_graphs()
    lod = {}
    cands = [(b, g, native.SL_LOD_WORKGROUP) for b, g in itertools.product(blocks, gpws)]
    cands.append((ctx.options["block"] or 128, ctx.options["gpw"] or 1, native.SL_LOD_LANES))
    for b, g, k in cands:
        try:
            ctx.set_options(b, g, ctx.options["msampler_mode"], k)
        except native.SlError:
            continue  # e.g. a block the LDS arena cannot hold
        lod[(b, g, k)] = _lod_ms(ctx, graphs, lod_reps)
    best_lod = min(lod, key=lod.get)
    msw = {}
    for m in m_modes:
        ctx.set_options(best_lod[0], best_lod[1], m, best_lod[2])
        try:
            msw[m] = _msweep_ms(ctx, 1, m_sweeps)
        except native.SlError:
            continue
    best_m = min(msw, key=msw.get) if msw else native.SL_MSAMPLER_AUTO
    ctx.set_options(best_lod[0], best_lod[1], best_m, best_lod[2])
    ctx.mcmc_init(graphs)  # the chains as they were
    ctx.lod_reset()
    name = lambda b, g, k: "lanes" if k == native.SL_LOD_LANES else f"{b}x{g}"
    return {"block": best_lod[0], "gpw": best_lod[1], "msampler_mode": best_m, "lod_kernel": best_lod[2],
            "lod_ms": {name(*key): v for key, v in lod.items()}, "msweep_ms": {int(k): v for k, v in msw.items()}}
