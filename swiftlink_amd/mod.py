"""MOD-score maximisation over the trait model with the device adjoint.

The sampled descent graphs depend on the marker data only (the samplers'
trait caches are marker priors), so the LOD of every trait model can be
scored on the same graphs.  `maximise` climbs the LOD at its best position
over (penetrance[0..2], disease allele frequency) with the gradient from
sl_lod_grad: projected gradient ascent in logit coordinates (scale-free for
a frequency of 1e-4 beside penetrances near 1) with a backtracking line
search, so every accepted step raises the MOD.  The reference fixes the model
at start-up and reports the LOD only (LinkageProgram, linkage_program.cc);
this is the MOD-score extension the adjoint exists for.
"""
from __future__ import annotations

import dataclasses
from typing import List, Sequence, Tuple

import numpy as np

EPS = 1e-9  # parameters are kept in [EPS, 1 - EPS]


@dataclasses.dataclass
class ModResult:
    mod: float                      # highest LOD found
    position: Tuple[int, int]       # (interval, position) it sits at
    penetrance: Tuple[float, float, float]
    disease_freq: float
    history: List[float]            # MOD after each accepted step (first entry: the start model)
    evaluations: int                # sl_lod_grad calls


def _logit(x):
    x = np.clip(x, EPS, 1 - EPS)
    return np.log(x) - np.log1p(-x)


def _sigmoid(z):
    return np.clip(1.0 / (1.0 + np.exp(-z)), EPS, 1 - EPS)


def maximise(ctx, dev_graphs: int, n_graphs: int, penetrance: Sequence[float], disease_freq: float,
             steps: int = 30, step0: float = 1.0, tol: float = 1e-7, affection=None) -> ModResult:
    """Projected gradient ascent of max over positions of the LOD, from the
    given model; ctx is a swiftlink_amd.native.Context and dev_graphs a device
    pointer to n_graphs packed graphs.  Leaves ctx on the best model."""
    x = np.array(list(penetrance) + [disease_freq], dtype=np.float64)
    evals = 0

    def score(params):
        nonlocal evals
        ctx.set_trait_model(params[:3], params[3], affection)
        lod, dlod = ctx.lod_grad_packed(dev_graphs, n_graphs)
        evals += 1
        flat = lod.reshape(-1)
        i = int(np.argmax(flat))
        return flat[i], np.unravel_index(i, lod.shape), dlod.reshape(-1, 6)[i, :4]

    best, pos, g = score(x)
    history = [float(best)]
    t = step0
    for _ in range(steps):
        z = _logit(x)
        gz = g * x * (1 - x)  # d LOD / d logit
        if np.linalg.norm(gz) < tol:
            break
        improved = False
        while t > 1e-8:
            xn = _sigmoid(z + t * gz)
            v, p2, g2 = score(xn)
            if v > best:
                x, best, pos, g = xn, v, p2, g2
                history.append(float(best))
                t *= 2.0
                improved = True
                break
            t *= 0.5
        if not improved:
            break
    ctx.set_trait_model(x[:3], x[3], affection)
    return ModResult(float(best), (int(pos[0]), int(pos[1])), (float(x[0]), float(x[1]), float(x[2])), float(x[3]),
                     history, evals)
