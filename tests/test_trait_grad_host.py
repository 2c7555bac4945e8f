"""CPU pins of the trait-peel gradient: the library's disease-model tables
against the reference's own dumped disease probabilities and against finite
differences, and the oracle's forward-mode gradient (the FP64 reference the
GPU adjoint is compared with) against central finite differences of the
oracle's trait peel."""
import dataclasses

import numpy as np
import pytest

from conftest import load_golden, GOLDEN_CASES
from oracle_lib import Oracle
from swiftlink_amd import native
from swiftlink_amd.linkage import disease_probs

RECESSIVE = ((0.0, 0.0, 1.0), 1e-4)        # the golden cases' model
MODEL = ((0.05, 0.3, 0.85), 0.12)           # no zero disease probability
H = 1e-9                                     # central-difference step (absolute, for a zero parameter)


def step(x):
    # relative steps: the likelihood is a high-degree polynomial in each parameter
    return 1e-5 * abs(x) if x != 0 else H
FD_RTOL = 1e-5                               # gradient relative-error bar


def tables(case, pen, freq):
    return native.trait_model_tables(pen, freq, case.affection, case.F)


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_model_tables_reproduce_reference_disease_probs(name):
    case, _ = load_golden(name)
    dp, _ = tables(case, *RECESSIVE)
    assert np.array_equal(dp, case.disease_prob.reshape(case.N, 4))


@pytest.mark.parametrize("pen,freq", [MODEL, RECESSIVE, ((0.2, 0.2, 0.2), 0.5)])
def test_model_tables_match_linkage_and_finite_differences(pen, freq):
    case, _ = load_golden("ped30_snp100")
    dp, jac = tables(case, pen, freq)
    ref = np.array([disease_probs(int(case.affection[i]), i < case.F, freq, pen) for i in range(case.N)])
    assert np.array_equal(dp, ref)
    params = list(pen) + [freq]
    for k in range(4):
        up, dn = list(params), list(params)
        up[k] += H
        dn[k] -= H
        fu = np.array([disease_probs(int(case.affection[i]), i < case.F, up[3], up[:3]) for i in range(case.N)])
        fd = np.array([disease_probs(int(case.affection[i]), i < case.F, dn[3], dn[:3]) for i in range(case.N)])
        assert np.allclose(jac[:, :, k], (fu - fd) / (2 * H), rtol=1e-6, atol=1e-9), k


def test_model_tables_reject_bad_models():
    case, _ = load_golden("loop")
    for pen, freq in [((0.0, 0.0, 1.5), 0.1), ((0.1, 0.2, 0.3), 0.0), ((0.1, 0.2, 0.3), 1.0)]:
        with pytest.raises(native.SlError):
            tables(case, pen, freq)


def with_model(case, pen, freq):
    dp = np.array([disease_probs(int(case.affection[i]), i < case.F, freq, pen) for i in range(case.N)])
    return dataclasses.replace(case, disease_prob=dp.reshape(case.disease_prob.shape))


def with_ptheta(case, scale):
    return dataclasses.replace(case, partial_theta=case.partial_theta * scale)


@pytest.mark.parametrize("name", GOLDEN_CASES)
@pytest.mark.parametrize("model", [MODEL, RECESSIVE])
def test_oracle_gradient_matches_finite_differences(name, model):
    case, fx = load_golden(name)
    pen, freq = model
    c0 = with_model(case, pen, freq)
    _, jac = tables(case, pen, freq)
    o = Oracle(c0)
    g = fx["graphs"][0]
    params = list(pen) + [freq]
    rng = np.random.default_rng(3)
    for locus in rng.choice(case.M - 1, size=min(3, case.M - 1), replace=False):
        for offset in (1, case.nlod):
            out = o.trait_grad(g, int(locus), offset, jac)
            # the value is the reference peel's, bit for bit
            assert out[0] == o.trait_likelihood(g, int(locus), offset)
            assert out[0] > 0
            # penetrances and frequency
            for k in range(4):
                up, dn = list(params), list(params)
                h = step(params[k])
                up[k] += h
                dn[k] -= h
                lu = Oracle(with_model(case, up[:3], up[3])).trait_likelihood(g, int(locus), offset)
                ld = Oracle(with_model(case, dn[:3], dn[3])).trait_likelihood(g, int(locus), offset)
                fd = (lu - ld) / (2 * h)
                assert abs(out[1 + k] - fd) <= FD_RTOL * max(abs(fd), 1e-3 * out[0]), (k, out[1 + k], fd)
            # theta_L = offset p, theta_R = (n + 1 - offset) p: perturb p by a
            # relative step (L is a high-degree polynomial in a small theta)
            hp = 1e-5 * case.partial_theta[locus]
            lu = Oracle(with_ptheta(c0, 1 + 1e-5)).trait_likelihood(g, int(locus), offset)
            ld = Oracle(with_ptheta(c0, 1 - 1e-5)).trait_likelihood(g, int(locus), offset)
            fd = (lu - ld) / (2 * hp)
            comb = offset * out[5] + (case.nlod + 1 - offset) * out[6]
            assert abs(comb - fd) <= FD_RTOL * max(abs(fd), 1e-3 * out[0]), (comb, fd)


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_oracle_trait_prob_gradient_matches_finite_differences(name):
    case, _ = load_golden(name)
    pen, freq = MODEL
    _, jac = tables(case, pen, freq)
    out = Oracle(with_model(case, pen, freq)).trait_grad(None, 0, 1, jac)
    assert out[5] == 0.0 and out[6] == 0.0
    params = list(pen) + [freq]
    for k in range(4):
        up, dn = list(params), list(params)
        h = step(params[k])
        up[k] += h
        dn[k] -= h
        lu = Oracle(with_model(case, up[:3], up[3])).trait_likelihood(None, 0, 1)
        ld = Oracle(with_model(case, dn[:3], dn[3])).trait_likelihood(None, 0, 1)
        fd = (lu - ld) / (2 * h)
        assert abs(out[1 + k] - fd) <= FD_RTOL * max(abs(fd), 1e-3 * out[0]), (k, out[1 + k], fd)
