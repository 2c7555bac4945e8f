"""Pins for the oracle that do not come from the reference's outputs:
known answers from an independent Philox implementation, brute-force
enumeration of the trait likelihood and the founder-allele-graph likelihood,
peel-order invariance, the closed form of the log-sum accumulator, and the
exact sampling distributions of the M- and L-samplers on tiny problems
(enumerated, then compared with oracle draws by a G-test)."""
import dataclasses
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest
from scipy import stats

from conftest import load_golden
from oracle_lib import Oracle, philox, uniform
import torch

PT_UU, PT_AA, PT_AU, PT_UA = 0, 1, 2, 3


def affected(pt, allele):
    # maternal allele is allele 0; AU = maternal affected (trait.h, rfunction.h)
    return {PT_UU: False, PT_AA: True, PT_AU: allele == 0, PT_UA: allele == 1}[pt]


def kid_of(m, p, i, j):
    ma, pa = affected(m, i), affected(p, j)
    return {(True, True): PT_AA, (True, False): PT_AU, (False, True): PT_UA, (False, False): PT_UU}[(ma, pa)]


KID = np.array([[[[kid_of(m, p, i, j) for j in range(2)] for i in range(2)] for p in range(4)] for m in range(4)])


# ---------------------------------------------------------------- Philox
def test_philox_matches_aten_engine():
    inc = os.path.join(os.path.dirname(torch.__file__), "include")
    if not shutil.which("g++") or not os.path.exists(os.path.join(inc, "ATen/core/PhiloxRNGEngine.h")):
        pytest.skip("g++ or ATen headers unavailable")
    cases = [((0, 0, 0, 0), (0, 0)), ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF)),
             ((7, 3, 11, 0xDEADBEEF), (0x12345678, 0x9ABCDEF0)), ((1, 0, 0, 0), (99, 0))]
    src = "#include <ATen/core/PhiloxRNGEngine.h>\n#include <cstdio>\nint main(){\n"
    for ctr, key in cases:
        seed = key[0] | (key[1] << 32)
        sub = ctr[2] | (ctr[3] << 32)
        off = ctr[0] | (ctr[1] << 32)
        src += (f"{{ at::Philox4_32 e({seed}ULL, {sub}ULL, 0); e.set_offset({off}ULL);"
                f" unsigned a=e(),b=e(),c=e(),d=e(); printf(\"%u %u %u %u\\n\",a,b,c,d); }}\n")
    src += "}\n"
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "k.cpp"), "w").write(src)
        subprocess.run(["g++", "-std=c++17", "-I", inc, "-o", os.path.join(d, "k"), os.path.join(d, "k.cpp")],
                       check=True, capture_output=True)
        out = subprocess.run([os.path.join(d, "k")], check=True, capture_output=True, text=True).stdout.split("\n")
    for (ctr, key), line in zip(cases, out):
        assert philox(ctr, key) == [int(x) for x in line.split()]


def test_uniform_range_and_mean():
    u = np.array([uniform(5, 0, s, 3, d) for s in range(50) for d in range(200)])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01
    assert stats.kstest(u, "uniform").pvalue > 1e-3


# ---------------------------------------------------------------- trait peel
def brute_trait_likelihood(case, dg, locus, offset):
    """Sum over all 4^N phased trait assignments of the product of disease
    probabilities and Mendelian transmissions with the flanking-theta
    recombination weights (the likelihood the trait peel factorises)."""
    N = case.N
    S = np.arange(4 ** N, dtype=np.int64)
    a = (S[:, None] >> (2 * np.arange(N))) & 3
    w = np.prod(case.disease_prob[np.arange(N), a], axis=1)
    if dg is not None:
        tl = case.partial_theta[locus] * offset
        tr = case.partial_theta[locus] * (case.nlod + 1 - offset)
    for c in range(case.F, N):
        m, p, kid = a[:, case.mother[c]], a[:, case.father[c]], a[:, c]
        s = np.zeros(len(S))
        for i in range(2):
            for j in range(2):
                r = 0.25
                if dg is not None:
                    b = [dg[locus, 2 * c], dg[locus + 1, 2 * c], dg[locus, 2 * c + 1], dg[locus + 1, 2 * c + 1]]
                    r *= (1 - tl if b[0] == i else tl) * (1 - tr if b[1] == i else tr)
                    r *= (1 - tl if b[2] == j else tl) * (1 - tr if b[3] == j else tr)
                s += (KID[m, p, i, j] == kid) * r
        w = w * s
    return w.sum()


def test_trait_peel_brute_force_loop():
    case, fx = load_golden("loop")  # 9 people, one marriage loop
    o = Oracle(case)
    np.testing.assert_allclose(o.trait_prob(), np.log(brute_trait_likelihood(case, None, 0, 1)), rtol=1e-13)
    for g in fx["graphs"][:2]:
        for l in range(case.M - 1):
            for off in range(1, case.nlod + 1):
                np.testing.assert_allclose(o.trait_likelihood(g, l, off), brute_trait_likelihood(case, g, l, off),
                                           rtol=1e-12)


def test_trait_peel_order_invariance():
    """Two different legitimate peel orders give the same likelihood."""
    from swiftlink_amd import peelseq
    case, fx = load_golden("ped30_snp100")
    o1 = Oracle(case)
    alt = dataclasses.replace(case, ops=peelseq.build(case.N, case.F, case.mother, case.father, seed=5))
    assert not np.array_equal(alt.ops.peel, case.ops.peel)
    o2 = Oracle(alt)
    np.testing.assert_allclose(o2.trait_prob(), o1.trait_prob(), rtol=1e-13)
    for l in (0, 41, 98):
        np.testing.assert_allclose(o2.lod_terms(fx["graphs"][2], l), o1.lod_terms(fx["graphs"][2], l), rtol=1e-12)


@pytest.mark.parametrize("name", ["loop", "east", "ped30_snp100"])
def test_scaled_trait_peel_is_the_plain_one_times_a_power_of_two(name):
    """The rescaled peel multiplies tables by exact powers of two only, so
    where the plain peel stays in range L = ldexp(rescaled L, e) bit for bit,
    and the rescaled LOD terms equal the plain ones to rounding of the log."""
    case, fx = load_golden(name)
    o = Oracle(case)
    for g in fx["graphs"][:2]:
        for l in range(0, case.M - 1, max(1, (case.M - 1) // 7)):
            for off in range(1, case.nlod + 1):
                v, e = o.trait_likelihood_scaled(g, l, off)
                assert 0.5 <= v < 1.0 or v == 0.0
                assert np.ldexp(v, e) == o.trait_likelihood(g, l, off)
            np.testing.assert_allclose(o.lod_terms_scaled(g, l), o.lod_terms(g, l), rtol=1e-13, atol=0)
    v, e = o.trait_likelihood_scaled(None, 0, 1)
    assert np.ldexp(v, e) == o.trait_likelihood(None, 0, 1)


def test_scaled_trait_peel_past_fp64_underflow():
    """Multiplying every disease probability by 2^-40 multiplies L by
    exactly 2^(-40 N) (every term of L holds one disease probability per
    person): past N = 26 the plain FP64 peel underflows to 0, the rescaled
    one gives ln L - 40 N ln 2."""
    case, fx = load_golden("ped30_snp100")
    tiny = dataclasses.replace(case, disease_prob=np.ldexp(case.disease_prob, -40))
    o, ot = Oracle(case), Oracle(tiny)
    shift = 40 * case.N * np.log(2.0)
    g = fx["graphs"][0]
    for l in (0, 37, 98):
        assert ot.trait_likelihood(g, l, 1) == 0.0
        with pytest.raises(RuntimeError):
            ot.lod_terms(g, l)
        np.testing.assert_allclose(ot.lod_terms_scaled(g, l), o.lod_terms(g, l) - shift, rtol=1e-13, atol=0)
        v, e = ot.trait_likelihood_scaled(g, l, 2)
        v0, e0 = o.trait_likelihood_scaled(g, l, 2)
        assert v == v0 and e == e0 - 40 * case.N


def test_lod_accumulator_closed_form():
    case, fx = load_golden("east")
    o = Oracle(case)
    acc = o.lod_accumulate(fx["graphs"])
    terms = fx["lod_terms"].reshape(len(fx["graphs"]), -1)
    np.testing.assert_allclose(acc, np.logaddexp.reduce(terms, axis=0), rtol=1e-13)


# ---------------------------------------------------------------- FAG
def founder_alleles(case, col):
    fa = np.zeros((case.N, 2), np.int64)
    done = np.zeros(case.N, bool)
    for i in range(case.F):
        fa[i] = (2 * i, 2 * i + 1)
        done[i] = True
    while not done.all():
        for i in range(case.F, case.N):
            if done[i] or not (done[case.mother[i]] and done[case.father[i]]):
                continue
            fa[i, 0] = fa[case.mother[i], col[2 * i]]
            fa[i, 1] = fa[case.father[i], col[2 * i + 1]]
            done[i] = True
    return fa


def brute_fag(case, dg, locus):
    """Sum over allele assignments (A or B) of the founder alleles carried by
    typed persons: product of allele frequencies times genotype consistency."""
    fa = founder_alleles(case, dg[locus])
    typed = [i for i in range(case.N) if case.typed[i] and case.geno[i, locus] != 0]
    if not typed:
        return 1.0
    S = sorted({int(x) for i in typed for x in fa[i]})
    pos = {s: k for k, s in enumerate(S)}
    X = (np.arange(2 ** len(S))[:, None] >> np.arange(len(S))) & 1  # 1 = allele B (minor)
    q = case.minor[locus]
    w = np.prod(np.where(X == 1, q, 1.0 - q), axis=1)
    for i in typed:
        x, y = X[:, pos[int(fa[i, 0])]], X[:, pos[int(fa[i, 1])]]
        g = case.geno[i, locus]
        ok = (x != y) if g == 1 else ((x == 0) & (y == 0)) if g == 2 else ((x == 1) & (y == 1))
        w = w * ok
    return w.sum()


@pytest.mark.parametrize("name,nloci", [("loop", None), ("ped30_snp100", None), ("east", 6)])
def test_fag_brute_force(name, nloci):
    case, fx = load_golden(name)
    o = Oracle(case)
    loci = range(case.M) if nloci is None else np.linspace(0, case.M - 1, nloci).astype(int)
    for g in fx["graphs"][:2]:
        for l in loci:
            np.testing.assert_allclose(o.fag(g, int(l)), brute_fag(case, g, int(l)), rtol=1e-12, atol=0)


def test_fag_detects_illegal_column():
    case, fx = load_golden("ped30_snp100")
    o = Oracle(case)
    g = fx["graphs"][0].copy()
    zeros = 0
    for l in range(case.M):
        for b in range(2 * case.F, 2 * case.N):
            h = g.copy()
            h[l, b] ^= 1
            f = o.fag(h, l)
            assert (f == 0.0) == (brute_fag(case, h, l) == 0.0)
            zeros += f == 0.0
        if zeros > 20:
            break
    assert zeros > 0


# ---------------------------------------------------------------- samplers
def slice_markers(case, l0, l1):
    """The same pedigree restricted to markers [l0, l1)."""
    return dataclasses.replace(case, M=l1 - l0, geno=case.geno[:, l0:l1], geno_prob=case.geno_prob[:, l0:l1],
                               legal=case.legal[:, l0:l1], theta=case.theta[l0:l1 - 1],
                               partial_theta=case.partial_theta[l0:l1 - 1], minor=case.minor[l0:l1])


def gtest(counts, probs):
    keep = probs > 0
    assert counts[~keep].sum() == 0, "sampled an impossible configuration"
    exp = probs[keep] * counts.sum()
    obs = counts[keep]
    # merge sparse bins so the asymptotic test holds
    order = np.argsort(exp)
    e2, o2, ea, oa = [], [], 0.0, 0
    for k in order:
        ea += exp[k]
        oa += obs[k]
        if ea >= 5:
            e2.append(ea), o2.append(oa)
            ea, oa = 0.0, 0
    if ea > 0:
        e2[-1] += ea
        o2[-1] += oa
    return stats.chisquare(o2, e2).pvalue


def test_msampler_path_distribution():
    """One M-step on meiosis m samples the indicator path along the markers
    with P(path) proportional to prod_l FAG(column with bit=b_l) times the
    recombination factors (meiosis_sampler.cc:105-191)."""
    case, fx = load_golden("ped30_snp100")
    sub = slice_markers(case, 40, 46)
    g0 = np.ascontiguousarray(fx["graphs"][1][40:46])
    o = Oracle(sub)
    m = int(sub.meioses()[3])
    col = 2 * (sub.F + m // 2) + m % 2
    M = sub.M
    raw = np.zeros((M, 2))
    for l in range(M):
        for b in range(2):
            h = g0.copy()
            h[l, col] = b
            raw[l, b] = o.fag(h, l)
    paths = (np.arange(2 ** M)[:, None] >> np.arange(M)) & 1
    p = np.prod(raw[np.arange(M), paths], axis=1)
    for l in range(M - 1):
        p = p * np.where(paths[:, l] != paths[:, l + 1], sub.theta[l], 1 - sub.theta[l])
    p /= p.sum()
    n = 6000
    counts = np.zeros(2 ** M)
    for s in range(n):
        h = o.msweep(g0, [m], 1234, 0, s)
        counts[int((h[:, col] << np.arange(M)).sum())] += 1
    assert gtest(counts, p) > 1e-4


def test_lsampler_locus_distribution():
    """One L-step at a locus samples phased genotypes from the peeled model
    (sampler trait cache x transmissions with flanking thetas) and then the
    meiosis indicators; enumerate all 4^N genotype assignments of the 9-person
    loop pedigree for the exact distribution of the locus's indicator vector."""
    case, fx = load_golden("loop")
    o = Oracle(case)
    N, F, l = case.N, case.F, 1
    g0 = fx["graphs"][0]
    th_l, th_r = case.theta[l - 1], case.theta[l]

    def ind_prob(c, s, b):  # neighbours-only probability of indicator b
        left = 1 - th_l if g0[l - 1, 2 * c + s] == b else th_l
        right = 1 - th_r if g0[l + 1, 2 * c + s] == b else th_r
        return left * right

    S = np.arange(4 ** N, dtype=np.int64)
    a = (S[:, None] >> (2 * np.arange(N))) & 3
    w = np.prod(case.geno_prob[np.arange(N), l, a], axis=1)
    nb = 2 * (N - F)
    bitdist = []  # per meiosis: P(bit=1 | genotypes) for each configuration
    for c in range(F, N):
        for s, par in ((0, case.mother[c]), (1, case.father[c])):
            pg = a[:, par]
            kid_allele = np.array([affected(int(k), s) for k in range(4)])[a[:, c]]   # True = A
            p0, p1 = ind_prob(c, s, 0), ind_prob(c, s, 1)
            het = (pg == PT_AU) | (pg == PT_UA)
            # hetero parent: indicator 0 passes the parent's maternal allele
            par_mat_aff = np.array([affected(int(k), 0) for k in range(4)])[pg]
            ind = np.where(kid_allele == par_mat_aff, 0, 1)
            # transmission probability of the kid's allele from this parent
            homo_ok = np.where(pg == PT_AA, kid_allele, np.where(pg == PT_UU, ~kid_allele, True))
            t = np.where(het, np.where(ind == 0, p0, p1) / (p0 + p1), homo_ok.astype(float))
            w = w * t
            bitdist.append(np.where(het, ind.astype(float), p1 / (p0 + p1)))
    w /= w.sum()
    nz = np.nonzero(w)[0]
    probs = np.zeros(2 ** nb)
    for k in nz:
        pk = np.array([1.0])
        for b in range(nb):
            q1 = bitdist[b][k]
            pk = np.concatenate([pk * (1 - q1), pk * q1])  # index bit b = meiosis b
        probs += w[k] * pk
    n = 8000
    counts = np.zeros(2 ** nb)
    for s in range(n):
        h, _ = o.lstep(g0, l, 77, 0, s)
        bits = h[l, 2 * F:]
        counts[int((bits << np.arange(nb)).sum())] += 1
    assert gtest(counts, probs) > 1e-4


def test_locus_by_locus_weight_brute_force():
    """With neither neighbour every meiosis transmits either parental allele
    with probability 1/2, so a locus step's likelihood is the sum over all
    4^N phased-trait assignments of prod_p q_p(t_p) (the sampler trait
    probabilities) times prod_kids P(t_kid | t_mother, t_father).  The weight
    of locus_by_locus is the log-sum over loci; enumerated here for the
    9-person loop pedigree with simulated genotypes at 6 loci (30% untyped)."""
    from swiftlink_amd import linkage
    case, fx = load_golden("loop")
    N, M = case.N, 6
    rng = np.random.default_rng(11)
    minor = rng.uniform(0.1, 0.5, M)
    al = np.zeros((N, M, 2), np.int64)
    al[:case.F] = rng.random((case.F, M, 2)) < minor[None, :, None]
    for k in range(case.F, N):
        for side, par in ((0, case.mother[k]), (1, case.father[k])):
            al[k, :, side] = al[par, np.arange(M), rng.integers(0, 2, M)]
    geno = np.where(al[..., 0] != al[..., 1], linkage.HETERO, np.where(al[..., 0] == 1, linkage.HOMOZ_B,
                                                                         linkage.HOMOZ_A)).astype(np.uint8)
    geno[rng.random((N, M)) < 0.3] = linkage.UNTYPED
    gp = linkage.sampler_cache(geno, minor)
    case = dataclasses.replace(case, M=M, geno=geno, minor=minor, geno_prob=gp,
                               legal=linkage.eliminate(geno, case.mother, case.father, case.F),
                               theta=np.full(M - 1, 0.05), partial_theta=np.full(M - 1, 0.025))
    assert len(np.unique(gp.reshape(-1, 4), axis=0)) > 3
    # phased traits UU, AA, AU, UA: (maternal allele, paternal allele), 1 = A
    alleles = [(0, 0), (1, 1), (1, 0), (0, 1)]
    T = np.zeros((4, 4, 4))
    for tm in range(4):
        for tf in range(4):
            for tk in range(4):
                pm = sum(a == alleles[tk][0] for a in alleles[tm]) / 2.0
                pf = sum(a == alleles[tk][1] for a in alleles[tf]) / 2.0
                T[tm, tf, tk] = pm * pf
    want = 0.0
    for l in range(case.M):
        full = np.ones((4,) * N)
        for p in range(N):
            shape = [1] * N
            shape[p] = 4
            full = full * case.geno_prob[p, l].reshape(shape)
        for k in range(case.F, N):
            m, f = int(case.mother[k]), int(case.father[k])
            shape = [1] * N
            shape[m], shape[f], shape[k] = 4, 4, 4
            full = full * np.moveaxis(T, [0, 1, 2], np.argsort(np.argsort([m, f, k]))).reshape(shape)
        want += np.log(full.sum())
    _, w = Oracle(case).impute(5, 0, by_locus=True)
    assert np.isfinite(want)
    assert abs(w - want) <= 1e-12 * abs(want)


@pytest.mark.parametrize("side", ["left", "right"])
def test_one_sided_locus_likelihood_brute_force(side):
    """The sequential-imputation walk (start_from, locus_sampler2.cc:209-241)
    steps away from its starting locus conditioning on one neighbour only: a
    heterozygous parent passes the allele its indicator b selects with
    probability (1-theta) when b equals the neighbour's indicator of that
    meiosis and theta otherwise (sampler_rfunction.cc:102-157 normalises
    0.5 x that factor over b), a homozygous one its only allele.  The
    likelihood of the step is the sum over all 4^N phased-trait assignments
    of prod_p q_p(t_p) times the kids' transmissions, enumerated here for the
    9-person loop pedigree at every locus that has the neighbour, with the
    two intervals given different thetas so that the interval taken for each
    side is pinned too."""
    case, fx = load_golden("loop")
    case = dataclasses.replace(case, theta=np.array([0.05, 0.2]))
    o = Oracle(case)
    N, F, M = case.N, case.F, case.M
    g0 = fx["graphs"][1]
    alleles = [(0, 0), (1, 1), (1, 0), (0, 1)]  # UU, AA, AU, UA: (maternal, paternal), 1 = A
    loci = range(1, M) if side == "left" else range(0, M - 1)
    for l in loci:
        nb, th = (l - 1, case.theta[l - 1]) if side == "left" else (l + 1, case.theta[l])
        full = np.ones((4,) * N)
        for p in range(N):
            shape = [1] * N
            shape[p] = 4
            full = full * case.geno_prob[p, l].reshape(shape)
        for k in range(F, N):
            m, f = int(case.mother[k]), int(case.father[k])
            T = np.zeros((4, 4, 4))
            for tm in range(4):
                for tf in range(4):
                    for tk in range(4):
                        pr = 1.0
                        for s, tp in ((0, tm), (1, tf)):
                            nbit = int(g0[nb, 2 * k + s])
                            # indicator b passes the parent's allele b (0 maternal, 1 paternal)
                            pr *= sum((1 - th if b == nbit else th) if alleles[tp][0] != alleles[tp][1] else 0.5
                                      for b in (0, 1) if alleles[tp][b] == alleles[tk][s])
                        T[tm, tf, tk] = pr
            shape = [1] * N
            shape[m], shape[f], shape[k] = 4, 4, 4
            full = full * np.moveaxis(T, [0, 1, 2], np.argsort(np.argsort([m, f, k]))).reshape(shape)
        want = full.sum()
        got = o.locus_likelihood(g0, l, use_left=side == "left", use_right=side == "right")
        assert want > 0
        assert abs(got - want) <= 1e-12 * want, (l, got, want)


def test_chain_is_its_schedule_of_steps():
    """MarkovChain::run (markov_chain.cc:335-389): sweep s is an L-sweep when
    the schedule stream's first uniform is below lprob (its second picks the
    parity done first; every locus of that parity, then of the other, is
    stepped) and otherwise an M-sweep over the meioses shuffled by
    random_shuffle on the shuffle stream (swap position i with floor(u (i+1)),
    i from the last down); a sweep at or past burnin whose index is a multiple
    of the period is scored, as one merge of every interval's terms.  The
    oracle's chain equals these steps composed by hand, sweep by sweep, and
    its accumulators the log-sum of the scored graphs' terms."""
    case, fx = load_golden("ped30_snp100")
    o = Oracle(case)
    seed, lprob, burnin, period, sweeps = 91, 0.5, 2, 3, 9
    g = fx["graph_init"].copy()
    ml = o.meioses()
    kinds, scored = [], []
    for sw in range(sweeps):
        if uniform(seed, 0, sw, 0xFFFFFFFF, 0) < lprob:
            kinds.append(0)
            p0 = 0 if uniform(seed, 0, sw, 0xFFFFFFFF, 1) < 0.5 else 1
            for par in (p0, 1 - p0):
                for l in range(par, case.M, 2):
                    g, _ = o.lstep(g, l, seed, 0, sw)
        else:
            kinds.append(1)
            order = ml.copy()
            for d, i in enumerate(range(len(order) - 1, 0, -1)):
                j = min(int(uniform(seed, 0, sw, 0xFFFFFFFE, d) * (i + 1)), i)
                order[i], order[j] = order[j], order[i]
            g = o.msweep(g, order, seed, 0, sw)
        if sw >= burnin and sw % period == 0:
            scored.append(np.concatenate([o.lod_terms(g, l) for l in range(case.M - 1)]))
    assert 0 < sum(kinds) < sweeps
    out, acc, cnt, k = o.chain(fx["graph_init"], seed, 0, sweeps, burnin, period, lprob)
    assert list(k) == kinds
    assert np.array_equal(out, g)
    assert cnt == len(scored) == len([s for s in range(burnin, sweeps) if s % period == 0])
    np.testing.assert_allclose(acc, np.logaddexp.reduce(np.array(scored), axis=0), rtol=1e-12, atol=0)


def test_recombination_term_closed_form():
    """Peeler::process returns ln L - R - T (peeler.cc:97-99): with T = 2(N-F)
    ln 1/2 and R = k ln theta_l + (2(N-F) - k) ln(1 - theta_l) for k
    meioses recombining across interval l.  A graph with identical columns
    (k = 0 everywhere) and then 3 meioses flipped at every locus past l
    (k = 3 at l, 0 elsewhere) against the trait likelihood alone; every
    interval gets its own theta, so the interval R is taken from is pinned."""
    case, fx = load_golden("ped30_snp100")
    case = dataclasses.replace(case, theta=np.random.default_rng(5).uniform(0.01, 0.3, case.M - 1))
    o = Oracle(case)
    N, F, M = case.N, case.F, case.M
    nm = 2 * (N - F)
    base = np.repeat(fx["graphs"][0][:1], M, axis=0)
    for l in (5, 40):
        g = base.copy()
        flip = [2 * F + 1, 2 * F + 6, 2 * N - 2]
        g[l + 1:, flip] ^= 1
        for li, k in ((l, 3), (l - 1, 0), (l + 1, 0)):
            th = case.theta[li]
            R = k * np.log(th) + (nm - k) * np.log1p(-th)
            want = [np.log(o.trait_likelihood(g, li, off)) - R - nm * np.log(0.5) for off in range(1, case.nlod + 1)]
            np.testing.assert_allclose(o.lod_terms(g, li), want, rtol=1e-13, atol=1e-12)


def test_impute_walk_covers_every_locus_once():
    """A replicate's locus walk: every column is written (no column keeps the
    zero it started with wherever a heterozygous parent fixes the bit), and
    replicates on different chains draw different starting loci."""
    case, fx = load_golden("ped30_snp100")
    o = Oracle(case)
    starts = {int(uniform(55, r, 0, 0xFFFFFFFC, 0) * case.M) for r in range(16)}
    assert len(starts) > 8
    g, w = o.impute(55, 0)
    assert np.isfinite(w) and w < 0
    assert all(o.fag(g, l) > 0 for l in range(case.M))
