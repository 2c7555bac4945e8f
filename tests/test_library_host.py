"""Library checks that need no GPU: it loads, exports every entry point the
header declares, rejects bad descriptors, and its compiled peel program (the
indexing the LOD kernel uses) reproduces the oracle when executed on the CPU."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT, load_golden, GOLDEN_CASES
from oracle_lib import Oracle
from program_emu import Program
from swiftlink_amd import native


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "swiftlink_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(sl_\w+)\s*\(", txt, re.M)))


def test_library_exports_header_symbols():
    lib = native.lib()
    names = header_symbols()
    assert len(names) >= 14
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(native.exported_symbols()) == names


def test_library_shares_torch_hip_runtime():
    """Loading the library before anything imports torch still leaves one HIP
    runtime in the process (two copies each open the GPU, and torch's own
    device initialisation then fails)."""
    import subprocess
    import sys
    code = ("import re\nfrom swiftlink_amd import native\nnative.lib()\nimport torch\n"
            "m = open('/proc/self/maps').read()\n"
            "print(len(set(re.findall(r'/\\S*libamdhip64\\S*', m))), len(set(re.findall(r'/\\S*libhsa-runtime64\\S*', m))))")
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert out.stdout.split() == ["1", "1"]


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_program_bounds_and_fp64_bit_exact(name):
    case, fx = load_golden(name)
    P = Program(native.program_export(case))
    P.check_bounds(case.nlod)
    P.check_schedule()
    P.check_terms(case)
    o = Oracle(case)
    assert P.run(case, None, 0)[0] == o.trait_prob()
    g = fx["graphs"][0]
    loci = range(case.M - 1) if case.M < 8 else (0, 1, case.M // 2, case.M - 2)
    for l in loci:
        lnL = P.run(case, g, l)
        ref = [np.log(o.trait_likelihood(g, l, off)) for off in range(1, case.nlod + 1)]
        assert np.array_equal(lnL, ref)


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_program_fp32_rescaled(name):
    case, fx = load_golden(name)
    P = Program(native.program_export(case))
    o = Oracle(case)
    g = fx["graphs"][1]
    for l in (0, case.M - 2):
        lnL = P.run(case, g, l, fp32=True)
        ref = [np.log(o.trait_likelihood(g, l, off)) for off in range(1, case.nlod + 1)]
        assert np.abs(lnL - ref).max() / np.log(10) < 1e-5


def test_program_rejects_bad_tables():
    case, fx = load_golden("loop")
    bad = case.ops.prev.copy()
    if bad.size:
        bad[0] = case.ops.n + 3
        case.ops.prev = bad
        with pytest.raises(native.SlError):
            native.program_export(case)


def test_create_rejects_founder_only_pedigree():
    """A pedigree of founders only has no meioses; sl_create refuses it with
    SL_E_UNSUPPORTED before touching a device."""
    import dataclasses
    case, fx = load_golden("loop")
    F = case.F
    cut = {k: getattr(case, k)[:F] for k in ("mother", "father", "sex", "affection", "disease_prob", "ids", "typed",
                                              "geno", "geno_prob", "legal")}
    sub = dataclasses.replace(case, N=F, **cut)
    with pytest.raises(native.SlError, match="no non-founders"):
        native.Context(sub)


def test_pack_graphs_layout():
    from swiftlink_amd.case import pack_graphs
    case, fx = load_golden("east")
    g = fx["graphs"][:2]
    p = pack_graphs(case, g)
    for gi in range(2):
        for l in (0, 57):
            for b in range(2 * (case.N - case.F)):
                assert ((p[gi, l, b // 32] >> (b % 32)) & 1) == g[gi, l, 2 * case.F + b]


def test_program_ped150_terms_and_bounds(tmp_path):
    """150 people, 3 loops: the compiled program's schedule, precompiled terms
    and arena bounds, and two intervals executed against the oracle."""
    from swiftlink_amd import synth
    case, data = synth.make_case("ped150_snp1000", str(tmp_path))
    P = Program(native.program_export(case))
    P.check_schedule()
    P.check_bounds(case.nlod)
    P.check_terms(case)
    g = synth.true_graph(data, case)
    o = Oracle(case)
    for l in (0, 600):
        ref = [np.log(o.trait_likelihood(g, l, off)) for off in range(1, case.nlod + 1)]
        assert np.array_equal(P.run(case, g, l), ref)


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_sequential_program_fp64_bit_exact_and_fp32_within_bar(name):
    """The lane kernel's sequential form: every table alive at once has its
    own arena cells, and executed with the workgroup kernel's arithmetic in
    that order it gives the oracle's likelihood bit for bit in FP64 and within
    1e-5 LOD in FP32 with power-of-two renormalisation; the lane kernel's
    descriptor stream, executed as the kernel reads it (blocks, padding cells,
    transmission table), is within 1e-5 LOD."""
    case, fx = load_golden(name)
    P = Program(native.program_export(case))
    P.check_seq()
    o = Oracle(case)
    for g in fx["graphs"][:2]:
        for l in (0, case.M // 2, case.M - 2):
            ref = [np.log(o.trait_likelihood(g, l, off)) for off in range(1, case.nlod + 1)]
            assert np.array_equal(P.run_seq(case, g, l), ref)
            assert np.abs(P.run_seq(case, g, l, fp32=True) - ref).max() / np.log(10) < 1e-5
            assert np.abs(P.run_lane_stream(case, g, l) - ref).max() / np.log(10) < 1e-5


def test_sequential_program_ped150_and_underflowing_scale(tmp_path):
    """150 people: the sequential form's arena layout and FP64 values, and the
    lane stream within the LOD bar; with every disease probability scaled by
    2^-40 (a likelihood of about 2^-6000 times the original, far below FP32
    and FP64 range) the FP32 renormalised execution of both still gives
    ln L - 40 N ln 2 within the LOD bar."""
    import dataclasses
    from swiftlink_amd import synth
    case, data = synth.make_case("ped150_snp1000", str(tmp_path))
    P = Program(native.program_export(case))
    P.check_seq()
    g = synth.true_graph(data, case)
    o = Oracle(case)
    ref = [np.log(o.trait_likelihood(g, 300, off)) for off in range(1, case.nlod + 1)]
    assert np.array_equal(P.run_seq(case, g, 300), ref)
    assert np.abs(P.run_lane_stream(case, g, 300) - ref).max() / np.log(10) < 1e-4
    tiny = dataclasses.replace(case, disease_prob=np.ldexp(case.disease_prob, -40))
    got = P.run_seq(tiny, g, 300, fp32=True)
    assert np.abs(got - (np.array(ref) - 40 * case.N * np.log(2.0))).max() / np.log(10) < 1e-4
    got = Program(native.program_export(tiny)).run_lane_stream(tiny, g, 300)
    assert np.abs(got - (np.array(ref) - 40 * case.N * np.log(2.0))).max() / np.log(10) < 1e-4
