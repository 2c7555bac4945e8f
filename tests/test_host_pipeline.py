"""Host pre-computation (swiftlink_amd.linkage / peelseq) against the tables
the compiled reference built for the same LINKAGE input (golden fixtures),
matched by person id because the reference's founders-first sort is not
stable."""
import os
import tempfile

import numpy as np
import pytest

from conftest import load_golden
from oracle_lib import Oracle
from swiftlink_amd import linkage, synth
from swiftlink_amd.case import Case


def _compare(mine: Case, ref: Case):
    perm = np.array([mine.ids.index(i) for i in ref.ids])
    assert np.array_equal(mine.disease_prob[perm], ref.disease_prob)
    assert np.array_equal(mine.geno_prob[perm], ref.geno_prob)
    assert np.array_equal(mine.legal[perm], ref.legal)
    assert np.array_equal(mine.geno[perm], ref.geno)
    assert np.array_equal(mine.typed[perm], ref.typed)
    assert np.array_equal(mine.theta, ref.theta)
    assert np.array_equal(mine.partial_theta, ref.partial_theta)
    assert np.array_equal(mine.minor, ref.minor)
    # any legitimate peel order gives the same likelihood
    assert abs(Oracle(mine).trait_prob() - ref.trait_prob) < 1e-12


def test_synthetic_pedigree_matches_reference_tables():
    ref, fx = load_golden("ped30_snp100")
    with tempfile.TemporaryDirectory() as d:
        p, m, dat = synth.write_config("ped30_snp100", d)
        mine = linkage.load(p, m, dat)
    _compare(mine, ref)


@pytest.mark.parametrize("ex", ["loop", "east"])
def test_examples_match_reference_tables(ex):
    src = os.path.join("/root/reference/examples", ex)
    if not os.path.exists(src + ".ped"):
        pytest.skip("reference examples not mounted (golden fixtures still cover the oracle)")
    ref, fx = load_golden(ex)
    _compare(linkage.load(src + ".ped", src + ".map", src + ".dat"), ref)


def test_true_graph_is_legal():
    with tempfile.TemporaryDirectory() as d:
        case, data = synth.make_case("ped30_snp100", d)
    g = synth.true_graph(data, case)
    o = Oracle(case)
    assert min(o.fag(g, l) for l in range(case.M)) > 0


def test_synthetic_configs_shapes():
    for name, (N, gens, loops) in {"ped30_snp100": (30, 4, 1), "ped150_snp1000": (150, 6, 3)}.items():
        spec = synth.CONFIGS[name]
        ped = synth.make_pedigree(spec, np.random.default_rng(spec.seed))
        assert len(ped.ids) == N and max(ped.generation) + 1 == gens and ped.loops == loops


def test_lod_table_text_matches_reference(golden):
    """format_lod_table against LinkageWriter::write's file, for one family
    and for two (per-family columns after the summed LOD)."""
    case, fx = golden
    names, gpos = list(fx["writer_names"]), fx["writer_gpos"]
    lod = fx["writer_lod"]
    assert linkage.format_lod_table(names, gpos, case.partial_theta, [lod[0]]) == str(fx["writer_text1"][0])
    assert linkage.format_lod_table(names, gpos, case.partial_theta, [lod[1], lod[2]]) == str(fx["writer_text2"][0])


def test_lod_table_file_and_formats(tmp_path):
    """write_lod_table writes the same text; values print as std::ostream
    does by default (6 significant digits, exponent form outside 1e-4..1e6,
    -inf for an empty accumulator)."""
    p = tmp_path / "out.txt"
    lods = np.array([[1234567.0, 0.000012345, -np.inf]])
    linkage.write_lod_table(str(p), ["a", "b"], np.array([0.0, 0.01]), np.array([0.002]), [lods])
    lines = p.read_text().splitlines()
    assert lines[0] == "marker\tposition\tlod"
    assert lines[1] == "a\t0"
    assert [l.split("\t")[2] for l in lines[2:5]] == ["1.23457e+06", "1.2345e-05", "-inf"]
    assert lines[5] == "b\t1"
