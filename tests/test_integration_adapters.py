"""The C++ adapters in INTEGRATION.md (the reference-side binding a
maintainer adds: GPULodscores, GPUMarkovChain, and the ELOD, peel-sequence,
batch and trait-model calls) compile against the reference's own headers
(gpu_lodscores.h:63-120, gpu_markov_chain.h:76-125 and the accessors of
pedigree.h, person.h, genetic_map.h, peel_sequence_generator.h,
descent_graph.h, lod_score.h, disease_model.h, types.h) and this library's
header, so a signature drift on either side fails here.  Skipped where the
reference sources are not mounted."""
import os
import re
import shutil
import subprocess

import pytest

from conftest import ROOT

REF_SRC = "/root/reference/src"


def _blocks():
    txt = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    return re.findall(r"```cpp\n(.*?)```", txt, re.S)


def test_integration_has_the_adapters():
    code = "\n".join(_blocks())
    for name in ("class GPULodscores", "class GPUMarkovChain", "struct SlInputs", "sl_elod(", "sl_batch_run(",
                 "sl_lod_grad(", "sl_peel_sequence(", "sl_sequential_imputation("):
        assert name in code, name


@pytest.mark.skipif(not os.path.isdir(REF_SRC) or shutil.which("g++") is None,
                    reason="reference sources or g++ not available")
def test_integration_adapters_compile_against_reference_headers(tmp_path):
    src = tmp_path / "adapters.cpp"
    src.write_text("\n".join(_blocks()))
    out = subprocess.run(["g++", "-std=c++11", "-fsyntax-only", "-Wall", "-I", REF_SRC,
                          "-I", os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-4000:]
    assert "warning" not in out.stderr, out.stderr[-4000:]
