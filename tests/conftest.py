import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(TESTS, "golden")
GOLDEN_CASES = ("loop", "east", "ped30_snp100")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs kernels on an AMD GPU (MI355X)")


def load_golden(name):
    """(Case, fixture dict) for a golden vector file."""
    from swiftlink_amd.case import Case
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    d = {k: z[k] for k in z.files}
    case = Case.from_npz({k[5:]: v for k, v in d.items() if k.startswith("case_")})
    return case, d


@pytest.fixture(params=GOLDEN_CASES)
def golden(request):
    return load_golden(request.param)
