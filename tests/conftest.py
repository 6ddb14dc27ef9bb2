import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
CASES = ["lig_eq1", "lig_fep2", "lig_fep2_nolrf", "lig_fep2_arith", "lig_fep2_anyatom", "lig_fep2_anyatom_nolrf"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run on the GPU box)")


def golden_path(name):
    return os.path.join(GOLDEN, name + ".npz")
