import os
import sys

import pytest

# The reference build races in OpenMP loops (unreduced counter in
# checkForCellsAboveTc, shared ghost-cell indices); goldens and comparisons run
# it on one thread.  Set before any OpenMP library is loaded.
os.environ["OMP_NUM_THREADS"] = "1"

_HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, _HERE)
sys.path.insert(0, os.path.join(os.path.dirname(_HERE), "cpu-vh_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
