import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (REPO, os.path.join(REPO, "mpi-msfec_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmsfec_hip.so)")


def box_vertices(dim, lo, hi):
    """2^dim vertices of an axis-aligned box in deal.II vertex order."""
    import numpy as np
    nv = 1 << dim
    return np.array([[hi[d] if (v >> d) & 1 else lo[d] for d in range(dim)] for v in range(nv)],
                    dtype=np.float64)


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O
    O.build()
    return O
