import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: full-size workload checks")


# micro-kernel id -> (KERNEL, layer) of the reference oracle that pins it (Test/Test.cpp:12-23)
UKER_REF = {2: (1, 0), 3: (1, 1), 4: (2, 0), 5: (2, 1), 6: (8, 0), 11: (16, 0), 9: (32, 0), 7: (128, 0), 8: (128, 1),
            14: (256, 0), 15: (256, 1), 16: (512, 0), 17: (512, 1), 18: (1024, 0), 19: (1024, 1), 20: (64, 0),
            22: (4, 0)}

# micro-kernels whose last source value is a radius / regularisation (kept small and positive)
RADIUS_UKERS = {9, 10, 11, 20, 24}


def source_values(uk, n, rng, ks=None):
    """Random source values as the reference's test generates them (Test/Test.cpp:318-328)."""
    from oracle import oracle as O
    ks = ks or O.src_dim(uk)
    v = rng.uniform(-1.0, 1.0, (n, ks))
    if uk in RADIUS_UKERS:
        v[:, -1] = np.abs(v[:, -1]) * 0.01
    return v


@pytest.fixture(scope="session")
def golden():
    path = os.path.join(ROOT, "tests", "golden", "simplekernel_fixtures.npz")
    return np.load(path, allow_pickle=False)
