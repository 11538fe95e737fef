"""GPU parity through the C++ mirror of STKFMM's interface (include/stkfmm_p2p.hpp): the test
program tests/cxx/direct_driver follows the reference test's direct mode (Test/Test.cpp:542-572) and every
kernel's result is compared with the oracle on that program's own inputs."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import ROOT
from tests.gpu_util import FP64_TOL, say

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "tests", "cxx", "direct_driver")
# KERNEL -> (kdimSL, kdimDL, kdimTrg, SL micro-kernel, DL micro-kernel or None)
KERNELS = {1: (1, 3, 4, 2, 3), 2: (1, 3, 10, 4, 5), 8: (3, 6, 3, 6, None), 16: (4, 7, 6, 11, None),
           32: (4, 7, 3, 9, None), 128: (4, 9, 4, 7, 8)}


def test_cxx_direct_driver(tmp_path):
    assert os.path.exists(DRIVER), "build first (make -C stkfmm_amd/csrc)"
    n = 613
    comb = sum(KERNELS)
    say(f"cxx driver n={n} comb={comb}")
    r = subprocess.run([DRIVER, str(tmp_path), str(n), str(comb)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    # kernels without a double layer print the reference's message and return (FMMData.cpp:270-273)
    assert r.stdout.count("PPKernel 2 not found for direct evaluation") == 3
    for k, (ks, kd, kt, uk_sl, uk_dl) in KERNELS.items():
        rd = lambda name: np.fromfile(os.path.join(tmp_path, f"K{k}_{name}.bin"), dtype=np.float64)  # noqa: E731
        trg = rd("trg").reshape(n, 3)
        got = rd("result").reshape(n, kt)
        ref = O.direct(uk_sl, trg, rd("valSL").reshape(n, ks), trg)
        if uk_dl is not None:
            ref = ref + O.direct(uk_dl, trg, rd("valDL").reshape(n, kd), trg)
        l2, mx = O.max_error(got, ref)
        assert l2 < FP64_TOL and mx < FP64_TOL, (k, l2, mx)
