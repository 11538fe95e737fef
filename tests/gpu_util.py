"""Helpers shared by the GPU tests."""
import numpy as np

DEV = "cuda:0"
FP64_TOL = 1e-12   # FP64 mode vs the FP64 oracle: errorL2 and errorMaxRel per component
MIXED_TOL = 1e-6   # mixed mode vs the FP64 oracle


def torch():
    import torch as t
    return t


def to_dev(a, dtype=None):
    t = torch()
    x = t.from_numpy(np.ascontiguousarray(a))
    return x.to(DEV) if dtype is None else x.to(DEV, dtype=dtype)


def zeros(n, k):
    t = torch()
    return t.zeros((n, k), dtype=t.float64, device=DEV)


def host(x):
    torch().cuda.synchronize()
    return x.detach().cpu().numpy()


def say(*a):
    print(*a, flush=True)
