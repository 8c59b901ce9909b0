"""Helpers shared by the GPU parity tests (not a test module)."""
import sys

import numpy as np


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def to_dev(torch, a):
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint16:
        return torch.from_numpy(a.view(np.int16)).cuda()
    if a.dtype == np.uint32:
        return torch.from_numpy(a.view(np.int32)).cuda()
    return torch.from_numpy(a).cuda()


def to_host(t, dtype=None):
    a = t.detach().cpu().numpy()
    if dtype is not None:
        a = a.view(dtype)
    return a


def laplace_symbols(rng, n, scale=4.0, density=0.3, maxsym=None, dtype=np.uint16):
    q = np.rint(rng.laplace(0, scale, size=n)).astype(np.int64)
    q[rng.random(n) > density] = 0
    s = np.where(q >= 0, 2 * q, -2 * q - 1)
    if maxsym is not None:
        s = np.minimum(s, maxsym)
    return s.astype(dtype)
