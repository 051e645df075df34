"""Shared test helpers: host<->device copies and the oracle comparisons."""
import numpy as np

import oracle


def seeded_complex(seed, count):
    """count complex64 values, re/im uniform[-1,1) from SplitMix64 (re then im)."""
    v = oracle.uniform_f32(seed, 2 * count)
    return v.view(np.complex64)


def to_dev(arr):
    import torch
    a = np.ascontiguousarray(arr)
    t = torch.from_numpy(a.view(np.uint8).copy()).to("cuda")
    return t


def from_dev(t, dtype, count=None):
    a = t.cpu().numpy().view(dtype)
    return a if count is None else a[:count]


def empty_dev(nbytes):
    import torch
    return torch.empty(int(nbytes), dtype=torch.uint8, device="cuda")


def c2c_expected(x, shape, axes=None, batch=1, inverse=False, norm=oracle.NORM_INVERSE, nthreads=8):
    return oracle.c2c_nd(np.asarray(x, dtype=np.complex128), shape, axes, batch, inverse, norm, nthreads)
