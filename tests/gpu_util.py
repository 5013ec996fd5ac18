"""Helpers for GPU tests: device arrays through torch (plumbing only)."""
import ctypes as C

import numpy as np
import torch

from electrostatic_amd import _lib


def dev_array(a, dtype=None):
    a = np.ascontiguousarray(a)
    if np.iscomplexobj(a):
        t = torch.from_numpy(a.view(np.float64).copy())
    else:
        t = torch.from_numpy(a.astype(dtype or a.dtype, copy=True))
    return t.cuda()


def dev_zeros(n, complex_=False):
    return torch.zeros(2 * n if complex_ else n, dtype=torch.float64, device="cuda")


def ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def to_np(t, complex_=False):
    a = t.detach().cpu().numpy()
    return a.view(np.complex128) if complex_ else a


def chain_params(N, t_even, t_odd, Uc, use_decimation=0, layers=10, max_iter=30,
                 rel_error=0, segment_length=0):
    p = _lib.negf_chain_params()
    p.N = N
    p.t_even = t_even
    p.t_odd = t_odd
    p.Uc[0], p.Uc[1] = Uc
    p.use_decimation = use_decimation
    p.decimation_layers = layers
    p.decimation_max_iter = max_iter
    p.decimation_rel_error = rel_error
    p.segment_length = segment_length
    return p


def smooth_potential(N, seed, lo=-0.5, hi=0.5, kmax=8):
    """Seeded smooth field: sum of kmax random Fourier modes rescaled to [lo, hi]."""
    rng = np.random.default_rng(seed)
    x = np.arange(N) / N
    U = np.zeros(N)
    for k in range(1, kmax + 1):
        U += rng.normal() * np.cos(2 * np.pi * k * x + rng.uniform(0, 2 * np.pi))
    U = (U - U.min()) / (U.max() - U.min())
    return lo + (hi - lo) * U
