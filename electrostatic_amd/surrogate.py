"""Python side of the BF16 charge surrogate (negf_surrogate_* in libnegf_amd).

The surrogate maps a tube's potential U (eV) and one per-sample feature
(the drain bias) to an estimate of the NEGF induced charge; it is trained on
the device from pairs that the GPU NEGF solver produced and used only to
seed and precondition the batched SCF.  Device arrays are torch tensors
(float64, sample-major), used for device memory only.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _dev(a, device):
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=torch.float64).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a, np.float64), device=device)


class Surrogate:
    def __init__(self, N, seed=0, device="cuda"):
        self.lib = _lib.load()
        self.N = N
        self.device = device
        h = C.c_void_p()
        _lib.check(self.lib.negf_surrogate_create(N, seed, C.byref(h)))
        self.h = h
        self.nparam = self.lib.negf_surrogate_num_params(self.h)

    def close(self):
        if self.h:
            self.lib.negf_surrogate_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # parameters ---------------------------------------------------------
    def get_params(self):
        w = np.zeros(self.nparam, np.float32)
        norm = np.zeros(2)
        _lib.check(self.lib.negf_surrogate_get_params(
            self.h, w.ctypes.data_as(C.POINTER(C.c_float)), norm.ctypes.data_as(C.POINTER(C.c_double))))
        return w, norm

    def set_params(self, w=None, norm=None):
        wp = None
        if w is not None:
            w = np.ascontiguousarray(w, np.float32)
            if w.size != self.nparam:
                raise ValueError(f"expected {self.nparam} parameters, got {w.size}")
            wp = w.ctypes.data_as(C.POINTER(C.c_float))
        npn = None
        if norm is not None:
            norm = np.ascontiguousarray(norm, np.float64)
            npn = norm.ctypes.data_as(C.POINTER(C.c_double))
        _lib.check(self.lib.negf_surrogate_set_params(self.h, wp, npn))

    def save(self, path):
        w, norm = self.get_params()
        np.savez(path, w=w, norm=norm, N=self.N)

    def load(self, path):
        z = np.load(path)
        self.set_params(z["w"], z["norm"])

    # evaluation ---------------------------------------------------------
    def forward(self, U, feat=None):
        """Q estimate [B, N] (device tensor) for U [B, N]."""
        dU = _dev(U, self.device)
        B = dU.numel() // self.N
        df = None if feat is None else _dev(feat, self.device)
        dQ = torch.empty_like(dU)
        _lib.check(self.lib.negf_surrogate_forward(self.h, B, _ptr(dU), _ptr(df), _ptr(dQ), None))
        torch.cuda.synchronize()
        return dQ.view(B, self.N)

    def jvp(self, U, V, feat=None):
        """dQ/dU . V [B, N] (device tensor)."""
        dU = _dev(U, self.device)
        dV = _dev(V, self.device)
        B = dU.numel() // self.N
        df = None if feat is None else _dev(feat, self.device)
        out = torch.empty_like(dU)
        _lib.check(self.lib.negf_surrogate_jvp(self.h, B, _ptr(dU), _ptr(df), _ptr(dV), _ptr(out), None))
        torch.cuda.synchronize()
        return out.view(B, self.N)

    def gradient(self, U, Q, feat=None):
        """(loss, flat parameter gradient) for one batch."""
        dU = _dev(U, self.device)
        dQ = _dev(Q, self.device)
        B = dU.numel() // self.N
        df = None if feat is None else _dev(feat, self.device)
        g = np.zeros(self.nparam, np.float32)
        loss = C.c_double()
        _lib.check(self.lib.negf_surrogate_gradient(self.h, B, _ptr(dU), _ptr(df), _ptr(dQ),
                                                    g.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss)))
        return loss.value, g

    # training -----------------------------------------------------------
    def fit_output_scale(self, Q):
        dQ = _dev(Q, self.device)
        _lib.check(self.lib.negf_surrogate_fit_output_scale(self.h, dQ.numel() // self.N, _ptr(dQ)))

    def train(self, U, Q, feat=None, epochs=10, batch=32, lr=2e-3, lr_end=1e-4, seed=0, fit_scale=True):
        """Adam on (U, f, Q) samples (device or host arrays [S, N]); returns
        the last minibatch loss (normalised MSE)."""
        dU = _dev(U, self.device)
        dQ = _dev(Q, self.device)
        S = dU.numel() // self.N
        df = None if feat is None else _dev(feat, self.device)
        if fit_scale:
            self.fit_output_scale(dQ)
        loss = C.c_double()
        _lib.check(self.lib.negf_surrogate_train(self.h, S, _ptr(dU), _ptr(df), _ptr(dQ), epochs, batch, lr,
                                                 lr_end, seed, C.byref(loss)))
        return loss.value


def sample_potentials(n, N, seed, gate=(-0.2, 1.2), bump_amp=0.5, field_amp=0.3):
    """Seeded smooth potentials (eV) [n, N] of the kind an SCF visits: a
    uniform shift -V (V uniform in gate), four Gaussian bumps (amplitudes up
    to bump_amp, widths N/20..N/5) and a random smooth field of 8 Fourier
    modes (amplitude field_amp)."""
    rng = np.random.default_rng(seed)
    x = np.arange(N)
    U = np.empty((n, N))
    for k in range(n):
        u = -rng.uniform(*gate) * np.ones(N)
        for _ in range(4):
            c, w, a = rng.uniform(0, N), rng.uniform(N / 20, N / 5), rng.uniform(-bump_amp, bump_amp)
            u += a * np.exp(-0.5 * ((x - c) / w) ** 2)
        f = np.zeros(N)
        for m in range(1, 9):
            f += rng.standard_normal() / m * np.sin(np.pi * m * x / N + rng.uniform(0, 2 * np.pi))
        u += field_amp * f / max(np.abs(f).max(), 1e-12)
        U[k] = u
    return U


def negf_pairs(batch, U, device="cuda"):
    """Exact NEGF charges for the potentials U [S, N], computed T at a time
    with the T equal tubes of a negf.Batch (the contact potentials follow
    U[0], U[N-1]); returns (U, Q) as device tensors [S, N]."""
    tubes = batch.tubes
    T, N = len(tubes), tubes[0].N
    S = U.shape[0]
    dU = torch.as_tensor(np.ascontiguousarray(U, np.float64), device=device)
    dQ = torch.empty_like(dU)
    buf = torch.empty(T * N, dtype=torch.float64, device=device)
    for s0 in range(0, S, T):
        n = min(T, S - s0)
        for t in range(T):
            tubes[t].set_potential(U[s0 + min(t, n - 1)])
        batch.solve_device(1, buf.data_ptr())
        dQ[s0:s0 + n] = buf.view(T, N)[:n]
    torch.cuda.synchronize()
    return dU, dQ
