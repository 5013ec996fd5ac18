"""GPU checks of negf_rgf_diag_accumulate_noalloc, the entry point whose
workspace is the library's fixed device scratch array: two calls in a row
(the second smaller) against negf_rgf_diag_accumulate, bit for bit (same
kernels, same fixed-order sums), and the size limit reported before any
launch."""
import ctypes as C

import numpy as np
import pytest

from electrostatic_amd import _lib
from tests.gpu_util import chain_params, dev_array, dev_zeros, ptr, smooth_potential, to_np

pytestmark = pytest.mark.gpu

BETA17 = 2.2286917788826917
GAMMA = 2.5


def case(N, nE, seed):
    rng = np.random.default_rng(seed)
    U = smooth_potential(N, seed=seed)
    th = np.sort(rng.uniform(0.05, np.pi - 0.05, nE // 2))
    E = np.concatenate([-4.5 + 5.5 * np.exp(1j * th),
                        np.sort(rng.uniform(-3, 3, nE - nE // 2)) + 1e-5j])
    coef = rng.normal(size=nE) + 1j * rng.normal(size=nE)
    return U, E, coef


def run(fn, U, E, coef, want_G, segment_length=0):
    lib = _lib.load()
    N = U.size
    p = chain_params(N, BETA17, GAMMA, (U[0], U[-1]), segment_length=segment_length)
    dU, dE, dc = dev_array(U), dev_array(E), dev_array(coef)
    acc = dev_zeros(N)
    G = dev_zeros(E.size * N, complex_=True) if want_G else None
    _lib.check(getattr(lib, fn)(p, ptr(dU), E.size, ptr(dE), ptr(dc), ptr(acc), ptr(G), None))
    return to_np(acc), (to_np(G, True) if want_G else None)


@pytest.mark.parametrize("want_G", [False, True])
def test_noalloc_matches_allocating_entry_point(want_G):
    # first call: 8192 x 4096 charge-only (45 MB of scratch), or 1000 x 2048
    # with diag G; second call smaller, with a partial last segment
    sizes = [(4096, 8192, 11), (1001, 300, 12)] if not want_G else [(2048, 1000, 13), (777, 130, 14)]
    for N, nE, seed in sizes:
        print(f"noalloc case N={N} nE={nE} want_G={want_G}", flush=True)
        U, E, coef = case(N, nE, seed)
        a_ref, g_ref = run("negf_rgf_diag_accumulate", U, E, coef, want_G)
        a, g = run("negf_rgf_diag_accumulate_noalloc", U, E, coef, want_G)
        assert np.all(np.isfinite(a))
        assert np.array_equal(a, a_ref)
        if want_G:
            assert np.array_equal(g, g_ref)


def test_noalloc_other_segment_length_and_decimation():
    U, E, coef = case(500, 200, 15)
    lib = _lib.load()
    for kw in (dict(segment_length=48), dict(use_decimation=1)):
        print(f"noalloc case N=500 nE=200 {kw}", flush=True)
        p = chain_params(U.size, BETA17, GAMMA, (U[0], U[-1]), **kw)
        dU, dE, dc = dev_array(U), dev_array(E), dev_array(coef)
        outs = []
        for fn in ("negf_rgf_diag_accumulate", "negf_rgf_diag_accumulate_noalloc"):
            acc = dev_zeros(U.size)
            _lib.check(getattr(lib, fn)(p, ptr(dU), E.size, ptr(dE), ptr(dc), ptr(acc), None, None))
            outs.append(to_np(acc))
        assert np.array_equal(outs[0], outs[1])


def test_noalloc_rejects_oversize_before_launch():
    lib = _lib.load()
    N, nE = 8192, 65536   # about 0.35 GB of workspace
    p = chain_params(N, BETA17, GAMMA, (0.0, 0.0))
    dU = dev_zeros(N)
    dE = dev_zeros(nE, complex_=True)
    acc = dev_zeros(N)
    rc = lib.negf_rgf_diag_accumulate_noalloc(p, ptr(dU), nE, ptr(dE), None, ptr(acc), None, None)
    assert rc == 1  # NEGF_EARG
    assert b"fixed scratch" in lib.negf_last_error()
