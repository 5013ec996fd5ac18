"""GPU parity of the dense-block path (negf_block_*: Sancho-Rubio leads and
block RGF on FP64 MFMA) against oracle/block_oracle.py on the same inputs.

Tolerances: diag G^R and the accumulated charge 1e-10 relative (max |d| /
max |ref|); the GPU's pivoted Gauss-Jordan and MFMA sums round differently
from numpy's LAPACK inverse and matmul.
"""
import numpy as np
import pytest

from electrostatic_amd import negf
from oracle import block_oracle as B

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def chain(nb, seed, sigma=0.05, Uc=(0.1, -0.05)):
    H0, T = B.zigzag_blocks(16, 2.5)
    rng = np.random.default_rng(seed)
    smooth = 0.2 * np.sin(np.linspace(0, np.pi, nb))[:, None]
    v = rng.normal(0.0, sigma, (nb, 32)) + smooth
    return H0, T, v, Uc


def test_sancho_rubio_self_energies_match_oracle(dev):
    H0, T = B.zigzag_blocks(16, 2.5)
    c = negf.BlockChain(3, H0, T)
    Uc = (0.1, -0.2)
    c.set_potential(np.zeros((3, 32)), Uc)
    E = np.array([0.37 + 1e-5j, -0.8 + 1e-5j, 1.9 + 1e-3j, -2.6 + 0.05j, 0.02 + 1e-5j, 5.5 + 1e-5j])
    SL, SR, it = c.self_energies(E)
    for k, z in enumerate(E):
        rl, rr, its = B.lead_self_energies(z, H0, T, Uc)
        assert rel(SL[k], rl) < 1e-10, (k, rel(SL[k], rl))
        assert rel(SR[k], rr) < 1e-10, (k, rel(SR[k], rr))
        assert abs(it[k, 0] - its[0]) <= 1 and abs(it[k, 1] - its[1]) <= 1
    c.close()


@pytest.mark.parametrize("keep_x", [True, False])
@pytest.mark.parametrize("nb", [1, 2, 7])
def test_block_rgf_diag_matches_oracle(dev, nb, keep_x):
    # keep_x False: a scratch bound of one g_b set per energy makes the
    # backward sweep recompute g_b T instead of reading it back
    H0, T, v, Uc = chain(nb, seed=nb)
    E = np.array([0.37 + 1e-5j, -0.8 + 1e-5j, 1.9 + 1e-3j, -2.6 + 0.05j, 0.02 + 1e-5j])
    c = negf.BlockChain(nb, H0, T)
    c.set_potential(v, Uc)
    if not keep_x:
        c.set_scratch_limit(E.size * max(nb - 1, 1) * 2 * 32 * 32 * 8)
    acc, dg, it = c.rgf(E, want_diag=True)
    assert np.all(it > 0)
    ref_acc = np.zeros((nb, 32))
    for k, z in enumerate(E):
        SL, SR, its = B.lead_self_energies(z, H0, T, Uc)
        ref = B.block_rgf_diag(z, H0, T, v, SL, SR)
        assert rel(dg[k], ref) < 1e-10, (k, rel(dg[k], ref))
        assert abs(it[k, 0] - its[0]) <= 1 and abs(it[k, 1] - its[1]) <= 1
        ref_acc += np.imag(ref)
    assert rel(acc, ref_acc) < 1e-10
    c.close()


def test_block_rgf_long_chain_charge_and_coefficients(dev):
    nb = 300
    H0, T, v, Uc = chain(nb, seed=11)
    rng = np.random.default_rng(12)
    E = rng.uniform(-3, 3, 6) + 1j * rng.uniform(1e-5, 1e-2, 6)
    coef = rng.standard_normal(6) + 1j * rng.standard_normal(6)
    c = negf.BlockChain(nb, H0, T)
    c.set_potential(v, Uc)
    acc, _, it = c.rgf(E, coef)
    ref = B.accumulate(E, coef, H0, T, v, Uc)
    assert rel(acc, ref) < 1e-10
    c.close()


def test_block_rgf_ring_traces_match_mode_space(dev):
    # a ring-uniform potential and no disorder: ring traces of the block
    # path equal the sum over angular modes of scalar chains (oracle)
    nb, m, gamma = 6, 16, 2.5
    H0, T = B.zigzag_blocks(m, gamma)
    Uring = np.linspace(-0.2, 0.25, 2 * nb)
    v = np.repeat(Uring.reshape(nb, 2), m, axis=1)
    c = negf.BlockChain(nb, H0, T)
    c.set_potential(v, (Uring[0], Uring[-1]))
    E = np.array([0.41 + 1e-3j, -1.3 + 1e-3j])
    _, dg, _ = c.rgf(E, want_diag=True)
    from tests.test_oracle import _mode_chain_diag
    for k, z in enumerate(E):
        ring = dg[k].reshape(nb, 2, m).sum(axis=2).ravel()
        modes = sum(_mode_chain_diag(z, 2 * gamma * abs(np.cos(np.pi * J / m)), gamma, Uring,
                                     (Uring[0], Uring[-1])) for J in range(m))
        assert rel(ring, modes) < 1e-10
    c.close()


def test_block_rgf_energy_chunks_are_bit_identical(dev):
    # a scratch limit of two energies' g_b splits 7 energies into sweeps of
    # 2, 2, 2, 1 (the last workgroup partly empty); every energy is computed
    # independently and the charge is summed in energy order, so the result
    # equals the single sweep bit for bit
    nb = 9
    H0, T, v, Uc = chain(nb, 5)
    E = np.array([0.37 + 1e-5j, -0.8 + 1e-5j, 1.9 + 1e-3j, -2.6 + 0.05j, 0.02 + 1e-5j, 0.9 + 1e-2j,
                  -1.4 + 1e-5j])
    coef = np.exp(1j * np.linspace(0.0, 1.0, E.size))
    c = negf.BlockChain(nb, H0, T)
    c.set_potential(v, Uc)
    acc1, d1, _ = c.rgf(E, coef, want_diag=True)
    c.set_scratch_limit(2 * (nb - 1) * 2 * 32 * 32 * 8)
    acc2, d2, _ = c.rgf(E, coef, want_diag=True)
    with pytest.raises(Exception):
        c.set_scratch_limit(0)
    c.close()
    assert np.array_equal(acc1, acc2)
    assert np.array_equal(d1, d2)


def test_block_rgf_rejects_bad_arguments(dev):
    H0, T = B.zigzag_blocks(16)
    with pytest.raises(Exception):
        negf.BlockChain(4, H0 + np.triu(np.ones((32, 32)), 1), T)   # not symmetric
    c = negf.BlockChain(2, H0, T)
    with pytest.raises(Exception):
        c.rgf(np.array([0.1 + 1e-5j]), tol=-1.0)
    c.close()


@pytest.mark.parametrize("op", [0, 1, 2, 3, 4, 5])
def test_block_building_blocks(dev, op):
    # the MFMA quadrant products (operand and result lane layout) and the
    # pivoted Gauss-Jordan inverse on one matrix pair, against numpy; ops 0-2
    # by a workgroup of four wavefronts, ops 3-5 by a single wavefront
    import ctypes as C
    from electrostatic_amd import _lib
    rng = np.random.default_rng(op)
    A = rng.standard_normal((32, 32)) + 1j * rng.standard_normal((32, 32))
    A[:, 3] *= 1e-3                     # forces row exchanges in the inverse
    Bm = rng.standard_normal((32, 32)) + 1j * rng.standard_normal((32, 32))
    planes = lambda M: np.ascontiguousarray(np.stack([M.real, M.imag]).ravel())
    a, b = planes(A), planes(Bm)
    out = np.zeros(2048)
    _lib.check(_lib.load().negf_block_probe(op, a.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p),
                                            out.ctypes.data_as(C.c_void_p)))
    Cm = out[:1024].reshape(32, 32) + 1j * out[1024:].reshape(32, 32)
    ref = {0: A @ Bm, 1: np.linalg.inv(A), 2: A @ Bm.real}[op % 3]
    assert rel(Cm, ref) < 1e-12


def _pivot_case(kind, rng):
    """(A, A^-1 reference, tolerance)."""
    n = 32
    z = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
    if kind == "antidiagonal":
        # every column has one nonzero, in a different row: each pivot is forced
        A = np.diag(z(n))[::-1].copy()
        return A, np.linalg.inv(A), 1e-13
    if kind == "permutation":
        A = np.zeros((n, n), complex)
        A[rng.permutation(n), np.arange(n)] = z(n)
        A = A + 1e-9 * z(n, n)
        return A, np.linalg.inv(A), 1e-13
    if kind == "graded":
        # rows spanning twelve decades (the pivot order follows the scales);
        # A = D M with M well conditioned, so A^-1 = M^-1 D^-1 column by column
        d = 10.0 ** rng.uniform(-6, 6, n)
        M = z(n, n)
        return d[:, None] * M, np.linalg.inv(M) / d[None, :], 1e-12
    # complex symmetric E + i eta - H with E near an eigenvalue of H (the
    # block sweep's case): large, nearly real, indefinite
    H = rng.standard_normal((n, n))
    H = H + H.T
    w = np.linalg.eigvalsh(H)
    A = (w[n // 2] + 1e-3 + 1e-5j) * np.eye(n) - H
    return A, np.linalg.inv(A), 1e-15 * np.linalg.cond(A) + 1e-13


@pytest.mark.parametrize("kind", ["antidiagonal", "permutation", "graded", "near_resonance"])
@pytest.mark.parametrize("op", [1, 4])
def test_block_inverse_pivoting_cases(dev, op, kind):
    # the four-wavefront (op 1) and one-wavefront (op 4) pivoted Gauss-Jordan
    # inverses on matrices whose pivots are forced or graded, column by column
    # (each column of A^-1 has its own scale)
    import ctypes as C
    from electrostatic_amd import _lib
    A, ref, tol = _pivot_case(kind, np.random.default_rng(11))
    planes = lambda M: np.ascontiguousarray(np.stack([M.real, M.imag]).ravel())
    a = planes(A)
    out = np.zeros(2048)
    _lib.check(_lib.load().negf_block_probe(op, a.ctypes.data_as(C.c_void_p), a.ctypes.data_as(C.c_void_p),
                                            out.ctypes.data_as(C.c_void_p)))
    Cm = out[:1024].reshape(32, 32) + 1j * out[1024:].reshape(32, 32)
    err = max(np.abs(Cm[:, j] - ref[:, j]).max() / np.abs(ref[:, j]).max() for j in range(32))
    assert err < tol, (err, tol)
