"""GPU checks of the BF16 charge surrogate (negf_surrogate_*) against the CPU
model in oracle/surrogate_oracle.py with the same parameters and inputs.

The CPU model rounds to bf16 at the points where the GPU stores bf16 and
sums exactly; the GPU sums in FP32 on MFMA, so a few bf16 roundings (2^-8
relative) can differ.  Tolerances: forward and tangent 2e-2 of the output
scale, parameter gradient 5e-2 relative (2-norm) per layer.  An operand or
lane-layout mistake in an MFMA tile gives O(1) errors.
"""
import numpy as np
import pytest
import torch

from electrostatic_amd.surrogate import Surrogate
from oracle import surrogate_oracle as S

pytestmark = pytest.mark.gpu


def smooth_U(B, N, seed):
    rng = np.random.default_rng(seed)
    n = np.arange(N)
    U = np.zeros((B, N))
    for b in range(B):
        for _ in range(4):
            c, w, a = rng.uniform(0, N), rng.uniform(N / 20, N / 5), rng.uniform(-0.4, 0.4)
            U[b] += a * np.exp(-0.5 * ((n - c) / w) ** 2)
        U[b] += rng.uniform(-0.3, 0.3)
    return U


@pytest.mark.parametrize("N,B", [(200, 3), (20, 2), (64, 1)])
def test_forward_matches_cpu_model(dev, N, B):
    sg = Surrogate(N, seed=1)
    p = S.init_params(seed=N)
    norm = np.array([0.02, 0.3])
    sg.set_params(p, norm)
    U = smooth_U(B, N, seed=N)
    f = np.linspace(0.0, 0.5, B)
    Qg = sg.forward(U, f).cpu().numpy()
    q, _ = S.forward(p, U, f, bf=True)
    Qc = norm[0] + norm[1] * q
    err = np.abs(Qg - Qc).max() / np.abs(Qc - norm[0]).max()
    assert err < 2e-2, err
    # without the feature
    Qg0 = sg.forward(U).cpu().numpy()
    Qc0 = norm[0] + norm[1] * S.forward(p, U, None, bf=True)[0]
    assert np.abs(Qg0 - Qc0).max() / np.abs(Qc0 - norm[0]).max() < 2e-2
    sg.close()


def test_gradient_matches_cpu_model(dev):
    N, B = 150, 4
    sg = Surrogate(N, seed=2)
    p = S.init_params(seed=7)
    norm = np.array([0.01, 0.2])
    sg.set_params(p, norm)
    U = smooth_U(B, N, seed=3)
    f = np.array([0.0, 0.1, 0.3, 0.5])
    rng = np.random.default_rng(4)
    Q = norm[0] + norm[1] * rng.standard_normal((B, N))
    lg, gg = sg.gradient(U, Q, f)
    lc, gc = S.loss_and_grad(p, U, Q, norm[0], norm[1], f, bf=True)
    assert abs(lg - lc) / lc < 1e-2
    lay, _ = S.layout()
    for l, (wo, bo) in enumerate(lay):
        end = bo + S.cout(l)
        a, b = gg[wo:end].astype(np.float64), gc[wo:end]
        e = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert e < 5e-2, (l, e)
        # padding entries carry no gradient
        W = np.zeros((S.cout(l), S.KW, S.cin(l)), bool)
        W[S.real_cout(l):] = True
        W[:, :, S.real_cin(l):] = True
        assert np.all(gg[wo:bo][W.ravel()] == 0.0)
    sg.close()


def test_jvp_matches_cpu_model_and_differences(dev):
    N, B = 180, 2
    sg = Surrogate(N, seed=3)
    p = S.init_params(seed=11)
    norm = np.array([0.0, 0.25])
    sg.set_params(p, norm)
    U = smooth_U(B, N, seed=5)
    V = smooth_U(B, N, seed=6)
    jg = sg.jvp(U, V).cpu().numpy()
    jc = S.jvp(p, U, V, norm[1], None, bf=True)
    assert np.abs(jg - jc).max() / np.abs(jc).max() < 2e-2
    # a central difference of the float64 model (bf16 rounding aside)
    h = 1e-4
    fd = norm[1] * (S.forward(p, U + h * V, None, bf=False)[0] - S.forward(p, U - h * V, None, bf=False)[0]) / (2 * h)
    assert np.abs(jg - fd).max() / np.abs(fd).max() < 5e-2
    sg.close()


def test_training_reduces_loss(dev):
    # a local, smooth target: Q = -0.2 tanh(3 (G * U)) + 0.05 f, G a Gaussian window
    N, Sn = 256, 256
    U = smooth_U(Sn, N, seed=8)
    f = np.random.default_rng(9).uniform(0, 0.5, Sn)
    k = np.exp(-0.5 * (np.arange(-12, 13) / 4.0) ** 2)
    k /= k.sum()
    Q = np.stack([-0.2 * np.tanh(3 * np.convolve(u, k, mode="same")) for u in U]) + 0.05 * f[:, None]
    sg = Surrogate(N, seed=4)
    sg.fit_output_scale(Q)
    l0, _ = sg.gradient(U[:64], Q[:64], f[:64])
    sg.train(U, Q, f, epochs=40, batch=16, lr=3e-3, lr_end=3e-4, seed=1, fit_scale=False)
    l1, _ = sg.gradient(U[:64], Q[:64], f[:64])
    w, _ = sg.get_params()
    assert np.all(np.isfinite(w))
    assert l1 < 0.05 * l0, (l0, l1)
    sg.close()
