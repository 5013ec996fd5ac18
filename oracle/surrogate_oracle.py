"""CPU model of the BF16 charge surrogate, test infrastructure only: the
GPU surrogate's forward, tangent and parameter gradient are checked against
it.  It is never part of the product path.

The surrogate has no counterpart in the reference (it is a new component
next to the NEGF charge path); this module fixes its mathematics:

  x0[n] = (2 U_n, exp(-min(n, N-1-n)/16), f, 0 ...)          32 channels
  x_{l+1} = tanh(conv9(x_l; W_l) + b_l)   l = 0..4          64 channels
  q = conv9(x_5; W_5)[:, 0] + b_5[0]
  conv9(x; W)[n, co] = sum_{tap, ci} W[co, tap, ci] x[n + tap - 4, ci]
  (x = 0 outside 0..N-1), loss = mean((q - (Q - mean_Q)/scale_Q)^2)

Parameters are one flat float32 vector, per layer W [cout][9][cin] then
b [cout], with cin = 32 for layer 0 and cout = 32 for layer 5 (padding
rows/columns are ignored: only 3 input and 1 output channel are real).

With bf16=True the values the GPU stores in bf16 are rounded to bf16
(round to nearest even) at the same points: weights, the input features,
each layer's activation, and the back-propagated / tangent signals.  Sums
are exact (float64) here and FP32 on the GPU.

Pinned by: a central finite-difference check of the float64 gradient and
tangent (tests/test_oracle.py), so that a dropped term or a wrong index in
the back-propagation formulas fails on the CPU.
"""
import numpy as np

KW = 9
HALO = 4
C = 64
CP = 32
LAYERS = 6
IN_REAL = 3


def cin(l):
    return CP if l == 0 else C


def cout(l):
    return CP if l == LAYERS - 1 else C


def real_cin(l):
    return IN_REAL if l == 0 else C


def real_cout(l):
    return 1 if l == LAYERS - 1 else C


def layout():
    """[(w_offset, b_offset)] per layer and the total parameter count."""
    out, off = [], 0
    for l in range(LAYERS):
        wo = off
        off += cout(l) * KW * cin(l)
        out.append((wo, off))
        off += cout(l)
    return out, off


def num_params():
    return layout()[1]


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return r.astype(np.uint32).view(np.float32).astype(np.float64).reshape(a.shape)


def unpack(params):
    """Real-channel weights W_l [cout_r, 9, cin_r] and biases b_l (float64)."""
    lay, n = layout()
    p = np.asarray(params)
    assert p.size == n
    Ws, bs = [], []
    for l, (wo, bo) in enumerate(lay):
        W = p[wo:bo].reshape(cout(l), KW, cin(l))[:real_cout(l), :, :real_cin(l)]
        Ws.append(W.astype(np.float64))
        bs.append(p[bo:bo + cout(l)][:real_cout(l)].astype(np.float64))
    return Ws, bs


def pack(Ws, bs, dtype=np.float32):
    lay, n = layout()
    p = np.zeros(n, dtype)
    for l, (wo, bo) in enumerate(lay):
        W = np.zeros((cout(l), KW, cin(l)), dtype)
        W[:real_cout(l), :, :real_cin(l)] = Ws[l]
        p[wo:bo] = W.ravel()
        b = np.zeros(cout(l), dtype)
        b[:real_cout(l)] = bs[l]
        p[bo:bo + cout(l)] = b
    return p


def conv(x, W):
    """x [B, N, ci], W [co, 9, ci] -> [B, N, co] (zero padding)."""
    B, N, _ = x.shape
    xp = np.zeros((B, N + 2 * HALO, x.shape[2]))
    xp[:, HALO:HALO + N] = x
    z = np.zeros((B, N, W.shape[0]))
    for tap in range(KW):
        z += xp[:, tap:tap + N, :] @ W[:, tap, :].T
    return z


def conv_T(dz, W):
    """Adjoint of conv: dx[m, ci] = sum_{co, tap} W[co, tap, ci] dz[m - tap + 4, co]."""
    B, N, _ = dz.shape
    dp = np.zeros((B, N + 2 * HALO, dz.shape[2]))
    dp[:, HALO:HALO + N] = dz
    dx = np.zeros((B, N, W.shape[2]))
    for tap in range(KW):
        dx += dp[:, KW - 1 - tap:KW - 1 - tap + N, :] @ W[:, tap, :]
    return dx


def features(U, feat=None):
    U = np.atleast_2d(np.asarray(U, np.float64))
    B, N = U.shape
    n = np.arange(N)
    d = np.minimum(n, N - 1 - n)
    x = np.zeros((B, N, IN_REAL))
    x[:, :, 0] = 2.0 * U
    x[:, :, 1] = np.exp(-d / 16.0)[None, :]
    if feat is not None:
        x[:, :, 2] = np.asarray(feat, np.float64)[:, None]
    return x


def forward(params, U, feat=None, bf=True):
    """Normalised output q [B, N] and the activations x_0..x_5."""
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    Ws, bs = unpack(params)
    Ws = [r(W) for W in Ws]
    x = r(features(U, feat))
    acts = [x]
    for l in range(LAYERS - 1):
        x = r(np.tanh(conv(x, Ws[l]) + bs[l]))
        acts.append(x)
    q = conv(x, Ws[-1])[:, :, 0] + bs[-1][0]
    return q, acts


def loss_and_grad(params, U, Q, mean_Q, scale_Q, feat=None, bf=True):
    """Loss and flat parameter gradient (float64, layout of the parameters)."""
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    Ws, bs = unpack(params)
    Wr = [r(W) for W in Ws]
    q, acts = forward(params, U, feat, bf)
    t = (np.asarray(Q, np.float64) - mean_Q) / scale_Q
    n = q.size
    d = q - t
    loss = float(np.mean(d * d))
    dz = np.zeros(q.shape + (1,))
    dz[:, :, 0] = r(2.0 * d / n)
    gW, gb = [None] * LAYERS, [None] * LAYERS
    for l in range(LAYERS - 1, -1, -1):
        x = acts[l]
        B, N, ci = x.shape
        xp = np.zeros((B, N + 2 * HALO, ci))
        xp[:, HALO:HALO + N] = x
        g = np.zeros(Ws[l].shape)
        for tap in range(KW):
            g[:, tap, :] = np.einsum('bnc,bnd->cd', dz, xp[:, tap:tap + N, :])
        gW[l] = g
        gb[l] = dz.sum(axis=(0, 1))
        if l == 0:
            break
        dz = r(conv_T(dz, Wr[l]) * (1.0 - x * x))
    return loss, pack_grad(gW, gb)


def pack_grad(gW, gb):
    lay, n = layout()
    p = np.zeros(n)
    for l, (wo, bo) in enumerate(lay):
        W = np.zeros((cout(l), KW, cin(l)))
        W[:real_cout(l), :, :real_cin(l)] = gW[l]
        p[wo:bo] = W.ravel()
        b = np.zeros(cout(l))
        b[:real_cout(l)] = gb[l]
        p[bo:bo + cout(l)] = b
    return p


def jvp(params, U, V, scale_Q, feat=None, bf=True):
    """dQ/dU . V [B, N] (Q = mean + scale q)."""
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    Ws, bs = unpack(params)
    Wr = [r(W) for W in Ws]
    _, acts = forward(params, U, feat, bf)
    V = np.atleast_2d(np.asarray(V, np.float64))
    tng = np.zeros(V.shape + (IN_REAL,))
    tng[:, :, 0] = r(2.0 * V)
    for l in range(LAYERS - 1):
        tng = r(conv(tng, Wr[l]) * (1.0 - acts[l + 1] ** 2))
    return scale_Q * conv(tng, Wr[-1])[:, :, 0]


def init_params(seed=0, dtype=np.float32):
    """Random parameters with the GPU initialiser's scale (not its values)."""
    rng = np.random.default_rng(seed)
    Ws, bs = [], []
    for l in range(LAYERS):
        lim = np.sqrt(3.0 / (real_cin(l) * KW))
        Ws.append(rng.uniform(-lim, lim, (real_cout(l), KW, real_cin(l))))
        bs.append(rng.uniform(-0.1, 0.1, real_cout(l)))
    return pack(Ws, bs, dtype)
