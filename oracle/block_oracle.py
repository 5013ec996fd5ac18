"""CPU restatement of the dense-block NEGF path, test infrastructure only:
block recursive Green's functions (diag G^R of a chain of dense m x m
blocks) and Sancho-Rubio decimation for the lead surface Green's function.

The reference names this path but leaves it stubbed: the 2-D MatrixBlock
multiply/inverse/dagger of Matrix_Block_Util.H:123-135, 431-438, 684-692
are empty, c_Graphene aborts (Transport.cpp:302-306), and the decimation
loop DecimationTechnique (NEGF_Common.cpp:4095-4153) only works for the
1-D (diagonal) block.  What is restated here is therefore the textbook
mathematics that those stubs would compute:

  * Sancho-Rubio (M. P. Lopez Sancho, J. M. Lopez Sancho, J. Rubio,
    J. Phys. F 15, 851 (1985)): eps_s, eps = H0; alpha = H_{0,1},
    beta = H_{1,0}; repeat g = (E - eps)^-1, eps_s += alpha g beta,
    eps += alpha g beta + beta g alpha, alpha = alpha g alpha,
    beta = beta g beta until max|alpha|, max|beta| < tol; the surface
    Green's function is (E - eps_s)^-1, with the same loop structure as
    DecimationTechnique (NEGF_Common.cpp:4116-4148) on dense blocks.
  * Block RGF: left-connected g_0 = (E - H_00 - Sigma_L)^-1,
    g_b = (E - H_bb - H_{b,b-1} g_{b-1} H_{b-1,b})^-1 (Sigma_R added on the
    last block), then G_{N-1} = g_{N-1} and
    G_bb = g_b + g_b H_{b,b+1} G_{b+1,b+1} H_{b+1,b} g_b, the block form of
    the scalar sweep of Compute_RhoEq (NEGF_Common.cpp:3667-3714).

Pinned by (tests/test_oracle.py): the mode-space reduction of a zigzag
tube (ring-summed block G equals the sum over angular modes of scalar
chains whose leads use the reference's closed-form surface Green's
function, CNT.cpp:354-393), and a brute-force dense inverse of the whole
device with the leads' self-energies.
"""
import numpy as np


def zigzag_blocks(m=16, gamma=2.5):
    """Atomistic (m,0) tube in blocks of two rings (2m orbitals).

    Ring angles (units 2 pi / m): ring 4k at j, 4k+1 and 4k+2 at j + 1/2,
    4k+3 at j.  Even blocks (rings 4k, 4k+1) index atoms by angle; odd
    blocks (rings 4k+2, 4k+3) by the reflected angle (1/2 - j, then -j), so
    that every block has the same on-site matrix H0 (the zigzag bonds of a
    ring pair: atom j of the first ring to atoms j and j-1 of the second)
    and every block-to-block coupling T = H_{b,b+1} is the same (an axial
    bond from atom j of the second ring to atom -j mod m of the next
    block's first ring).  Hoppings are -gamma.
    """
    C = np.zeros((m, m))
    for j in range(m):
        C[j, j] = 1.0
        C[j, (j - 1) % m] = 1.0
    H0 = np.zeros((2 * m, 2 * m))
    H0[:m, m:] = -gamma * C
    H0[m:, :m] = -gamma * C.T
    T = np.zeros((2 * m, 2 * m))
    for j in range(m):
        T[m + j, (-j) % m] = -gamma
    return H0, T


def sancho_rubio(E, H0, alpha, beta, tol=1e-12, max_iter=200):
    """Surface Green's function of a semi-infinite lead whose surface
    layer couples to the next layer by alpha (= H_{0,1}); returns (g, iters)."""
    n = H0.shape[0]
    I = np.eye(n)
    eps_s = H0.astype(np.complex128)
    eps = eps_s.copy()
    a = alpha.astype(np.complex128)
    b = beta.astype(np.complex128)
    it = 0
    for it in range(1, max_iter + 1):
        g = np.linalg.inv(E * I - eps)
        ga = g @ a
        gb = g @ b
        agb = a @ gb
        bga = b @ ga
        eps_s = eps_s + agb
        eps = eps + agb + bga
        a = a @ ga
        b = b @ gb
        if max(np.abs(a).max(), np.abs(b).max()) < tol:
            break
    return np.linalg.inv(E * I - eps_s), it


def lead_self_energies(E, H0, T, Uc=(0.0, 0.0), tol=1e-12, max_iter=200):
    """Sigma_L = T^T g_L T (lead to the left, surface = its last block) and
    Sigma_R = T g_R T^T (lead to the right, surface = its first block), for
    leads with on-site H0 shifted by the contact potentials Uc."""
    gL, itL = sancho_rubio(E - Uc[0], H0, T.T, T, tol, max_iter)
    gR, itR = sancho_rubio(E - Uc[1], H0, T, T.T, tol, max_iter)
    return T.T @ gL @ T, T @ gR @ T.T, (itL, itR)


def block_rgf_diag(E, H0, T, v, SigL, SigR):
    """diag G^R per block [nb, n] of the chain with on-site H0 + diag(v[b])."""
    nb, n = v.shape
    I = np.eye(n)
    gs = []
    gp = None
    for b in range(nb):
        A = E * I - H0 - np.diag(v[b])
        A = A - (SigL if b == 0 else T.T @ gp @ T)
        if b == nb - 1:
            A = A - SigR
        gp = np.linalg.inv(A)
        gs.append(gp)
    out = np.zeros((nb, n), np.complex128)
    G = gs[-1]
    out[-1] = np.diag(G)
    for b in range(nb - 2, -1, -1):
        X = gs[b] @ T
        G = gs[b] + X @ G @ X.T
        out[b] = np.diag(G)
    return out


def dense_diag(E, H0, T, v, SigL, SigR):
    """Brute force: diag of (E - H - Sigma)^-1 over the whole device."""
    nb, n = v.shape
    M = nb * n
    H = np.zeros((M, M), np.complex128)
    for b in range(nb):
        H[b * n:(b + 1) * n, b * n:(b + 1) * n] = H0 + np.diag(v[b])
        if b + 1 < nb:
            H[b * n:(b + 1) * n, (b + 1) * n:(b + 2) * n] = T
            H[(b + 1) * n:(b + 2) * n, b * n:(b + 1) * n] = T.T
    H[:n, :n] += SigL
    H[-n:, -n:] += SigR
    G = np.linalg.inv(E * np.eye(M) - H)
    return np.diag(G).reshape(nb, n)


def accumulate(E_list, coef, H0, T, v, Uc=(0.0, 0.0), tol=1e-12, max_iter=200):
    """sum_e Im(coef_e G_e,oo) per orbital [nb, n], the block analogue of
    the scalar charge accumulation (NEGF_Common.cpp:3742-3754)."""
    acc = np.zeros(v.shape)
    for E, c in zip(E_list, coef):
        SL, SR, _ = lead_self_energies(E, H0, T, Uc, tol, max_iter)
        acc += np.imag(c * block_rgf_diag(E, H0, T, v, SL, SR))
    return acc
