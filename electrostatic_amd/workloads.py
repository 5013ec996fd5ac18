"""Seeded synthetic workloads (potential profiles, energy grids, tube batches).

These define the inputs of the benchmark and test configurations; the same
arrays feed the GPU path and the CPU checker.
"""
import ctypes as C

import numpy as np

from . import _lib

KB_EV_PER_K = 8.617333262e-5
PI = 3.14159265358979323846


def smooth_potential(N, seed, lo=-0.5, hi=0.5, kmax=8):
    """Sum of kmax random Fourier modes (k = 1..kmax), rescaled to [lo, hi]."""
    rng = np.random.default_rng(seed)
    x = np.arange(N) / N
    U = np.zeros(N)
    for k in range(1, kmax + 1):
        U += rng.normal() * np.cos(2 * np.pi * k * x + rng.uniform(0, 2 * np.pi))
    U = (U - U.min()) / (U.max() - U.min())
    return lo + (hi - lo) * U


def gaussian_bumps(N, seed, n_bumps=4, amp=0.3):
    """Sum of n_bumps Gaussians: centres U(0,N), widths U(N/20,N/5),
    amplitudes U(-amp, amp)."""
    rng = np.random.default_rng(seed)
    n = np.arange(N)
    U = np.zeros(N)
    for _ in range(n_bumps):
        c = rng.uniform(0, N)
        w = rng.uniform(N / 20, N / 5)
        a = rng.uniform(-amp, amp)
        U += a * np.exp(-0.5 * ((n - c) / w) ** 2)
    return U


def gl_path(a, b, n, kind):
    """Gauss-Legendre path from the product library (line: kind 0, arc: 1)."""
    lib = _lib.load()
    E = np.zeros(2 * n)
    w = np.zeros(n)
    mul = np.zeros(2 * n)
    a = complex(a)
    b = complex(b)
    _lib.check(lib.negf_gl_path(_lib.negf_complex(a.real, a.imag),
                                _lib.negf_complex(b.real, b.imag), n, kind,
                                E.ctypes.data_as(C.c_void_p),
                                w.ctypes.data_as(C.POINTER(C.c_double)),
                                mul.ctypes.data_as(C.c_void_p)))
    return E.view(np.complex128), w, mul.view(np.complex128)


def fermi(z, kT):
    return 1.0 / (np.exp(z / kT) + 1.0)


def batched_rgf_workload(N=4096, n_energy=8192, seed=2, T=298.0,
                         E_valence_min=-10.0, E_pole_max=3.0, eta=1e-5,
                         tails=(14.0, 14.0), real_range=(-3.0, 3.0)):
    """Batched-RGF workload: scalar chain of N sites with hoppings alternating
    2.5*cos(pi/17) (even links) and 2.5 eV (odd links), seeded smooth
    potential in [-0.5, 0.5] eV; half the energies on the equilibrium contour
    arc (E_eta -> E_left for mu = 0 at T), half on the real axis with
    Im E = eta.  Charge coefficient per energy:
    coef = -i * (-2/pi) * w * mul * f(E; kT), so sum_k Re(coef G_nn) is
    Im of the reference's RhoEq accumulation (NEGF_Common.cpp:3742-3754).
    """
    U = smooth_potential(N, seed)
    t_even = 2.5 * np.cos(PI / 17)
    t_odd = 2.5
    kT = KB_EV_PER_K * T
    n_poles = int((E_pole_max - PI * kT) / (2 * PI * kT) + 1)
    zeta_im = 2 * n_poles * PI * kT
    E_eta = complex(0.0 - max(tails) * kT, zeta_im)
    E_left = complex(E_valence_min, eta)
    n_arc = n_energy // 2
    n_real = n_energy - n_arc
    Ea, wa, ma = gl_path(E_eta, E_left, n_arc, 1)
    Er, wr, mr = gl_path(complex(real_range[0], eta), complex(real_range[1], eta),
                         n_real, 0)
    E = np.concatenate([Ea, Er])
    w = np.concatenate([wa, wr])
    mul = np.concatenate([ma, mr])
    coef = -1j * (-2.0 / PI) * w * mul * fermi(E, kT)
    return dict(N=N, U=U, t_even=t_even, t_odd=t_odd, Uc=(U[0], U[-1]), E=E,
                coef=coef, kT=kT)


def tube_batch(n_tubes=64, lo=500, hi=20000, seed=3, bias_max=0.5):
    """Uneven tube batch: lengths log-uniform in [lo, hi] (rounded), per-tube
    bias uniform in [0, bias_max] V."""
    rng = np.random.default_rng(seed)
    N = np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), n_tubes))).astype(int)
    bias = rng.uniform(0, bias_max, n_tubes)
    return N, bias


def zigzag_ring_blocks(m=16, gamma=2.5):
    """Nearest-neighbour pi-orbital Hamiltonian of an (m,0) zigzag tube cut
    into blocks of two rings (2m orbitals), hoppings -gamma.

    Within a block, atom j of the first ring bonds to atoms j and j-1 of the
    second ring (the zigzag bonds).  Between blocks, atom j of a block's
    second ring bonds axially to one atom of the next block's first ring;
    numbering the atoms of every other block by the mirrored angle makes
    that partner atom -j mod m, so every block has the same on-site matrix
    H0 and the same coupling T = H_{b,b+1}.  Returns (H0, T), real 2m x 2m.
    """
    zig = np.eye(m) + np.roll(np.eye(m), -1, axis=1)     # j -> j and j-1
    H0 = np.zeros((2 * m, 2 * m))
    H0[:m, m:] = -gamma * zig
    H0[m:, :m] = -gamma * zig.T
    T = np.zeros((2 * m, 2 * m))
    j = np.arange(m)
    T[m + j, (-j) % m] = -gamma
    return H0, T
