"""Independent pins of the oracle functions that the reference's own
fixtures do not reach: the non-equilibrium charge assembly
(Compute_RhoNonEq, NEGF_Common.cpp:3377-3497), the adaptive intercuts
(Find_NonEq_Percent_Intercuts_Adaptively, :2062-2185) and the Broyden
history beyond the first step (Broyden_Second_Parallel.cpp:9-211).

Each check is built from something the oracle's code does not share:
numpy dense inverses, numpy Gauss-Legendre nodes, the quadratic equation
of the surface Green's function solved by numpy.roots, a textbook
inverse-Jacobian form of Broyden's second method, and numbers derived by
hand from the reference's formulas.
"""
import numpy as np
import pytest

from oracle import oracle as O

KB = 8.617333262e-5          # WarpXConst.H:61-62, eV/K
BETA, GAMMA = 2.2286917788826917, 2.5   # (17,0) lowest mode, CNT.cpp:288-292


def lead_sigma_numpy(E, Uc, beta=BETA, gamma=GAMMA):
    """Sigma = gamma^2 g with g the root of gamma^2 z g^2 - F g + z = 0
    (z = E - Uc, F = z^2 + gamma^2 - beta^2) that has Im g < 0."""
    z = E - Uc
    F = z * z + gamma ** 2 - beta ** 2
    r = np.roots([gamma ** 2 * z, -F, z])
    g = [x for x in r if x.imag < 0]
    assert len(g) == 1
    return gamma ** 2 * g[0]


def dense_G(E, U, Uc, beta=BETA, gamma=GAMMA):
    N = U.size
    t = np.where(np.arange(N - 1) % 2 == 0, beta, gamma)
    H = np.diag(U.astype(complex)) - np.diag(t, 1) - np.diag(t, -1)
    A = E * np.eye(N) - H
    SL, SR = lead_sigma_numpy(E, Uc[0], beta, gamma), lead_sigma_numpy(E, Uc[1], beta, gamma)
    A[0, 0] -= SL
    A[-1, -1] -= SR
    return np.linalg.inv(A), SL, SR


def fermi(z, kT):
    return 1.0 / (np.exp(z / kT) + 1.0)


def test_noneq_assembly_matches_dense_inverse_integral():
    """RhoNonEq_n = sum_E (-s/2pi) w mul g sum_k |G_nk|^2 Gamma_k f_k(E) on a
    biased 14-site chain, with G from a dense inverse, Sigma from the
    quadratic, GL nodes from numpy, and the window [mu_min - 14 kT,
    mu_max + 14 kT] + i eta of Define_EnergyLimits (NEGF_Common.cpp:1735-1745).
    Also pins the integrand samples at N/2, N/4, 3N/4 (:3457-3495)."""
    N, deg, spin, eta, Tk = 14, 2, 2, 1e-5, 298.0
    mu = (-0.9, -1.25)
    U = 0.15 * np.cos(np.arange(N) * 0.7) - 0.1
    Uc = (U[0], U[-1])
    npts = 96
    t = O.make_tube(N, [BETA], [deg], mu=mu, eq_pts=(30, 30, 64), noneq_pts=(npts,), eta=eta)
    got, integ, st = O.rho_noneq(t, U, Uc, with_integrand=True)
    assert st == 0
    kT = KB * Tk
    a = min(mu) - 14 * kT + 1j * eta
    b = max(mu) + 14 * kT + 1j * eta
    x, w = np.polynomial.legendre.leggauss(npts)
    E = (b - a) / 2 * x + (b + a) / 2
    mul = (b - a) / 2
    ref = np.zeros(N, complex)
    ref_int = np.zeros((3, npts))
    for e in range(npts):
        G, SL, SR = dense_G(E[e], U, Uc)
        GamL, GamR = -2 * SL.imag, -2 * SR.imag          # i (Sigma - Sigma^+)
        fL, fR = fermi(E[e] - mu[0], kT), fermi(E[e] - mu[1], kT)
        AnF = np.abs(G[:, 0]) ** 2 * GamL * fL + np.abs(G[:, -1]) ** 2 * GamR * fR
        ref += (-spin / (2 * np.pi)) * AnF * w[e] * mul * deg
        v = ((-spin / (2 * np.pi)) * AnF * deg).real
        for k, n in enumerate((N // 2, N // 4, 3 * N // 4)):
            ref_int[k, e] = v[n]
    # 1e-10: the reference's Gauss-Legendre weights stop their Newton
    # iteration at 3e-11 (CodeUtil.cpp:640-684), numpy's do not; the nodes
    # are in the same ascending order
    assert np.abs(got.real - ref.real).max() < 1e-10 * np.abs(ref.real).max()
    assert np.abs(integ - ref_int).max() < 1e-10 * np.abs(ref_int).max()


def test_noneq_split_reproduces_equilibrium_charge():
    """With mu_R = mu_L + 2e-8 eV the reference switches to its
    non-equilibrium layout (flag_noneq, NEGF_Common.cpp:1721): contour only
    up to E_cr = mu - 14 kT without poles, then the real axis with
    sum_k A^k f_k.  At equal chemical potentials sum_k A^k = i(G - G^+) -
    2 eta G G^+, so the split must reproduce the equilibrium charge (contour
    + poles) up to O(eta) and the Fermi tail of the contour leg; at the
    reference's default eta = 1e-8 (NEGF_Common.H:386) both are below 1e-7."""
    N, mu = 24, -1.0
    U = 0.2 * np.sin(np.arange(N) / N * np.pi)
    Uc = (U[0], U[-1])
    te = O.make_tube(N, [BETA], [2], mu=(mu, mu), eq_pts=(30, 30, 64), noneq_pts=(200,), eta=1e-8)
    tn = O.make_tube(N, [BETA], [2], mu=(mu, mu + 2e-8), eq_pts=(30, 30, 64), noneq_pts=(200,), eta=1e-8)
    L = O.energy_limits(tn, (mu, mu + 2e-8))
    assert L.flag_noneq == 1 and L.num_poles == 0
    r0, _ = O.rho0(te)
    _, ce, _, _ = O.solve_negf(te, U, Uc, r0)
    _, cn, _, _ = O.solve_negf(tn, U, Uc, r0)
    assert np.abs(ce[1]).max() == 0.0            # no real-axis part in equilibrium
    assert np.abs(cn[1]).max() > 0.05             # the split puts charge there
    eq = ce[0] + ce[1]
    split = cn[0] + cn[1]
    assert np.abs(eq - split).max() < 1e-7
    assert np.abs(eq).max() > 1.0


# ---------------------------------------------------------------- adaptive intercuts

def _intercut_tube(N, mu, npts, intercuts, window):
    t = O.make_tube(N, [BETA], [2], mu=mu, eq_pts=(30, 30, 64), noneq_pts=npts,
                    intercuts=intercuts, eta=1e-5, kT_window=window)
    return t


def _state(t):
    n = t.num_noneq_paths
    return ([t.noneq_pts[i] for i in range(n)], [t.noneq_intercuts[i] for i in range(n - 1)])


def test_adaptive_intercuts_window_clamped_hand_derived():
    """Windows wider than the integration range clamp to E_min and E_max
    (NEGF_Common.cpp:2134-2140): cuts (0, 100); the middle path's points are
    (E_max - E_min) * d1 rounded to even with d1 = 40 / (0.5 (E_max - E_min))
    per kT from the first call, stored in the reference's int vector
    (NEGF_Common.H:404), i.e. truncated: E_max - E_min = 28 + 0.2/kT = 35.79 kT,
    d1 = int(2.235) = 2, points 2 round(35.79) = 72 (80 without the
    truncation); the outer paths get 0."""
    N = 32
    U = np.zeros(N)
    kT = KB * 298.0
    t = _intercut_tube(N, (0.0, -0.2), (20, 40, 20), (25.0, 75.0), (1000.0, 1000.0))
    O.adaptive_intercuts(t, U, (0.0, 0.0))
    pts, cuts = _state(t)
    span_kT = 28 + 0.2 / kT
    assert int(40 / (0.5 * span_kT)) == 2
    assert pts == [0, 2 * round(span_kT * 2 / 2), 0] == [0, 72, 0]
    assert cuts[0] == 0.0 and abs(cuts[1] - 100.0) < 1e-12


def test_adaptive_intercuts_given_density():
    """noneq_integration_pts_density given as input (NEGF_Common.cpp:537-539)
    replaces the densities the first correction would compute: with (3, 5, 3)
    points per kT and the clamped window of the test above, the middle path
    gets 2 round(35.79 * 5 / 2) = 178 points."""
    N = 32
    U = np.zeros(N)
    kT = KB * 298.0
    t = O.make_tube(N, [BETA], [2], mu=(0.0, -0.2), eq_pts=(30, 30, 64), noneq_pts=(20, 40, 20),
                    intercuts=(25.0, 75.0), eta=1e-5, kT_window=(1000.0, 1000.0), density=(3, 5, 3))
    O.adaptive_intercuts(t, U, (0.0, 0.0))
    pts, _ = _state(t)
    assert pts == [0, 2 * round((28 + 0.2 / kT) * 5 / 2), 0] == [0, 178, 0]


def test_adaptive_intercuts_interior_window_hand_derived():
    """Window of +-1 kT around the integrand maximum at site N/2.  The
    maximum is taken from a dense-inverse integrand (independent of the
    oracle's RGF); cuts and points follow from the reference's formulas
    written here in absolute energies: cut = 100 (E_cut - E_cr) /
    (E_rm - E_cr), points = 2 round(len * n_p / (2 len_p)) with the lengths
    of the first call's paths."""
    N, mu, npts, cuts0 = 16, (0.0, -0.2), (20, 40, 20), (25.0, 75.0)
    U = 0.1 * np.cos(np.arange(N))
    Uc = (U[0], U[-1])
    t = _intercut_tube(N, mu, npts, cuts0, (1.0, 1.0))
    O.adaptive_intercuts(t, U, Uc)
    pts, cuts = _state(t)
    kT = KB * 298.0
    Ecr, Erm = min(mu) - 14 * kT, max(mu) + 14 * kT
    span = Erm - Ecr
    # the reference samples the integrand on the present (initial) paths
    bounds = [Ecr, Ecr + 0.25 * span, Ecr + 0.75 * span, Erm]
    best, E_star = -1.0, None
    for p in range(3):
        x, _ = np.polynomial.legendre.leggauss(npts[p])
        for xe in x:
            E = (bounds[p + 1] - bounds[p]) / 2 * xe + (bounds[p + 1] + bounds[p]) / 2 + 1e-5j
            G, SL, SR = dense_G(E, U, Uc)
            v = abs(((-2 / (2 * np.pi)) * 2 * (np.abs(G[N // 2, 0]) ** 2 * (-2 * SL.imag) * fermi(E - mu[0], kT)
                                               + np.abs(G[N // 2, -1]) ** 2 * (-2 * SR.imag)
                                               * fermi(E - mu[1], kT))).real)
            if v > best:
                best, E_star = v, E.real
    assert abs(t.E_at_max - E_star) < 1e-12
    lo = max(E_star - kT, Ecr)      # E* measured from mu_0 = 0 (kT0 = kT)
    hi = min(E_star + kT, Erm)
    assert abs(cuts[0] - 100 * (lo - Ecr) / span) < 1e-9
    assert abs(cuts[1] - 100 * (hi - Ecr) / span) < 1e-9
    # densities per kT of the first call's paths, truncated to int as the
    # reference stores them (NEGF_Common.H:404)
    fr = [0.25, 0.5, 0.25]
    d = [int(npts[i] / (fr[i] * span / kT)) for i in range(3)]
    want = [2 * round((lo - Ecr) / kT * d[0] / 2),
            2 * round((hi - lo) / kT * d[1] / 2),
            2 * round((Erm - hi) / kT * d[2] / 2)]
    assert pts == want
    assert all(p % 2 == 0 for p in pts)


def test_adaptive_intercuts_densities_frozen_at_first_call():
    """The point densities per kT are set at the first correction and kept
    (NEGF_Common.cpp:2078-2100): after the bias grows from 0.2 to 0.4 V the
    clamped middle path covers a range 28 + 0.4/kT kT wide at the first
    call's density int(40 / (0.5 (28 + 0.2/kT))) = 2 per kT, not 72 points
    again."""
    N = 32
    U = np.zeros(N)
    kT = KB * 298.0
    t = _intercut_tube(N, (0.0, -0.2), (20, 40, 20), (25.0, 75.0), (1000.0, 1000.0))
    O.adaptive_intercuts(t, U, (0.0, 0.0))
    assert _state(t)[0] == [0, 72, 0]
    t.mu[1] = -0.4
    O.adaptive_intercuts(t, U, (0.0, 0.0))
    d1 = int(40 / (0.5 * (28 + 0.2 / kT)))
    want_mid = 2 * round((28 + 0.4 / kT) * d1 / 2)
    assert _state(t)[0] == [0, want_mid, 0]
    assert want_mid != 72


def test_adaptive_intercuts_two_paths_hand_derived():
    """Two paths: only the upper window edge moves the cut, unclamped
    (NEGF_Common.cpp:2160-2183): with a window of 1e6 kT the cut lands far
    beyond 100 % and the second path's points come out negative-even."""
    N = 32
    U = np.zeros(N)
    kT = KB * 298.0
    t = _intercut_tube(N, (0.0, -0.2), (30, 30), (50.0,), (1.0, 1e6))
    O.adaptive_intercuts(t, U, (0.0, 0.0))
    pts, cuts = _state(t)
    span_kT = 28 + 0.2 / kT
    E_star = t.E_at_max / kT
    E_min = -0.2 / kT - 14
    E_up = E_star + 1e6
    assert abs(cuts[0] - 100 * (E_up - E_min) / span_kT) < 1e-6 * abs(cuts[0])
    d0 = int(30 / (0.5 * span_kT))
    d1 = int(30 / (0.5 * span_kT))
    assert pts[0] == 2 * round((E_up - E_min) * d0 / 2)
    assert pts[1] == 2 * round((14 - E_up) * d1 / 2)
    assert pts[1] < 0 and pts[1] % 2 == 0


# ---------------------------------------------------------------- Broyden

def _textbook_broyden(n0, G, alpha, steps):
    """Broyden's second method in inverse-Jacobian form with H_0 = alpha I:
    F = n - G(n);  H_m = H_{m-1} + (dn - H_{m-1} dF) dF^T / (dF^T dF);
    n <- n - H_m F."""
    n = n0.copy()
    H = alpha * np.eye(n.size)
    F_old = n_old = None
    hist = [n.copy()]
    for _ in range(steps):
        F = n - G(n)
        if F_old is not None:
            dF = F - F_old
            dn = n - n_old
            H = H + np.outer(dn - H @ dF, dF) / (dF @ dF)
        F_old, n_old = F, n.copy()
        n = n - H @ F
        hist.append(n.copy())
    return hist


def _linear_map(S, seed, lam):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(S, S)))
    A = Q @ np.diag(lam) @ Q.T
    c = rng.normal(size=S)
    return A, c


def test_broyden_iterates_match_textbook_inverse_update():
    """The reference keeps the rank-one terms as V/W histories; the iterates
    must equal the explicit-matrix form of the same method on a nonlinear
    map, step by step (a slipped history index or sign changes them)."""
    S = 6
    rng = np.random.default_rng(11)
    B = 0.3 * rng.normal(size=(S, S))
    c = rng.normal(size=S)

    def G(n):
        return np.tanh(B @ n) + c

    alpha = 0.3
    b = O.Broyden(S, alpha, 50, 0)
    b.set_nin(np.zeros(S))
    got = [b.get_nin()]
    for _ in range(9):
        b.step(G(got[-1]))
        got.append(b.get_nin())
    want = _textbook_broyden(np.zeros(S), G, alpha, 9)
    for k, (g, w) in enumerate(zip(got, want)):
        assert np.abs(g - w).max() < 1e-12 * max(1.0, np.abs(w).max()), k


def test_broyden_terminates_where_simple_mixing_diverges():
    """Linear map n -> A n + c with eigenvalues of A in {-30, -12, 0.5, 2.5}:
    simple mixing with fraction 0.1 diverges (|1 - 0.1 (1 - lam)| > 1 for
    lam = -30, -12), Broyden's method on a linear map of dimension S reaches
    the fixed point within 2S steps (Gay 1979)."""
    S = 4
    A, c = _linear_map(S, 5, [-30.0, -12.0, 0.5, 2.5])
    x_star = np.linalg.solve(np.eye(S) - A, c)
    alpha = 0.1
    n = np.zeros(S)
    for _ in range(40):
        n = n - alpha * (n - (A @ n + c))
    assert np.abs(n - x_star).max() > 1e3          # simple mixing diverges
    b = O.Broyden(S, alpha, 50, 0)
    b.set_nin(np.zeros(S))
    reached = None
    for k in range(2 * S + 1):
        nin = b.get_nin()
        mx, _, _, st = b.step(A @ nin + c)
        assert st == 0
        if mx < 1e-10:
            reached = k + 1
            break
    assert reached is not None and reached <= 2 * S + 1
    assert np.abs(b.get_nin() - x_star).max() < 1e-8


@pytest.mark.parametrize("norm_type", [0, 1])
def test_broyden_secant_condition_on_linear_map(norm_type):
    """After update m the implied inverse Jacobian maps the latest residual
    difference onto the latest input difference: with the next step
    n_{m+1} = n_m - H_m F_m and H_m dF_m = dn_m, the same H_m applied to
    F_m - dF_m = F_{m-1} gives n_m - n_{m+1} - dn_m.  On a linear map the
    textbook form checks this for every m."""
    S = 5
    A, c = _linear_map(S, 9, [-3.0, -1.0, 0.2, 0.6, 1.4])
    alpha = 0.2
    b = O.Broyden(S, alpha, 50, norm_type)
    b.set_nin(np.full(S, 0.5))
    ns = [b.get_nin()]
    for _ in range(6):
        b.step(A @ ns[-1] + c)
        ns.append(b.get_nin())
    want = _textbook_broyden(np.full(S, 0.5), lambda n: A @ n + c, alpha, 6)
    for g, w in zip(ns, want):
        assert np.abs(g - w).max() < 1e-11 * max(1.0, np.abs(w).max())
    # secant condition of the textbook matrices (built from the oracle's iterates)
    H = alpha * np.eye(S)
    F = [n - (A @ n + c) for n in ns]
    for m in range(1, 6):
        dF, dn = F[m] - F[m - 1], ns[m] - ns[m - 1]
        H = H + np.outer(dn - H @ dF, dF) / (dF @ dF)
        assert np.abs(H @ dF - dn).max() < 1e-10 * max(1.0, np.abs(dn).max())
        assert np.abs(ns[m + 1] - (ns[m] - H @ F[m])).max() < 1e-10 * max(1.0, np.abs(ns[m]).max())


def test_product_block_generator_equals_oracle_blocks():
    """The benchmark builds its dense-block inputs with the product's own
    generator; it must give the matrices the block oracle is pinned on."""
    from electrostatic_amd.workloads import zigzag_ring_blocks
    from oracle.block_oracle import zigzag_blocks
    for m in (4, 16):
        a, b = zigzag_ring_blocks(m, 2.5), zigzag_blocks(m, 2.5)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
