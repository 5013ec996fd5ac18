"""Pins for the CPU oracle (runs without a GPU).

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign or index, a transposed operand) fails it: golden outputs
committed in the reference, library routines (numpy), closed forms, dense
brute force, and physical invariants.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
GAMMA = 2.5
BETA6 = 2 * GAMMA * np.cos(6 * np.pi / 17)          # (17,0) mode J=6
KT = 8.617333262e-5 * 298.0


def load(name):
    return np.loadtxt(os.path.join(GOLD, name), skiprows=1)


def exact_grid(col):
    """The harnesses sample linspace grids but print E to 6 significant
    digits; rebuild the exact energies (a 3e-6 eV shift matters at band
    edges)."""
    E = np.linspace(col[0], col[-1], col.size)
    assert np.abs(E - col).max() < 5e-6
    return E


def dense_G(E, U, te, to, sL, sR):
    N = U.size
    H = np.zeros((N, N))
    for k in range(N - 1):
        H[k, k + 1] = H[k + 1, k] = te if k % 2 == 0 else to
    A = E * np.eye(N) - H - np.diag(U)
    A = A.astype(complex)
    A[0, 0] -= sL
    A[-1, -1] -= sR
    return np.linalg.inv(A)


# ---------------------------------------------------------------- quadrature
@pytest.mark.parametrize("n", [1, 2, 3, 7, 30, 64, 101, 200])
def test_gauss_legendre_matches_numpy(n):
    x, w = O.gauss_legendre(n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert np.abs(x - xr).max() < 1e-14
    # the reference stops Newton at |dz| <= 3e-11 and takes the weight from
    # the derivative at the previous iterate: relative weight errors reach
    # ~1e-9, and ~6e-7 for the outermost nodes at n = 200
    assert (np.abs(w - wr) / wr).max() < (2e-6 if n > 101 else 2e-9)
    assert abs(w.sum() - 2.0) < 1e-9


def test_line_path_integrates_polynomials():
    # GL of degree n integrates z^k exactly for k < 2n along the segment
    a, b = -2.0 + 0.5j, 1.0 + 2.0j
    E, w, mul = O.gl_path(a, b, 12, 0)
    for k in range(0, 20):
        exact = (b ** (k + 1) - a ** (k + 1)) / (k + 1)
        got = np.sum(w * mul * E ** k)
        # exact up to the ~1e-10 weight error of the Newton stopping rule
        assert abs(got - exact) < 1e-9 * max(1.0, abs(exact))


def test_arc_path_geometry_and_integral():
    L = O.energy_limits(O.make_tube(10, [BETA6], [2]), [0.0, 0.0])
    a = complex(L.E_eta.re, L.E_eta.im)
    b = complex(L.E_contour_left.re, L.E_contour_left.im)
    E, w, mul = O.gl_path(a, b, 40, 1)
    # all nodes on the circle centred on the line Im = Im(b) through a and b
    cre = (b.real ** 2 - a.real ** 2 - a.imag ** 2) / (2 * (b.real - a.real))
    c = complex(cre, b.imag)
    R = cre - b.real
    assert np.abs(np.abs(E - c) - R).max() < 1e-12
    # the reference starts the arc at the angle asin(|Im a| / R) measured from
    # the centre's height Im b (IntegrationPath.cpp:97-98): the start point sits
    # Im b above a, on the circle
    th1 = np.arcsin(abs(a.imag) / R)
    if cre > a.real:
        th1 = np.pi - th1
    a0 = c + R * np.exp(1j * th1)
    assert abs(a0.imag - (a.imag + b.imag)) < 1e-12 and abs(a0 - a) < 2e-5
    # analytic integrand: the quadrature reproduces F(b) - F(a0)
    for f, F in ((lambda z: z ** 3, lambda z: z ** 4 / 4),
                 (lambda z: np.exp(z / 3), lambda z: 3 * np.exp(z / 3))):
        got = np.sum(w * mul * f(E))
        assert abs(got - (F(b) - F(a0))) < 1e-9 * abs(F(b) - F(a0))


# ---------------------------------------------------------------- CNT modes
def test_cnt_modes_17_0():
    mode, gap, deg = O.cnt_sorted_modes(17, 0, 1.42e-10, 2.5)
    assert mode[0] == 6 and deg[0] == 2          # J=6 and J=11 share the gap
    assert deg.sum() == 17
    R = 1.42e-10 * np.sqrt(3 * 17 ** 2) / (2 * np.pi)
    assert abs(gap[0] - 1.42e-10 * 2.5 / R) < 1e-12
    assert abs(O.cnt_beta(6, 17, 2.5) - 2.2286917788826917) < 1e-15


def test_ring_positions_zigzag():
    ptd = O.cnt_ring_positions(17, 4, 1.42e-10)
    d = np.diff(ptd)
    acc = 0.142
    # ring spacings acc/2, acc, acc/2, acc, ... (nm)
    assert np.allclose(d[0::2], acc / 2) and np.allclose(d[1::2], acc)
    assert abs(ptd[0] + ptd[-1]) < 0.15  # centred


# ---------------------------------------------------------------- surface GF
@pytest.mark.parametrize("E", [-5.1 + 1e-5j, -2.0 + 1e-5j, -0.1 + 1e-5j, 0.7 + 1e-3j,
                               1.5 + 1e-5j, 4.0 + 1e-5j, -6.0 + 2.0j, 3 + 0.2j])
def test_surface_gf_root_and_sign(E):
    g, st = O.surface_gf(E, 0.1, BETA6, GAMMA)
    z = E - 0.1
    F = z * z + GAMMA ** 2 - BETA6 ** 2
    assert st == 0
    assert abs(GAMMA ** 2 * z * g * g - F * g + z) < 1e-12 * (abs(F * g) + abs(z))
    assert g.imag < 0     # retarded


def test_surface_gf_is_the_semi_infinite_lead():
    # g is the corner element of a long (truncated) lead with damping eta
    # (the truncation error decays like exp(-2 N Im k); N = 800 at Im E = 0.1
    # leaves < 1e-14)
    E = -1.3 + 0.1j
    g, _ = O.surface_gf(E, 0.0, BETA6, GAMMA)
    N = 800
    G = dense_G(E, np.zeros(N), BETA6, GAMMA, 0.0, 0.0)
    assert abs(G[0, 0] - g) < 1e-12
    # the other end of an even chain ends on the same link: same corner
    assert abs(G[-1, -1] - g) < 1e-12


def test_sigma_pointcontact_fixture():
    d = load("pointcontact_Sigma.txt")
    for E, row in zip(exact_grid(d[:, 0]), d):
        g, st = O.surface_gf(E + 1e-8j, 0.0, BETA6, GAMMA)
        s = GAMMA ** 2 * g
        ref = complex(row[1], row[2])
        # six printed digits per component
        assert abs(s - ref) <= 1e-5 * abs(ref) + 1e-12


def test_sigma_decimation_harness_fixture_below_band():
    # that harness always took the (F-S)/D root; it equals the retarded
    # (Im<0) root for E <= -4 eV only
    d = load("decimation_Sigma.txt")
    for E, row in zip(exact_grid(d[:, 0]), d):
        if E > -4.0 + 1e-9:
            continue
        g, st = O.surface_gf(E + 1e-8j, 0.0, BETA6, GAMMA)
        s = GAMMA ** 2 * g
        ref = complex(row[1], row[2])
        assert abs(s - ref) <= 1e-5 * abs(ref) + 1e-12


@pytest.mark.parametrize("E", [-7.0, -5.5, -5.0, 5.0, 6.5])
def test_decimation_matches_closed_form_outside_band(E):
    # integer tolerance 0 in the reference: fixed max_iter + 2 updates.  The
    # reference takes the dagger (complex conjugate) of the condensed coupling
    # Pi (NEGF_Common.cpp:4113), which is exact only for real Pi: deviations
    # scale with Im E; with 30 layers they are below 1e-10 outside the band
    z = E + 1e-5j
    g_cf, _ = O.surface_gf(z, 0.0, BETA6, GAMMA)
    assert abs(O.decimation_gf(z, BETA6, GAMMA, 30, 30, 0) - g_cf) < 1e-10 * abs(g_cf)
    assert abs(O.decimation_gf(z, BETA6, GAMMA, 2, 30, 0) - g_cf) < 1e-5 * abs(g_cf)


@pytest.mark.parametrize("E", [-4.6, -3.0, -1.0, 0.6, 2.2, 4.5])
def test_decimation_two_layers_in_band(E):
    # in the band the conjugated coupling leaves an O(Im E / level spacing)
    # deviation; with two layers it stays below 1e-3 on this grid, while a
    # dropped or mis-signed term is an O(1) change
    z = E + 1e-5j
    g_cf, _ = O.surface_gf(z, 0.0, BETA6, GAMMA)
    assert abs(O.decimation_gf(z, BETA6, GAMMA, 2, 30, 0) - g_cf) < 1e-3 * abs(g_cf)


# ---------------------------------------------------------------- RGF and G
@pytest.mark.parametrize("N", [2, 3, 8, 33, 64])
def test_rgf_diag_matches_dense_inverse(N):
    rng = np.random.default_rng(N)
    U = rng.uniform(-0.4, 0.4, N)
    for E in (-2.1 + 1e-5j, 0.9 + 0.02j, -7.0 + 3.0j):
        sL = GAMMA ** 2 * O.surface_gf(E, U[0], BETA6, GAMMA)[0]
        sR = GAMMA ** 2 * O.surface_gf(E, U[-1], BETA6, GAMMA)[0]
        _, dg, _ = O.accumulate_diagG(U, BETA6, GAMMA, (U[0], U[-1]), BETA6, GAMMA,
                                      [E], [1.0], want_diag=True)
        G = dense_G(E, U, BETA6, GAMMA, sL, sR)
        assert np.abs(dg[0] - np.diag(G)).max() < 1e-11 * np.abs(np.diag(G)).max()


def test_noneq_walk_matches_dense_offdiagonals():
    N = 24
    rng = np.random.default_rng(3)
    U = rng.uniform(-0.3, 0.3, N)
    E = -1.7 + 1e-5j
    sL = GAMMA ** 2 * O.surface_gf(E, U[0], BETA6, GAMMA)[0]
    sR = GAMMA ** 2 * O.surface_gf(E, U[-1], BETA6, GAMMA)[0]
    G = dense_G(E, U, BETA6, GAMMA, sL, sR)
    acc, _ = O.accumulate_noneq(U, BETA6, GAMMA, (U[0], U[-1]), BETA6, GAMMA, [E],
                                [1.0], [0.0])
    assert np.allclose(acc, -2 * sL.imag * np.abs(G[:, 0]) ** 2, rtol=1e-11, atol=0)
    acc, _ = O.accumulate_noneq(U, BETA6, GAMMA, (U[0], U[-1]), BETA6, GAMMA, [E],
                                [0.0], [1.0])
    assert np.allclose(acc, -2 * sR.imag * np.abs(G[:, -1]) ** 2, rtol=1e-11, atol=0)


# ---------------------------------------------------------------- DOS / T
def test_dos_transmission_pointcontact_fixture_in_band():
    t = O.make_tube(4, [BETA6], [1], spin=1)
    dos_f = load("pointcontact_DOS.txt")
    tr_f = load("pointcontact_Transmission.txt")
    E = exact_grid(dos_f[:, 0])
    band = np.abs(E) > 0.30           # gap edge at |gamma - beta| = 0.2713 eV
    dos, tr, st = O.dos_transmission(t, np.zeros(4), np.zeros(2), E[band] + 1e-8j, 1)
    assert np.allclose(dos, dos_f[band, 1], rtol=2e-5, atol=0)
    assert np.allclose(tr, tr_f[band, 1], rtol=2e-5, atol=1e-12)


def test_dos_decimation_harness_fixture_below_band():
    t = O.make_tube(4, [BETA6], [1], spin=1)
    dos_f = load("decimation_DOS.txt")
    tr_f = load("decimation_Transmission.txt")
    Eg = exact_grid(dos_f[:, 0])
    sel = Eg <= -4.0 + 1e-9
    E = Eg[sel] + 1e-8j
    dos, tr, st = O.dos_transmission(t, np.zeros(4), np.zeros(2), E, 1)
    inband = dos_f[sel, 1] > 1e-3
    assert np.allclose(dos[inband], dos_f[sel, 1][inband], rtol=2e-5)
    assert np.allclose(tr[inband], tr_f[sel, 1][inband], rtol=2e-5)


def test_dos_all_modes_fixture():
    mode, gap, deg = O.cnt_sorted_modes(17, 0, 1.42e-10, 2.5)
    betas = [O.cnt_beta(int(J), 17, 2.5) for J in mode]
    t = O.make_tube(4, betas, list(deg), spin=2)
    ld = load("doscompare_LDOS.txt")
    tf = load("doscompare_Transmission.txt")
    E = exact_grid(ld[:, 0])
    # in band of the lowest modes and away from every mode's band edge
    edges = np.concatenate([[g / 2, -g / 2] for g in gap])
    far = np.min(np.abs(E[:, None] - edges[None, :]), axis=1) > 0.03
    sel = (np.abs(E) > 0.30) & far
    dos, tr, st = O.dos_transmission(t, np.zeros(4), np.zeros(2), E[sel] + 1e-8j, 68)
    assert np.allclose(dos, ld[sel, 1], rtol=2e-5)
    assert np.allclose(tr, tf[sel, 1], rtol=2e-5, atol=1e-9)
    assert abs(tr[0] - 4.0) < 1e-6       # four (degenerate) modes conduct at -1 eV


# ---------------------------------------------------------------- charge
def test_rho0_is_half_filling():
    t = O.make_tube(40, [BETA6], [2], eta=1e-5)
    r0, st = O.rho0(t)
    # Im Rho0 = -(spin * degen / 2) per site at the midgap reference (mu = 0)
    # (30/30/30-point paths integrate the band to ~1.4e-6)
    assert np.abs(r0.imag + 2.0).max() < 5e-6


def test_particle_hole_symmetry():
    t = O.make_tube(40, [BETA6], [2], mu=(0.0, 0.0))
    r0, _ = O.rho0(t)
    U = O.cnt_ring_positions(17, 10) * 0.0
    Q0, _, _, _ = O.solve_negf(t, U, np.zeros(2), r0)
    assert np.abs(Q0).max() < 1e-7
    rng = np.random.default_rng(0)
    U = rng.uniform(-0.2, 0.2, 40)
    tp = O.make_tube(40, [BETA6], [2], mu=(0.15, 0.15))
    tm = O.make_tube(40, [BETA6], [2], mu=(-0.15, -0.15))
    Qp, _, _, _ = O.solve_negf(tp, U, (U[0], U[-1]), r0)
    Qm, _, _, _ = O.solve_negf(tm, -U, (-U[0], -U[-1]), r0)
    assert np.abs(Qp + Qm).max() < 1e-6 * max(1.0, np.abs(Qp).max())


def test_uniform_chain_gives_uniform_charge():
    t = O.make_tube(48, [BETA6], [2], mu=(-1.0, -1.0))
    r0, _ = O.rho0(t)
    Q, _, _, _ = O.solve_negf(t, np.zeros(48), np.zeros(2), r0)
    assert np.abs(Q - 0.26140897).max() < 1e-6


def test_contour_equals_real_axis_integration():
    # equilibrium contour + poles vs a real-axis Gauss-Legendre integral of
    # (s g/pi) f Im G on the same chain (Im RhoEq counts electrons with a
    # minus sign), split at the contact-shifted band edges and around mu
    N = 12
    rng = np.random.default_rng(1)
    U = rng.uniform(-0.2, 0.2, N)
    mu = -0.6
    t = O.make_tube(N, [BETA6], [2], mu=(mu, mu), eq_pts=(60, 60, 120), eta=1e-5)
    eq, st = O.rho_eq(t, U, (U[0], U[-1]))
    edges = [-(GAMMA + BETA6) + u for u in (U[0], U[-1])]
    edges += [-(GAMMA - BETA6) + u for u in (U[0], U[-1])]
    cuts = sorted([-10.0, mu - 40 * KT, mu + 40 * KT] + edges)
    xs, ws = np.polynomial.legendre.leggauss(500)
    E = np.concatenate([0.5 * (b - a) * xs + 0.5 * (b + a)
                        for a, b in zip(cuts[:-1], cuts[1:])]) + 1e-9j
    w = np.concatenate([0.5 * (b - a) * ws for a, b in zip(cuts[:-1], cuts[1:])])
    f = 1.0 / (np.exp((E.real - mu) / KT) + 1.0)
    coef = (2.0 * 2 / np.pi) * w * f
    # accumulate_diagG sums Re(c G); Re(-i c G) = c Im G
    acc, _ = O.accumulate_diagG(U, BETA6, GAMMA, (U[0], U[-1]), BETA6, GAMMA, E,
                                -1j * coef)
    assert np.abs(acc - eq.imag).max() < 1e-5
    # and the sign: electrons below mu give Im RhoEq < 0
    assert np.all(eq.imag < 0)


def test_solve_components_combine():
    t = O.make_tube(30, [BETA6], [2], mu=(-0.2, -0.5), noneq_pts=(40,))
    r0, _ = O.rho0(t)
    rng = np.random.default_rng(2)
    U = rng.uniform(-0.1, 0.1, 30)
    Q, comps, npts, st = O.solve_negf(t, U, (U[0], U[-1]), r0)
    assert npts == 40
    assert np.allclose(Q, comps[0] + comps[1] - comps[2], rtol=0, atol=1e-15)
    assert np.abs(comps[1]).max() > 0     # non-equilibrium part present


# ---------------------------------------------------------------- Broyden
def test_broyden_first_step_is_simple_mixing():
    b = O.Broyden(5, 0.1, 50)
    n0 = np.linspace(0.1, 0.5, 5)
    b.set_nin(n0)
    nout = n0 * 0.5 + 0.02
    mx, l2, nps, st = b.step(nout)
    assert np.allclose(b.get_nin(), n0 - 0.1 * (n0 - nout))
    assert abs(mx - np.abs(n0 - nout).max()) < 1e-15


def test_broyden_converges_to_fixed_point_of_linear_map():
    rng = np.random.default_rng(0)
    S = 20
    A = rng.normal(size=(S, S)) * 0.05
    c = rng.normal(size=S)
    b = O.Broyden(S, 0.2, 200)
    b.set_nin(np.zeros(S))
    for it in range(120):
        n = b.get_nin()
        mx, l2, _, st = b.step(A @ n + c)
        if mx < 1e-12:
            break
    x = np.linalg.solve(np.eye(S) - A, c)
    assert np.abs(n - x).max() < 1e-10


def test_broyden_reset_keeps_or_resets_input():
    b = O.Broyden(3, 0.1, 10)
    b.set_nin(np.array([1.0, 2.0, 3.0]))
    b.step(np.zeros(3))
    kept = b.get_nin()
    b.reset(1)
    assert np.allclose(b.get_nin(), kept) and b.step_count == 1
    b.reset(0, 0.5)
    assert np.allclose(b.get_nin(), 0.5)


# ---------------------------------------------------------------- current
def test_current_is_landauer():
    # the reference's contact-site form (Gamma_k f_k A_kk - Gamma_k Gn_kk)
    # equals Landauer s q/h sum w T (f_L - f_R), and I_1 = -I_0
    mu = (-1.0, -1.2)
    N = 24
    rng = np.random.default_rng(4)
    U = rng.uniform(-0.2, 0.2, N)
    t = O.make_tube(N, [BETA6], [2], mu=mu, noneq_pts=(150,), eta=1e-5)
    I, st = O.current(t, U, (U[0], U[-1]))
    assert abs(I[0] + I[1]) < 1e-12 * abs(I[0])
    L = O.energy_limits(t, mu)
    a = complex(L.E_contour_right.re, L.E_contour_right.im)
    b = complex(L.E_rightmost.re, L.E_rightmost.im)
    E, w, mul = O.gl_path(a, b, 150, 0)
    dos, tr, _ = O.dos_transmission(t, U, (U[0], U[-1]), E, 68)
    kT = [8.617333262e-5 * 298.0] * 2
    fL = np.array([O.fermi(e - mu[0], kT[0]) for e in E])
    fR = np.array([O.fermi(e - mu[1], kT[1]) for e in E])
    land = 2 * 1.602176634e-19 / 4.135667696e-15 * np.sum(w * tr * ((fL - fR) * mul).real)
    assert abs(I[0] - land) < 1e-10 * abs(land)
    # a clean chain at small bias conducts g = 2 modes: I ~ (2 s q/h) dmu
    t0 = O.make_tube(N, [BETA6], [2], mu=(-1.0, -1.01), noneq_pts=(400,), eta=1e-5)
    I0, _ = O.current(t0, np.zeros(N), (0.0, 0.0))
    G0 = 2 * 2 * 1.602176634e-19 / 4.135667696e-15
    assert abs(I0[0] / (G0 * 0.01) - 1.0) < 1e-3


def test_current_is_zero_in_equilibrium():
    t = O.make_tube(12, [BETA6], [2], mu=(-1.0, -1.0))
    I, _ = O.current(t, np.zeros(12), (0.0, 0.0))
    assert I[0] == 0.0 and I[1] == 0.0


# ---------------------------------------------------------------- surrogate CPU model
def test_surrogate_model_gradient_matches_central_differences():
    from oracle import surrogate_oracle as S
    p = S.init_params(1, np.float64)
    rng = np.random.default_rng(0)
    B, N = 2, 40
    U = rng.uniform(-0.5, 0.5, (B, N))
    Q = rng.standard_normal((B, N))
    f = rng.uniform(0, 0.5, B)
    _, g = S.loss_and_grad(p, U, Q, 0.1, 2.0, f, bf=False)
    lay, n = S.layout()
    h = 1e-6
    for l, (wo, bo) in enumerate(lay):
        W = np.zeros((S.cout(l), S.KW, S.cin(l)), bool)
        W[:S.real_cout(l), :, :S.real_cin(l)] = True
        idx = list(wo + rng.choice(np.flatnonzero(W.ravel()), 3, replace=False)) + [bo]
        for j in idx:
            e = np.zeros(n)
            e[j] = h
            fd = (S.loss_and_grad(p + e, U, Q, 0.1, 2.0, f, bf=False)[0]
                  - S.loss_and_grad(p - e, U, Q, 0.1, 2.0, f, bf=False)[0]) / (2 * h)
            assert abs(g[j] - fd) <= 1e-6 * max(abs(fd), 1e-3), (l, j, g[j], fd)


def test_surrogate_model_tangent_matches_central_differences():
    from oracle import surrogate_oracle as S
    p = S.init_params(2, np.float64)
    rng = np.random.default_rng(1)
    U = rng.uniform(-0.5, 0.5, (2, 30))
    V = rng.standard_normal((2, 30))
    jv = S.jvp(p, U, V, 0.7, None, bf=False)
    h = 1e-6
    fd = 0.7 * (S.forward(p, U + h * V, None, bf=False)[0] - S.forward(p, U - h * V, None, bf=False)[0]) / (2 * h)
    assert np.abs(jv - fd).max() < 1e-7 * np.abs(fd).max()


def test_surrogate_model_conv_is_a_banded_matrix_product():
    # conv9 against an explicit (N*ci) x (N*co) band matrix: zero padding,
    # tap order n + tap - 4, and the adjoint as its transpose
    from oracle import surrogate_oracle as S
    rng = np.random.default_rng(2)
    N, ci, co = 12, 3, 2
    W = rng.standard_normal((co, S.KW, ci))
    x = rng.standard_normal((1, N, ci))
    M = np.zeros((N * co, N * ci))
    for n in range(N):
        for tap in range(S.KW):
            m = n + tap - S.HALO
            if 0 <= m < N:
                M[n * co:(n + 1) * co, m * ci:(m + 1) * ci] += W[:, tap, :]
    z = S.conv(x, W)
    assert np.allclose(z.ravel(), M @ x.ravel())
    dz = rng.standard_normal((1, N, co))
    assert np.allclose(S.conv_T(dz, W).ravel(), M.T @ dz.ravel())


def test_bf16_rounding_is_nearest_even():
    from oracle import surrogate_oracle as S
    one = 1.0
    ulp = 2.0 ** -7          # bf16 spacing at 1
    assert S.bf16(one + ulp / 2) == one            # tie -> even (1.0)
    assert S.bf16(one + 1.5 * ulp) == one + 2 * ulp  # tie -> even
    assert S.bf16(one + 0.51 * ulp) == one + ulp
    assert S.bf16(-3.0) == -3.0


# ---------------------------------------------------------------- dense-block path
def _mode_chain_diag(E, beta, gamma, U, Uc):
    """Scalar mode-space chain (hoppings beta on even bonds, gamma on odd
    bonds) with the reference's closed-form lead self-energies."""
    N = len(U)
    H = np.diag(np.asarray(U, np.complex128))
    for n in range(N - 1):
        t = beta if n % 2 == 0 else gamma
        H[n, n + 1] = H[n + 1, n] = -t
    H[0, 0] += gamma ** 2 * O.surface_gf(E, Uc[0], beta, gamma)[0]
    H[-1, -1] += gamma ** 2 * O.surface_gf(E, Uc[1], beta, gamma)[0]
    return np.diag(np.linalg.inv(E * np.eye(N) - H))


@pytest.mark.parametrize("E", [0.31 + 1e-3j, -1.7 + 1e-3j, 0.05 + 1e-3j, 2.9 + 1e-3j])
def test_block_sancho_rubio_matches_closed_form_per_mode(E):
    from oracle import block_oracle as B
    m, gamma = 16, 2.5
    H0, T = B.zigzag_blocks(m, gamma)
    gR, it = B.sancho_rubio(E - 0.1, H0, T, T.T)
    assert it < 200
    j = np.arange(m)
    F = np.exp(2j * np.pi * np.outer(j, j) / m) / np.sqrt(m)
    gm = F.conj().T @ gR[:m, :m] @ F
    ref = np.array([O.surface_gf(E, 0.1, 2 * gamma * abs(np.cos(np.pi * J / m)), gamma)[0] for J in range(m)])
    assert np.abs(np.diag(gm) - ref).max() < 1e-10 * np.abs(ref).max()
    assert np.abs(gm - np.diag(np.diag(gm))).max() < 1e-10 * np.abs(ref).max()


@pytest.mark.parametrize("E", [0.31 + 1e-3j, -1.2 + 1e-3j])
def test_block_rgf_ring_traces_match_mode_space_chains(E):
    from oracle import block_oracle as B
    m, gamma, nb = 16, 2.5, 4
    H0, T = B.zigzag_blocks(m, gamma)
    rng = np.random.default_rng(3)
    Uring = rng.uniform(-0.3, 0.3, 2 * nb)
    v = np.repeat(Uring.reshape(nb, 2), m, axis=1)
    Uc = (Uring[0], Uring[-1])
    SL, SR, _ = B.lead_self_energies(E, H0, T, Uc)
    Gb = B.block_rgf_diag(E, H0, T, v, SL, SR)
    ring = Gb.reshape(nb, 2, m).sum(axis=2).ravel()
    modes = sum(_mode_chain_diag(E, 2 * gamma * abs(np.cos(np.pi * J / m)), gamma, Uring, Uc) for J in range(m))
    assert np.abs(ring - modes).max() < 1e-10 * np.abs(modes).max()


def test_block_rgf_matches_dense_inverse_with_disorder():
    from oracle import block_oracle as B
    m, nb = 16, 5
    H0, T = B.zigzag_blocks(m)
    rng = np.random.default_rng(4)
    v = rng.normal(0.0, 0.05, (nb, 2 * m)) + 0.1
    for E in (0.4 + 1e-5j, -0.9 + 1e-5j, 0.2 + 0.3j):
        SL, SR, _ = B.lead_self_energies(E, H0, T, (0.1, 0.1))
        a = B.block_rgf_diag(E, H0, T, v, SL, SR)
        b = B.dense_diag(E, H0, T, v, SL, SR)
        assert np.abs(a - b).max() < 1e-9 * np.abs(b).max()
