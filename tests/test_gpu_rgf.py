"""GPU parity of the batched RGF entry points (negf_rgf_*) against the oracle.

Tolerances: site charge 1e-10 relative (max-norm over sites, relative to
max |Q|); diag G^R 1e-9 relative per element where |G| > 1e-6 max|G|.
"""
import numpy as np
import pytest

from electrostatic_amd import _lib
from oracle import oracle as O
from tests.gpu_util import (chain_params, dev_array, dev_zeros, ptr, to_np,
                            smooth_potential)

pytestmark = pytest.mark.gpu

BETA17 = 2.2286917788826917   # 2*2.5*cos(6 pi/17): (17,0) lowest mode
GAMMA = 2.5


def energies(n_arc, n_real, seed=0):
    """Arc-contour energies (large Im) plus real-axis energies (eta=1e-5)."""
    rng = np.random.default_rng(seed)
    arc = []
    if n_arc:
        th = np.sort(rng.uniform(0.05, np.pi - 0.05, n_arc))
        arc = -4.5 + 5.5 * np.exp(1j * th)
    real = np.sort(rng.uniform(-3, 3, n_real)) + 1e-5j if n_real else []
    return np.concatenate([np.asarray(arc, complex), np.asarray(real, complex)])


def run_diag(U, te, to, Uc, E, coef, want_G=True, **dec):
    lib = _lib.load()
    N = U.size
    p = chain_params(N, te, to, Uc, **dec)
    dU = dev_array(U)
    dE = dev_array(E)
    dc = dev_array(coef)
    acc = dev_zeros(N)
    G = dev_zeros(E.size * N, complex_=True) if want_G else None
    rc = lib.negf_rgf_diag_accumulate(p, ptr(dU), E.size, ptr(dE), ptr(dc),
                                      ptr(acc), ptr(G), None)
    _lib.check(rc)
    out = to_np(acc)
    g = to_np(G, True).reshape(E.size, N) if want_G else None
    return out, g


def rel_err(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("N", [2, 5, 16, 40, 257, 400, 1001, 2048, 4096])
def test_diag_accumulate_matches_oracle(N):
    U = smooth_potential(N, seed=N)
    E = energies(40, 40, seed=N)
    rng = np.random.default_rng(N + 1)
    coef = rng.normal(size=E.size) + 1j * rng.normal(size=E.size)
    Uc = (U[0], U[-1])
    acc_ref, G_ref, st = O.accumulate_diagG(U, BETA17, GAMMA, Uc, BETA17, GAMMA,
                                            E, coef, want_diag=True)
    assert st == 0
    acc, G = run_diag(U, BETA17, GAMMA, Uc, E, coef)
    assert rel_err(acc, acc_ref) < 1e-10
    big = np.abs(G_ref) > 1e-6 * np.abs(G_ref).max()
    assert (np.abs(G - G_ref)[big] / np.abs(G_ref)[big]).max() < 1e-9


def test_diag_config2_shape():
    """Batched-RGF workload shape: N=4096, contour arc + real axis."""
    N = 4096
    U = smooth_potential(N, seed=2)
    te, to = 2.5 * np.cos(np.pi / 17), 2.5
    E = energies(512, 512, seed=3)
    coef = np.exp(1j * np.linspace(0, 3, E.size))
    Uc = (U[0], U[-1])
    acc_ref, st = O.accumulate_diagG(U, te, to, Uc, te, to, E, coef)
    acc, _ = run_diag(U, te, to, Uc, E, coef, want_G=False)
    assert rel_err(acc, acc_ref) < 1e-10


@pytest.mark.parametrize("N,L", [(64, 0), (257, 0), (1000, 16), (1000, 48), (1000, 128),
                                 (4096, 96), (20011, 0), (20011, 128)])
def test_charge_only_path_matches_oracle(N, L):
    """Charge without diag G^R: full segments take the unrolled sweep, the
    last (partial) one the general sweep; chains far beyond 8192 sites."""
    U = smooth_potential(N, seed=N + L)
    E = energies(48, 48, seed=N)
    rng = np.random.default_rng(N)
    coef = rng.normal(size=E.size) + 1j * rng.normal(size=E.size)
    Uc = (U[0], U[-1])
    acc_ref, st = O.accumulate_diagG(U, BETA17, GAMMA, Uc, BETA17, GAMMA, E, coef)
    acc, _ = run_diag(U, BETA17, GAMMA, Uc, E, coef, want_G=False, segment_length=L)
    assert rel_err(acc, acc_ref) < 1e-10


def test_segment_length_does_not_change_results():
    N = 1234
    U = smooth_potential(N, seed=7)
    E = energies(70, 70, seed=7)
    coef = np.exp(1j * np.linspace(0, 2, E.size))
    Uc = (U[0], U[-1])
    outs = [run_diag(U, BETA17, GAMMA, Uc, E, coef, want_G=False, segment_length=L)[0]
            for L in (16, 32, 64, 80, 128)]
    for o in outs[1:]:
        assert rel_err(o, outs[0]) < 1e-12


def test_diag_decimation_matches_oracle():
    N = 64
    U = smooth_potential(N, seed=5)
    E = energies(10, 10, seed=5)
    coef = np.ones(E.size, complex)
    Uc = (U[0], U[-1])
    t = O.make_tube(N, [BETA17], [1], use_decimation=1, decimation_layers=10,
                    decimation_max_iter=30, decimation_rel_error=0)
    # oracle: alpha with decimated Sigma via the generic per-energy path
    G_ref = []
    for e in E:
        g0 = O.decimation_gf(e - Uc[0], BETA17, GAMMA, 10, 30, 0)
        g1 = O.decimation_gf(e - Uc[1], BETA17, GAMMA, 10, 30, 0)
        H = np.zeros((N, N), complex)
        for k in range(N - 1):
            H[k, k + 1] = H[k + 1, k] = BETA17 if k % 2 == 0 else GAMMA
        A = e * np.eye(N) - H - np.diag(U)
        A[0, 0] -= GAMMA ** 2 * g0
        A[-1, -1] -= GAMMA ** 2 * g1
        G_ref.append(np.diag(np.linalg.inv(A)))
    G_ref = np.array(G_ref)
    del t
    acc, G = run_diag(U, BETA17, GAMMA, Uc, E, coef, use_decimation=1,
                      layers=10, max_iter=30, rel_error=0)
    assert np.abs(G - G_ref).max() / np.abs(G_ref).max() < 1e-10


@pytest.mark.parametrize("N", [3, 64, 400, 999])
def test_noneq_accumulate_matches_oracle_walk(N):
    lib = _lib.load()
    U = smooth_potential(N, seed=7 + N, lo=-0.3, hi=0.2)
    E = np.linspace(-2.5, 1.5, 60) + 1e-5j
    rng = np.random.default_rng(N)
    cL = rng.normal(size=E.size)
    cR = rng.normal(size=E.size)
    Uc = (U[0], U[-1])
    ref, st = O.accumulate_noneq(U, BETA17, GAMMA, Uc, BETA17, GAMMA, E, cL, cR)
    p = chain_params(N, BETA17, GAMMA, Uc)
    dU, dE = dev_array(U), dev_array(E)
    dL, dR = dev_array(cL), dev_array(cR)
    acc = dev_zeros(N)
    integ = dev_zeros(3 * E.size)
    _lib.check(lib.negf_rgf_noneq_accumulate(p, ptr(dU), E.size, ptr(dE), ptr(dL),
                                             ptr(dR), None, None, ptr(acc),
                                             ptr(integ), None))
    assert rel_err(to_np(acc), ref) < 1e-10
    # integrand samples at N/2, N/4, 3N/4: A_nn for one energy at a time
    got = to_np(integ).reshape(E.size, 3)
    for k in (0, 17, 59):
        for w, n in enumerate((N // 2, N // 4, N * 3 // 4)):
            one_L = np.zeros(E.size)
            one_R = np.zeros(E.size)
            one_L[k] = cL[k]
            one_R[k] = cR[k]
            r, _ = O.accumulate_noneq(U, BETA17, GAMMA, Uc, BETA17, GAMMA, E,
                                      one_L, one_R)
            assert abs(got[k, w] - r[n]) <= 1e-10 * max(abs(r[n]), 1e-300) + 1e-300


@pytest.mark.parametrize("N,L", [(128, 0), (1000, 0), (1000, 64), (4096, 128), (20011, 0)])
def test_noneq_without_samples_matches_oracle(N, L):
    """Without integrand samples full segments take the unrolled
    non-equilibrium sweep (signed weights folded into the states)."""
    lib = _lib.load()
    U = smooth_potential(N, seed=N + 3, lo=-0.3, hi=0.2)
    E = np.concatenate([np.linspace(-2.5, 1.5, 50), [0.05, 0.1]]) + 1e-5j  # in-gap too
    rng = np.random.default_rng(N + 1)
    cL = rng.normal(size=E.size)
    cR = rng.normal(size=E.size)
    Uc = (U[0], U[-1])
    p = chain_params(N, BETA17, GAMMA, Uc, segment_length=L)
    dU, dE = dev_array(U), dev_array(E)
    dL, dR = dev_array(cL), dev_array(cR)
    acc = dev_zeros(N)
    _lib.check(lib.negf_rgf_noneq_accumulate(p, ptr(dU), E.size, ptr(dE), ptr(dL),
                                             ptr(dR), None, None, ptr(acc), None, None))
    if N <= 4096:
        ref, st = O.accumulate_noneq(U, BETA17, GAMMA, Uc, BETA17, GAMMA, E, cL, cR)
        assert rel_err(to_np(acc), ref) < 1e-10
    else:
        # the oracle's O(N^2) walk is too slow at this length: compare with the
        # general sweep (integrand samples requested)
        acc2 = dev_zeros(N)
        integ = dev_zeros(3 * E.size)
        _lib.check(lib.negf_rgf_noneq_accumulate(p, ptr(dU), E.size, ptr(dE), ptr(dL),
                                                 ptr(dR), None, None, ptr(acc2),
                                                 ptr(integ), None))
        assert rel_err(to_np(acc), to_np(acc2)) < 1e-12


def test_dos_transmission_matches_oracle():
    lib = _lib.load()
    N = 120
    U = smooth_potential(N, seed=11, lo=-0.2, hi=0.2)
    E = np.linspace(-2, 2, 81) + 1e-5j
    Uc = (U[0], U[-1])
    t = O.make_tube(N, [BETA17], [1], spin=1)
    dos_ref, tr_ref, st = O.dos_transmission(t, U, np.array(Uc), E, 1)
    p = chain_params(N, BETA17, GAMMA, Uc)
    dU, dE = dev_array(U), dev_array(E)
    dd, dt = dev_zeros(E.size), dev_zeros(E.size)
    _lib.check(lib.negf_rgf_dos(p, ptr(dU), E.size, ptr(dE), ptr(dd), ptr(dt), None))
    dos = to_np(dd) / (2 * np.pi)
    assert rel_err(dos, dos_ref) < 1e-10
    assert rel_err(to_np(dt), tr_ref) < 1e-10


def test_bad_arguments_raise():
    lib = _lib.load()
    p = chain_params(1, BETA17, GAMMA, (0, 0))
    rc = lib.negf_rgf_diag_accumulate(p, None, 0, None, None, None, None, None)
    assert rc == 1
    assert b"invalid" in lib.negf_last_error()


@pytest.mark.parametrize("N", [1000, 2048])
def test_scaled_and_unscaled_recurrences_agree(N):
    """The transfer and equilibrium sweep kernels divide the recurrence by
    Im E when every chain of a 64-chain wave has t / |Im E| <= 2^24, and keep
    the plain recurrence otherwise.  Wave 0 (eta = 1e-5) and wave 2
    (eta = 2e-7, just above the bound 2.5 / 2^24 = 1.5e-7) take the scaled
    form; wave 1 repeats wave 0's energies but one lane has eta = 1e-9, so the
    whole wave falls back to the plain form (that lane's weight is 0 in both
    waves); wave 3 is a contour arc.  Each wave alone and all four together
    match the oracle to 1e-10, and the scaled and plain runs of the same
    energies agree to 1e-11 (measured 1.3e-12: rounding near the band edges
    is amplified by about 1/eta, and the plain recurrence alone moves the
    charge by ~1e-12 between segment lengths)."""
    rng = np.random.default_rng(7 + N)
    U = smooth_potential(N, seed=N + 3)
    Uc = (U[0], U[-1])
    re = np.sort(rng.uniform(-3, 3, 64))
    w0 = re + 1e-5j
    w1 = w0.copy()
    w1[17] = 1.5 + 1e-9j  # in the band of both leads
    w2 = np.sort(rng.uniform(-3, 3, 64)) + 2e-7j
    th = np.sort(rng.uniform(0.05, np.pi - 0.05, 64))
    w3 = -4.5 + 5.5 * np.exp(1j * th)
    c = rng.normal(size=64) + 1j * rng.normal(size=64)
    c[17] = 0.0
    waves = [w0, w1, w2, w3]
    accs = []
    for w in waves:
        ref, st = O.accumulate_diagG(U, BETA17, GAMMA, Uc, BETA17, GAMMA, w, c)
        assert st == 0
        acc, _ = run_diag(U, BETA17, GAMMA, Uc, w, c, want_G=False)
        assert rel_err(acc, ref) < 1e-10
        accs.append(acc)
    assert rel_err(accs[1], accs[0]) < 1e-11
    E = np.concatenate(waves)
    coef = np.tile(c, 4)
    ref, st = O.accumulate_diagG(U, BETA17, GAMMA, Uc, BETA17, GAMMA, E, coef)
    acc, _ = run_diag(U, BETA17, GAMMA, Uc, E, coef, want_G=False)
    assert rel_err(acc, ref) < 1e-10


def test_scaled_and_unscaled_noneq_sweeps_agree():
    """Non-equilibrium full-segment sweep: a wave of real-axis energies at
    eta = 1e-5 takes the recurrence divided by Im E (contributions
    Q_n |w_n|^2 + P_n |w'_n|^2 with carried scale factors); the same energies
    with one lane at eta = 1e-9 (weight 0) fall back to the plain recurrence.
    Both match the oracle's O(N^2) contact walk to 1e-10 and each other to
    1e-11."""
    lib = _lib.load()
    N = 1536
    U = smooth_potential(N, seed=11, lo=-0.3, hi=0.2)
    Uc = (U[0], U[-1])
    rng = np.random.default_rng(12)
    E0 = np.concatenate([np.linspace(-2.5, 1.5, 62), [0.05, 0.1]]) + 1e-5j  # in-gap too
    E1 = E0.copy()
    E1[17] = 1.5 + 1e-9j
    cL = rng.normal(size=64)
    cR = rng.normal(size=64)
    cL[17] = cR[17] = 0.0
    p = chain_params(N, BETA17, GAMMA, Uc)
    dU, dL, dR = dev_array(U), dev_array(cL), dev_array(cR)
    accs = []
    for E in (E0, E1):
        ref, st = O.accumulate_noneq(U, BETA17, GAMMA, Uc, BETA17, GAMMA, E, cL, cR)
        dE = dev_array(E)
        acc = dev_zeros(N)
        _lib.check(lib.negf_rgf_noneq_accumulate(p, ptr(dU), E.size, ptr(dE), ptr(dL),
                                                 ptr(dR), None, None, ptr(acc), None, None))
        got = to_np(acc)
        assert rel_err(got, ref) < 1e-10
        accs.append(got)
    assert rel_err(accs[1], accs[0]) < 1e-11


def test_full_batched_workload_matches_oracle():
    """The benchmark's own workload at full size (8192 energies x 4096
    sites: reference contour arc with Fermi-weighted Gauss-Legendre
    coefficients plus the real axis at eta = 1e-5) through the plan entry
    points the benchmark times, against the oracle on all host cores."""
    import ctypes as C
    import os
    from electrostatic_amd.workloads import batched_rgf_workload
    lib = _lib.load()
    wl = batched_rgf_workload()
    N, nE = wl["N"], wl["E"].size
    p = chain_params(N, wl["t_even"], wl["t_odd"], wl["Uc"])
    dU, dE, dc = dev_array(wl["U"]), dev_array(wl["E"]), dev_array(wl["coef"])
    acc = dev_zeros(N)
    plan = C.c_void_p()
    _lib.check(lib.negf_rgf_plan_create(p, ptr(dU), nE, ptr(dE), ptr(dc), 0, C.byref(plan)))
    try:
        for _ in range(2):   # a second execute reuses the plan's workspace
            _lib.check(lib.negf_rgf_plan_execute(plan, ptr(acc), None, None))
        _lib.check(lib.negf_rgf_plan_status(plan, None))
    finally:
        lib.negf_rgf_plan_destroy(plan)
    got = to_np(acc)
    ncpu = len(os.sched_getaffinity(0))
    ref, st = O.accumulate_diagG_par(wl["U"], wl["t_even"], wl["t_odd"], wl["Uc"],
                                     wl["t_even"], wl["t_odd"], wl["E"], wl["coef"], ncpu)
    assert st == 0
    assert rel_err(got, ref) < 1e-10


def bulk_dimer_G(E, te, to):
    """Exact G_nn of the infinite two-hopping chain, the value every site of a
    flat chain with matched leads has: E / sqrt((E^2 - (te+to)^2)(E^2 - (te-to)^2))
    with Im G < 0, evaluated in 40-digit arithmetic at the double inputs."""
    import mpmath as mp
    mp.mp.dps = 40
    out = []
    for z in E:
        zz = mp.mpc(z.real, z.imag)
        a, b = mp.mpf(te) + mp.mpf(to), mp.mpf(to) - mp.mpf(te)
        g = zz / mp.sqrt((zz - a) * (zz + a) * (zz - b) * (zz + b))
        if mp.im(g) > 0:
            g = -g
        out.append(complex(g))
    return np.array(out)


def band_edge_energies(te, to, eta):
    edges = [-(to + te), -(to - te), to - te, to + te]
    offs = np.array([-30, -10, -3, -1, 0, 1, 3, 10, 30]) * eta
    return np.array([e + o for e in edges for o in offs]) + 1j * eta


@pytest.mark.parametrize("eta,flat", [(1e-5, True), (1e-5, False), (1e-8, True), (1e-8, False)])
def test_band_edge_energies_match_oracle(eta, flat):
    """Real-axis energies within a few eta of the four band edges
    +-(gamma - beta), +-(gamma + beta) of the two-hopping chain, where the
    1-D density of states diverges and the recurrences barely decay; N = 4096.
    eta = 1e-8 is the reference's default E_zPlus (NEGF_Common.H:386).

    On the flat chain (matched leads) every site has the closed-form bulk G,
    so both this library and the oracle are measured against the exact value:
    - eta = 1e-5: both within 2e-10 of it and of each other (the oracle, a
      serial Riccati sweep, is at about 1e-11, this library at about 1e-10:
      its segment transfer matrices are near defective there);
    - eta = 1e-8: the problem's own condition number near the edges (about
      1/eta) puts even the oracle 3e-8 from the exact value; this library
      must stay within 10x the oracle's own error.
    With a potential (smooth, +-0.05 eV) the same energies are off the local
    band edges on most of the chain and the 1e-10 parity with the oracle
    holds.  Both the general sweep (diag G written) and the full-segment
    sweep (charge only) are checked."""
    N = 4096
    te, to = BETA17, GAMMA
    U = np.zeros(N) if flat else smooth_potential(N, seed=9, lo=-0.05, hi=0.05)
    E = band_edge_energies(te, to, eta)
    rng = np.random.default_rng(4)
    coef = rng.normal(size=E.size) + 1j * rng.normal(size=E.size)
    Uc = (U[0], U[-1])
    acc_ref, G_ref, st = O.accumulate_diagG(U, te, to, Uc, te, to, E, coef, want_diag=True)
    acc, G = run_diag(U, te, to, Uc, E, coef)
    acc_f, _ = run_diag(U, te, to, Uc, E, coef, want_G=False)
    assert np.all(np.isfinite(acc)) and np.all(np.isfinite(acc_f))
    if not flat:
        assert rel_err(acc, acc_ref) < 1e-10
        assert rel_err(acc_f, acc_ref) < 1e-10
        return
    Gx = bulk_dimer_G(E, te, to)
    acc_x = np.full(N, np.sum((coef * Gx).real))
    g_err = (np.abs(G - Gx[:, None]) / np.abs(Gx[:, None])).max()
    g_err_ref = (np.abs(G_ref - Gx[:, None]) / np.abs(Gx[:, None])).max()
    errs = dict(gpu=rel_err(acc, acc_x), gpu_full=rel_err(acc_f, acc_x), oracle=rel_err(acc_ref, acc_x))
    print(f"eta={eta:g}: charge vs exact {errs}; G vs exact gpu {g_err:.2e} oracle {g_err_ref:.2e}")
    if eta >= 1e-5:
        assert max(errs.values()) < 2e-10
        assert rel_err(acc, acc_ref) < 2e-10 and rel_err(acc_f, acc_ref) < 2e-10
        assert g_err < 2e-10
    else:
        assert errs["gpu"] < 10 * errs["oracle"] and errs["gpu_full"] < 10 * errs["oracle"]
        assert g_err < 10 * g_err_ref


@pytest.mark.parametrize("eta,flat", [(1e-5, True), (1e-5, False), (1e-8, False)])
def test_band_edge_noneq_matches_oracle(eta, flat):
    """Non-equilibrium (contact-resolved) accumulation at the band-edge
    energies against the oracle's O(N^2) contact walk, N = 4096.  |G_n0|^2 and
    |G_n,N-1|^2 carry the normalisation of the states along the whole chain,
    which the segment products of the scan hold to about 2e-10 next to the
    outer band edges of the flat chain (eta = 1e-5); with a potential the
    1e-10 parity holds."""
    lib = _lib.load()
    N = 4096
    te, to = BETA17, GAMMA
    U = np.zeros(N) if flat else smooth_potential(N, seed=9, lo=-0.05, hi=0.05)
    E = band_edge_energies(te, to, eta)
    rng = np.random.default_rng(5)
    cL, cR = rng.normal(size=E.size), rng.normal(size=E.size)
    Uc = (U[0], U[-1])
    ref, st = O.accumulate_noneq(U, te, to, Uc, te, to, E, cL, cR)
    p = chain_params(N, te, to, Uc)
    dU, dE, dL, dR = dev_array(U), dev_array(E), dev_array(cL), dev_array(cR)
    acc = dev_zeros(N)
    _lib.check(lib.negf_rgf_noneq_accumulate(p, ptr(dU), E.size, ptr(dE), ptr(dL), ptr(dR),
                                             None, None, ptr(acc), None, None))
    err = rel_err(to_np(acc), ref)
    print(f"eta={eta:g} flat={flat}: noneq vs oracle {err:.2e}")
    assert err < (3e-10 if flat else 1e-10)


def with_fused(flag, fn):
    """fn() with the fused transfer+scan launch switched on or off (the
    library reads NEGF_RGF_FUSED whenever it builds a launch plan)."""
    import os
    old = os.environ.get("NEGF_RGF_FUSED")
    os.environ["NEGF_RGF_FUSED"] = flag
    try:
        return fn()
    finally:
        if old is None:
            os.environ.pop("NEGF_RGF_FUSED", None)
        else:
            os.environ["NEGF_RGF_FUSED"] = old


@pytest.mark.parametrize("N,L", [(2100, 0), (4096, 0), (3072, 0), (2176, 0), (1000, 48), (777, 32)])
def test_fused_transfer_scan_matches_separate_launches(N, L):
    """Tubes of 17..32 segments take one launch for transfer matrices and
    walks (matrices kept in LDS), and when every segment is a full 128-site
    one the equilibrium charge without diag G^R takes one launch for the
    sweep too (states kept in LDS, 24 and 17 segments leave some wavefronts'
    halves idle); both must give what the separate launches give for every
    launch kind: charge with and without diag G^R, lead self-energies in
    closed form and by decimation, non-equilibrium charge with and without
    integrand samples, DOS and transmission.  Tolerance 1e-13 relative: the
    partial rows of the fused sweep sum 32 chains each instead of 64."""
    lib = _lib.load()
    U = smooth_potential(N, seed=N + 1, lo=-0.3, hi=0.3)
    E = energies(40, 40, seed=N)
    rng = np.random.default_rng(N)
    coef = rng.normal(size=E.size) + 1j * rng.normal(size=E.size)
    Uc = (U[0], U[-1])
    Er = np.linspace(-2.5, 1.5, 48) + 1e-5j
    cL = rng.normal(size=Er.size)
    cR = rng.normal(size=Er.size)

    def run():
        out = {}
        out["eqG"] = run_diag(U, BETA17, GAMMA, Uc, E, coef, segment_length=L)
        out["eq"] = run_diag(U, BETA17, GAMMA, Uc, E, coef, want_G=False, segment_length=L)[0]
        out["dec"] = run_diag(U, BETA17, GAMMA, Uc, E, coef, want_G=False, segment_length=L,
                              use_decimation=1, layers=10, max_iter=30, rel_error=0)[0]
        p = chain_params(N, BETA17, GAMMA, Uc, segment_length=L)
        dU, dE, dL, dR = dev_array(U), dev_array(Er), dev_array(cL), dev_array(cR)
        acc, acc2, integ = dev_zeros(N), dev_zeros(N), dev_zeros(3 * Er.size)
        _lib.check(lib.negf_rgf_noneq_accumulate(p, ptr(dU), Er.size, ptr(dE), ptr(dL), ptr(dR),
                                                 None, None, ptr(acc), None, None))
        _lib.check(lib.negf_rgf_noneq_accumulate(p, ptr(dU), Er.size, ptr(dE), ptr(dL), ptr(dR),
                                                 None, None, ptr(acc2), ptr(integ), None))
        out["neq"], out["neq2"], out["integ"] = to_np(acc), to_np(acc2), to_np(integ)
        dd, dt = dev_zeros(Er.size), dev_zeros(Er.size)
        _lib.check(lib.negf_rgf_dos(p, ptr(dU), Er.size, ptr(dE), ptr(dd), ptr(dt), None))
        out["dos"], out["tr"] = to_np(dd), to_np(dt)
        return out

    fused = with_fused("1", run)
    sep = with_fused("0", run)
    assert rel_err(fused["eqG"][0], sep["eqG"][0]) < 1e-13
    assert np.abs(fused["eqG"][1] - sep["eqG"][1]).max() < 1e-13 * np.abs(sep["eqG"][1]).max()
    for key in ("eq", "dec", "neq", "neq2", "integ", "dos", "tr"):
        assert rel_err(fused[key], sep[key]) < 1e-13, key
    # and against the oracle
    acc_ref, st = O.accumulate_diagG(U, BETA17, GAMMA, Uc, BETA17, GAMMA, E, coef)
    assert st == 0 and rel_err(fused["eq"], acc_ref) < 1e-10


@pytest.mark.parametrize("N", [4096, 1000])
def test_plan_timing_events(N):
    """negf_rgf_plan_enable_timing / negf_rgf_plan_kernel_time: with N = 4096
    the plan is the single fused launch, whose start and stop events the launch
    records itself; with N = 1000 (a partial last segment) the events bracket
    several launches.  One timed interval per execute, each positive and below
    the wall time of the executes; the charge is the same bit for bit with and
    without timing."""
    import ctypes as C
    import time
    import torch
    lib = _lib.load()
    U = smooth_potential(N, seed=7)
    E = energies(256, 256, seed=8)
    coef = np.exp(1j * np.linspace(0, 2, E.size))
    p = chain_params(N, BETA17, GAMMA, (U[0], U[-1]))
    dU, dE, dc = dev_array(U), dev_array(E), dev_array(coef)
    plan = C.c_void_p()
    _lib.check(lib.negf_rgf_plan_create(p, ptr(dU), E.size, ptr(dE), ptr(dc), 0, C.byref(plan)))
    try:
        acc0, acc1 = dev_zeros(N), dev_zeros(N)
        _lib.check(lib.negf_rgf_plan_execute(plan, ptr(acc0), None, None))
        _lib.check(lib.negf_rgf_plan_enable_timing(plan, 1))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            _lib.check(lib.negf_rgf_plan_execute(plan, ptr(acc1), None, None))
        torch.cuda.synchronize()
        wall_ms = (time.perf_counter() - t0) * 1e3
        ms, n = C.c_double(), C.c_int()
        _lib.check(lib.negf_rgf_plan_kernel_time(plan, C.byref(ms), C.byref(n)))
        assert n.value == 3
        assert 0.0 < ms.value < wall_ms
        _lib.check(lib.negf_rgf_plan_kernel_time(plan, C.byref(ms), C.byref(n)))
        assert n.value == 0   # reading the times resets them
        _lib.check(lib.negf_rgf_plan_status(plan, None))
    finally:
        lib.negf_rgf_plan_destroy(plan)
    assert np.array_equal(to_np(acc0), to_np(acc1))
