"""ctypes wrapper of the CPU oracle (oracle/lib/libnegf_oracle.so).

TEST INFRASTRUCTURE ONLY.  The oracle restates the reference NEGF charge
path on the CPU so that tests, smoke() and the CPU-baseline leg of bench.py
can check the HIP product against it.  Product code never imports this
module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libnegf_oracle.so")

MAX_MODES = 32
MAX_PATHS = 8
MAX_POLES = 256


class orc_c(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class orc_tube(C.Structure):
    _fields_ = [
        ("N", C.c_int),
        ("nmodes", C.c_int),
        ("beta", C.c_double * MAX_MODES),
        ("degen", C.c_int * MAX_MODES),
        ("gamma", C.c_double),
        ("spin_degen", C.c_int),
        ("E_valence_min", C.c_double),
        ("E_pole_max", C.c_double),
        ("E_zPlus_imag", C.c_double),
        ("tail_lower", C.c_double),
        ("tail_upper", C.c_double),
        ("eq_pts", C.c_int * 3),
        ("num_noneq_paths", C.c_int),
        ("noneq_pts", C.c_int * MAX_PATHS),
        ("noneq_intercuts", C.c_double * MAX_PATHS),
        ("mu", C.c_double * 2),
        ("T", C.c_double * 2),
        ("use_decimation", C.c_int),
        ("decimation_layers", C.c_int),
        ("decimation_max_iter", C.c_int),
        ("decimation_rel_error", C.c_int),
        ("kT_window", C.c_double * 2),
        ("has_density", C.c_int),
        ("noneq_density", C.c_double * MAX_PATHS),
        ("E_at_max", C.c_double),
    ]


class orc_limits(C.Structure):
    _fields_ = [
        ("kT_contact", C.c_double * 2),
        ("kT_min", C.c_double),
        ("kT_max", C.c_double),
        ("mu_min", C.c_double),
        ("mu_max", C.c_double),
        ("flag_noneq", C.c_int),
        ("E_contour_left", orc_c),
        ("E_contour_right", orc_c),
        ("E_rightmost", orc_c),
        ("E_zeta", orc_c),
        ("E_eta", orc_c),
        ("E_zPlus", orc_c),
        ("num_poles", C.c_int),
        ("num_poles_formula", C.c_int),
        ("poles", orc_c * MAX_POLES),
    ]


class orc_coulomb(C.Structure):
    _fields_ = [("eps_r", C.c_double), ("R_ring_nm", C.c_double),
                ("R_image_nm", C.c_double)]


class orc_scf_params(C.Structure):
    _fields_ = [("fraction", C.c_double), ("max_norm", C.c_double),
                ("norm_type", C.c_int), ("max_step", C.c_int),
                ("reset_with_previous", C.c_int),
                ("initial_value", C.c_double),
                ("U_ext", C.POINTER(C.c_double))]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    cp = C.POINTER(orc_c)
    L.orc_gauss_legendre.argtypes = [C.c_int, dp, dp]
    L.orc_gl_path.argtypes = [orc_c, orc_c, C.c_int, C.c_int, cp, dp, cp]
    L.orc_cnt_sorted_modes.argtypes = [C.c_int, C.c_int, C.c_double,
                                       C.c_double, ip, dp, ip]
    L.orc_cnt_sorted_modes.restype = C.c_int
    L.orc_cnt_beta.argtypes = [C.c_int, C.c_int, C.c_double]
    L.orc_cnt_beta.restype = C.c_double
    L.orc_cnt_ring_positions.argtypes = [C.c_int, C.c_int, C.c_double, dp]
    L.orc_surface_gf.argtypes = [orc_c, C.c_double, C.c_double, C.c_double,
                                 cp]
    L.orc_surface_gf.restype = C.c_int
    L.orc_decimation_gf.argtypes = [orc_c, C.c_double, C.c_double, C.c_int,
                                    C.c_int, C.c_int]
    L.orc_decimation_gf.restype = orc_c
    L.orc_fermi.argtypes = [orc_c, C.c_double]
    L.orc_fermi.restype = orc_c
    L.orc_energy_limits.argtypes = [C.POINTER(orc_tube), dp,
                                    C.POINTER(orc_limits)]
    L.orc_rgf.argtypes = [C.c_int, cp, cp, cp, cp, cp, cp, cp]
    L.orc_rho0.argtypes = [C.POINTER(orc_tube), cp]
    L.orc_rho0.restype = C.c_int
    L.orc_rho_eq.argtypes = [C.POINTER(orc_tube), dp, dp, cp, ip]
    L.orc_rho_eq.restype = C.c_int
    L.orc_rho_noneq.argtypes = [C.POINTER(orc_tube), dp, dp, cp, dp, ip]
    L.orc_rho_noneq.restype = C.c_int
    L.orc_solve_negf.argtypes = [C.POINTER(orc_tube), dp, dp, cp, dp, dp, ip]
    L.orc_solve_negf.restype = C.c_int
    L.orc_adaptive_intercuts.argtypes = [C.POINTER(orc_tube), dp, dp]
    L.orc_adaptive_intercuts.restype = C.c_int
    L.orc_accumulate_diagG.argtypes = [C.c_int, dp, C.c_double, C.c_double,
                                       dp, C.c_double, C.c_double, C.c_int,
                                       cp, cp, dp, cp]
    L.orc_accumulate_diagG.restype = C.c_int
    L.orc_accumulate_diagG_par.argtypes = [C.c_int, C.c_int, dp, C.c_double, C.c_double,
                                           dp, C.c_double, C.c_double, C.c_int,
                                           C.c_void_p, C.c_void_p, dp]
    L.orc_accumulate_diagG_par.restype = C.c_int
    L.orc_accumulate_noneq.argtypes = [C.c_int, dp, C.c_double, C.c_double,
                                       dp, C.c_double, C.c_double, C.c_int,
                                       cp, dp, dp, dp]
    L.orc_accumulate_noneq.restype = C.c_int
    L.orc_dos_transmission.argtypes = [C.POINTER(orc_tube), dp, dp, C.c_int,
                                       cp, C.c_int, dp, dp, dp]
    L.orc_dos_transmission.restype = C.c_int
    L.orc_broyden_create.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int]
    L.orc_broyden_create.restype = C.c_void_p
    L.orc_broyden_free.argtypes = [C.c_void_p]
    L.orc_broyden_set_nin.argtypes = [C.c_void_p, dp]
    L.orc_broyden_get_nin.argtypes = [C.c_void_p, dp]
    L.orc_broyden_step.argtypes = [C.c_void_p, dp, dp, dp, dp]
    L.orc_broyden_step.restype = C.c_int
    L.orc_broyden_reset.argtypes = [C.c_void_p, C.c_int, C.c_double]
    L.orc_broyden_get_step.argtypes = [C.c_void_p]
    L.orc_broyden_get_step.restype = C.c_int
    L.orc_coulomb_kernel.argtypes = [C.c_int, dp, C.POINTER(orc_coulomb), dp]
    L.orc_scf_bias_point.argtypes = [C.POINTER(orc_tube), cp, dp, C.c_double,
                                     C.POINTER(orc_scf_params), C.c_void_p,
                                     dp, dp, ip, dp]
    L.orc_scf_bias_point.restype = C.c_int


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _cp(a):
    return a.ctypes.data_as(C.POINTER(orc_c))


def _c(z):
    z = complex(z)
    return orc_c(z.real, z.imag)


def _carr(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex128))


# --------------------------------------------------------------------------
def gauss_legendre(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauss_legendre(n, _dp(x), _dp(w))
    return x, w


def gl_path(a, b, n, kind):
    E = np.zeros(n, np.complex128)
    w = np.zeros(n)
    mul = np.zeros(n, np.complex128)
    lib().orc_gl_path(_c(a), _c(b), n, kind, _cp(E), _dp(w), _cp(mul))
    return E, w, mul


def cnt_sorted_modes(m, n=0, acc=1.42e-10, gamma=2.5):
    k = m + 1
    mode = np.zeros(k, np.int32)
    gap = np.zeros(k)
    deg = np.zeros(k, np.int32)
    cnt = lib().orc_cnt_sorted_modes(
        m, n, acc, gamma, mode.ctypes.data_as(C.POINTER(C.c_int)), _dp(gap),
        deg.ctypes.data_as(C.POINTER(C.c_int)))
    return mode[:cnt].copy(), gap[:cnt].copy(), deg[:cnt].copy()


def cnt_beta(J, m, gamma=2.5):
    return lib().orc_cnt_beta(J, m, gamma)


def cnt_ring_positions(m, num_unitcells, acc=1.42e-10):
    out = np.zeros(4 * num_unitcells)
    lib().orc_cnt_ring_positions(m, num_unitcells, acc, _dp(out))
    return out


def surface_gf(E, U, beta, gamma):
    g = orc_c()
    st = lib().orc_surface_gf(_c(E), U, beta, gamma, C.byref(g))
    return complex(g.re, g.im), st


def decimation_gf(EmU, beta, gamma, layers, max_iter, rel_error=0):
    r = lib().orc_decimation_gf(_c(EmU), beta, gamma, layers, max_iter,
                                rel_error)
    return complex(r.re, r.im)


def fermi(z, kT):
    r = lib().orc_fermi(_c(z), kT)
    return complex(r.re, r.im)


def make_tube(N, betas, degens, gamma=2.5, mu=(0.0, 0.0), T=(298.0, 298.0),
              E_valence_min=-10.0, E_pole_max=3.0, eta=1e-5,
              tails=(14.0, 14.0), eq_pts=(30, 30, 30), noneq_pts=(100,),
              intercuts=(), spin=2, use_decimation=0, decimation_layers=10,
              decimation_max_iter=30, decimation_rel_error=0,
              kT_window=(1.0, 1.0), density=None):
    t = orc_tube()
    t.N = N
    t.nmodes = len(betas)
    for i, (b, d) in enumerate(zip(betas, degens)):
        t.beta[i] = b
        t.degen[i] = d
    t.gamma = gamma
    t.spin_degen = spin
    t.E_valence_min = E_valence_min
    t.E_pole_max = E_pole_max
    t.E_zPlus_imag = eta
    t.tail_lower, t.tail_upper = tails
    for i in range(3):
        t.eq_pts[i] = eq_pts[i]
    t.num_noneq_paths = len(noneq_pts)
    for i, p in enumerate(noneq_pts):
        t.noneq_pts[i] = p
    ic = list(intercuts)
    if not ic and len(noneq_pts) > 1:
        ic = [(c + 1) * 100.0 / len(noneq_pts) for c in range(len(noneq_pts) - 1)]
    for i, p in enumerate(ic):
        t.noneq_intercuts[i] = p
    t.mu[0], t.mu[1] = mu
    t.T[0], t.T[1] = T
    t.use_decimation = use_decimation
    t.decimation_layers = decimation_layers
    t.decimation_max_iter = decimation_max_iter
    t.decimation_rel_error = decimation_rel_error
    t.kT_window[0], t.kT_window[1] = kT_window
    if density is not None:   # noneq_integration_pts_density given (ints)
        t.has_density = 1
        for i, d in enumerate(density):
            t.noneq_density[i] = int(d)
    return t


def energy_limits(t, mu):
    L = orc_limits()
    m = np.asarray(mu, dtype=np.float64)
    lib().orc_energy_limits(C.byref(t), _dp(m), C.byref(L))
    return L


def rgf(alpha, Hb, Hc):
    alpha = _carr(alpha)
    N = alpha.size
    Hb = _carr(Hb)
    Hc = _carr(Hc)
    out = [np.zeros(N, np.complex128) for _ in range(4)]
    lib().orc_rgf(N, _cp(alpha), _cp(Hb), _cp(Hc), *[_cp(o) for o in out])
    return out  # Xt, X, Yt, Y


def rho0(t):
    out = np.zeros(t.N, np.complex128)
    st = lib().orc_rho0(C.byref(t), _cp(out))
    return out, st


def rho_eq(t, U, Uc):
    U = np.ascontiguousarray(U, np.float64)
    Uc = np.ascontiguousarray(Uc, np.float64)
    out = np.zeros(t.N, np.complex128)
    n = C.c_int()
    st = lib().orc_rho_eq(C.byref(t), _dp(U), _dp(Uc), _cp(out), C.byref(n))
    return out, st


def rho_noneq(t, U, Uc, with_integrand=False):
    U = np.ascontiguousarray(U, np.float64)
    Uc = np.ascontiguousarray(Uc, np.float64)
    out = np.zeros(t.N, np.complex128)
    n = C.c_int()
    total = sum(t.noneq_pts[i] for i in range(t.num_noneq_paths))
    integ = np.zeros(3 * max(total, 1)) if with_integrand else None
    st = lib().orc_rho_noneq(C.byref(t), _dp(U), _dp(Uc), _cp(out),
                             _dp(integ) if with_integrand else None,
                             C.byref(n))
    if with_integrand:
        return out, integ[:3 * n.value].reshape(3, n.value), st
    return out, st


def solve_negf(t, U, Uc, Rho0):
    U = np.ascontiguousarray(U, np.float64)
    Uc = np.ascontiguousarray(Uc, np.float64)
    Rho0 = _carr(Rho0)
    Q = np.zeros(t.N)
    comps = np.zeros(3 * t.N)
    n = C.c_int()
    st = lib().orc_solve_negf(C.byref(t), _dp(U), _dp(Uc), _cp(Rho0), _dp(Q),
                              _dp(comps), C.byref(n))
    return Q, comps.reshape(3, t.N), n.value, st


def adaptive_intercuts(t, U, Uc):
    U = np.ascontiguousarray(U, np.float64)
    Uc = np.ascontiguousarray(Uc, np.float64)
    return lib().orc_adaptive_intercuts(C.byref(t), _dp(U), _dp(Uc))


def accumulate_diagG(U, t_even, t_odd, Uc, lead_beta, lead_gamma, E, coef,
                     want_diag=False):
    U = np.ascontiguousarray(U, np.float64)
    Uc = np.ascontiguousarray(Uc, np.float64)
    E = _carr(E)
    coef = _carr(coef)
    N = U.size
    acc = np.zeros(N)
    dg = np.zeros((E.size, N), np.complex128) if want_diag else None
    st = lib().orc_accumulate_diagG(N, _dp(U), t_even, t_odd, _dp(Uc),
                                    lead_beta, lead_gamma, E.size, _cp(E),
                                    _cp(coef), _dp(acc),
                                    _cp(dg) if want_diag else None)
    return (acc, dg, st) if want_diag else (acc, st)


def accumulate_diagG_par(U, t_even, t_odd, Uc, lead_beta, lead_gamma, E, coef, nthreads):
    """accumulate_diagG over nthreads OpenMP threads (energy chunks)."""
    U = np.ascontiguousarray(U, np.float64)
    Uc = np.ascontiguousarray(Uc, np.float64)
    E = _carr(E)
    coef = _carr(coef)
    acc = np.zeros(U.size)
    st = lib().orc_accumulate_diagG_par(int(nthreads), U.size, _dp(U), t_even, t_odd, _dp(Uc),
                                        lead_beta, lead_gamma, E.size, _cp(E), _cp(coef),
                                        _dp(acc))
    return acc, st


def accumulate_noneq(U, t_even, t_odd, Uc, lead_beta, lead_gamma, E, cL, cR):
    U = np.ascontiguousarray(U, np.float64)
    Uc = np.ascontiguousarray(Uc, np.float64)
    E = _carr(E)
    cL = np.ascontiguousarray(cL, np.float64)
    cR = np.ascontiguousarray(cR, np.float64)
    acc = np.zeros(U.size)
    st = lib().orc_accumulate_noneq(U.size, _dp(U), t_even, t_odd, _dp(Uc),
                                    lead_beta, lead_gamma, E.size, _cp(E),
                                    _dp(cL), _dp(cR), _dp(acc))
    return acc, st


def dos_transmission(t, U, Uc, E, atoms_per_unitcell, want_ldos=False):
    U = np.ascontiguousarray(U, np.float64)
    Uc = np.ascontiguousarray(Uc, np.float64)
    E = _carr(E)
    dos = np.zeros(E.size)
    tr = np.zeros(E.size)
    ld = np.zeros((E.size, t.N)) if want_ldos else None
    st = lib().orc_dos_transmission(C.byref(t), _dp(U), _dp(Uc), E.size,
                                    _cp(E), atoms_per_unitcell, _dp(dos),
                                    _dp(tr), _dp(ld) if want_ldos else None)
    return (dos, tr, ld, st) if want_ldos else (dos, tr, st)


class Broyden:
    def __init__(self, S, fraction, max_step, norm_type=0):
        self.S = S
        self.h = lib().orc_broyden_create(S, fraction, max_step, norm_type)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_broyden_free(self.h)
            self.h = None

    def set_nin(self, n):
        n = np.ascontiguousarray(n, np.float64)
        lib().orc_broyden_set_nin(self.h, _dp(n))

    def get_nin(self):
        n = np.zeros(self.S)
        lib().orc_broyden_get_nin(self.h, _dp(n))
        return n

    def step(self, n_out):
        n_out = np.ascontiguousarray(n_out, np.float64)
        mx = C.c_double()
        l2 = C.c_double()
        nps = np.zeros(self.S)
        st = lib().orc_broyden_step(self.h, _dp(n_out), C.byref(mx),
                                    C.byref(l2), _dp(nps))
        return mx.value, l2.value, nps, st

    def reset(self, keep, init_value=0.0):
        lib().orc_broyden_reset(self.h, int(keep), init_value)

    @property
    def step_count(self):
        return lib().orc_broyden_get_step(self.h)


def coulomb_kernel(ptd_nm, eps_r=1.0, R_ring_nm=0.666, R_image_nm=3.0):
    ptd = np.ascontiguousarray(ptd_nm, np.float64)
    N = ptd.size
    K = np.zeros((N, N))
    c = orc_coulomb(eps_r, R_ring_nm, R_image_nm)
    lib().orc_coulomb_kernel(N, _dp(ptd), C.byref(c), _dp(K))
    return K


def scf_bias_point(t, Rho0, K, V_gate, broyden, fraction=0.1, max_norm=1e-6,
                   norm_type=0, max_step=200, reset_with_previous=1,
                   initial_value=0.0, U_ext=None):
    ext = None if U_ext is None else np.ascontiguousarray(U_ext, np.float64)
    sp = orc_scf_params(fraction, max_norm, norm_type, max_step,
                        reset_with_previous, initial_value,
                        _dp(ext) if ext is not None else None)
    Rho0 = _carr(Rho0)
    K = np.ascontiguousarray(K, np.float64)
    U = np.zeros(t.N)
    Q = np.zeros(t.N)
    it = C.c_int()
    fn = C.c_double()
    st = lib().orc_scf_bias_point(C.byref(t), _cp(Rho0), _dp(K), V_gate,
                                  C.byref(sp), broyden.h, _dp(U), _dp(Q),
                                  C.byref(it), C.byref(fn))
    return U, Q, it.value, fn.value, st


def write_transport_char(filename, E, mu, kT, dos, tran, cond):
    arrs = [np.ascontiguousarray(a, np.float64) for a in (E, mu, kT, dos, tran, cond)]
    lib().orc_write_transport_char(filename.encode(), arrs[0].size, *[_dp(a) for a in arrs])


def write_ldos(folder, E, ptd, ldos):
    E = np.ascontiguousarray(E, np.float64)
    ptd = np.ascontiguousarray(ptd, np.float64)
    ldos = np.ascontiguousarray(ldos, np.float64)
    lib().orc_write_ldos(folder.encode(), E.size, _dp(E), ptd.size, _dp(ptd), _dp(ldos))


def write_step_files(prefix, ptd, Q=None, U=None, norm=None):
    ptd = np.ascontiguousarray(ptd, np.float64)
    arrs = [None if a is None else np.ascontiguousarray(a, np.float64) for a in (Q, U, norm)]
    lib().orc_write_step_files(prefix.encode(), ptd.size, _dp(ptd),
                               *[None if a is None else _dp(a) for a in arrs])


def current(t, U, Uc):
    U = np.ascontiguousarray(U, np.float64)
    Uc = np.ascontiguousarray(Uc, np.float64)
    I = np.zeros(2)
    st = lib().orc_current(C.byref(t), _dp(U), _dp(Uc), _dp(I))
    return I, st


def write_qin(prefix, ptd, Qin):
    ptd = np.ascontiguousarray(ptd, np.float64)
    q = np.ascontiguousarray(Qin, np.float64)
    lib().orc_write_qin(prefix.encode(), ptd.size, _dp(ptd), _dp(q))


def write_charge_components(filename, ptd, eq, neq, r0, Q):
    a = [np.ascontiguousarray(x, np.float64) for x in (ptd, eq, neq, r0, Q)]
    lib().orc_write_charge_components(filename.encode(), a[0].size, *[_dp(x) for x in a])


def write_current_header(filename):
    lib().orc_write_current_header(filename.encode())


def write_current_line(filename, step, Vds, Vgs, I, avg_pts, iters, conductance):
    I = np.ascontiguousarray(I, np.float64)
    lib().orc_write_current_line(filename.encode(), C.c_int(step), C.c_double(Vds),
                                 C.c_double(Vgs), _dp(I), C.c_double(avg_pts),
                                 C.c_int(iters), C.c_double(conductance))
