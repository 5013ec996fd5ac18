"""Python face of the C ABI for tests, benchmarks and scripting.

Mirrors the reference's NEGF material interface (c_NEGF_Common / c_CNT,
Source/Solver/Transport/NEGF/NEGF_Common.H:478-605) with its input key names;
every call goes to libnegf_amd.so (HIP, gfx950).  No computation happens in
Python beyond argument marshalling.
"""
import ctypes as C

import numpy as np

from . import _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def tube_params(**keys):
    """negf_tube_params with the reference defaults, overridden by input keys,
    e.g. num_unitcells=100, eq_integration_pts=(30, 30, 64),
    noneq_integration_pts=(200,), contact_mu=(-1, -1), E_zPlus_imag=1e-5.
    noneq_percent_intercuts defaults to uniform cuts as in the reference."""
    lib = _lib.load()
    p = _lib.negf_tube_params()
    lib.negf_tube_params_default(C.byref(p))
    for k, v in keys.items():
        if not hasattr(p, k):
            raise KeyError(f"unknown tube parameter {k}")
        cur = getattr(p, k)
        if hasattr(cur, "__len__"):
            vals = list(v)
            if len(vals) > len(cur):
                raise ValueError(f"{k}: at most {len(cur)} values")
            for i, x in enumerate(vals):
                cur[i] = x
        else:
            setattr(p, k, v)
    if "noneq_integration_pts" in keys and "num_noneq_paths" not in keys:
        p.num_noneq_paths = len(list(keys["noneq_integration_pts"]))
    return p


class Tube:
    """One NEGF tube (c_CNT in mode space) on the GPU."""

    def __init__(self, params=None, **keys):
        self.lib = _lib.load()
        self.params = params if params is not None else tube_params(**keys)
        h = C.c_void_p()
        _lib.check(self.lib.negf_tube_create(C.byref(self.params), C.byref(h)))
        self.h = h
        self.N = self.lib.negf_tube_num_sites(h)

    def close(self):
        if self.h:
            self.lib.negf_tube_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def positions(self):
        out = np.zeros(self.N)
        _lib.check(self.lib.negf_tube_positions(self.h, _dp(out)))
        return out

    def modes(self):
        ids = np.zeros(32, np.int32)
        beta = np.zeros(32)
        deg = np.zeros(32, np.int32)
        n = self.lib.negf_tube_modes(self.h, _ip(ids), _dp(beta), _ip(deg))
        if n < 0:
            _lib.check(n)
        return ids[:n].copy(), beta[:n].copy(), deg[:n].copy()

    def set_potential(self, U, Uc=None):
        U = np.ascontiguousarray(U, np.float64)
        Uc = np.ascontiguousarray(Uc if Uc is not None else (U[0], U[-1]), np.float64)
        _lib.check(self.lib.negf_tube_set_potential(self.h, _dp(U), _dp(Uc)))

    def set_contacts(self, mu, T=(298.0, 298.0)):
        m = np.ascontiguousarray(mu, np.float64)
        t = np.ascontiguousarray(T, np.float64)
        _lib.check(self.lib.negf_tube_set_contacts(self.h, _dp(m), _dp(t)))

    def rho0(self):
        out = np.zeros(self.N)
        _lib.check(self.lib.negf_compute_rho0(self.h, _dp(out)))
        return out

    def solve(self, iter=0):
        """Solve_NEGF: induced charge per site (e) and integration points."""
        Q = np.zeros(self.N)
        n = C.c_int()
        _lib.check(self.lib.negf_solve(self.h, iter, _dp(Q), C.byref(n)))
        return Q, n.value

    def solve_device(self, iter, dQ_ptr, stream=None):
        n = C.c_int()
        _lib.check(self.lib.negf_solve_device(self.h, iter, C.c_void_p(dQ_ptr), C.byref(n),
                                              C.c_void_p(stream) if stream else None))
        return n.value

    def status(self):
        _lib.check(self.lib.negf_tube_status(self.h))

    def charge_components(self):
        eq, neq, r0 = np.zeros(self.N), np.zeros(self.N), np.zeros(self.N)
        _lib.check(self.lib.negf_get_charge_components(self.h, _dp(eq), _dp(neq), _dp(r0)))
        return eq, neq, r0

    def limits(self):
        out = np.zeros(10)
        _lib.check(self.lib.negf_tube_limits(self.h, _dp(out)))
        keys = ["mu_min", "mu_max", "kT_min", "kT_max", "flag_noneq", "num_poles",
                "E_contour_right", "E_zeta_imag", "E_eta", "E_rightmost"]
        return dict(zip(keys, out))

    def noneq_layout(self):
        n = C.c_int()
        pts = np.zeros(8, np.int32)
        cuts = np.zeros(8)
        _lib.check(self.lib.negf_tube_noneq_layout(self.h, C.byref(n), _ip(pts), _dp(cuts)))
        return pts[:n.value].copy(), cuts[:max(0, n.value - 1)].copy()

    def noneq_integrand(self):
        n = self.lib.negf_tube_noneq_integrand(self.h, 0, None, None)
        if n < 0:
            _lib.check(-n)
        E = np.zeros(max(n, 1))
        I = np.zeros(3 * max(n, 1))
        n = self.lib.negf_tube_noneq_integrand(self.h, n, _dp(E), _dp(I))
        return E[:n], I[:3 * n].reshape(3, n)

    def current(self):
        """Compute_Current: terminal currents (A) at contacts 0 and 1."""
        I = np.zeros(2)
        _lib.check(self.lib.negf_tube_current(self.h, _dp(I)))
        return I

    def dos(self, E_real):
        E = np.ascontiguousarray(E_real, np.float64)
        d = np.zeros(E.size)
        t = np.zeros(E.size)
        _lib.check(self.lib.negf_tube_dos(self.h, E.size, _dp(E), _dp(d), _dp(t)))
        return d, t

    def dos_complex(self, E, want_ldos=False):
        """DOS, transmission (and LDOS [nE][N]) at complex energies."""
        Ec = np.ascontiguousarray(E, np.complex128)
        d = np.zeros(Ec.size)
        t = np.zeros(Ec.size)
        ld = np.zeros((Ec.size, self.N)) if want_ldos else None
        _lib.check(self.lib.negf_tube_dos_complex(self.h, Ec.size, Ec.ctypes.data, _dp(d), _dp(t),
                                                  _dp(ld) if want_ldos else None))
        return (d, t, ld) if want_ldos else (d, t)

    def dos_path(self, which):
        """Energies of the reference's DOS path: 0 flat-band, 1 per step."""
        n = C.c_int()
        _lib.check(self.lib.negf_tube_dos_path(self.h, int(which), 0, None, C.byref(n)))
        E = np.zeros(n.value, np.complex128)
        _lib.check(self.lib.negf_tube_dos_path(self.h, int(which), n.value, E.ctypes.data, C.byref(n)))
        return E


class Batch:
    """Many tubes solved in one launch sequence (negf_batch_*)."""

    def __init__(self, tubes):
        self.lib = _lib.load()
        self.tubes = list(tubes)
        arr = (C.c_void_p * len(self.tubes))(*[t.h for t in self.tubes])
        h = C.c_void_p()
        _lib.check(self.lib.negf_batch_create(arr, len(self.tubes), C.byref(h)))
        self.h = h
        off = (C.c_longlong * (len(self.tubes) + 1))()
        _lib.check(self.lib.negf_batch_offsets(h, off))
        self.offsets = np.array(off[:], dtype=np.int64)

    def close(self):
        if self.h:
            self.lib.negf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve_device(self, iter, dQ_ptr, stream=None):
        _lib.check(self.lib.negf_batch_solve(self.h, iter, C.c_void_p(dQ_ptr),
                                             C.c_void_p(stream) if stream else None))

    def use_graph(self, on=True):
        _lib.check(self.lib.negf_batch_use_graph(self.h, int(on)))

    def set_segment_length(self, L):
        _lib.check(self.lib.negf_batch_set_segment_length(self.h, L))

    def autotune(self, dQ_ptr):
        L = C.c_int()
        ms = C.c_double()
        _lib.check(self.lib.negf_batch_autotune(self.h, C.c_void_p(dQ_ptr), C.byref(L), C.byref(ms)))
        return L.value, ms.value

    def autotune_cached(self, dQ_ptr, cache_path):
        """Segment length from the tuning cache at cache_path (tuned and
        appended on a miss); returns (L, ms, from_cache)."""
        L = C.c_int()
        ms = C.c_double()
        hit = C.c_int()
        _lib.check(self.lib.negf_batch_autotune_cached(self.h, C.c_void_p(dQ_ptr), str(cache_path).encode(),
                                                       C.byref(L), C.byref(ms), C.byref(hit)))
        return L.value, ms.value, bool(hit.value)

    def enable_timing(self, on=True):
        _lib.check(self.lib.negf_batch_enable_timing(self.h, int(on)))

    def last_time_ms(self):
        ms = C.c_double()
        _lib.check(self.lib.negf_batch_last_time(self.h, C.byref(ms)))
        return ms.value

    def energy_sites(self):
        v = C.c_double()
        _lib.check(self.lib.negf_batch_energy_sites(self.h, C.byref(v)))
        return v.value

    def status(self):
        rc = self.lib.negf_batch_status(self.h)
        if rc not in (0, 3):
            _lib.check(rc)
        return rc


class BatchScf:
    """Batched SCF of equal tubes sharing one electrostatic kernel."""

    def __init__(self, batch, params=None, **keys):
        self.lib = _lib.load()
        self.batch = batch
        self.T = len(batch.tubes)
        self.N = batch.tubes[0].N
        self.params = params if params is not None else scf_params(**keys)
        arr = (C.c_void_p * self.T)(*[t.h for t in batch.tubes])
        h = C.c_void_p()
        _lib.check(self.lib.negf_batch_scf_create(batch.h, arr, self.T, C.byref(self.params), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.negf_batch_scf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_external_potential(self, U_ext):
        """U_ext: [T][N] in eV, or None to clear it (zeros)."""
        if U_ext is None:
            self._U_ext = None
            _lib.check(self.lib.negf_batch_scf_set_external_potential(self.h, None))
            return
        self._U_ext = np.ascontiguousarray(U_ext, np.float64).reshape(self.T * self.N)
        _lib.check(self.lib.negf_batch_scf_set_external_potential(self.h, _dp(self._U_ext)))

    def set_input_charge(self, n_in):
        n = np.ascontiguousarray(n_in, np.float64).reshape(self.T * self.N)
        _lib.check(self.lib.negf_batch_scf_set_input_charge(self.h, _dp(n)))

    def get_input_charge(self):
        n = np.zeros(self.T * self.N)
        _lib.check(self.lib.negf_batch_scf_get_input_charge(self.h, _dp(n)))
        return n.reshape(self.T, self.N)

    def use_graph(self, on=True):
        _lib.check(self.lib.negf_batch_scf_use_graph(self.h, int(on)))

    def _gates(self, V_gate):
        V = np.ascontiguousarray(V_gate, np.float64)
        if V.size != self.T:
            raise ValueError(f"V_gate needs one value per tube ({self.T}), got {V.size}")
        return V.reshape(self.T)

    def run(self, V_gate):
        V = self._gates(V_gate)
        U = np.zeros(self.T * self.N)
        Q = np.zeros(self.T * self.N)
        it = np.zeros(self.T, np.int32)
        nrm = np.zeros(self.T)
        tot = C.c_int()
        _lib.check(self.lib.negf_batch_scf_run(self.h, _dp(V), _dp(U), _dp(Q), _ip(it), _dp(nrm),
                                               C.byref(tot)))
        return U.reshape(self.T, self.N), Q.reshape(self.T, self.N), it, nrm, tot.value


    # ---- surrogate-seeded, preconditioned runs
    def kernel(self):
        K = np.zeros(self.N * self.N)
        _lib.check(self.lib.negf_batch_scf_get_kernel(self.h, _dp(K)))
        return K.reshape(self.N, self.N)

    def set_surrogate(self, surrogate, feat=None):
        """Attach a charge surrogate (electrostatic_amd.surrogate.Surrogate of
        the same N) and the eigenbasis of the shared kernel (host eigen
        decomposition, once per geometry)."""
        lam, Z = np.linalg.eigh(self.kernel())
        Z = np.ascontiguousarray(Z)
        _lib.check(self.lib.negf_batch_scf_set_preconditioner(self.h, _dp(Z), _dp(np.ascontiguousarray(lam))))
        f = None if feat is None else np.ascontiguousarray(feat, np.float64)
        _lib.check(self.lib.negf_batch_scf_set_surrogate(self.h, surrogate.h, None if f is None else _dp(f)))
        self._surrogate = surrogate

    def run_surrogate(self, V_gate, seed_iters=6, refresh_iter=0):
        V = self._gates(V_gate)
        U = np.zeros(self.T * self.N)
        Q = np.zeros(self.T * self.N)
        it = np.zeros(self.T, np.int32)
        nrm = np.zeros(self.T)
        cap = np.zeros(self.T)
        tot = C.c_int()
        _lib.check(self.lib.negf_batch_scf_run_surrogate(self.h, _dp(V), seed_iters, refresh_iter, _dp(U), _dp(Q),
                                                         _ip(it), _dp(nrm), C.byref(tot), _dp(cap)))
        return U.reshape(self.T, self.N), Q.reshape(self.T, self.N), it, nrm, tot.value, cap


class Broyden:
    """Broyden modified-second mixer on the GPU (Broyden_Second_Parallel.cpp)."""

    def __init__(self, S, fraction=0.1, max_step=100, norm_type=0):
        self.lib = _lib.load()
        h = C.c_void_p()
        _lib.check(self.lib.negf_broyden_create(S, fraction, max_step, norm_type, C.byref(h)))
        self.h = h
        self.S = S

    def close(self):
        if self.h:
            self.lib.negf_broyden_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_nin(self, n):
        n = np.ascontiguousarray(n, np.float64)
        _lib.check(self.lib.negf_broyden_set_nin(self.h, _dp(n)))

    def get_nin(self):
        out = np.zeros(self.S)
        _lib.check(self.lib.negf_broyden_get_nin(self.h, _dp(out)))
        return out

    def step(self, n_out):
        n_out = np.ascontiguousarray(n_out, np.float64)
        mx, l2 = C.c_double(), C.c_double()
        per = np.zeros(self.S)
        _lib.check(self.lib.negf_broyden_step(self.h, _dp(n_out), C.byref(mx), C.byref(l2), _dp(per)))
        return mx.value, l2.value, per

    def reset(self, keep_n_in=False, init_value=0.0):
        _lib.check(self.lib.negf_broyden_reset(self.h, int(keep_n_in), init_value))

    @property
    def step_count(self):
        return self.lib.negf_broyden_get_step(self.h)


class MultiScf:
    """SCF over the tubes of a batch (any lengths) with one Broyden system over
    their concatenated sites (negf_multi_scf_*)."""

    def __init__(self, batch, params=None, **keys):
        self.lib = _lib.load()
        self.batch = batch
        self.T = len(batch.tubes)
        self.params = params if params is not None else scf_params(**keys)
        arr = (C.c_void_p * self.T)(*[t.h for t in batch.tubes])
        h = C.c_void_p()
        _lib.check(self.lib.negf_multi_scf_create(batch.h, arr, self.T, C.byref(self.params), C.byref(h)))
        self.h = h
        S = C.c_longlong()
        _lib.check(self.lib.negf_multi_scf_num_sites(h, C.byref(S)))
        self.S = S.value

    def close(self):
        if self.h:
            self.lib.negf_multi_scf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_external_potential(self, U_ext):
        U = np.ascontiguousarray(U_ext, np.float64)
        if U.size != self.S:
            raise ValueError(f"U_ext needs {self.S} values, got {U.size}")
        self._U_ext = U.reshape(self.S)
        _lib.check(self.lib.negf_multi_scf_set_external_potential(self.h, _dp(self._U_ext)))

    def use_graph(self, on=True):
        _lib.check(self.lib.negf_multi_scf_use_graph(self.h, int(on)))

    def bias_point(self, V_gate, max_iters=0):
        """Returns (U, Q, iterations, final norm, ms per iteration)."""
        V = np.ascontiguousarray(V_gate, np.float64)
        if V.size != self.T:
            raise ValueError(f"V_gate needs one value per tube ({self.T}), got {V.size}")
        U = np.zeros(self.S)
        Q = np.zeros(self.S)
        it = C.c_int()
        nrm = C.c_double()
        ms = C.c_double()
        _lib.check(self.lib.negf_multi_scf_bias_point(self.h, _dp(V), int(max_iters), _dp(U), _dp(Q),
                                                      C.byref(it), C.byref(nrm), C.byref(ms)))
        return U, Q, it.value, nrm.value, ms.value

    def input_charge(self):
        n = np.zeros(self.S)
        _lib.check(self.lib.negf_multi_scf_get_input_charge(self.h, _dp(n)))
        return n


def scf_params(eps_r=3.9, R_ring_nm=0.665, R_image_nm=10.0, Broyden_fraction=0.1,
               Broyden_max_norm=1e-6, Broyden_norm_type=0, Broyden_threshold_maxstep=200,
               reset_with_previous_charge_distribution=1, NS_initial_deposit_value=0.0):
    p = _lib.negf_scf_params()
    p.eps_r = eps_r
    p.R_ring_nm = R_ring_nm
    p.R_image_nm = R_image_nm
    p.Broyden_fraction = Broyden_fraction
    p.Broyden_max_norm = Broyden_max_norm
    p.Broyden_norm_type = Broyden_norm_type
    p.Broyden_threshold_maxstep = Broyden_threshold_maxstep
    p.reset_with_previous_charge_distribution = reset_with_previous_charge_distribution
    p.NS_initial_deposit_value = NS_initial_deposit_value
    return p


class Scf:
    """Self-consistent loop of one tube (c_TransportSolver::Solve) with the
    dense electrostatic stand-in."""

    def __init__(self, tube, params=None, **keys):
        self.lib = _lib.load()
        self.tube = tube
        self.params = params if params is not None else scf_params(**keys)
        h = C.c_void_p()
        _lib.check(self.lib.negf_scf_create(tube.h, C.byref(self.params), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.negf_scf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def kernel(self):
        N = self.tube.N
        K = np.zeros(N * N)
        _lib.check(self.lib.negf_scf_kernel(self.h, _dp(K)))
        return K.reshape(N, N)

    def set_external_potential(self, U_ext):
        if U_ext is None:
            _lib.check(self.lib.negf_scf_set_external_potential(self.h, None))
            return
        self._U_ext = np.ascontiguousarray(U_ext, np.float64)
        _lib.check(self.lib.negf_scf_set_external_potential(self.h, _dp(self._U_ext)))

    def use_graph(self, on=True):
        _lib.check(self.lib.negf_scf_use_graph(self.h, int(on)))

    def bias_point(self, V_gate):
        N = self.tube.N
        U, Q = np.zeros(N), np.zeros(N)
        it, nrm = C.c_int(), C.c_double()
        _lib.check(self.lib.negf_scf_bias_point(self.h, V_gate, _dp(U), _dp(Q), C.byref(it), C.byref(nrm)))
        return U, Q, it.value, nrm.value


def write_step_files(prefix, ptd, Q=None, U=None, norm=None):
    lib = _lib.load()
    ptd = np.ascontiguousarray(ptd, np.float64)
    arrs = [None if a is None else np.ascontiguousarray(a, np.float64) for a in (Q, U, norm)]
    ptrs = [None if a is None else _dp(a) for a in arrs]
    _lib.check(lib.negf_write_step_files(prefix.encode(), ptd.size, _dp(ptd), *ptrs))


def write_qin(prefix, ptd, Qin):
    lib = _lib.load()
    ptd = np.ascontiguousarray(ptd, np.float64)
    q = np.ascontiguousarray(Qin, np.float64)
    _lib.check(lib.negf_write_qin(prefix.encode(), ptd.size, _dp(ptd), _dp(q)))


def write_charge_components(filename, ptd, eq, neq, rho0, Q):
    lib = _lib.load()
    a = [np.ascontiguousarray(x, np.float64) for x in (ptd, eq, neq, rho0, Q)]
    _lib.check(lib.negf_write_charge_components(filename.encode(), a[0].size, *[_dp(x) for x in a]))


def write_current_header(filename):
    _lib.check(_lib.load().negf_write_current_header(filename.encode()))


def write_current_line(filename, step, Vds, Vgs, I, avg_intg_pts, total_iterations, conductance):
    I = np.ascontiguousarray(I, np.float64)
    _lib.check(_lib.load().negf_write_current_line(filename.encode(), step, Vds, Vgs, _dp(I),
                                                   avg_intg_pts, total_iterations, conductance))


def read_charge_distribution(filename, N):
    Q = np.zeros(N)
    _lib.check(_lib.load().negf_read_charge_distribution(filename.encode(), N, _dp(Q)))
    return Q


def conductance(E, T, mu0, kT0):
    """Compute_Conductance: per-energy conductance and the summed total."""
    E = np.ascontiguousarray(E, np.float64)
    T = np.ascontiguousarray(T, np.float64)
    cond = np.zeros(E.size)
    tot = C.c_double()
    _lib.check(_lib.load().negf_conductance(E.size, _dp(E), _dp(T), mu0, kT0, _dp(cond), C.byref(tot)))
    return cond, tot.value


class BlockChain:
    """Chain of dense m x m blocks, m <= 32 (negf_block_*): on-site H0 + diag(v_b),
    coupling T = H_{b,b+1}, leads continuing the chain; Sancho-Rubio lead
    surface Green's functions and the block RGF sweep on FP64 MFMA."""

    def __init__(self, nblocks, H0, T):
        self.lib = _lib.load()
        self.nb = nblocks
        H0 = np.ascontiguousarray(H0, np.float64)
        T = np.ascontiguousarray(T, np.float64)
        self.M = H0.shape[0]   # orbitals per block (at most 32)
        if H0.shape != (self.M, self.M) or T.shape != (self.M, self.M):
            raise ValueError("H0 and T must be square of the same size")
        h = C.c_void_p()
        _lib.check(self.lib.negf_block_chain_create(self.M, nblocks, _dp(H0), _dp(T), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.negf_block_chain_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_potential(self, v, Uc=(0.0, 0.0)):
        v = np.ascontiguousarray(v, np.float64).reshape(self.nb * self.M)
        uc = np.ascontiguousarray(Uc, np.float64)
        _lib.check(self.lib.negf_block_chain_set_potential(self.h, _dp(v), _dp(uc)))

    def set_scratch_limit(self, nbytes):
        """Bound on the device scratch for the stored g_b (the sweep runs the
        energies in chunks that fit; results do not depend on it)."""
        _lib.check(self.lib.negf_block_chain_set_scratch_limit(self.h, int(nbytes)))

    def rgf(self, E, coef=None, tol=1e-12, max_iter=200, want_diag=False):
        """(acc [nb, 32], diagG [nE, nb, 32] or None, Sancho-Rubio iterations [nE, 2])."""
        E = np.ascontiguousarray(np.atleast_1d(E), np.complex128)
        nE = E.size
        coef = np.ones(nE, np.complex128) if coef is None else np.ascontiguousarray(coef, np.complex128)
        acc = np.zeros(self.nb * self.M)
        diag = np.zeros(nE * self.nb * self.M, np.complex128) if want_diag else None
        it = np.zeros(2 * nE, np.int32)
        _lib.check(self.lib.negf_block_rgf(self.h, nE, E.ctypes.data_as(C.c_void_p), coef.ctypes.data_as(C.c_void_p),
                                           tol, max_iter, _dp(acc),
                                           None if diag is None else diag.ctypes.data_as(C.c_void_p), _ip(it)))
        return (acc.reshape(self.nb, self.M),
                None if diag is None else diag.reshape(nE, self.nb, self.M), it.reshape(nE, 2))

    def self_energies(self, E, tol=1e-12, max_iter=200):
        """(Sigma_L, Sigma_R) [nE, 32, 32] complex and iterations [nE, 2]."""
        E = np.ascontiguousarray(np.atleast_1d(E), np.complex128)
        nE = E.size
        sig = np.zeros(nE * 4 * self.M * self.M)
        it = np.zeros(2 * nE, np.int32)
        _lib.check(self.lib.negf_block_self_energies(self.h, nE, E.ctypes.data_as(C.c_void_p), tol, max_iter,
                                                     _dp(sig), _ip(it)))
        s = sig.reshape(nE, 2, 2, self.M, self.M)
        z = s[:, :, 0] + 1j * s[:, :, 1]
        return z[:, 0], z[:, 1], it.reshape(nE, 2)

    def last_time(self):
        a = C.c_float()
        b = C.c_float()
        _lib.check(self.lib.negf_block_last_time(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value


class BlockMaterial:
    """Equilibrium charge and self-consistent loop of a BlockChain with the
    tube material's contour, poles and Rho0 (negf_block_material_*).  The
    contour settings are tube input keys (contact_mu, contact_T,
    E_valence_min, E_pole_max, E_zPlus_imag, eq_integration_pts, ...)."""

    def __init__(self, chain, spin_degen=2.0, sr_tol=1e-12, sr_max_iter=200, params=None, **keys):
        self.lib = _lib.load()
        self.chain = chain
        self.S = chain.nb * chain.M
        self.params = params if params is not None else tube_params(**keys)
        h = C.c_void_p()
        _lib.check(self.lib.negf_block_material_create(chain.h, C.byref(self.params), float(spin_degen),
                                                       float(sr_tol), int(sr_max_iter), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.negf_block_material_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_potential(self, v, Uc=(0.0, 0.0)):
        v = np.ascontiguousarray(v, np.float64).reshape(self.S)
        uc = np.ascontiguousarray(Uc, np.float64)
        _lib.check(self.lib.negf_block_material_set_potential(self.h, _dp(v), _dp(uc)))

    def set_contacts(self, mu, T=(298.0, 298.0)):
        m = np.ascontiguousarray(mu, np.float64)
        t = np.ascontiguousarray(T, np.float64)
        _lib.check(self.lib.negf_block_material_set_contacts(self.h, _dp(m), _dp(t)))

    def rho0(self):
        out = np.zeros(self.S)
        _lib.check(self.lib.negf_block_material_rho0(self.h, _dp(out)))
        return out.reshape(self.chain.nb, self.chain.M)

    def solve(self):
        """(Q [nb, 32] = RhoEq - Rho0, number of integration points)."""
        Q = np.zeros(self.S)
        n = C.c_int()
        _lib.check(self.lib.negf_block_solve(self.h, _dp(Q), C.byref(n)))
        return Q.reshape(self.chain.nb, self.chain.M), n.value

    def scf_bias_point(self, K, V_gate, U_ext=None, contact_orbitals=None, max_iters=100, params=None, **keys):
        """Self-consistent (U, Q, iterations, final norm) with
        U = U_ext - V_gate - K n (K: S x S); the leads take U of
        contact_orbitals (first block, last block) or the end blocks' mean."""
        sp = params if params is not None else scf_params(**keys)
        co = None if contact_orbitals is None else (C.c_int * 2)(*[int(x) for x in contact_orbitals])
        K = np.ascontiguousarray(K, np.float64).reshape(self.S, self.S)
        ue = None if U_ext is None else np.ascontiguousarray(U_ext, np.float64).reshape(self.S)
        U = np.zeros(self.S)
        Q = np.zeros(self.S)
        it = C.c_int()
        nrm = C.c_double()
        _lib.check(self.lib.negf_block_scf_bias_point(self.h, _dp(K), None if ue is None else _dp(ue), float(V_gate),
                                                      co, C.byref(sp), int(max_iters), _dp(U), _dp(Q), C.byref(it),
                                                      C.byref(nrm)))
        return U, Q, it.value, nrm.value
