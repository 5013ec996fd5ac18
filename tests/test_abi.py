"""CPU checks of the C-ABI library (no GPU): it loads, exports every symbol
include/negf_amd.h declares, and its host-only entry points (parameter
defaults, quadrature paths, output writers) agree with the oracle."""
import os
import re

import numpy as np
import pytest

from electrostatic_amd import _lib, negf
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "negf_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    names = re.findall(r"\b(negf_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 40
    missing = [n for n in names if getattr(lib, n, None) is None]
    assert not missing, missing
    # and the ctypes table covers them all
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_header_is_plain_c(tmp_path):
    # a C caller (cgo, a C material class) includes the header as C99
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "use.c"
    src.write_text('#include "negf_amd.h"\nint (*probe)(int, const double *, const double *, double *) = '
                   'negf_block_probe;\nint main(void) { return 0; }\n')
    subprocess.run([cc, "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.dirname(HEADER), str(src)],
                   check=True)


def test_version_and_limits():
    lib = _lib.load()
    assert lib.negf_version() >= 100
    assert lib.negf_max_sites() == 8192 * 128


def test_tube_param_defaults_follow_the_reference():
    p = negf.tube_params()
    assert tuple(p.type_id) == (17, 0)
    assert p.acc == 1.42e-10 and p.gamma == 2.5
    assert p.E_valence_min == -14.0 and p.E_pole_max == 10.0 and p.E_zPlus_imag == 1e-8
    assert tuple(p.Fermi_tail_factors) == (14.0, 14.0)
    assert p.num_noneq_paths == 1
    assert p.decimation_layers == 10 and p.decimation_max_iter == 30
    assert p.decimation_rel_error == 0           # int(1.e-5) in the reference
    assert p.integrand_correction_interval == 500
    assert tuple(p.kT_window_around_singularity) == (1.0, 1.0)
    assert tuple(p.contact_T) == (298.0, 298.0)
    q = negf.tube_params(noneq_integration_pts=(50, 60, 70), contact_mu=(-1, -1.1))
    assert q.num_noneq_paths == 3 and tuple(q.contact_mu) == (-1.0, -1.1)
    with pytest.raises(KeyError):
        negf.tube_params(no_such_key=1)


@pytest.mark.parametrize("kind,a,b,n", [(0, -1 + 0.5j, 2 + 1j, 30), (1, -0.3 + 3.1j, -10 + 1e-5j, 64),
                                        (0, -10 + 1e-8j, 0.4 + 1e-8j, 200)])
def test_gl_path_matches_oracle(kind, a, b, n):
    from electrostatic_amd.workloads import gl_path
    E, w, mul = gl_path(a, b, n, kind)
    Eo, wo, mo = O.gl_path(a, b, n, kind)
    # same Newton iteration; compilers contract multiply-adds differently, and
    # the outermost weights (taken at the previous Newton iterate) amplify it
    assert (np.abs(w - wo) / wo).max() < 1e-11
    assert np.abs(E - Eo).max() < 1e-14 * max(1.0, np.abs(Eo).max())
    assert np.abs(mul - mo).max() < 1e-14 * max(1.0, np.abs(mo).max())


def test_gl_path_rejects_bad_arguments():
    lib = _lib.load()
    z = _lib.negf_complex(0, 0)
    rc = lib.negf_gl_path(z, z, 0, 0, None, None, None)
    assert rc == 1 and b"invalid" in lib.negf_last_error()


def test_step_files_match_reference_format(tmp_path):
    rng = np.random.default_rng(0)
    N = 37
    ptd = np.sort(rng.normal(size=N)) * 7.3
    Q = rng.normal(size=N) * 1e-3
    Q[3] = 0.0
    Q[5] = 1.25e-17
    U = rng.normal(size=N)
    nrm = np.abs(rng.normal(size=N)) * 1e-7
    a = str(tmp_path / "prod")
    b = str(tmp_path / "orc")
    negf.write_step_files(a, ptd, Q=Q, U=U, norm=nrm)
    O.write_step_files(b, ptd, Q=Q, U=U, norm=nrm)
    for suffix in ("_Qout.dat", "_U.dat", "_norm.dat"):
        pa = open(a + suffix, "rb").read()
        pb = open(b + suffix, "rb").read()
        assert pa == pb, suffix
    lines = open(a + "_Qout.dat").read().splitlines()
    assert lines[0] == "'axial location / (nm)', 'Induced charge per site / (e)'"
    assert len(lines) == N + 1 and len(lines[1]) == 70


def test_qin_and_charge_components_match_reference_format(tmp_path):
    rng = np.random.default_rng(5)
    N = 21
    ptd = np.linspace(-3.2, 3.2, N)
    q = rng.normal(size=N) * 1e-3
    comps = [rng.normal(size=N) for _ in range(3)]
    a, b = str(tmp_path / "p"), str(tmp_path / "o")
    negf.write_qin(a, ptd, q)
    O.write_qin(b, ptd, q)
    assert open(a + "_Qin.dat", "rb").read() == open(b + "_Qin.dat", "rb").read()
    negf.write_charge_components(a + "_chargeComp.dat", ptd, *comps, q)
    O.write_charge_components(b + "_chargeComp.dat", ptd, *comps, q)
    assert open(a + "_chargeComp.dat", "rb").read() == open(b + "_chargeComp.dat", "rb").read()


def test_current_file_matches_reference_format(tmp_path):
    a, b = str(tmp_path / "I_p.dat"), str(tmp_path / "I_o.dat")
    for f, mod in ((a, negf), (b, O)):
        mod.write_current_header(f)
        mod.write_current_line(f, 3, 0.1, 0.5, np.array([1.234e-7, -1.234e-7]), 212.5, 17, 0.731)
        mod.write_current_line(f, 4, 0.2, 0.5, np.array([2.5e-6, -2.5e-6]), 200.0, 9, 1.25)
    assert open(a, "rb").read() == open(b, "rb").read()


def test_restart_reads_qout(tmp_path):
    rng = np.random.default_rng(6)
    N = 40
    ptd = np.linspace(-5, 5, N)
    Q = rng.normal(size=N) * 1e-3
    pre = str(tmp_path / "step0007")
    negf.write_step_files(pre, ptd, Q=Q)
    back = negf.read_charge_distribution(pre + "_Qout.dat", N)
    # the file carries 15 significant digits, as the reference writes it
    assert np.abs(back - Q).max() <= 1e-14 * np.abs(Q).max()
    with pytest.raises(_lib.NegfError):
        negf.read_charge_distribution(pre + "_Qout.dat", N + 1)


def test_conductance_of_a_perfect_channel():
    # T = 1 over a grid spanning mu +- 30 kT: the two sign flips of the
    # reference (dE = E_0 - E_1 < 0 in both the difference and the sum)
    # cancel and the total is the integral of -dF/dE = 1
    kT = 8.617333262e-5 * 298.0
    E = np.linspace(-1.0 - 30 * kT, -1.0 + 30 * kT, 2001)
    cond, tot = negf.conductance(E, np.ones_like(E), -1.0, kT)
    assert abs(tot - 1.0) < 1e-5
    assert np.all(cond <= 0.0)                # -T dF/dE with dE < 0
    # two conducting modes double it
    _, tot2 = negf.conductance(E, 2 * np.ones_like(E), -1.0, kT)
    assert abs(tot2 - 2 * tot) < 1e-14


def test_transport_char_and_ldos_writers_match_oracle(tmp_path):
    """transport_char.dat (Write_Eql_Characteristics NEGF_Common.cpp:1931-1970)
    and Ept_<e>.dat (:2778-2796 via Write_Table1D) byte for byte against the
    oracle's iostream restatement, including Fermi factors that underflow."""
    lib = _lib.load()
    rng = np.random.default_rng(7)
    nE, N = 37, 23
    E = np.sort(rng.uniform(-3.0, 3.0, nE))
    E[0] = -40.0           # exp(-39/kT) underflows to 0: f = 1
    E[-1] = 45.0           # exp overflows: f = 0
    mu = np.array([-1.0, -1.2])
    kT = np.array([0.025680, 0.02])
    dos = rng.uniform(0, 5, nE) * 10.0 ** rng.integers(-12, 3, nE)
    tr = rng.uniform(0, 2, nE)
    cond = rng.normal(size=nE) * 1e-3
    a = str(tmp_path / "a.dat")
    b = str(tmp_path / "b.dat")
    rc = lib.negf_write_transport_char(a.encode(), nE, _ptr(E), _ptr(mu), _ptr(kT), _ptr(dos), _ptr(tr), _ptr(cond))
    assert rc == 0
    O.write_transport_char(b, E, mu, kT, dos, tr, cond)
    assert open(a, "rb").read() == open(b, "rb").read()
    ptd = np.linspace(-2.0, 2.0, N)
    ldos = rng.uniform(0, 1, (nE, N)) * 10.0 ** rng.integers(-8, 2, (nE, N))
    (tmp_path / "g").mkdir()
    (tmp_path / "o").mkdir()
    rc = lib.negf_write_ldos(str(tmp_path / "g").encode(), nE, _ptr(E), N, _ptr(ptd), _ptr(ldos))
    assert rc == 0
    O.write_ldos(str(tmp_path / "o"), E, ptd, ldos)
    for e in range(nE):
        name = f"Ept_{e}.dat"
        assert (tmp_path / "g" / name).read_bytes() == (tmp_path / "o" / name).read_bytes()


def _ptr(a):
    import ctypes as C
    return a.ctypes.data_as(C.c_void_p)
