"""GPU tests of the block material (negf_block_material_*, negf_block_solve,
negf_block_scf_bias_point): the tube material's contour, poles, Rho0 and
charge combine on dense 32 x 32 block chains.

Independent check: a block chain built from 16 decoupled two-site dimers
(one orbital pair per transverse mode of a zigzag tube, each mode as often
as its degeneracy, orbitals permuted at random so that the dense block code
is exercised) is the tube's mode-space Hamiltonian, so the orbital charges
summed over modes must equal the oracle's site charge of the tube
(NEGF_Common.cpp Solve_NEGF restated) on the same contour.  The block path
computes the lead surface Green's functions by Sancho-Rubio decimation
(tolerance 1e-12) where the oracle uses the closed form, and inverts 32 x 32
blocks: tolerance 1e-8 of max |Q|.
"""
import numpy as np
import pytest

from electrostatic_amd import _lib, negf
from electrostatic_amd.workloads import gaussian_bumps
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GAMMA = 2.5
M = 32   # orbitals per block in the SCF test


def dimer_blocks(betas, degens, seed=0):
    """(H0, T, perm): one orbital pair per mode and degenerate copy (mode k
    repeated degens[k] times), pair q on orbitals (2q, 2q+1) before the
    permutation perm (orbital i of the block is old orbital perm[i])."""
    pairs = [b for b, d in zip(betas, degens) for _ in range(int(d))]
    M = 2 * len(pairs)
    H0 = np.zeros((M, M))
    T = np.zeros((M, M))
    for q, b in enumerate(pairs):
        H0[2 * q, 2 * q + 1] = H0[2 * q + 1, 2 * q] = b   # even link inside the block
        T[2 * q + 1, 2 * q] = GAMMA                         # odd link to the next block
    perm = np.random.default_rng(seed).permutation(M)
    return H0[np.ix_(perm, perm)], T[np.ix_(perm, perm)], perm


def site_potential_to_orbitals(U, perm):
    nb = U.size // 2
    M = perm.size
    v = np.empty((nb, M))
    old = np.empty((nb, M))
    for q in range(M // 2):
        old[:, 2 * q] = U[0::2]
        old[:, 2 * q + 1] = U[1::2]
    v[:, :] = old[:, perm]
    return v


def orbitals_to_sites(Q, perm):
    nb = Q.shape[0]
    old = np.empty_like(Q)
    old[:, perm] = Q
    site = np.zeros(2 * nb)
    site[0::2] = old[:, 0::2].sum(axis=1)
    site[1::2] = old[:, 1::2].sum(axis=1)
    return site


def cnt_modes(num_modes=8):
    # (17, 0) zigzag modes of the tube material, each of degeneracy 2 here
    t = negf.Tube(num_unitcells=4, num_modes=num_modes)
    _, betas, degs = t.modes()
    return list(betas), list(degs)


CONTOUR = dict(contact_mu=(-0.4, -0.4), eq_integration_pts=(24, 24, 32), E_zPlus_imag=1e-5,
               E_valence_min=-10.0, E_pole_max=3.0)


@pytest.mark.parametrize("num_modes,dup", [(8, True), (8, False), (5, False)])
def test_block_solve_matches_tube_oracle_by_mode_decomposition(dev, num_modes, dup):
    """Blocks of 32 (8 modes, each twice), 16 and 10 orbitals (the last two
    run padded to 32 on the device)."""
    betas, degs = cnt_modes(num_modes)
    if not dup:
        degs = [1] * len(betas)
    nb = 24
    N = 2 * nb
    H0, T, perm = dimer_blocks(betas, degs, seed=3)
    chain = negf.BlockChain(nb, H0, T)
    mat = negf.BlockMaterial(chain, **CONTOUR)
    U = gaussian_bumps(N, seed=4, amp=0.3)
    Uc = (U[0], U[-1])
    mat.set_potential(site_potential_to_orbitals(U, perm), Uc)
    Q, npts = mat.solve()
    ot = O.make_tube(N, betas, degs, gamma=GAMMA, mu=CONTOUR["contact_mu"],
                     eq_pts=CONTOUR["eq_integration_pts"], eta=1e-5,
                     E_valence_min=-10.0, E_pole_max=3.0)
    r0, st0 = O.rho0(ot)
    Qo, comps, n_o, st = O.solve_negf(ot, U, np.array(Uc), r0)
    assert st0 == 0 and st == 0
    # Rho0 alone, and the combined charge
    rho0_sites = orbitals_to_sites(mat.rho0(), perm)
    assert np.abs(rho0_sites - r0.imag).max() < 1e-8 * np.abs(r0.imag).max()
    Qs = orbitals_to_sites(Q, perm)
    assert np.abs(Qs - Qo).max() < 1e-8 * np.abs(Qo).max()
    assert npts == n_o
    old = np.empty_like(Q)
    old[:, perm] = Q
    if dup:   # every pair of the same mode carries the same charge
        assert np.abs(old[:, 0:2] - old[:, 2:4]).max() < 1e-10 * np.abs(old).max()
    # diag G^R of each orbital against the mode's chain at a few energies
    E = np.array([-1.3 + 1e-5j, 0.2 + 0.05j, -4.0 + 1.0j])
    _, dg, _ = chain.rgf(E, want_diag=True)
    pairs = [b for b, d in zip(betas, degs) for _ in range(int(d))]
    for q, b in enumerate(pairs):
        _, g1, st1 = O.accumulate_diagG(U, b, GAMMA, np.array(Uc), b, GAMMA, E, np.ones(E.size), want_diag=True)
        assert st1 == 0
        for j in range(2):
            i = int(np.where(perm == 2 * q + j)[0][0])
            ref = g1[:, j::2]
            assert np.abs(dg[:, :, i] - ref).max() < 1e-9 * np.abs(ref).max(), (q, j)
    chain.close()


def test_block_material_rejects_bias_and_keeps_equilibrium(dev):
    betas, degs = cnt_modes()
    H0, T, perm = dimer_blocks(betas, degs)
    chain = negf.BlockChain(6, H0, T)
    mat = negf.BlockMaterial(chain, **CONTOUR)
    Q0, _ = mat.solve()
    with pytest.raises(_lib.NegfError, match="equilibrium"):
        mat.set_contacts((-0.4, -0.6))
    Q1, _ = mat.solve()   # the previous contacts stay in force
    assert np.array_equal(Q0, Q1)
    chain.close()


def test_block_scf_matches_tube_scf(dev):
    """The block SCF over the decomposed tube, with the tube SCF's kernel
    spread over the orbitals of each site and the leads at the end sites'
    orbitals, must reach the tube SCF's self-consistent potential."""
    num_unitcells = 6
    tube = negf.Tube(num_unitcells=num_unitcells, num_modes=8, **CONTOUR)
    _, betas, degs = tube.modes()
    N = tube.N
    assert N % 2 == 0
    nb = N // 2
    sp = dict(Broyden_max_norm=1e-9, Broyden_fraction=0.1)
    scf = negf.Scf(tube, **sp)
    Vg = 0.3
    U_t, Q_t, it_t, nrm_t = scf.bias_point(Vg)[:4]
    K = scf.kernel()
    H0, T, perm = dimer_blocks(betas, degs, seed=1)
    chain = negf.BlockChain(nb, H0, T)
    mat = negf.BlockMaterial(chain, **CONTOUR)
    # orbital i of block b sits on site 2b + (old orbital index of i) % 2
    site_of = np.array([2 * b + perm[i] % 2 for b in range(nb) for i in range(M)])
    Kb = K[np.ix_(site_of, site_of)]
    left = int(np.where(perm == 0)[0][0])      # old orbital 0: site 0 of block 0
    right = int(np.where(perm == 1)[0][0])     # old orbital 1: site N-1 of the last block
    U_b, Q_b, it_b, nrm_b = mat.scf_bias_point(Kb, Vg, contact_orbitals=(left, right), max_iters=200,
                                               Broyden_max_norm=1e-10, Broyden_fraction=0.1,
                                               Broyden_threshold_maxstep=200)
    assert nrm_b <= 1e-10
    U_sites = np.empty(N)
    U_sites[site_of] = U_b
    assert np.abs(U_sites - U_t).max() < 1e-6
    Qs = orbitals_to_sites(Q_b.reshape(nb, M), perm)
    assert np.abs(Qs - Q_t).max() < 1e-6 * max(np.abs(Q_t).max(), 1e-3)
    # the returned charge is self-consistent: U = -V_gate - K Q
    assert np.abs(U_b - (-Vg - Kb @ Q_b)).max() < 1e-6
    chain.close()
