"""GPU tests of the batched multi-tube solve (negf_batch_*): per-tube charge
equals each tube's own solve and the oracle; graph replay and the segment
length autotuner do not change results.

Tolerance: 1e-10 relative to the oracle; 1e-12 between launch forms (the
batch groups chains differently, so partial sums round differently).
"""
import numpy as np
import pytest
import torch

from electrostatic_amd import negf
from electrostatic_amd.workloads import gaussian_bumps
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def make_tubes(cells, biases, modes=2, seed=0):
    tubes, Us = [], []
    for i, (nc, V) in enumerate(zip(cells, biases)):
        t = negf.Tube(num_unitcells=nc, num_modes=modes, contact_mu=(-1.0, -1.0 - V),
                      eq_integration_pts=(24, 24, 32), noneq_integration_pts=(48,),
                      E_zPlus_imag=1e-5, E_valence_min=-10.0, E_pole_max=3.0)
        U = gaussian_bumps(t.N, seed=seed + i, amp=0.25)
        t.set_potential(U)
        tubes.append(t)
        Us.append(U)
    return tubes, Us


def test_batch_matches_single_tube_solves_and_oracle(dev):
    cells = [10, 37, 3, 120, 64]
    biases = [0.0, 0.2, 0.35, 0.0, 0.1]
    tubes, Us = make_tubes(cells, biases)
    singles = [t.solve(1)[0] for t in tubes]
    batch = negf.Batch(tubes)
    dQ = torch.zeros(int(batch.offsets[-1]), dtype=torch.float64, device=dev)
    batch.solve_device(1, dQ.data_ptr())
    torch.cuda.synchronize()
    Q = dQ.cpu().numpy()
    for i, t in enumerate(tubes):
        q = Q[batch.offsets[i]:batch.offsets[i + 1]]
        assert rel(q, singles[i]) < 1e-12, i
    # one tube against the oracle
    i = 1
    _, betas, degs = tubes[i].modes()
    ot = O.make_tube(tubes[i].N, list(betas), list(degs), mu=(-1.0, -1.2),
                     eq_pts=(24, 24, 32), noneq_pts=(48,), eta=1e-5,
                     E_valence_min=-10.0, E_pole_max=3.0)
    r0, _ = O.rho0(ot)
    Qo, _, _, _ = O.solve_negf(ot, Us[i], (Us[i][0], Us[i][-1]), r0)
    assert rel(Q[batch.offsets[i]:batch.offsets[i + 1]], Qo) < 1e-10
    batch.close()


def test_batch_graph_replay_and_autotune(dev):
    tubes, Us = make_tubes([30, 7, 80], [0.15, 0.0, 0.3], modes=1, seed=5)
    batch = negf.Batch(tubes)
    n = int(batch.offsets[-1])
    d1 = torch.zeros(n, dtype=torch.float64, device=dev)
    batch.solve_device(1, d1.data_ptr())
    torch.cuda.synchronize()
    ref = d1.cpu().numpy()
    batch.use_graph(True)
    d2 = torch.empty(n, dtype=torch.float64, device=dev)
    for rep in range(3):
        d2.fill_(float("nan"))            # the replay must rewrite every site
        batch.solve_device(2 + rep, d2.data_ptr())
        torch.cuda.synchronize()
        assert rel(d2.cpu().numpy(), ref) < 1e-14
        # eager work between replays
        d1.add_(1.0)
        torch.cuda.synchronize()
    L, ms = batch.autotune(d1.data_ptr())
    assert L in (32, 48, 64, 96, 128) and ms > 0
    d3 = torch.zeros(n, dtype=torch.float64, device=dev)
    batch.solve_device(5, d3.data_ptr())
    torch.cuda.synchronize()
    assert rel(d3.cpu().numpy(), ref) < 1e-12
    # a contact change in one tube rebuilds the batch (and its graph)
    tubes[1].set_contacts((-1.0, -1.25))
    batch.solve_device(6, d3.data_ptr())
    torch.cuda.synchronize()
    q1, _ = tubes[1].solve(7)
    assert rel(d3.cpu().numpy()[batch.offsets[1]:batch.offsets[2]], q1) < 1e-12
    batch.close()


def test_batch_long_uneven_tubes(dev):
    # lengths far beyond one segment scan of 64 segments, up to 20000 sites
    # (the 20000-site tube is in equilibrium: the oracle's O(N^2) contact walk
    # would take minutes at that length)
    tubes, Us = make_tubes([125, 5000, 1210], [0.3, 0.0, 0.1], modes=1, seed=9)
    batch = negf.Batch(tubes)
    dQ = torch.zeros(int(batch.offsets[-1]), dtype=torch.float64, device=dev)
    batch.solve_device(1, dQ.data_ptr())
    torch.cuda.synchronize()
    Q = dQ.cpu().numpy()
    i = 1   # N = 20000
    _, betas, degs = tubes[i].modes()
    ot = O.make_tube(tubes[i].N, list(betas), list(degs), mu=(-1.0, -1.0),
                     eq_pts=(24, 24, 32), noneq_pts=(48,), eta=1e-5,
                     E_valence_min=-10.0, E_pole_max=3.0)
    r0, _ = O.rho0(ot)
    Qo, _, _, _ = O.solve_negf(ot, Us[i], (Us[i][0], Us[i][-1]), r0)
    assert rel(Q[batch.offsets[i]:batch.offsets[i + 1]], Qo) < 1e-10
    # the biased long tube against its own single-tube solve
    for j in (0, 2):
        qj, _ = tubes[j].solve(1)
        assert rel(Q[batch.offsets[j]:batch.offsets[j + 1]], qj) < 1e-12
    batch.close()


def test_batched_scf_matches_oracle_per_tube(dev):
    T, cells = 4, 40
    tubes = [negf.Tube(num_unitcells=cells, num_modes=1, contact_mu=(-1.0, -1.0),
                       eq_integration_pts=(30, 30, 64), noneq_integration_pts=(200,),
                       E_zPlus_imag=1e-5, E_valence_min=-10.0, E_pole_max=3.0) for _ in range(T)]
    N = tubes[0].N
    batch = negf.Batch(tubes)
    sp = dict(eps_r=3.9, R_ring_nm=0.665, R_image_nm=10.0)
    scf = negf.BatchScf(batch, Broyden_fraction=0.1, Broyden_max_norm=1e-8,
                        Broyden_threshold_maxstep=200, **sp)
    U_ext = np.stack([gaussian_bumps(N, seed=20 + t, amp=0.3) for t in range(T)])
    V = np.array([0.0, 0.25, 0.5, 0.8])
    scf.set_external_potential(U_ext)
    U, Q, it, nrm, tot = scf.run(V)
    assert np.all(nrm <= 1e-8) and tot == it.max()
    _, betas, degs = tubes[0].modes()
    K = O.coulomb_kernel(tubes[0].positions(), **sp)
    for t in range(T):
        ot = O.make_tube(N, list(betas), list(degs), mu=(-1.0, -1.0), eq_pts=(30, 30, 64),
                         noneq_pts=(200,), eta=1e-5, E_valence_min=-10.0, E_pole_max=3.0)
        r0, _ = O.rho0(ot)
        ob = O.Broyden(N, 0.1, 200, 0)
        Uo, Qo, ito, nrmo, st = O.scf_bias_point(ot, r0, K, V[t], ob, max_norm=1e-8,
                                                 U_ext=U_ext[t])
        assert st == 0
        assert np.abs(U[t] - Uo).max() < 1e-6, t
        assert abs(int(it[t]) - ito) <= 1, (t, it[t], ito)
    # the graph path and the eager path agree
    scf.use_graph(False)
    scf.set_input_charge(np.zeros((T, N)))
    U2, _, it2, _, _ = scf.run(V)
    assert np.abs(U2 - U).max() < 1e-10
    scf.close()
    batch.close()


def test_surrogate_preconditioned_scf_matches_plain(dev):
    # a small surrogate trained on NEGF pairs of these tubes; the seeded,
    # preconditioned run must reach the same fixed point as the plain run
    from electrostatic_amd.surrogate import Surrogate, negf_pairs, sample_potentials
    T, cells = 4, 40
    tubes = [negf.Tube(num_unitcells=cells, num_modes=1, contact_mu=(-1.0, -1.0),
                       eq_integration_pts=(30, 30, 64), noneq_integration_pts=(200,),
                       E_zPlus_imag=1e-5, E_valence_min=-10.0, E_pole_max=3.0) for _ in range(T)]
    N = tubes[0].N
    batch = negf.Batch(tubes)
    U_tr = sample_potentials(256, N, seed=1)
    dU, dQ = negf_pairs(batch, U_tr, device=dev)
    sg = Surrogate(N, seed=3)
    loss = sg.train(dU, dQ, epochs=30, batch=16, lr=3e-3, lr_end=3e-4, seed=2)
    assert np.isfinite(loss)
    sp = dict(eps_r=3.9, R_ring_nm=0.665, R_image_nm=10.0)
    scf = negf.BatchScf(batch, Broyden_fraction=0.1, Broyden_max_norm=1e-8,
                        Broyden_threshold_maxstep=200, **sp)
    U_ext = np.stack([gaussian_bumps(N, seed=40 + t, amp=0.3) for t in range(T)])
    V = np.array([0.0, 0.3, 0.6, 1.0])
    scf.set_external_potential(U_ext)
    scf.set_input_charge(np.zeros((T, N)))
    U0, Q0, it0, nrm0, tot0 = scf.run(V)
    assert np.all(nrm0 <= 1e-8)
    scf.set_surrogate(sg)
    for seed_iters, graph in ((0, True), (1, True), (1, False)):
        scf.use_graph(graph)
        scf.set_input_charge(np.zeros((T, N)))
        U1, Q1, it1, nrm1, tot1, cap = scf.run_surrogate(V, seed_iters=seed_iters, refresh_iter=3)
        assert np.all(nrm1 <= 1e-8)
        assert np.all(cap > 0)
        assert np.abs(U1 - U0).max() < 1e-6, (seed_iters, np.abs(U1 - U0).max())
        assert np.abs(Q1 - Q0).max() < 1e-6
    scf.close()
    sg.close()
    batch.close()


def _small_scf(dev, T=3, cells=20):
    tubes = [negf.Tube(num_unitcells=cells, num_modes=1, contact_mu=(-1.0, -1.0),
                       eq_integration_pts=(24, 24, 32), noneq_integration_pts=(48,),
                       E_zPlus_imag=1e-5, E_valence_min=-10.0, E_pole_max=3.0) for _ in range(T)]
    batch = negf.Batch(tubes)
    scf = negf.BatchScf(batch, Broyden_fraction=0.1, Broyden_max_norm=1e-8,
                        Broyden_threshold_maxstep=200, eps_r=3.9, R_ring_nm=0.665, R_image_nm=10.0)
    return tubes, batch, scf


def test_batched_scf_graph_recaptured_after_batch_rebuild(dev):
    # a rebuild of the batch's plans from outside the SCF (segment length,
    # autotune) moves their buffers; the next graph replay must use the new ones
    tubes, batch, scf = _small_scf(dev)
    T, N = len(tubes), tubes[0].N
    U_ext = np.stack([gaussian_bumps(N, seed=60 + t, amp=0.3) for t in range(T)])
    V = np.array([0.0, 0.4, 0.9])
    scf.set_external_potential(U_ext)
    scf.set_input_charge(np.zeros((T, N)))
    U_g1, _, it1, _, _ = scf.run(V)
    batch.set_segment_length(32)          # larger workspace: plan buffers reallocated
    scf.set_input_charge(np.zeros((T, N)))
    U_g2, _, it2, nrm2, _ = scf.run(V)
    scf.use_graph(False)
    scf.set_input_charge(np.zeros((T, N)))
    U_e, _, it_e, _, _ = scf.run(V)
    assert np.all(nrm2 <= 1e-8)
    assert np.abs(U_g2 - U_e).max() < 1e-10
    assert np.abs(U_g1 - U_e).max() < 1e-10
    assert list(it2) == list(it_e)
    dQ = torch.zeros(int(batch.offsets[-1]), dtype=torch.float64, device=dev)
    batch.autotune(dQ.data_ptr())
    scf.use_graph(True)
    scf.set_input_charge(np.zeros((T, N)))
    U_g3, _, _, _, _ = scf.run(V)
    assert np.abs(U_g3 - U_e).max() < 1e-10
    scf.close()
    batch.close()


def test_batched_scf_external_potential_change_between_graph_runs(dev):
    tubes, batch, scf = _small_scf(dev)
    T, N = len(tubes), tubes[0].N
    V = np.array([0.1, 0.5, 0.7])
    results = {}
    for graph in (True, False):
        scf.use_graph(graph)
        out = []
        for U_ext in (None, np.stack([gaussian_bumps(N, seed=70 + t, amp=0.3) for t in range(T)]), None):
            scf.set_external_potential(U_ext)
            scf.set_input_charge(np.zeros((T, N)))
            out.append(scf.run(V)[0])
        results[graph] = out
    for a, b in zip(results[True], results[False]):
        assert np.abs(a - b).max() < 1e-10
    # the runs with and without U_ext differ, and clearing U_ext restores the first
    assert np.abs(results[True][0] - results[True][1]).max() > 1e-3
    assert np.abs(results[True][0] - results[True][2]).max() < 1e-10
    scf.close()
    batch.close()


def test_batched_scf_rejects_bad_host_inputs(dev):
    tubes, batch, scf = _small_scf(dev)
    T, N = len(tubes), tubes[0].N
    bad = np.zeros((T, N))
    bad[1, 3] = np.nan
    with pytest.raises(Exception):
        scf.set_external_potential(bad)
    with pytest.raises(ValueError):
        scf.run(0.2)                      # one gate value for three tubes
    scf.close()
    batch.close()


def test_batch_autotune_cache(dev, tmp_path):
    # a miss tunes and appends one line; a second batch of the same workload
    # takes the stored length without timing; a different workload misses
    cache = tmp_path / "tune.txt"
    tubes, _ = make_tubes([30, 7, 80], [0.15, 0.0, 0.3], modes=1, seed=5)
    batch = negf.Batch(tubes)
    n = int(batch.offsets[-1])
    d = torch.zeros(n, dtype=torch.float64, device=dev)
    L1, ms1, hit1 = batch.autotune_cached(d.data_ptr(), cache)
    assert not hit1 and L1 in (32, 48, 64, 96, 128) and ms1 > 0
    lines = cache.read_text().splitlines()
    assert len(lines) == 1 and lines[0].split()[-2] == str(L1)
    batch.solve_device(1, d.data_ptr())
    torch.cuda.synchronize()
    ref = d.cpu().numpy()
    batch.close()
    tubes2, _ = make_tubes([30, 7, 80], [0.15, 0.0, 0.3], modes=1, seed=5)
    batch2 = negf.Batch(tubes2)
    L2, ms2, hit2 = batch2.autotune_cached(d.data_ptr(), cache)
    assert hit2 and L2 == L1 and ms2 == pytest.approx(ms1, rel=1e-5)
    d2 = torch.zeros(n, dtype=torch.float64, device=dev)
    batch2.solve_device(2, d2.data_ptr())
    torch.cuda.synchronize()
    assert rel(d2.cpu().numpy(), ref) < 1e-14
    batch2.close()
    tubes3, _ = make_tubes([30, 8, 80], [0.15, 0.0, 0.3], modes=1, seed=5)
    batch3 = negf.Batch(tubes3)
    d3 = torch.zeros(int(batch3.offsets[-1]), dtype=torch.float64, device=dev)
    _, _, hit3 = batch3.autotune_cached(d3.data_ptr(), cache)
    assert not hit3 and len(cache.read_text().splitlines()) == 2
    batch3.close()
