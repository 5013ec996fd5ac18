#!/usr/bin/env python3
"""Benchmark of the batched RGF hot path on one MI355X.

Workload (one "step"): 8192 complex energies (4096 Gauss-Legendre points on
the equilibrium contour arc, 4096 on the real axis with eta = 1e-5 eV) over a
4096-site chain with hoppings alternating 2.5*cos(pi/17) / 2.5 eV and a seeded
smooth potential in [-0.5, 0.5] eV; output: the energy-integrated site charge
(fused variant, diag G^R not materialised).  Inputs are resident in HBM when
the timed region starts.

Prints one JSON line with the throughput (energy*site/s), the RGF kernel's
roofline figures (FP64 flops per launch over its hipEvent-timed duration),
the parity of the timed charge against the oracle on the same arrays, and a
CPU baseline (the oracle's serial RGF on the same arrays: one thread, and
energy chunks over every host core).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RGF energy·site/s (% FP64 peak); SCF bias points/s surrogate vs plain"
FLOP_PER_ENERGY_SITE = 31          # algorithmic FP64 flops (add/mul/div = 1, FMA = 2)
FP64_PEAK_TFLOPS = 78.6            # MI355X FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0              # MI355X HBM3E
HOT_PATH_KERNELS = "rgf_eq_fused_kernel (RGF hot path: transfer matrices, walks and equilibrium sweep in one launch)"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "rgf_seg_batched_traffic.json")


def host_cpu():
    """CPU model (from /proc/cpuinfo) and the cores this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return model, cores


def cpu_baseline(wl, max_seconds=15.0):
    """The oracle (serial host RGF per energy, the reference's loop order) on
    the full workload: once on one thread (the reference's per-rank
    behaviour, USE_OMP=FALSE) on at most ~max_seconds of work, and once with
    the energy list split over every core this process may use.  Returns the
    baseline record and the all-core charge, which is also the parity
    reference of the GPU result."""
    from oracle import oracle as O
    E, coef = wl["E"], wl["coef"]
    n = E.size
    args = (wl["U"], wl["t_even"], wl["t_odd"], wl["Uc"], wl["t_even"], wl["t_odd"])
    model, cores = host_cpu()
    # all cores, full workload
    t0 = time.perf_counter()
    ref, _ = O.accumulate_diagG_par(*args, E, coef, cores)
    dt_all = time.perf_counter() - t0
    # one thread: size the sample from a 256-energy prefix
    k0 = min(256, n)
    t0 = time.perf_counter()
    O.accumulate_diagG(*args, E[:k0], coef[:k0])
    dt0 = time.perf_counter() - t0
    k = n if dt0 * n / k0 <= max_seconds else max(k0, int(k0 * max_seconds / dt0))
    t0 = time.perf_counter()
    O.accumulate_diagG(*args, E[:k], coef[:k])
    dt1 = time.perf_counter() - t0
    rate_all = n * wl["N"] / dt_all
    rate_1 = k * wl["N"] / dt1
    sample = (f"full workload {n} energies x {wl['N']} sites on {cores} threads "
              f"(energy chunks, OpenMP) in {dt_all:.2f} s; one thread: {k} of {n} energies "
              f"in {dt1:.2f} s (serial host RGF per energy, oracle restatement of the "
              f"reference loop order)")
    rec = {"value": rate_all, "unit": "energy·site/s", "cores": cores, "kind": "port",
           "sample": sample, "value_1core": rate_1, "cpu_model": model, "nproc": cores}
    return rec, ref


def scf_sweep(with_cpu=True):
    """SCF gate sweep of one (17,0) tube, N = 400 sites (100 unit cells), 1 mode,
    contour 30/30/64 + 19 poles, eta = 1e-5 eV, contacts at -1 eV, external
    potential = 4 seeded Gaussians (amplitude <= 0.3 eV), metal voltage 0..1 V
    in 0.1 V steps, Broyden fraction 0.1, absolute norm <= 1e-6, warm start.
    Reports bias points per second on the GPU (whole SCF loop per bias point,
    device-resident, each iteration replayed from one captured graph) and for
    the oracle's serial SCF on one host thread."""
    from electrostatic_amd import negf
    from electrostatic_amd.workloads import gaussian_bumps
    keys = dict(num_unitcells=100, num_modes=1, contact_mu=(-1.0, -1.0),
                eq_integration_pts=(30, 30, 64), noneq_integration_pts=(200,),
                E_zPlus_imag=1e-5, E_valence_min=-10.0, E_pole_max=3.0)
    sp = dict(eps_r=3.9, R_ring_nm=0.665, R_image_nm=10.0)
    tube = negf.Tube(**keys)
    scf = negf.Scf(tube, Broyden_fraction=0.1, Broyden_max_norm=1e-6,
                   Broyden_threshold_maxstep=200, **sp)
    U_ext = gaussian_bumps(tube.N, seed=1, amp=0.3)
    scf.set_external_potential(U_ext)
    Vs = [round(0.1 * i, 10) for i in range(11)]
    scf.bias_point(0.0)                      # warm-up (plans, first solve)
    scf.close()
    scf = negf.Scf(tube, Broyden_fraction=0.1, Broyden_max_norm=1e-6,
                   Broyden_threshold_maxstep=200, **sp)
    scf.set_external_potential(U_ext)
    iters = []
    t0 = time.perf_counter()
    for V in Vs:
        U, Q, it, nrm = scf.bias_point(V)
        iters.append(it)
    dt = time.perf_counter() - t0
    out = {"workload": "SCF gate sweep 0..1 V (11 bias points), (17,0) tube N=400, "
                       "contour 30/30/64 + poles, Broyden abs norm 1e-6",
           "bias_points_per_s": len(Vs) / dt, "ms_per_bias_point": 1e3 * dt / len(Vs),
           "iterations": iters, "kind": "plain"}
    if with_cpu:
        from oracle import oracle as O
        _, betas, degs = tube.modes()
        ot = O.make_tube(tube.N, list(betas), list(degs), mu=(-1.0, -1.0),
                         eq_pts=(30, 30, 64), noneq_pts=(200,), eta=1e-5,
                         E_valence_min=-10.0, E_pole_max=3.0)
        r0, _ = O.rho0(ot)
        K = O.coulomb_kernel(tube.positions(), **sp)
        ob = O.Broyden(tube.N, 0.1, 200, 0)
        t0 = time.perf_counter()
        for V in Vs:
            Uo, Qo, ito, nrmo, st = O.scf_bias_point(ot, r0, K, V, ob, U_ext=U_ext)
        dtc = time.perf_counter() - t0
        out["cpu_bias_points_per_s"] = len(Vs) / dtc
        out["cpu_kind"] = "port (oracle serial SCF, reference loop order), 1 thread"
        out["max_abs_dU_last_bias_V"] = float(np.abs(U - Uo).max())
    scf.close()
    tube.close()
    return out


def copy_calibration_leg(dev, mib=256, reps=3):
    """A plain device-to-device copy of mib MiB, reps times: the known-size
    traffic that scales the FETCH_SIZE and WRITE_SIZE counters."""
    import torch
    n = (mib << 20) // 8
    x = torch.arange(n, dtype=torch.float64, device=dev)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    return {"bytes_read": n * 8, "bytes_written": n * 8, "reps": reps, "GB_per_s": 2 * n * 8 / dt / 1e9}


def batch_leg(dev, steps=10, segment_length=0):
    """64 (17,0) tubes, lengths log-uniform in [500, 20000] sites (seed 3),
    2 modes each, 256 energies per tube (contour 32/32/64 + 128 real-axis
    points for the tube's own bias V ~ U(0, 0.5) V, contacts at -1 and -1-V eV),
    smooth potentials; one batched solve per step (all tubes in one launch
    sequence), eager and replayed from a captured hipGraph, segment length
    from the autotuner."""
    import torch
    from electrostatic_amd import negf
    from electrostatic_amd.workloads import smooth_potential, tube_batch
    Ns, biases = tube_batch(n_tubes=64, lo=500, hi=20000, seed=3, bias_max=0.5)
    tubes = []
    for i, (N, V) in enumerate(zip(Ns, biases)):
        cells = max(1, int(round(N / 4)))
        t = negf.Tube(num_unitcells=cells, num_modes=2, contact_mu=(-1.0, -1.0 - V),
                      eq_integration_pts=(32, 32, 64), noneq_integration_pts=(128,),
                      E_zPlus_imag=1e-5, E_valence_min=-10.0, E_pole_max=3.0)
        t.set_potential(smooth_potential(t.N, seed=100 + i, lo=-0.3, hi=0.3))
        tubes.append(t)
    batch = negf.Batch(tubes)
    n = int(batch.offsets[-1])
    dQ = torch.zeros(n, dtype=torch.float64, device=dev)
    if segment_length > 0:
        batch.set_segment_length(segment_length)
        L = segment_length
    else:
        L, tune_ms = batch.autotune(dQ.data_ptr())
    es = batch.energy_sites()
    out = {"workload": "64 uneven (17,0) tubes, N log-uniform [500, 20000] (seed 3), 2 modes, "
                       "256 energies per tube (32/32/64 contour + 128 real axis), per-tube bias U(0, 0.5) V",
           "sites": n, "energy_sites": es, "segment_length": L}
    for graph in (False, True):
        batch.use_graph(graph)
        batch.solve_device(1, dQ.data_ptr())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            batch.solve_device(1, dQ.data_ptr())
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        key = "graph" if graph else "eager"
        out[f"ms_per_solve_{key}"] = ms
        out[f"energy_site_per_s_{key}"] = es / (ms * 1e-3)
    if not np.all(np.isfinite(dQ.cpu().numpy())):
        raise RuntimeError("non-finite batched charge")
    # every tube is biased (V > 0): 128 contour chains at 31 flop and 128
    # real-axis chains at 63 flop per mode and site
    if abs(es - 2 * 256 * n) > 0.5:
        raise RuntimeError(f"unexpected chain count: {es} energy-sites for {n} sites")
    flop = float(n) * 2 * (128 * 31 + 128 * 63)
    ms_solve = out["ms_per_solve_graph"]
    out["flop_per_solve"] = flop
    out["roofline"] = {"bound": "fp64_valu", "achieved": flop / (ms_solve * 1e-3) / 1e12, "peak": 78.6,
                       "unit": "TFLOP/s", "frac": flop / (ms_solve * 1e-3) / 1e12 / 78.6}
    # one SCF iteration over all tubes with one Broyden over the 317972 sites
    # (negf_multi_scf): electrostatic stand-in, batched solve, Broyden,
    # replayed from a captured graph; timed over scf_iters iterations
    scf_iters = 8
    ms_scf = negf.MultiScf(batch, Broyden_fraction=0.1, Broyden_max_norm=1e-12, eps_r=3.9,
                           R_ring_nm=0.665, R_image_nm=10.0)
    ms_scf.set_external_potential(np.concatenate([smooth_potential(t.N, seed=100 + i, lo=-0.3, hi=0.3)
                                                  for i, t in enumerate(tubes)]))
    ms_scf.bias_point(np.zeros(len(tubes)), max_iters=2)   # graph capture, warm-up
    t0 = time.perf_counter()
    U, Q, it, nrm, ms_it = ms_scf.bias_point(np.zeros(len(tubes)), max_iters=scf_iters)
    wall_ms = (time.perf_counter() - t0) * 1e3 / scf_iters
    if it != scf_iters or not np.all(np.isfinite(U)):
        raise RuntimeError("multi-tube SCF iterations failed")
    out["scf"] = {"what": "SCF iteration over all 64 tubes (one Broyden over all sites; dense 1/r-with-image "
                          "stand-in within each tube; batched NEGF solve), hipGraph replay",
                  "iterations_timed": it, "ms_per_iteration_device": ms_it,
                  "ms_per_iteration_wall": wall_ms, "scf_iterations_per_s": 1e3 / wall_ms,
                  "ms_negf_solve": ms_solve, "broyden_norm_after": nrm}
    ms_scf.close()
    batch.close()
    for t in tubes:
        t.close()
    return out


def surrogate_leg(dev, n_pairs=20000, epochs=5, with_cpu=True):
    """128 (17,0) tubes of N = 2048 sites (512 unit cells, 1 mode), contacts
    at -1 eV, contour 30/30/64 + poles, dense 1/r-with-image coupling
    (eps_r 3.9, R 0.665 nm, image 10 nm), U_ext = 4 seeded Gaussians
    (amplitude 0.3 eV, seed 100 + t), gates V_t = linspace(0, 1, 128) V;
    convergence at max |n_in - n_out| <= 1e-8.  The BF16 surrogate is trained
    first (not timed) on n_pairs exact NEGF (U, Q) pairs of seeded smooth
    potentials.  Timed: one batched run from n = 0, plain Broyden (fraction
    0.1) against the run preconditioned by (I + c_t K)^-1 with the
    capacitances c_t from the surrogate's tangent ("surrogate"), and against
    the same run after 6 preconditioned steps on the surrogate charge
    ("surrogate_seeded")."""
    import torch
    from electrostatic_amd import negf
    from electrostatic_amd.surrogate import Surrogate, negf_pairs, sample_potentials
    from electrostatic_amd.workloads import gaussian_bumps
    T, cells = 128, 512
    tubes = [negf.Tube(num_unitcells=cells, num_modes=1, contact_mu=(-1.0, -1.0),
                       eq_integration_pts=(30, 30, 64), noneq_integration_pts=(200,),
                       E_zPlus_imag=1e-5, E_valence_min=-10.0, E_pole_max=3.0) for _ in range(T)]
    N = tubes[0].N
    batch = negf.Batch(tubes)
    t0 = time.perf_counter()
    dU, dQ = negf_pairs(batch, sample_potentials(n_pairs, N, seed=5), device=dev)
    gen_s = time.perf_counter() - t0
    sg = Surrogate(N, seed=1)
    t0 = time.perf_counter()
    loss = sg.train(dU, dQ, epochs=epochs, batch=32, lr=2e-3, lr_end=1e-4, seed=3)
    train_s = time.perf_counter() - t0
    del dU, dQ
    torch.cuda.empty_cache()
    # segment length from the autotuner (both runs use it)
    dQb = torch.zeros(T * N, dtype=torch.float64, device=dev)
    seg_L, _ = batch.autotune(dQb.data_ptr())
    del dQb
    sp = dict(eps_r=3.9, R_ring_nm=0.665, R_image_nm=10.0)
    scf = negf.BatchScf(batch, Broyden_fraction=0.1, Broyden_max_norm=1e-8, Broyden_threshold_maxstep=300, **sp)
    U_ext = np.stack([gaussian_bumps(N, seed=100 + t, amp=0.3) for t in range(T)])
    V = np.linspace(0.0, 1.0, T)
    scf.set_external_potential(U_ext)
    scf.set_surrogate(sg)
    out = {"workload": "128 (17,0) tubes x 2048 sites, 1 mode, dense 1/r coupling, gates 0..1 V, norm <= 1e-8",
           "pairs": n_pairs, "epochs": epochs, "pair_generation_s": gen_s, "training_s": train_s,
           "train_loss": loss, "segment_length": seg_L}
    runs = {}
    variants = (("plain", lambda: scf.run(V)),
                ("surrogate", lambda: scf.run_surrogate(V, seed_iters=0, refresh_iter=0)),
                ("surrogate_seeded", lambda: scf.run_surrogate(V, seed_iters=6, refresh_iter=0)))
    for key, fn in variants:
        scf.set_input_charge(np.zeros((T, N)))
        fn()                      # warm-up: graph capture, path set-up
        scf.set_input_charge(np.zeros((T, N)))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        dt = time.perf_counter() - t0
        runs[key] = r
        out[key] = {"s": dt, "bias_points_per_s": T / dt, "iterations_max": int(r[2].max()),
                    "iterations_mean": float(r[2].mean()), "batched_iterations": int(r[4])}
    U0 = runs["plain"][0]
    for key in ("surrogate", "surrogate_seeded"):
        out[key]["max_abs_dU_vs_plain_V"] = float(np.abs(runs[key][0] - U0).max())
        out[key]["speedup_vs_plain"] = out["plain"]["s"] / out[key]["s"]
    out["capacitance_e_per_eV"] = [float(runs["surrogate"][5].min()), float(runs["surrogate"][5].max())]
    scf.close()
    sg.close()
    batch.close()
    if with_cpu:
        from oracle import oracle as O
        _, betas, degs = tubes[0].modes()
        K = O.coulomb_kernel(tubes[0].positions(), **sp)
        t0 = time.perf_counter()
        for t in (0, T - 1):
            ot = O.make_tube(N, list(betas), list(degs), mu=(-1.0, -1.0), eq_pts=(30, 30, 64),
                             noneq_pts=(200,), eta=1e-5, E_valence_min=-10.0, E_pole_max=3.0)
            r0, _ = O.rho0(ot)
            ob = O.Broyden(N, 0.1, 300, 0)
            O.scf_bias_point(ot, r0, K, V[t], ob, max_norm=1e-8, max_step=300, U_ext=U_ext[t])
        dt = time.perf_counter() - t0
        out["cpu_bias_points_per_s"] = 2 / dt
        out["cpu_kind"] = "port (oracle serial SCF, reference loop order), 1 thread, 2 tubes"
    for t in tubes:
        t.close()
    return out


def surrogate_kernels_leg(dev, B=128, N=2048, reps=10):
    """Forward pass and tangent of an (untrained) surrogate on B seeded
    smooth potentials of N sites, reps times each: the BF16 MFMA
    convolution kernels in isolation (for kernel traces and counters)."""
    import torch
    from electrostatic_amd.surrogate import Surrogate, sample_potentials
    sg = Surrogate(N, seed=1)
    U = torch.as_tensor(sample_potentials(B, N, seed=7), device=dev)
    ones = torch.ones_like(U)
    out = {"workload": f"{B} samples x {N} sites, 6 conv layers (3->64, 4 x 64->64, 64->1), k=9"}
    for name, fn in (("forward", lambda: sg.forward(U)), ("tangent", lambda: sg.jvp(U, ones))):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        out[f"ms_{name}"] = (time.perf_counter() - t0) * 1e3 / reps
    sg.close()
    return out


def block_leg(steps=3, with_cpu=True):
    """Dense-block chain: atomistic (16,0) tube in blocks of two rings (32
    orbitals), 2000 blocks, 1024 energies (512 on a contour arc, 512 on the
    real axis [-3, 3] eV with eta = 1e-5 eV), on-site disorder N(0, 0.05 eV)
    per orbital (seed 4) plus a smooth potential; Sancho-Rubio leads to
    1e-12.  Algorithmic flops per energy-block: 36 m^3 (forward: g T, T^T (gT),
    inverse; backward: g T, X G, Y X^T; real H0 and T).  The general-block
    count of the same sweep (complex T, a second inverse for G: 56 m^3) is
    reported beside it."""
    from electrostatic_amd import negf
    from electrostatic_amd.workloads import zigzag_ring_blocks
    nb, nE, m = 2000, 1024, 32
    H0, T = zigzag_ring_blocks(16, 2.5)
    rng = np.random.default_rng(4)
    v = rng.normal(0.0, 0.05, (nb, m)) + 0.2 * np.sin(np.linspace(0, np.pi, nb))[:, None]
    E = np.concatenate([np.linspace(-3, 3, nE // 2) + 1e-5j,
                        -1.0 + 1.5 * np.exp(1j * np.linspace(0.01, np.pi - 0.01, nE - nE // 2))])
    w = np.full(nE, 1.0 / nE, np.complex128)
    c = negf.BlockChain(nb, H0, T)
    c.set_potential(v, (0.0, 0.0))
    acc, _, it = c.rgf(E, w)
    sr_ms, rgf_ms = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        acc, _, it = c.rgf(E, w)
        a, b = c.last_time()
        sr_ms.append(a)
        rgf_ms.append(b)
    wall = (time.perf_counter() - t0) * 1e3 / steps
    c.close()
    if not np.all(np.isfinite(acc)):
        raise RuntimeError("non-finite block charge")
    flops = 36 * m ** 3 * nb * nE
    rg = float(np.median(rgf_ms))
    out = {"workload": "(16,0) tube, 2000 blocks of 32 orbitals, 1024 energies, disorder N(0, 0.05 eV) "
                       "+ smooth U, Sancho-Rubio to 1e-12",
           "ms_per_evaluation": wall, "ms_sancho_rubio": float(np.median(sr_ms)), "ms_block_rgf": rg,
           "sancho_rubio_iterations_max": int(it.max()),
           "energy_blocks_per_s": nb * nE / (wall * 1e-3),
           "flop_per_energy_block": 36 * m ** 3,
           "rgf_tflops": flops / (rg * 1e-3) / 1e12,
           "rgf_frac_fp64_peak": flops / (rg * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
           "note": "fraction at the executed 36 m^3 flops per energy-block (real H0 and T); "
                   "a general complex block would need 56 m^3"}
    if with_cpu:
        from oracle import block_oracle as B
        # numpy oracle on 4 energies of the same chain
        sample = E[::nE // 4][:4]
        t0 = time.perf_counter()
        B.accumulate(sample, np.ones(4), H0, T, v, (0.0, 0.0))
        dt = time.perf_counter() - t0
        out["cpu_energy_blocks_per_s"] = 4 * nb / dt
        out["cpu_kind"] = "port (numpy block oracle, 4 energies x 2000 blocks)"
    return out


def diag_g_leg(dev, wl, steps=20, warmup=3):
    """The headline workload with diag G^R written ([energy][site] complex FP64,
    16 B per energy·site): ms per evaluation, the hot path's launch time by HIP
    events on its stream, the G^R bytes against the HBM peak, and the written
    rows at four energies against the oracle."""
    import torch
    from electrostatic_amd import _lib
    lib = _lib.load()
    N, nE = wl["N"], wl["E"].size
    dU = torch.from_numpy(wl["U"]).to(dev)
    dE = torch.from_numpy(wl["E"].view(np.float64).copy()).to(dev)
    dc = torch.from_numpy(wl["coef"].view(np.float64).copy()).to(dev)
    acc = torch.zeros(N, dtype=torch.float64, device=dev)
    dG = torch.zeros(2 * nE * N, dtype=torch.float64, device=dev)
    p = _lib.negf_chain_params()
    p.N = N
    p.t_even = wl["t_even"]
    p.t_odd = wl["t_odd"]
    p.Uc[0], p.Uc[1] = wl["Uc"]
    plan = C.c_void_p()
    _lib.check(lib.negf_rgf_plan_create(p, C.c_void_p(dU.data_ptr()), nE, C.c_void_p(dE.data_ptr()),
                                        C.c_void_p(dc.data_ptr()), 1, C.byref(plan)))
    stream = torch.cuda.Stream(dev)
    sp = C.c_void_p(stream.cuda_stream)

    def step():
        _lib.check(lib.negf_rgf_plan_execute(plan, C.c_void_p(acc.data_ptr()), C.c_void_p(dG.data_ptr()), sp))

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    _lib.check(lib.negf_rgf_plan_enable_timing(plan, 1))
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for _ in range(steps):
        step()
    t1.record(stream)
    torch.cuda.synchronize()
    ms_step = t0.elapsed_time(t1) / steps
    kms = C.c_double()
    nl = C.c_int()
    _lib.check(lib.negf_rgf_plan_kernel_time(plan, C.byref(kms), C.byref(nl)))
    _lib.check(lib.negf_rgf_plan_status(plan, sp))
    kernel_ms = kms.value / max(1, nl.value)
    from oracle import oracle as O
    ks = np.array([0, nE // 3, nE // 2, nE - 1])
    _, dg_ref, _ = O.accumulate_diagG(wl["U"], wl["t_even"], wl["t_odd"], wl["Uc"], wl["t_even"], wl["t_odd"],
                                      wl["E"][ks], wl["coef"][ks], want_diag=True)
    g = dG.view(nE, N, 2)[torch.as_tensor(ks, device=dev)].cpu().numpy()
    g = g[..., 0] + 1j * g[..., 1]
    lib.negf_rgf_plan_destroy(plan)
    gbytes = 16.0 * N * nE
    return {
        "workload": "the headline workload with diag G^R written ([energy][site] complex FP64)",
        "ms_per_step": ms_step,
        "kernel_ms": kernel_ms,
        "gr_bytes": gbytes,
        "achieved_GBs": gbytes / (kernel_ms * 1e-3) / 1e9,
        "hbm_peak_GBs": HBM_PEAK_GBS,
        "frac_hbm_peak": gbytes / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "diag_G_max_rel_err_vs_oracle": float(np.abs(g - dg_ref).max() / np.abs(dg_ref).max()),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--segment-length", type=int, default=0,
                    help="RGF segment length (0: library default)")
    ap.add_argument("--no-diag-g", action="store_true",
                    help="skip the diag-G^R-writing leg of the default run")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the secondary 64-tube batched-solve measurement")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the surrogate-preconditioned 128-tube SCF leg")
    ap.add_argument("--no-block", action="store_true",
                    help="skip the dense-block (32-orbital) chain leg")
    ap.add_argument("--no-scf", action="store_true",
                    help="skip the secondary SCF gate-sweep measurement")
    ap.add_argument("--surrogate-kernels", action="store_true",
                    help="also time the surrogate's forward and tangent passes alone")
    ap.add_argument("--write-diag-g", action="store_true",
                    help="also materialise diag G^R ([energy][site], 512 MiB)")
    ap.add_argument("--only", choices=["batch", "block", "surrogate-kernels", "copy-calibration", "scf",
                                       "surrogate-scf"],
                    help="run one secondary leg alone and print its figures (for kernel traces and "
                         "counters; --segment-length then fixes the batch leg's segment length)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the energy-sharded multi-GPU form is a design note, not a run)",
              file=sys.stderr)
        sys.exit(2)
    if args.only:
        import torch
        dev = torch.device("cuda:0")
        legs = {"batch": lambda: batch_leg(dev, segment_length=args.segment_length),
                "block": lambda: block_leg(with_cpu=False),
                "surrogate-kernels": lambda: surrogate_kernels_leg(dev),
                "copy-calibration": lambda: copy_calibration_leg(dev),
                "scf": lambda: scf_sweep(with_cpu=False),
                "surrogate-scf": lambda: surrogate_leg(dev, with_cpu=False)}
        print(json.dumps({"only": args.only, "result": legs[args.only]()}), flush=True)
        return

    import torch
    from electrostatic_amd import _lib
    from electrostatic_amd.workloads import batched_rgf_workload

    lib = _lib.load()
    dev = torch.device("cuda:0")
    wl = batched_rgf_workload()
    N, nE = wl["N"], wl["E"].size
    dU = torch.from_numpy(wl["U"]).to(dev)
    dE = torch.from_numpy(wl["E"].view(np.float64).copy()).to(dev)
    dc = torch.from_numpy(wl["coef"].view(np.float64).copy()).to(dev)
    acc = torch.zeros(N, dtype=torch.float64, device=dev)
    dG = (torch.zeros(2 * nE * N, dtype=torch.float64, device=dev)
          if args.write_diag_g else None)

    p = _lib.negf_chain_params()
    p.N = N
    p.t_even = wl["t_even"]
    p.t_odd = wl["t_odd"]
    p.Uc[0], p.Uc[1] = wl["Uc"]
    p.segment_length = args.segment_length
    plan = C.c_void_p()
    _lib.check(lib.negf_rgf_plan_create(p, C.c_void_p(dU.data_ptr()), nE,
                                        C.c_void_p(dE.data_ptr()),
                                        C.c_void_p(dc.data_ptr()),
                                        1 if dG is not None else 0, C.byref(plan)))
    # a created stream, not the legacy default stream: launches on the latter
    # carry its implicit synchronisation (measured: about 8 us per step)
    stream = torch.cuda.Stream(dev)
    sp = C.c_void_p(stream.cuda_stream)
    gp = C.c_void_p(dG.data_ptr()) if dG is not None else None

    def step():
        _lib.check(lib.negf_rgf_plan_execute(plan, C.c_void_p(acc.data_ptr()), gp, sp))

    for _ in range(args.warmup):
        step()
    _lib.check(lib.negf_rgf_plan_status(plan, sp))
    torch.cuda.synchronize()

    _lib.check(lib.negf_rgf_plan_enable_timing(plan, 1))
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_start.record(stream)
    for _ in range(args.steps):
        step()
    t_end.record(stream)
    torch.cuda.synchronize()
    ms_total = t_start.elapsed_time(t_end)
    ms_per_step = ms_total / args.steps
    kms = C.c_double()
    nl = C.c_int()
    _lib.check(lib.negf_rgf_plan_kernel_time(plan, C.byref(kms), C.byref(nl)))
    _lib.check(lib.negf_rgf_plan_status(plan, sp))
    kernel_ms = kms.value / max(1, nl.value)

    energy_sites = N * nE
    value = energy_sites / (ms_per_step * 1e-3)
    flops = FLOP_PER_ENERGY_SITE * energy_sites
    achieved = flops / (kernel_ms * 1e-3) / 1e12
    q = acc.cpu().numpy()
    if not np.all(np.isfinite(q)):
        print("non-finite charge", file=sys.stderr)
        sys.exit(1)

    traffic = None
    if os.path.exists(TRAFFIC_FILE) and not args.write_diag_g:
        with open(TRAFFIC_FILE) as f:
            tr = json.load(f)
        traffic = tr.get("hbm_bytes_per_launch")

    if args.write_diag_g:
        # the G^R-writing variant is bound by HBM: 16 B per energy-site written
        gbytes = 16.0 * energy_sites
        roof = {
            "bound": "hbm",
            "kernel": "RGF hot path with diag G^R written ([energy][site] complex FP64)",
            "bytes_per_energy_site": 16,
            "kernel_ms": kernel_ms,
            "achieved": gbytes / (kernel_ms * 1e-3) / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": gbytes / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "traffic": None,
            "fp64_tflops": achieved,
        }
    else:
        roof = {
            "bound": "fp64_valu",
            "compute": "FP64 VALU, compute-bound (MI355X FP64 peak 78.6 TFLOP/s)",
            "kernel": HOT_PATH_KERNELS,
            "flop_per_energy_site": FLOP_PER_ENERGY_SITE,
            "kernel_ms": kernel_ms,
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": traffic,
            "traffic_source": ("HBM bytes per launch from rocprofv3 FETCH_SIZE and WRITE_SIZE of this "
                               "launch, copy-calibrated: profiles/rgf_batched_8192E_4096sites_traffic.md"),
        }

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "energy·site/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": ("batched RGF: 8192 complex energies (4096 contour arc + 4096 "
                         "real axis, eta=1e-5 eV) x 4096-site chain, hoppings "
                         "2.5cos(pi/17)/2.5 eV, smooth potential seed 2, complex FP64, "
                         + ("diag G^R written" if args.write_diag_g else
                            "fused site charge")),
            "N": N,
            "energies": nE,
            "energy_sites": energy_sites,
            "segment_length": args.segment_length or 128,
        },
        "roofline": roof,
    }
    # parity of the timed result against the oracle on the same arrays
    cpu_rec, ref = cpu_baseline(wl)
    err = float(np.abs(q - ref).max() / np.abs(ref).max())
    out["max_rel_err_vs_oracle"] = err
    out["parity"] = {"max_rel_err_vs_oracle": err, "tolerance": 1e-10,
                     "oracle": "oracle restatement of the reference RGF, full workload"}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_rec
    if args.write_diag_g:
        # spot check of the materialised G^R rows against the oracle
        from oracle import oracle as O
        ks = np.array([0, nE // 3, nE // 2, nE - 1])
        _, dg_ref, _ = O.accumulate_diagG(wl["U"], wl["t_even"], wl["t_odd"], wl["Uc"],
                                          wl["t_even"], wl["t_odd"], wl["E"][ks],
                                          wl["coef"][ks], want_diag=True)
        g = dG.view(nE, N, 2)[torch.as_tensor(ks, device=dev)].cpu().numpy()
        g = g[..., 0] + 1j * g[..., 1]
        out["diag_G_max_rel_err_vs_oracle"] = float(np.abs(g - dg_ref).max() / np.abs(dg_ref).max())
    if err > 1e-10:
        print(json.dumps(out), flush=True)
        print(f"parity failure: max rel err {err:.3e} vs oracle", file=sys.stderr)
        sys.exit(1)
    lib.negf_rgf_plan_destroy(plan)
    del dG
    if not args.no_diag_g and not args.write_diag_g:
        out["diag_g"] = diag_g_leg(dev, wl)
    if not args.no_scf:
        out["scf"] = scf_sweep(with_cpu=not args.no_cpu_baseline)
    if not args.no_batch:
        out["batch"] = batch_leg(dev)
    if not args.no_block:
        out["block"] = block_leg(with_cpu=not args.no_cpu_baseline)
    if args.surrogate_kernels:
        out["surrogate_kernels"] = surrogate_kernels_leg(dev)
    if not args.no_surrogate:
        out["surrogate_scf"] = surrogate_leg(dev, with_cpu=not args.no_cpu_baseline)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
