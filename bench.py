"""Throughput benchmark of the DMC propagation path on one MI355X.

Workload (the metric's headline case): plain DMC on the triatomic Morse+bend PES
(atoms [H, H, O], masses 1837.15 / 29156.95, De=0.1916, a=1.2, re=1.81, kb=0.16,
theta0=104.5 deg), 1M walkers started at the equilibrium geometry + N(0, 0.05 bohr),
dt = 5 a.u., discrete birth/death, FP32 coordinates and per-walker PES with FP64 weights
and reductions, Philox draws.  A step is one fused kernel launch (move, PES, weights,
branching, compaction, vref in the last tile); value = walker-steps/s = sum over the K timed steps of the
population entering each step, divided by the wall time between two device syncs.

Also reported: the same case with importance sampling (Gaussian-product psi_T, drift and
local energy from the trial kernels, Metropolis); the continuous-weighting workload (100k
walkers, 10 000 steps, exact split replay, descendant weighting, and the same at 1M walkers
for 2 000 steps) in FP32 and FP64; the
surrogate form (an MLP trained on the
GPU on 2M geometries an analytic importance-sampled run visits, then driving the same
importance-sampled DMC); ZPE runs of every form against FP64 runs; the dominant kernel's
roofline fraction (algorithmic bytes / average launch duration from HIP events, against
8 TB/s); and the CPU baseline (the oracle with the reference's process-pool PES manager on
the host cores, on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0  # dense
# algorithmic FLOP of one network evaluation: 2 (D H + H H + H H + H), D = 3, H = 128
MLP_FLOP_PER_WALKER = 2 * (3 * 128 + 128 * 128 + 128 * 128 + 128)
METRIC = "walker-steps/s at 1M walkers, plain vs IS+surrogate; ZPE err cm^-1 vs FP64"
# algorithmic HBM bytes per walker-step of the fused discrete step, FP32 triatomic:
# read 9 coordinates (36 B) + write 9 coordinates per surviving copy (36 B on average,
# since the mean copy count is 1); draws are generated in registers, energies and weights
# never leave registers.
BYTES_PER_WALKER_STEP = {('plain', 'fp32', 9): 72, ('plain', 'fp64', 9): 144, ('is', 'fp32', 9): 72}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_sim(n, steps, precision='fp32', imp=False, seed=1234, pot=None, train_every=None, dt=5):
    import pyvibdmc_amd as pv
    from pyvibdmc_amd.impsamp import water_trial
    t = np.deg2rad(104.5)
    eq = np.array([[1.81, 0.0, 0.0], [1.81 * np.cos(t), 1.81 * np.sin(t), 0.0], [0.0, 0.0, 0.0]])
    start = eq[None] + np.random.RandomState(seed).normal(0.0, 0.05, size=(n, 3, 3))
    sim = pv.DMC_Sim(num_walkers=n, num_timesteps=steps, atoms=['H', 'H', 'O'], masses=[1837.15, 1837.15, 29156.95],
                     delta_t=dt, potential=pot if pot is not None else pv.MorseBendTriatomic(), start_structures=start,
                     imp_samp=water_trial() if imp else None, precision=precision, rng='philox', seed=seed,
                     write_files=False, equil_steps=10 ** 9, wfn_every=10 ** 9, chkpt_every=10 ** 9,
                     DEBUG_save_training_every=train_every)
    return sim


WN = 4.556335281212229e-6  # hartree per cm^-1


def blocking_se(x):
    """Standard error of the mean of a correlated series: the largest block-average
    estimate over block sizes 1, 2, 4, ... with at least 16 blocks (Flyvbjerg-Petersen)."""
    x = np.asarray(x, np.float64)
    best, b = 0.0, 1
    while len(x) // b >= 16:
        m = len(x) // b
        blocks = x[:m * b].reshape(m, b).mean(axis=1)
        best = max(best, blocks.std(ddof=1) / np.sqrt(m))
        b *= 2
    return best


def zpe_of(vref, t=None):
    """The reference's ZPE estimate after t steps: mean of vref[t//4:t] (pyvibdmc.py:904-908),
    in cm^-1, with its blocking standard error."""
    v = np.asarray(vref[:t] if t else vref, np.float64) / WN
    tail = v[len(v) // 4:]
    return float(tail.mean()), float(blocking_se(tail))


def time_to_accuracy(vref, target, ms_per_step, tol=1.0, sigma=0.3, every=250):
    """First checkpoint t (a multiple of `every`) from which the running ZPE estimate stays
    within `tol` cm^-1 of `target` up to the end of the run and whose blocking error is at
    most `sigma`, with the wall time it costs at the measured step rate; None if the run never
    settles there.  Staying, not touching, is required: an estimate that crosses the band
    because of its time-step bias or a fluctuation has not converged."""
    v = np.asarray(vref, np.float64) / WN
    ts = np.arange(every, len(v) + 1, every)
    c = np.concatenate([[0.0], np.cumsum(v)])
    lo = ts // 4
    means = (c[ts] - c[lo]) / (ts - lo)
    inside = np.abs(means - target) <= tol
    if not inside[-1]:
        return None
    k = len(ts) - 1
    while k > 0 and inside[k - 1]:
        k -= 1
    for j in range(k, len(ts)):
        z, se = zpe_of(vref, int(ts[j]))
        if se <= sigma:
            t = int(ts[j])
            return {"steps": t, "seconds": t * ms_per_step * 1e-3, "zpe_cm": z, "se_cm": se}
    return None


def extrapolate_dt0(dts, zs, ses):
    """ZPE(dt) = a + b dt + c dt^2 through three time steps: a and its standard error (the
    Lagrange weights at dt = 0 applied to the independent runs' errors), plus a + b dt fitted
    by weighted least squares to the two smallest steps as a check of the curvature."""
    dts = np.asarray(dts, np.float64)
    zs = np.asarray(zs, np.float64)
    ses = np.asarray(ses, np.float64)
    w = np.array([np.prod([(0.0 - dts[j]) / (dts[i] - dts[j]) for j in range(3) if j != i]) for i in range(3)])
    a = float(w @ zs)
    se = float(np.sqrt(np.sum((w * ses) ** 2)))
    o = np.argsort(dts)[:2]
    lin = float((zs[o[0]] * dts[o[1]] - zs[o[1]] * dts[o[0]]) / (dts[o[1]] - dts[o[0]]))
    lin_se = float(np.hypot(ses[o[0]] * dts[o[1]], ses[o[1]] * dts[o[0]]) / (dts[o[1]] - dts[o[0]]))
    return {"zpe0_cm": a, "se_cm": se, "lagrange_weights": w.tolist(), "linear_two_smallest_cm": lin,
            "linear_se_cm": lin_se}


def timed_propagate(sim):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.propagate()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, float(np.sum(sim._pop_vs_tau))


def train_surrogate(n, epochs, geoms, seed=11):
    """The surrogate workflow: geometries visited by an analytic importance-sampled run
    (training-data hook, every 300 steps), labelled by the analytic surface, then Adam on
    the GPU.  Returns the model and a summary."""
    import pyvibdmc_amd as pv
    snaps = max(1, int(np.ceil(geoms / n)))
    sim = make_sim(n, 300 * snaps + 1, imp=True, seed=seed, train_every=300)
    t0 = time.perf_counter()
    sim.propagate()
    data = [(c, v) for st, c, v in sim.training if st > 0][:snaps]
    coords = np.concatenate([c for c, _ in data])[:geoms]
    pots = np.concatenate([v for _, v in data])[:geoms]
    t_data = time.perf_counter() - t0
    del sim
    m = pv.SurrogateModel(['H', 'H', 'O'], hidden=128)
    losses = m.fit(coords, pots, epochs=epochs, batch=8192, lr=1e-3)
    te = coords[::197]
    err = (m.energy(te) - pv.MorseBendTriatomic().getpot(te)) / WN
    info = {"geometries": int(len(coords)), "epochs": epochs, "batch": 8192, "lr": 1e-3,
            "data_seconds": t_data, "train_seconds": m.train_seconds, "final_loss": losses[-1],
            "rms_err_cm": float(np.sqrt(np.mean(err ** 2))), "max_err_cm": float(np.abs(err).max()),
            "checked_on": int(len(te))}
    return m, info


def run_timed(sim, warmup, steps, events_steps=0):
    """Warm up, then time exactly `steps` fused steps between two device syncs.  Optionally
    time `events_steps` more launches individually with HIP events on the launch stream."""
    import torch
    from pyvibdmc_amd import _native as N
    lib = N.load()
    st = N.stream_handle()
    sim._start()
    N.check(lib.pvd_sim_run(sim._h, warmup, st))
    sim._raise_if_failed()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    N.check(lib.pvd_sim_run(sim._h, steps, st))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    sim._raise_if_failed()
    sim._fetch_traj(0, warmup + steps)
    pops = sim._pop_vs_tau
    # population entering step t = population after step t-1
    entering = pops[warmup - 1:warmup + steps - 1]
    ws = float(np.sum(entering))
    kern_ms = None
    if events_steps:
        # HIP event pairs around every kernel of each step, recorded by the library on the
        # launch stream; kern_ms[k] = mean duration of the step's k-th kernel
        import ctypes
        avg = (ctypes.c_float * 8)()
        nk = ctypes.c_int32(0)
        N.check(lib.pvd_sim_run_timed(sim._h, events_steps, st, avg, ctypes.byref(nk)))
        sim._raise_if_failed()
        kern_ms = [float(avg[k]) for k in range(nk.value)]
        sim._fetch_traj(0, warmup + steps + events_steps)
        ev_n = float(np.mean(sim._pop_vs_tau[warmup + steps - 1:warmup + steps + events_steps - 1]))
    else:
        ev_n = None
    return ws, wall, float(np.mean(entering)), kern_ms, ev_n


def continuous_variant(n=100_000, T=10_000):
    """The continuous-weighting workload: the triatomic PES, n walkers started at the
    equilibrium geometry + N(0, 0.05 bohr) (seed 0), dt = 5, T steps, lower weight threshold
    0.01 (walkers below it are replaced by splitting the heaviest, pyvibdmc.py:340-356),
    descendant weighting every 500 steps over 50 steps, in FP32 and FP64.  Reports
    walker-steps/s over the whole propagate (host events of the wave-function saves
    included), each precision's ZPE, which split replay each step took, and the step's three
    kernels from HIP events over 200 further steps, with the weight kernel's HBM roofline."""
    import ctypes
    import torch
    import pyvibdmc_amd as pv
    from pyvibdmc_amd import _native as N
    t = np.deg2rad(104.5)
    eq = np.array([[1.81, 0.0, 0.0], [1.81 * np.cos(t), 1.81 * np.sin(t), 0.0], [0.0, 0.0, 0.0]])
    start = eq[None] + np.random.RandomState(0).normal(0.0, 0.05, size=(n, 3, 3))
    out = {"config": f"triatomic Morse+bend PES, {n} walkers, dt=5, {T} steps, continuous weights "
                     f"(lower threshold 0.01, exact replay of the sequential split loop), descendant "
                     f"weighting every 500 steps over 50 steps"}
    for prec in ('fp32', 'fp64'):
        sim = pv.DMC_Sim(num_walkers=n, num_timesteps=T, atoms=['H', 'H', 'O'], masses=[1837.15, 1837.15, 29156.95],
                         delta_t=5, potential=pv.MorseBendTriatomic(), start_structures=start,
                         weighting='continuous', cont_wt_thresh=0.01, equil_steps=500, wfn_every=500,
                         desc_wt_steps=50, chkpt_every=10 ** 9, precision=prec, seed=4, write_files=False)
        wall, _ = timed_propagate(sim)
        nb, how = sim.trajectory_counts(0, T)
        z, se = zpe_of(sim._vref_vs_tau)
        lib = N.load()
        avg = (ctypes.c_float * 8)()
        nk = ctypes.c_int32(0)
        N.check(lib.pvd_sim_run_timed(sim._h, 200, N.stream_handle(), avg, ctypes.byref(nk)))
        sim._raise_if_failed()
        kus = {nm: float(avg[k]) * 1e3 for k, nm in enumerate(['k_cont', 'k_cont_replay', 'k_cont_copy'][:nk.value])}
        bpw = 88 if prec == 'fp32' else 160  # coordinates read + written, weight read + written
        ach = bpw * n / (kus['k_cont'] * 1e-6) / 1e9
        out[prec] = {"value": n * T / wall, "unit": "walker-steps/s", "ms_per_step": wall * 1e3 / T,
                     "zpe_cm": z, "zpe_se_cm": se, "splits_per_step": float(np.mean(nb)),
                     "replay_paths": {str(int(k)): int(v) for k, v in zip(*np.unique(how, return_counts=True))},
                     "wfns": len(getattr(sim, 'wfns', [])), "step_kernels_us": kus,
                     "k_cont_roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                         "frac": ach / HBM_PEAK_GBS, "bytes_per_walker_step": bpw}}
        log(f'[bench] continuous {prec}: {n * T / wall:.4g} walker-steps/s, {wall * 1e6 / T:.1f} us/step, '
            f'ZPE {z:.2f} +- {se:.2f}, kernels {kus}')
        del sim
        torch.cuda.empty_cache()
    return out


DTS = (5.0, 2.5, 1.25)
TAU = 100_000.0  # imaginary time of every extrapolation run (a.u.)


def zpe_study(n, T, model):
    """The FP64 oracle and the FP32 forms against it, at n walkers.

    * FP64 plain and FP64 importance-sampled DMC at dt = 5, 2.5 and 1.25 over the same
      imaginary time (TAU), each extrapolated to dt -> 0 (``extrapolate_dt0``).  The two
      propagators carry different time-step errors but share the dt -> 0 limit, so their
      agreement checks the importance-sampled propagator, and the combined limit is the
      oracle ZPE.
    * FP32 plain, importance-sampled and surrogate runs at dt = 5 (T steps), each against the
      FP64 run of the same form and time step."""
    import torch
    import pyvibdmc_amd as pv
    from pyvibdmc_amd.surrogate import SurrogatePotential
    out = {"walkers": n, "estimator": "mean of vref over steps T/4..T (pyvibdmc.py:904-908), blocking error",
           "imaginary_time_au": TAU}
    fp64 = {}
    for form, imp in (('plain', False), ('is', True)):
        rows = []
        for dt in DTS:
            steps = int(round(TAU / dt))
            log(f'[bench] FP64 {form} dt={dt}: {n} walkers x {steps} steps')
            sim = make_sim(n, steps, precision='fp64', imp=imp, seed=2024, dt=dt)
            wall, ws = timed_propagate(sim)
            z, se = zpe_of(sim._vref_vs_tau)
            rows.append({"dt": dt, "steps": steps, "zpe_cm": z, "se_cm": se, "wall_s": wall})
            log(f'[bench]   ZPE {z:.2f} +- {se:.2f} cm-1 in {wall:.1f} s')
            del sim
            torch.cuda.empty_cache()
        ext = extrapolate_dt0([r['dt'] for r in rows], [r['zpe_cm'] for r in rows], [r['se_cm'] for r in rows])
        fp64[form] = {"runs": rows, "dt0": ext}
        log(f'[bench] FP64 {form}: dt->0 {ext["zpe0_cm"]:.2f} +- {ext["se_cm"]:.2f} cm-1')
    a, b = fp64['plain']['dt0'], fp64['is']['dt0']
    diff = a['zpe0_cm'] - b['zpe0_cm']
    comb = np.hypot(a['se_cm'], b['se_cm'])
    wa, wb = 1 / a['se_cm'] ** 2, 1 / b['se_cm'] ** 2
    oracle = (wa * a['zpe0_cm'] + wb * b['zpe0_cm']) / (wa + wb)
    out['fp64'] = fp64
    out['dt0_plain_minus_is_cm'] = diff
    out['dt0_plain_minus_is_sigma'] = diff / comb
    out['oracle_zpe_cm'] = float(oracle)
    out['oracle_se_cm'] = float(1 / np.sqrt(wa + wb))
    # FP32 forms at dt = 5 against the FP64 run of the same form and step
    runs = [('plain_fp32', False, None), ('is_fp32', True, None)]
    if model is not None:
        runs.append(('is_surrogate_fp32', True, 'surrogate'))
    for name, imp, kind in runs:
        pot = SurrogatePotential(model, pv.MorseBendTriatomic(), judge_every=100) if kind else None
        log(f'[bench] {name} dt=5: {n} walkers x {T} steps')
        sim = make_sim(n, T, precision='fp32', imp=imp, seed=2024, pot=pot)
        wall, ws = timed_propagate(sim)
        z, se = zpe_of(sim._vref_vs_tau)
        same = fp64['is' if imp else 'plain']['runs'][0]
        out[name] = {"dt": 5.0, "steps": T, "zpe_cm": z, "se_cm": se, "wall_s": wall, "walker_steps_per_s": ws / wall,
                     "err_vs_fp64_same_form_cm": z - same['zpe_cm'] if same['steps'] == T else None,
                     "err_vs_oracle_cm": z - oracle}
        log(f'[bench]   {name}: ZPE {z:.2f} +- {se:.2f} cm-1 in {wall:.1f} s')
        del sim
        torch.cuda.empty_cache()
    return out


def time_to_accuracy_study(model, oracle, n=100_000, tol=1.0, sigma=0.3):
    """Wall time for each FP32 form to settle within tol cm^-1 of the oracle with a blocking
    error <= sigma (``time_to_accuracy``), in the variance-limited regime (n walkers): every
    form runs at each time step of DTS (the importance-sampled forms also at 0.625, since their
    time-step bias at 1.25 is about 1 cm^-1) for a fixed imaginary time long enough for plain
    DMC to reach sigma; a form's time is that of its fastest time step.  The ratio plain /
    surrogate is the measured speed-up."""
    import torch
    import pyvibdmc_amd as pv
    from pyvibdmc_amd.surrogate import SurrogatePotential
    forms = [('plain_fp32', False, None, 500_000.0, DTS), ('is_fp32', True, None, 100_000.0, DTS + (0.625,))]
    if model is not None:
        forms.append(('is_surrogate_fp32', True, 'surrogate', 100_000.0, DTS + (0.625,)))
    out = {"walkers": n, "tol_cm": tol, "sigma_cm": sigma, "oracle_zpe_cm": oracle, "checkpoint_every": 250}
    for name, imp, kind, tau, dts in forms:
        rows = []
        for dt in dts:
            steps = int(round(tau / dt))
            pot = SurrogatePotential(model, pv.MorseBendTriatomic(), judge_every=100) if kind else None
            sim = make_sim(n, steps, precision='fp32', imp=imp, seed=77, pot=pot, dt=dt)
            wall, _ = timed_propagate(sim)
            ms = wall * 1e3 / steps
            z, se = zpe_of(sim._vref_vs_tau)
            tta = time_to_accuracy(sim._vref_vs_tau, oracle, ms, tol, sigma)
            rows.append({"dt": dt, "steps_run": steps, "ms_per_step": ms, "zpe_cm": z, "se_cm": se,
                         "err_vs_oracle_cm": z - oracle, "reached": tta})
            log(f'[bench] TTA {name} dt={dt}: {ms * 1e3:.1f} us/step, full-run ZPE {z:.2f} +- {se:.2f}, '
                f'reached {tta}')
            del sim
            torch.cuda.empty_cache()
        ok = [r for r in rows if r['reached']]
        best = min(ok, key=lambda r: r['reached']['seconds']) if ok else None
        out[name] = {"runs": rows, "best": None if best is None else
                     {"dt": best['dt'], "steps": best['reached']['steps'], "seconds": best['reached']['seconds']}}
    p, q = out['plain_fp32']['best'], (out.get('is_surrogate_fp32') or {}).get('best')
    out['speedup_is_surrogate_over_plain'] = (p['seconds'] / q['seconds']) if (p and q) else None
    r = (out.get('is_fp32') or {}).get('best')
    out['speedup_is_over_plain'] = (p['seconds'] / r['seconds']) if (p and r) else None
    out['target_speedup'] = 3.0
    return out


def ps_variant(n, steps, epochs, zpe_steps=2000):
    """The Partridge-Schwenke water surface (the reference's Fortran potential as a device
    kernel): importance-sampled DMC at n walkers with the analytic surface and with a
    3 -> 128^3 -> 1 surrogate trained on 2M geometries an analytic run visits (200 epochs),
    the analytic surface as its fallback and judge.  Throughput of both forms and their ZPEs
    over zpe_steps steps (same seed and start)."""
    import torch
    import pyvibdmc_amd as pv
    from pyvibdmc_amd.impsamp import water_trial
    from pyvibdmc_amd.surrogate import SurrogatePotential

    def sim(nw, T, pot, seed, train_every=None):
        start = pv.PartridgeSchwenkeWater.equilibrium()[None] + np.random.RandomState(seed).normal(0, 0.05, (nw, 3, 3))
        return pv.DMC_Sim(num_walkers=nw, num_timesteps=T, atoms=['H', 'H', 'O'], delta_t=5, potential=pot,
                          start_structures=start, imp_samp=water_trial(), precision='fp32', seed=seed,
                          write_files=False, equil_steps=10 ** 9, wfn_every=10 ** 9, chkpt_every=10 ** 9,
                          DEBUG_save_training_every=train_every)
    s = sim(1_000_000, 601, pv.PartridgeSchwenkeWater(), 11, train_every=300)
    s.propagate()
    data = [(c, v) for st, c, v in s.training if st > 0]
    coords = np.concatenate([c for c, _ in data])
    pots = np.concatenate([v for _, v in data])
    del s, data
    m = pv.SurrogateModel(['H', 'H', 'O'], hidden=128)
    m.fit(coords, pots, epochs=epochs, batch=8192, lr=1e-3)
    te = coords[::197]
    err = (m.energy(te) - pv.PartridgeSchwenkeWater().getpot(te)) / WN
    out = {"config": f"Partridge-Schwenke water, {n} walkers, IS (Gaussian-product psi_T), dt=5, FP32",
           "training": {"geometries": int(len(coords)), "epochs": epochs, "train_seconds": m.train_seconds,
                        "rms_err_cm": float(np.sqrt(np.mean(err ** 2))), "max_err_cm": float(np.abs(err).max())}}
    del coords, pots
    for name, mk in (('analytic', lambda: pv.PartridgeSchwenkeWater()),
                     ('surrogate', lambda: SurrogatePotential(m, pv.PartridgeSchwenkeWater(), judge_every=100))):
        sm = sim(n, 10 + steps + 2, mk(), 7)
        ws, wall, _, _, _ = run_timed(sm, 10, steps, 0)
        del sm
        pot = mk()
        sz = sim(n, zpe_steps, pot, 7)
        sz.propagate()
        z, se = zpe_of(sz._vref_vs_tau)
        out[name] = {"value": ws / wall, "unit": "walker-steps/s", "ms_per_step": wall * 1e3 / steps, "zpe_cm": z,
                     "zpe_se_cm": se}
        if name == 'surrogate':
            out[name]["judge_rms_cm"] = [r[1] / WN for r in pot.judge_log][-3:]
        log(f'[bench] PS {name}: {ws / wall:.4g} walker-steps/s, ZPE {z:.2f} +- {se:.2f}')
        del sz
        torch.cuda.empty_cache()
    # plain DMC (no trial function) on the analytic surface, for the time-to-accuracy comparison
    start = pv.PartridgeSchwenkeWater.equilibrium()[None] + np.random.RandomState(7).normal(0, 0.05, (n, 3, 3))
    sp = pv.DMC_Sim(num_walkers=n, num_timesteps=10 + steps + 2, atoms=['H', 'H', 'O'], delta_t=5,
                    potential=pv.PartridgeSchwenkeWater(), start_structures=start, precision='fp32', seed=7,
                    write_files=False, equil_steps=10 ** 9, wfn_every=10 ** 9, chkpt_every=10 ** 9)
    ws, wall, _, _, _ = run_timed(sp, 10, steps, 0)
    out['plain_analytic'] = {"value": ws / wall, "unit": "walker-steps/s", "ms_per_step": wall * 1e3 / steps}
    log(f'[bench] PS plain analytic: {ws / wall:.4g} walker-steps/s')
    del sp
    out["zpe_surrogate_minus_analytic_cm"] = out['surrogate']['zpe_cm'] - out['analytic']['zpe_cm']
    return out


CLUSTER_EL = ['O', 'H', 'H', 'O', 'H', 'H']
# algorithmic FLOP of the 15 -> 256 -> 256 -> 256 -> 1 network per walker
MLP256_FLOP_PER_WALKER = 2 * (15 * 256 + 256 * 256 + 256 * 256 + 256)


def cluster_sim(n, steps, precision, pot, seed, train_every=None):
    import pyvibdmc_amd as pv
    from oracle import pes as P
    start = P.cluster6_eq()[None] + np.random.RandomState(seed).normal(0, 0.05, (n, 6, 3))
    return pv.DMC_Sim(num_walkers=n, num_timesteps=steps, atoms=CLUSTER_EL,
                      masses=[P.M_O, P.M_H, P.M_H, P.M_O, P.M_H, P.M_H], delta_t=5, potential=pot,
                      start_structures=start, imp_samp=pv.cluster6_trial(), precision=precision, seed=seed,
                      write_files=False, equil_steps=10 ** 9, wfn_every=10 ** 9, chkpt_every=10 ** 9,
                      DEBUG_save_training_every=train_every)


def cluster_study(n, steps, epochs, zpe_steps):
    """The six-atom cluster workload: the 15 -> 256^3 -> 1 surrogate (DistIt monomer sorting)
    trained on 2M geometries an analytic importance-sampled run visits (1M walkers, snapshots
    at steps 300 and 600; the snapshot at step 1200 is held out), then n-walker importance-
    sampled DMC with the analytic surface and with the surrogate: throughput over `steps`
    timed steps, the network kernel's MFMA roofline, and (zpe_steps > 0) both ZPEs."""
    import torch
    import pyvibdmc_amd as pv
    from pyvibdmc_amd.surrogate import SurrogatePotential
    t0 = time.perf_counter()
    s = cluster_sim(1_000_000, 1201, 'fp32', pv.TwoMonomerCluster(), 11, train_every=300)
    s.propagate()
    data = [(c, v) for st, c, v in s.training if st > 0]
    coords = np.concatenate([c for c, _ in data[:2]])
    pots = np.concatenate([v for _, v in data[:2]])
    held_c, held_v = data[-1][0][::4], data[-1][1][::4]
    del s, data
    t_data = time.perf_counter() - t0
    m = pv.SurrogateModel(CLUSTER_EL, hidden=256, seed=3, sorted_atoms=[[0], [1, 2], [3], [4, 5]],
                          sorted_groups=[[0, 1, 2], [3, 4, 5]])
    losses = m.fit(coords, pots, epochs=epochs, batch=8192, lr=1e-3)
    err = (m.energy(held_c) - held_v) / WN
    out = {"config": f"six-atom cluster, {n} walkers, importance sampling (per-monomer Gaussian psi_T), dt=5, FP32",
           "training": {"geometries": int(len(coords)), "epochs": epochs, "batch": 8192, "lr": 1e-3,
                        "data_seconds": t_data, "train_seconds": m.train_seconds, "final_loss": losses[-1],
                        "held_out_rms_cm": float(np.sqrt(np.mean(err ** 2))),
                        "held_out_max_cm": float(np.abs(err).max()), "held_out_mean_cm": float(err.mean()),
                        "held_out": int(len(held_c))}}
    log(f'[bench] cluster surrogate: {epochs} epochs in {m.train_seconds:.1f} s, held-out rms '
        f'{out["training"]["held_out_rms_cm"]:.2f} cm-1')
    del coords, pots
    for name, mk in (('is_analytic', lambda: pv.TwoMonomerCluster()),
                     ('is_surrogate', lambda: SurrogatePotential(m, pv.TwoMonomerCluster(), judge_every=100,
                                                                 judge_stride=1000))):
        pot = mk()
        sim = cluster_sim(n, 10 + steps + 2, 'fp32', pot, 7)
        ws, wall, mean_n, _, _ = run_timed(sim, 10, steps, 0)
        row = {"value": ws / wall, "unit": "walker-steps/s", "ms_per_step": wall * 1e3 / steps}
        if name == 'is_surrogate':
            from pyvibdmc_amd import _native as N
            import ctypes
            planes, stride, ndev = ctypes.c_void_p(), ctypes.c_int64(), ctypes.c_void_p()
            N.check(N.load().pvd_sim_device_view(sim._h, ctypes.byref(planes), ctypes.byref(stride),
                                                 ctypes.byref(ndev)))
            vbuf = torch.empty(stride.value, dtype=torch.float64, device='cuda')
            m.energy_planes(planes.value, stride.value, ndev.value, stride.value, vbuf, N.PVD_FP32, pot.fallback)
            ea, eb = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ea.record()
            for _ in range(5):
                m.energy_planes(planes.value, stride.value, ndev.value, stride.value, vbuf, N.PVD_FP32, pot.fallback)
            eb.record()
            torch.cuda.synchronize()
            e_ms = ea.elapsed_time(eb) / 5
            tf = MLP256_FLOP_PER_WALKER * mean_n / (e_ms * 1e-3) / 1e12
            row["mlp_roofline"] = {"bound": "mfma", "achieved": tf, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                                   "frac": tf / MFMA_BF16_PEAK_TFLOPS, "avg_launch_ms": e_ms,
                                   "kernel": "k_mlp_energy_st<float,18,256> + k_mlp_fallback",
                                   "flop_per_walker": MLP256_FLOP_PER_WALKER}
            del vbuf
        out[name] = row
        log(f'[bench] cluster {name}: {row["value"]:.4g} walker-steps/s, {row["ms_per_step"]:.2f} ms/step')
        del sim
        torch.cuda.empty_cache()
        if zpe_steps > 0:
            sim = cluster_sim(n, zpe_steps, 'fp32', mk(), 7)
            sim.propagate()
            out[name]["zpe_cm"], out[name]["zpe_se_cm"] = zpe_of(sim._vref_vs_tau)
            log(f'[bench] cluster {name}: ZPE {out[name]["zpe_cm"]:.2f} +- {out[name]["zpe_se_cm"]:.2f}')
            del sim
            torch.cuda.empty_cache()
    return out


def cpu_baseline(n):
    from oracle.cpu_ref import cpu_model, time_plain_triatomic
    steps = 60 if n >= 500_000 else 200
    rate, sec, ws, cores = time_plain_triatomic(num_walkers=n, steps=steps, warmup=1)
    return {"value": rate, "unit": "walker-steps/s", "cores": cores, "kind": "port",
            "sample": f"plain discrete triatomic DMC, FP64 NumPy oracle with a {cores}-process pool PES "
                      f"(reference structure), {n} walkers x {steps} timed steps ({sec:.1f} s), "
                      f"CPU: {cpu_model()}"}


def copy_bandwidth(nbytes=1 << 30, reps=10):
    """Achievable HBM bandwidth on this box: a plain device-to-device copy of `nbytes`
    (read + write = 2 * nbytes per copy), timed with HIP events."""
    import torch
    src = torch.empty(nbytes // 4, dtype=torch.float32, device='cuda').fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        dst.copy_(src)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    del src, dst
    return 2 * nbytes / (ms * 1e-3) / 1e9


def plane_copy_bandwidth(n, np_=9, reps=50):
    """Read-then-write ceiling of the walker layout: the library's streaming copy of n walkers
    of np_ FP32 planes (16-byte loads and stores), bytes moved over the mean launch time."""
    import ctypes
    import torch
    from pyvibdmc_amd import _native as N
    cap = (int(1.5 * n) + 1 + 255) // 256 * 256
    src = torch.ones(np_ * cap, dtype=torch.float32, device='cuda')
    dst = torch.empty_like(src)
    ms = ctypes.c_float()
    N.check(N.load().pvd_bench_plane_copy(src.data_ptr(), dst.data_ptr(), np_, cap, n, reps, N.stream_handle(),
                                          ctypes.byref(ms)), 'pvd_bench_plane_copy')
    del src, dst
    return 2 * 4 * np_ * n / (ms.value * 1e-3) / 1e9


def load_traffic(key):
    path = os.path.join(ROOT, 'profiles', 'traffic.json')
    try:
        with open(path) as fh:
            return json.load(fh).get(key)
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=2000)
    ap.add_argument('--warmup', type=int, default=200)
    ap.add_argument('--walkers', type=int, default=1_000_000)
    ap.add_argument('--precision', default='fp32', choices=['fp32', 'fp64'])
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-variants', action='store_true')
    ap.add_argument('--event-steps', type=int, default=200)
    ap.add_argument('--copy-only', action='store_true', help='only the 1 GiB copy (counter calibration)')
    ap.add_argument('--continuous-only', action='store_true', help='only the continuous-weighting workload')
    ap.add_argument('--continuous-walkers', type=int, default=100_000)
    ap.add_argument('--continuous-steps', type=int, default=10_000)
    ap.add_argument('--no-surrogate', action='store_true')
    ap.add_argument('--epochs', type=int, default=200)
    ap.add_argument('--train-geoms', type=int, default=2_000_000)
    ap.add_argument('--zpe-steps', type=int, default=20000, help='0 skips the ZPE runs')
    ap.add_argument('--is-only', action='store_true', help='only the importance-sampled variant')
    ap.add_argument('--cluster-only', action='store_true',
                    help='only the six-atom cluster workload (16M walkers by default, --walkers to change)')
    ap.add_argument('--cluster-zpe-steps', type=int, default=2000)
    ap.add_argument('--ps-only', action='store_true', help='only the Partridge-Schwenke surface variant')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the DMC walkers of this workload live on a single MI355X "
              f"(--gpus {args.gpus} is not supported: this build runs on one GPU only).",
              file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), 'bench.py needs the GPU'
    if args.copy_only:
        print(json.dumps({"copy_GBs": copy_bandwidth()}), flush=True)
        return
    if args.continuous_only:
        print(json.dumps({"continuous": continuous_variant(args.continuous_walkers, args.continuous_steps)}),
              flush=True)
        return
    if args.ps_only:
        print(json.dumps({"ps_surface": ps_variant(args.walkers, args.steps, args.epochs)}), flush=True)
        return
    if args.cluster_only:
        n_c = args.walkers if args.walkers != 1_000_000 else 16_000_000
        print(json.dumps({"cluster": cluster_study(n_c, args.steps, args.epochs, args.cluster_zpe_steps)}), flush=True)
        return
    n = args.walkers
    K, W = args.steps, args.warmup
    total = W + K + args.event_steps + 2

    if args.is_only:
        sim = make_sim(n, total, precision=args.precision, imp=True)
        ws2, wall2, _, kern2, _ = run_timed(sim, W, K, min(args.event_steps, 50))
        print(json.dumps({"is": {"value": ws2 / wall2, "ms_per_step": wall2 * 1e3 / K,
                                 "step_kernels_us": {nm: t * 1e3 for nm, t in zip(['k_is_move', 'k_discrete'],
                                                                                  kern2 or [])}}}), flush=True)
        return
    log(f'[bench] plain {args.precision} {n} walkers: warmup {W}, timed {K}')
    sim = make_sim(n, total, precision=args.precision)
    ws, wall, mean_n, kern_ms, ev_n = run_timed(sim, W, K, args.event_steps)
    value = ws / wall
    ms_per_step = wall * 1e3 / K
    bpw = BYTES_PER_WALKER_STEP[('plain', args.precision, 9)]
    roof = None
    if kern_ms:
        # the plain discrete step is one kernel (k_discrete, finalised by its last tile)
        k_ms = kern_ms[0]
        bytes_per_launch = bpw * ev_n
        achieved = bytes_per_launch / (k_ms * 1e-3) / 1e9
        rname = 'float' if args.precision == 'fp32' else 'double'
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS,
                "traffic": load_traffic(f'k_discrete_{args.precision}_np9_{n}'),
                "traffic_source": "committed counter profile (profiles/k_discrete_water_1M_walkers_hbm_traffic.txt: "
                                  "rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, scaled by a copy kernel of known "
                                  "size); read from profiles/traffic.json, not measured in this run",
                "kernel": f"k_discrete<{rname},9> (fused move+PES+branch+compaction)",
                "avg_launch_us": k_ms * 1e3, "bytes_per_launch": bytes_per_launch,
                "bytes_per_walker_step": bpw,
                "step_kernels_us": {"k_discrete": kern_ms[0] * 1e3}}
    log(f'[bench] plain: {value:.4g} walker-steps/s, {ms_per_step * 1e3:.2f} us/step, kernels {kern_ms} ms')
    variants = {}
    del sim
    torch.cuda.empty_cache()

    copy_gbs = copy_bandwidth()
    pc_gbs = plane_copy_bandwidth(n)
    if roof is not None:
        roof["copy_GBs"] = copy_gbs
        roof["frac_of_copy"] = roof["achieved"] / copy_gbs
        roof["plane_copy_GBs"] = pc_gbs
        roof["frac_of_plane_copy"] = roof["achieved"] / pc_gbs
    log(f'[bench] device copy: {copy_gbs:.0f} GB/s; 9-plane streaming copy at {n} walkers: {pc_gbs:.0f} GB/s')

    if not args.no_variants:
        log('[bench] importance-sampled variant')
        sim = make_sim(n, total, precision=args.precision, imp=True)
        ws2, wall2, mean2, kern2, ev2 = run_timed(sim, W, K, min(args.event_steps, 50))
        variants['is'] = {"value": ws2 / wall2, "unit": "walker-steps/s", "ms_per_step": wall2 * 1e3 / K,
                          "config": "same PES and walkers, Gaussian-product psi_T (2 stretches + bend), drift and "
                                    "local energy from the trial kernels, Metropolis accept, dt_eff"}
        if kern2:
            names = ['k_is_move', 'k_discrete']
            variants['is']['step_kernels_us'] = {nm: t * 1e3 for nm, t in zip(names, kern2)}
            bpw_is = BYTES_PER_WALKER_STEP[('is', args.precision, 9)] if args.precision == 'fp32' else None
            if bpw_is:
                step_ms = sum(kern2)
                ach = bpw_is * ev2 / (step_ms * 1e-3) / 1e9
                variants['is']['roofline'] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                              "frac": ach / HBM_PEAK_GBS, "bytes_per_walker_step": bpw_is,
                                              "kernels": "k_is_move + k_discrete<float,9,IS> per step (algorithmic "
                                                         "bytes of the step over both kernels' time)"}
        log(f'[bench] IS: {ws2 / wall2:.4g} walker-steps/s')
        del sim
        torch.cuda.empty_cache()
        log('[bench] continuous-weighting variant')
        variants['continuous'] = continuous_variant()
        torch.cuda.empty_cache()
        log('[bench] continuous-weighting variant at 1M walkers (16 000 kills per step)')
        variants['continuous_1m'] = continuous_variant(1_000_000, 2000)

    if not args.no_variants and not args.no_surrogate:
        import pyvibdmc_amd as pv
        from pyvibdmc_amd.surrogate import SurrogatePotential
        log(f'[bench] surrogate: sampling {args.train_geoms} geometries, training {args.epochs} epochs')
        model, tinfo = train_surrogate(n, args.epochs, args.train_geoms)
        log(f'[bench] surrogate trained in {tinfo["train_seconds"]:.1f} s, rms {tinfo["rms_err_cm"]:.2f} cm-1')
        spot = SurrogatePotential(model, pv.MorseBendTriatomic(), judge_every=100, judge_stride=100)
        extra = max(0, 101 - (W + K))   # untimed steps after the timing, up to the first judge record
        sim = make_sim(n, W + K + extra + 2, imp=True, pot=spot)
        ws3, wall3, mean3, _, _ = run_timed(sim, W, K, 0)
        if extra:
            from pyvibdmc_amd import _native as N
            N.check(N.load().pvd_sim_run(sim._h, extra, N.stream_handle()))
            sim._raise_if_failed()
        # the network kernel alone, on the simulation's own walkers and stream
        from pyvibdmc_amd import _native as N
        import ctypes
        planes, stride, ndev = ctypes.c_void_p(), ctypes.c_int64(), ctypes.c_void_p()
        N.check(N.load().pvd_sim_device_view(sim._h, ctypes.byref(planes), ctypes.byref(stride), ctypes.byref(ndev)))
        vbuf = torch.empty(stride.value, dtype=torch.float64, device='cuda')
        for _ in range(3):
            model.energy_planes(planes.value, stride.value, ndev.value, stride.value, vbuf, N.PVD_FP32, spot.fallback)
        ea, eb = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ea.record()
        for _ in range(20):
            model.energy_planes(planes.value, stride.value, ndev.value, stride.value, vbuf, N.PVD_FP32, spot.fallback)
        eb.record()
        torch.cuda.synchronize()
        e_ms = ea.elapsed_time(eb) / 20
        flops = MLP_FLOP_PER_WALKER * mean3
        variants['is_surrogate'] = {
            "value": ws3 / wall3, "unit": "walker-steps/s", "ms_per_step": wall3 * 1e3 / K,
            "config": "importance sampling as above; V from the BF16-MFMA MLP (3 sorted inverse distances -> 128 -> "
                      "128 -> 128 -> 1, SiLU, FP32 quadratic baseline) with the analytic fallback outside the "
                      "training range and the analytic judge on every 100th walker every 100 steps",
            "training": tinfo,
            "mlp_roofline": {"bound": "mfma", "achieved": flops / (e_ms * 1e-3) / 1e12, "peak": MFMA_BF16_PEAK_TFLOPS,
                             "unit": "TFLOP/s", "frac": flops / (e_ms * 1e-3) / 1e12 / MFMA_BF16_PEAK_TFLOPS,
                             "kernel": "k_mlp_energy<float,9,128> + k_mlp_fallback", "avg_launch_us": e_ms * 1e3,
                             "flop_per_walker": MLP_FLOP_PER_WALKER},
            "judge_rms_cm": [r[1] / WN for r in spot.judge_log][-3:]}
        log(f'[bench] IS+surrogate: {ws3 / wall3:.4g} walker-steps/s, network {e_ms * 1e3:.1f} us/launch')
        del sim, vbuf
        torch.cuda.empty_cache()
    else:
        model = None

    if not args.no_variants and not args.no_surrogate:
        log('[bench] Partridge-Schwenke surface: analytic and surrogate')
        variants['ps_surface'] = ps_variant(n, K, args.epochs)

    zpe = tta = None
    if not args.no_variants and args.zpe_steps > 0:
        zpe = zpe_study(n, args.zpe_steps, model)
        tta = time_to_accuracy_study(model, zpe['oracle_zpe_cm'])

    cpu = None
    if not args.no_cpu:
        log('[bench] CPU baseline sample')
        cpu = cpu_baseline(n)
        log(f'[bench] CPU: {cpu["value"]:.4g} walker-steps/s on {cpu["cores"]} cores')

    out = {"metric": METRIC, "value": value, "unit": "walker-steps/s", "n_gpus": 1, "steps": K, "warmup": W,
           "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32" if args.precision == 'fp32' else "f64", "data": "synthetic",
           "config": {"workload": f"plain DMC, triatomic Morse+bend PES, {n} walkers, dt=5, discrete "
                                  f"birth/death, {args.precision} coordinates and branching weights "
                                  f"(exponent argument in FP64), FP64 reductions (vref, population), "
                                  f"Philox draws",
                      "walkers": n, "mean_population": mean_n, "parallelism": "single GPU"},
           "roofline": roof, "cpu_baseline": cpu, "variants": variants, "zpe": zpe, "time_to_accuracy": tta}
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
