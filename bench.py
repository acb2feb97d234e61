"""Throughput benchmark of the DMC propagation path on one MI355X.

Workload (the metric's headline case): plain DMC on the triatomic Morse+bend PES
(atoms [H, H, O], masses 1837.15 / 29156.95, De=0.1916, a=1.2, re=1.81, kb=0.16,
theta0=104.5 deg), 1M walkers started at the equilibrium geometry + N(0, 0.05 bohr),
dt = 5 a.u., discrete birth/death, FP32 coordinates and per-walker PES with FP64 weights
and reductions, Philox draws.  A step is one fused kernel launch (move, PES, weights,
branching, compaction, vref in the last tile); value = walker-steps/s = sum over the K timed steps of the
population entering each step, divided by the wall time between two device syncs.

Also reported: the same case with importance sampling (Gaussian-product psi_T, drift and
local energy from the trial kernels, Metropolis), the dominant kernel's roofline fraction
(algorithmic bytes / average launch duration from HIP events, against 8 TB/s) and the CPU
baseline (the oracle with the reference's process-pool PES manager on the host cores, on a
bounded sample of the same workload).
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
METRIC = "walker-steps/s at 1M walkers, plain vs IS+surrogate; ZPE err cm^-1 vs FP64"
# algorithmic HBM bytes per walker-step of the fused discrete step, FP32 triatomic:
# read 9 coordinates (36 B) + write 9 coordinates per surviving copy (36 B on average,
# since the mean copy count is 1); draws are generated in registers, energies and weights
# never leave registers.
BYTES_PER_WALKER_STEP = {('plain', 'fp32', 9): 72, ('plain', 'fp64', 9): 144, ('is', 'fp32', 9): 72}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_sim(n, steps, precision='fp32', imp=False, seed=1234):
    import pyvibdmc_amd as pv
    from pyvibdmc_amd.impsamp import water_trial
    t = np.deg2rad(104.5)
    eq = np.array([[1.81, 0.0, 0.0], [1.81 * np.cos(t), 1.81 * np.sin(t), 0.0], [0.0, 0.0, 0.0]])
    start = eq[None] + np.random.RandomState(seed).normal(0.0, 0.05, size=(n, 3, 3))
    sim = pv.DMC_Sim(num_walkers=n, num_timesteps=steps, atoms=['H', 'H', 'O'], masses=[1837.15, 1837.15, 29156.95],
                     delta_t=5, potential=pv.MorseBendTriatomic(), start_structures=start,
                     imp_samp=water_trial() if imp else None, precision=precision, rng='philox', seed=seed,
                     write_files=False, equil_steps=10 ** 9, wfn_every=10 ** 9, chkpt_every=10 ** 9)
    return sim


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


def cpu_baseline(n):
    from oracle.cpu_ref import cpu_model, time_plain_triatomic
    steps = 12 if n >= 500_000 else 40
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
    n = args.walkers
    K, W = args.steps, args.warmup
    total = W + K + args.event_steps + 2

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
                "kernel": f"k_discrete<{rname},9> (fused move+PES+branch+compaction)",
                "avg_launch_us": k_ms * 1e3, "bytes_per_launch": bytes_per_launch,
                "bytes_per_walker_step": bpw,
                "step_kernels_us": {"k_discrete": kern_ms[0] * 1e3}}
    log(f'[bench] plain: {value:.4g} walker-steps/s, {ms_per_step * 1e3:.2f} us/step, kernels {kern_ms} ms')
    del sim
    torch.cuda.empty_cache()

    copy_gbs = copy_bandwidth()
    if roof is not None:
        roof["copy_GBs"] = copy_gbs
        roof["frac_of_copy"] = roof["achieved"] / copy_gbs
    log(f'[bench] device copy: {copy_gbs:.0f} GB/s')

    variants = {}
    if not args.no_variants:
        log('[bench] importance-sampled variant')
        sim = make_sim(n, total, precision=args.precision, imp=True)
        ws2, wall2, mean2, kern2, _ = run_timed(sim, W, K, min(args.event_steps, 50))
        variants['is'] = {"value": ws2 / wall2, "unit": "walker-steps/s", "ms_per_step": wall2 * 1e3 / K,
                          "config": "same PES and walkers, Gaussian-product psi_T (2 stretches + bend), drift and "
                                    "local energy from the trial kernels, Metropolis accept, dt_eff"}
        if kern2:
            names = ['k_is_move', 'k_discrete']
            variants['is']['step_kernels_us'] = {nm: t * 1e3 for nm, t in zip(names, kern2)}
        log(f'[bench] IS: {ws2 / wall2:.4g} walker-steps/s')
        del sim
        torch.cuda.empty_cache()

    cpu = None
    if not args.no_cpu:
        log('[bench] CPU baseline sample')
        cpu = cpu_baseline(n)
        log(f'[bench] CPU: {cpu["value"]:.4g} walker-steps/s on {cpu["cores"]} cores')

    out = {"metric": METRIC, "value": value, "unit": "walker-steps/s", "n_gpus": 1, "steps": K, "warmup": W,
           "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32" if args.precision == 'fp32' else "f64", "data": "synthetic",
           "config": {"workload": f"plain DMC, triatomic Morse+bend PES, {n} walkers, dt=5, discrete "
                                  f"birth/death, {args.precision} coordinates with FP64 weights/reductions, "
                                  f"Philox draws",
                      "walkers": n, "mean_population": mean_n, "parallelism": "single GPU"},
           "roofline": roof, "cpu_baseline": cpu, "variants": variants}
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
