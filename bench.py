"""Benchmark of the pyrochlore Heisenberg Monte Carlo step on one MI355X.

Workload (the parallel-tempering configuration): L=16 (65,536 spins per
replica) x 64 replicas on a geometric temperature ladder from 0.01 to 1.0,
FP32 unit-vector spins, H = sum_<ij> s_i.s_j (J = 1, unit scale), each sweep
1 Metropolis pass (cone proposal, adapted toward 50% acceptance before
timing) + 10 over-relaxation passes, neighbour-replica swaps every 10 sweeps,
FP64 energy reductions. A step is one sweep over all 64 replicas. Spins start
uniform on the sphere from the Philox stream (synthetic input).

Prints ONE JSON line: spin updates/s over the timed sweeps, the
tau_int-corrected independent samples/s, the roofline record of the dominant
kernel (the over-relaxation colour phase) timed with HIP events on the
library's stream, and the CPU oracle timed on a bounded sample of the same
workload on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

L, R = 16, 64
T_MIN, T_MAX = 0.01, 1.0
N_METRO, N_OR, PT_EVERY = 1, 10, 10
SEED = 56081376
BYTES_PER_UPDATE = 60  # 12 N B read + 3 N B written per colour phase of N/4 updates
# Fused over-relaxation (two colours per launch between two spin buffers):
# 12 N B read + 6 N B written for N/2 updates.
BYTES_PER_UPDATE_FUSED = 36
# A whole over-relaxation pass per launch (L=16): 12 N B read + 12 N B
# written for N updates.
BYTES_PER_UPDATE_PASS = 24
HBM_PEAK_GBS = 8000.0
FLOP_PER_FORWARD = 2 * (6 * 64 + 64 * 64 + 64 * 4)  # MLP 6 -> 64 -> 64 -> 4
BF16_PEAK_TFLOPS = 2500.0  # dense bf16 MFMA


OR_KERNEL = "k_phase_cells<1, colour, 16, false, false>"
OR_PAIR_KERNEL = "k_or_pair_plane<pair, 3>"
OR_PASS_KERNEL = "k_or_pass_plane<16>"
OR_PASS_KERNEL_C5 = "k_or_pass_strip<64>"
OR_PAIR_KERNEL_C5 = "k_or_pair<pair, 64, false>"
OR_KERNEL_C5 = "k_phase_cells<1, colour, 64, false, false>"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic.json")


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    record (FETCH_SIZE and WRITE_SIZE passes scaled by a copy kernel of known
    size), or None when no record for this kernel exists."""
    try:
        records = json.load(open(TRAFFIC_FILE))["records"]
    except (OSError, ValueError, KeyError, TypeError):
        return None
    for rec in records:
        if rec.get("kernel") == kernel:
            return rec.get("traffic_bytes_per_launch")
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu():
    """Logical CPUs of the host, the CPUs this process may run on, the thread
    budget (OMP_NUM_THREADS when the environment sets one: a shared host's
    share for one GPU's job) and the CPU model."""
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    budget = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or allowed
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": allowed, "threads": min(budget, allowed),
            "model": model}


def cpu_baseline(target_s: float = 15.0):
    """The oracle on the host cores: the same workload for as many whole sweeps
    as fit in about target_s seconds (at least one) on every core this process
    may use (capped by OMP_NUM_THREADS where the environment sets one), timed
    end to end; then one thread for about a fifth of that, for the per-core
    rate."""
    from cudacode_amd.pyro import geometric_ladder
    from oracle import oracle
    cpu = host_cpu()
    oracle.set_num_threads(cpu["threads"])
    used = oracle.num_threads()
    temps = geometric_ladder(T_MIN, T_MAX, R)
    run = oracle.OracleRun(L, temps, seed=SEED, n_metro=N_METRO, n_or=N_OR, pt_every=PT_EVERY,
                           cos_cone=0.5)
    t0 = time.perf_counter()
    run.run(1)
    one = time.perf_counter() - t0
    n = max(1, min(100, int(target_s / max(one, 1e-3)) - 1))
    log(f"cpu oracle: 1 sweep {one:.2f} s on {used} threads; timing {n} more")
    t0 = time.perf_counter()
    run.run(n)
    dt = time.perf_counter() - t0
    per_sweep = (N_METRO + N_OR) * 16 * L ** 3 * R
    updates = n * per_sweep
    # One thread: whole sweeps, continuing the same chain.
    oracle.set_num_threads(1)
    t0 = time.perf_counter()
    run.run(1)
    one1 = time.perf_counter() - t0
    n1 = max(0, min(20, int(0.2 * target_s / max(one1, 1e-3)) - 1))
    t0 = time.perf_counter()
    if n1:
        run.run(n1)
    dt1 = time.perf_counter() - t0
    single = per_sweep * (n1 + 1) / (dt1 + one1)
    oracle.set_num_threads(used)
    return {
        "value": updates / dt,
        "unit": "spin updates/s",
        "cores": used,
        "kind": "port",
        "sample": f"{n} sweeps of the same L=16 x 64-replica workload "
                  f"({updates:.3g} spin updates, {dt:.1f} s), C oracle with OpenMP on {used} "
                  f"threads; single thread: {n1 + 1} sweeps",
        "single_thread_value": single,
        "host": cpu,
    }


def calibrate_copy(mib: int = 512, reps: int = 5, sweep=()):
    """A plain device-to-device copy of known size (torch's elementwise copy
    kernel), run under rocprofv3 --pmc to scale FETCH_SIZE / WRITE_SIZE; also
    timed (torch events on torch's stream, which runs the copy) as a practical
    streaming-bandwidth reference beside the 8 TB/s peak. `sweep`: further
    sizes in MiB, timed the same way (sizes whose read + write fit the 256 MB
    Infinity Cache show its rate)."""
    import torch
    rates = []
    for m in sweep:
        k = m * (1 << 20) // 4
        x = torch.ones(k, device="cuda", dtype=torch.float32)
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        f0, f1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        f0.record()
        for _ in range(50):
            y.copy_(x)
        f1.record()
        torch.cuda.synchronize()
        t = f0.elapsed_time(f1) / 50
        rates.append({"MiB": m, "copy_ms": t, "copy_GBps_read_plus_write": 2 * k * 4 / (t * 1e-3) / 1e9})
        del x, y
    n = mib * (1 << 20) // 4
    a = torch.ones(n, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    timed = 20
    e0.record()
    for _ in range(timed):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / timed
    del a, b
    print(json.dumps({"calibration_copy_bytes_read": n * 4, "bytes_written": n * 4, "reps": reps,
                      "copy_ms": ms, "copy_GBps_read_plus_write": 2 * n * 4 / (ms * 1e-3) / 1e9,
                      "size_sweep": rates, "streams": stream_rates()}))


def stream_rates(gib: int = 1, reps: int = 20):
    """Practical HBM rates of torch's own streaming kernels, far beyond the
    Infinity Cache: a read-only sum of 2 GiB, and c = a + b over 1 GiB each
    (2 bytes read per byte written, about the fused over-relaxation kernel's
    12 N : 6 N). GB/s of bytes moved, torch events on torch's stream."""
    import torch
    n = gib * (1 << 30) // 4
    x = torch.ones(2 * n, device="cuda", dtype=torch.float32)
    a, b = x[:n], x[n:]
    c = torch.empty(n, device="cuda", dtype=torch.float32)

    def timed(fn):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    t_sum = timed(lambda: x.sum())
    t_add = timed(lambda: torch.add(a, b, out=c))
    del x, a, b, c
    return {"read_only_sum_ms": t_sum, "read_only_sum_GBps": 8 * n / (t_sum * 1e-3) / 1e9,
            "add_2read_1write_ms": t_add, "add_2read_1write_GBps": 12 * n / (t_add * 1e-3) / 1e9}


def learned_leg(g, args):
    """The learned local proposal on the same ensemble (cudacode_amd.learned):
    environments (beta h, s) gathered from the plain sampler across the
    temperature ladder (16 snapshots x 64 replicas x 1024 sites = 1,048,576),
    the network trained on the GPU, the chain thermalised by the plain
    sampler, then alternating legs of a fixed length (plain, learned, plain,
    learned; --series-sweeps each, independent of --steps) that give each
    sampler's sweeps/s and per-rung tau_int of the energy with its error,
    hence the tau_int-corrected samples/s ratio with a bootstrap interval;
    observables legs (C/N, <|m_A2|^2>) of each sampler; the learned phase's
    MFMA roofline; and which sampler the fallback rule keeps.
    Returns (record, pooled plain leg)."""
    from cudacode_amd import learned, pyro
    log("learned: gathering environments and training on the GPU")
    info = learned.train(g, steps=args.train_steps, init_seed=SEED)
    log(f"learned: {args.therm_sweeps} plain thermalisation sweeps")
    g.Mc(args.therm_sweeps)
    log(f"learned: 2 x (plain, learned) legs of {args.series_sweeps} timed sweeps")
    plain_legs, learned_legs = learned.paired_legs(g, args.series_sweeps, rounds=2)
    log(f"learned: observables, {args.obs_sweeps} measured sweeps per sampler")
    prev = learned.use_learned(g, 1)
    learned_obs = learned.observables_leg(g, args.obs_sweeps)
    learned.use_plain(g, prev)
    plain_obs = learned.observables_leg(g, args.obs_sweeps)
    # Timed last: back-to-back launches of one phase reuse the sweep counter's
    # draws, which would leave the chain out of equilibrium for a measurement.
    log("learned: phase timing, 200 learned launches")
    learned_ms, _ = g.time_phase(2, 200)
    n_spins = 16 * L ** 3
    rung_T = np.sort(pyro.geometric_ladder(T_MIN, T_MAX, R))[::-1]  # rungs: descending T
    verdict = learned.judge_paired(plain_legs, learned_legs, n_spins, rung_T, plain_obs,
                                   learned_obs)
    choice = learned.choose(g, verdict, prev)
    learned.use_plain(g, prev)
    # Dense work of one learned colour phase: two forward passes per proposal
    # of 2 (6*64 + 64*64 + 64*4) = 9,472 FLOP each, N/4 proposals per replica.
    proposals = n_spins // 4 * R
    flop = proposals * 2 * FLOP_PER_FORWARD
    out = {"sweep": "1 learned pass + 10 over-relaxation passes (plain: 1 Metropolis + 10)",
           "thermalisation_sweeps": args.therm_sweeps}
    out.update(info)
    out["phase"] = {
        "kernel": "k_learned_phase<16, 1>", "avg_launch_ms": learned_ms,
        "proposals_per_launch": proposals, "flop_per_launch": flop,
        "achieved_tflops": flop / (learned_ms * 1e-3) / 1e12, "peak_tflops": BF16_PEAK_TFLOPS,
        "frac_mfma": flop / (learned_ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS,
        "achieved_gbs": proposals * BYTES_PER_UPDATE / (learned_ms * 1e-3) / 1e9,
    }
    out["spin_updates_per_s"] = verdict["sweeps_per_s"] * (1 + N_OR) * n_spins * R
    out.update(verdict)
    out["worst_rungs_T"] = [float(rung_T[k]) for k in verdict.pop("worst_rungs")]
    out["fallback_choice"] = choice
    return out, learned.pool(plain_legs)


def plain_tau_leg(g, args):
    """Without the learned leg: the plain sampler's tau_int from two legs of
    --series-sweeps after --therm-sweeps thermalisation sweeps."""
    from cudacode_amd import learned
    g.Mc(args.therm_sweeps)
    return learned.pool([learned.timed_leg(g, args.series_sweeps, warmup=20) for _ in range(2)])


def phase_only(args):
    """The L=16 x 64 ensemble after the cone adaptation and --warmup sweeps;
    for the learned phase also the network trained as in learned_leg; then
    --phase-launches launches of one phase kernel over all 64 replicas, each
    timed by start/stop events in its own dispatch (pyro_time_phase)."""
    from cudacode_amd import learned, pyro
    g = pyro.Pyrochlore(L, pyro.geometric_ladder(T_MIN, T_MAX, R), seed=SEED, n_metro=N_METRO,
                        n_or=N_OR, pt_every=PT_EVERY, cos_cone=0.5)
    g.CalculateSolidAngle(target=0.5, rounds=10, sweeps_per_round=10)
    g.Mc(args.warmup)
    kind = {"learned": 2, "or_pair": 3, "metropolis": 0, "metropolis_pair": 4,
            "or_pass": 5}[args.phase_only]
    if kind == 2:
        learned.train(g, steps=args.train_steps, init_seed=SEED)
    avg_ms, span_ms = g.time_phase(kind, args.phase_launches)
    g.close()
    n_total = 16 * L ** 3 * R
    out = {"phase": args.phase_only, "launches": args.phase_launches, "avg_launch_ms": avg_ms,
           "span_ms": span_ms}
    if kind == 2:
        flop = n_total // 4 * 2 * FLOP_PER_FORWARD
        out.update({"flop_per_launch": flop,
                    "achieved_tflops": flop / (avg_ms * 1e-3) / 1e12,
                    "frac_bf16_peak": flop / (avg_ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS})
    else:
        b = {3: BYTES_PER_UPDATE_FUSED * n_total // 2, 4: BYTES_PER_UPDATE_FUSED * n_total // 2,
             5: BYTES_PER_UPDATE_PASS * n_total}.get(kind, BYTES_PER_UPDATE * n_total // 4)
        out.update({"bytes_per_launch": b, "achieved_gbs": b / (avg_ms * 1e-3) / 1e9,
                    "frac_hbm_peak": b / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS})
    return out


def line(metric_value, steps, warmup, dt, workload, extra):
    out = {
        "metric": "spin updates/s (plain sweep); tau_int-corrected independent samples/s",
        "value": metric_value, "unit": "spin updates/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": 1e3 * dt / steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": workload},
    }
    out.update(extra)
    return out


def workload_c1(args):
    """L=4 (1,024 spins), one replica at T=0.5, 1 Metropolis (cone adapted to
    50%) + 5 over-relaxation per sweep, 10^4 sweeps."""
    from cudacode_amd import presets, stats
    g = presets.preset("c1").build()
    g.CalculateSolidAngle(0.5, 20, 10)
    g.Mc(args.warmup)
    tune = g.autotune(5)
    mode = g.exec_mode()
    g.record_start(10000)
    g.sync()
    t0 = time.perf_counter()
    g.Mc(10000)
    g.sync()
    dt = time.perf_counter() - t0
    e = g.record_fetch()[:, 0] / 1024
    tau = stats.tau_int(e)[0]
    c, c_err = stats.specific_heat(e * 1024, 0.5, 1024)
    cone = float(g.cos_cone()[0])
    g.close()
    extra = {"exec_mode": {"mode": mode[0], "block": mode[1], "bins": mode[2]},
             "autotune_ms_per_sweep": tune.tolist(),
             "energy_per_spin": float(e.mean()), "energy_err": stats.binning_error(e),
             "specific_heat": c, "specific_heat_err": c_err, "tau_int_sweeps": tau,
             "indep_samples_per_s": stats.independent_samples_per_s(10000 / dt, tau)}
    if not args.no_cpu:
        # The whole workload on one host core (the phases of a 1,024-spin
        # lattice are too small for threads to pay), with the adapted cone.
        from oracle import oracle
        oracle.set_num_threads(1)
        run = oracle.OracleRun(4, [0.5], seed=SEED, n_metro=1, n_or=5, cos_cone=cone)
        t0 = time.perf_counter()
        run.run(10000)
        cpu_dt = time.perf_counter() - t0
        extra["cpu_baseline"] = {
            "value": 10000 * 6 * 1024 / cpu_dt, "unit": "spin updates/s", "cores": 1,
            "kind": "port", "sample": f"the full workload: 10^4 sweeps, 6.1e7 spin updates, "
                                      f"{cpu_dt:.1f} s, C oracle on one thread"}
    return line(10000 * 6 * 1024 / dt, 10000, args.warmup, dt,
                "L=4 (1,024 spins), 1 replica, T=0.5, 1 Metropolis + 5 over-relaxation, 10^4 sweeps",
                extra)


def fixed_exec(g, spec, mode2_tile=None):
    """Set the execution named by "STREAMS,GROUPS" (--exec) in place of the
    autotuner's choice; returns an empty result table."""
    streams, groups = (int(v) for v in spec.split(","))
    if mode2_tile:
        g.set_exec_mode(2, mode2_tile)
    g.set_streams(streams)
    g.set_group_blocking(groups)
    return np.zeros((0, 6))


def workload_c3(args):
    """256 replicas, L drawn from the seed over {8,12,16,24}, 32 temperatures
    from 0.05 to 1.0 (8 replicas each), 1 Metropolis + 10 over-relaxation,
    swaps every 10 sweeps within equal-size groups; execution chosen by the
    autotuner (persistent per-sweep kernel or tile-list phases)."""
    from cudacode_amd import presets
    cfg = presets.preset("c3")
    Ls = cfg.sizes()
    g = cfg.build()
    g.CalculateSolidAngle(0.5, 5, 10)
    g.Mc(args.warmup)
    tune = fixed_exec(g, args.exec, mode2_tile=256) if args.exec else g.autotune(5)
    mode = g.exec_mode()
    streams, groups = g.streams(), g.group_blocking()
    g.sync()
    steps = max(args.steps // 4, 10)
    t0 = time.perf_counter()
    g.Mc(steps)
    g.sync()
    dt = time.perf_counter() - t0
    n_spins = int(g.N.sum())
    g.close()
    return line(steps * 11 * n_spins / dt, steps, args.warmup, dt,
                "256 replicas, L from the seed over {8,12,16,24}, 32 temperatures 0.05..1.0, "
                "1 Metropolis + 10 over-relaxation, swaps every 10 within equal-size groups",
                {"spins": n_spins, "exec_mode": {"mode": mode[0], "block": mode[1], "bins": mode[2],
                                                 "streams": streams, "group_blocking": groups},
                 "autotune_ms_per_sweep": tune.tolist(),
                 "L_counts": {int(k): int((Ls == k).sum()) for k in (8, 12, 16, 24)}})


def workload_c5(args):
    """L=64 (4,194,304 spins) x 32 replicas, T from 0.02 to 0.5, swaps every 10
    sweeps (1.61 GB resident). Plain sweeps of 1 Metropolis + 10
    over-relaxation passes give the line's value; then the learned proposal,
    trained on 1,048,576 environments of this ensemble, is interleaved
    (1 Metropolis + 1 learned + 10 over-relaxation per sweep)."""
    from cudacode_amd import learned, presets, pyro
    cfg = presets.preset("c5")
    g = cfg.build()
    g.Mc(max(args.warmup // 20, 2))
    tune = fixed_exec(g, args.exec) if args.exec else g.autotune(3)
    steps = max(args.steps // 40, 10)
    g.sync()
    t0 = time.perf_counter()
    g.Mc(steps)
    g.sync()
    dt = time.perf_counter() - t0
    # the sweeps' over-relaxation kernel: a whole pass per launch where the
    # tuned execution runs one (the strips of k_or_pass_strip), else the pair
    try:
        pass_ms, _ = g.time_phase(5, 20)
    except pyro.PyroError:
        pass_ms = None
    pair_ms, _ = g.time_phase(3, 40)
    or_ms, _ = g.time_phase(1, 40)
    n_total = 16 * 64 ** 3 * 32
    achieved = BYTES_PER_UPDATE * n_total / 4 / (or_ms * 1e-3) / 1e9
    pair_achieved = BYTES_PER_UPDATE_FUSED * n_total / 2 / (pair_ms * 1e-3) / 1e9
    pair = {"kernel": OR_PAIR_KERNEL_C5, "achieved": pair_achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": pair_achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(OR_PAIR_KERNEL_C5), "avg_launch_ms": pair_ms,
            "bytes_per_launch": BYTES_PER_UPDATE_FUSED * n_total // 2}
    if pass_ms is not None:
        pass_achieved = BYTES_PER_UPDATE_PASS * n_total / (pass_ms * 1e-3) / 1e9
        main = {"kernel": OR_PASS_KERNEL_C5, "achieved": pass_achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": pass_achieved / HBM_PEAK_GBS,
                "traffic": measured_traffic(OR_PASS_KERNEL_C5), "avg_launch_ms": pass_ms,
                "bytes_per_launch": BYTES_PER_UPDATE_PASS * n_total, "pair": pair}
        or_bytes = BYTES_PER_UPDATE_PASS
    else:
        main, or_bytes = pair, BYTES_PER_UPDATE_FUSED
    sweep_bytes = (BYTES_PER_UPDATE + or_bytes * 10) * n_total
    extra = {"roofline": {"bound": "hbm", **main,
                          "single_phase": {
                              "kernel": OR_KERNEL_C5, "achieved": achieved,
                              "frac": achieved / HBM_PEAK_GBS,
                              "traffic": measured_traffic(OR_KERNEL_C5), "avg_launch_ms": or_ms,
                              "bytes_per_launch": BYTES_PER_UPDATE * n_total // 4},
                          "sweep_achieved": sweep_bytes * steps / dt / 1e9,
                          "sweep_frac": sweep_bytes * steps / dt / 1e9 / HBM_PEAK_GBS},
             "streams": g.streams(), "group_blocking": g.group_blocking(),
             "autotune_ms_per_sweep": tune.tolist()}
    if not args.no_learned:
        log("c5: training the learned proposal on this ensemble")
        info = learned.train(g, snapshots=16, per_replica=2048, steps=args.train_steps,
                             init_seed=SEED)
        g.set_learned_passes(cfg.n_learned)  # interleaved: Metropolis, learned, over-relaxation
        g.Mc(2)
        g.reset_acceptance()
        g.sync()
        t0 = time.perf_counter()
        g.Mc(steps)
        g.sync()
        dl = time.perf_counter() - t0
        a, p = g.acceptance_learned()
        learned_ms, _ = g.time_phase(2, 8)
        info.update({
            "sweep": "1 Metropolis + 1 learned + 10 over-relaxation passes",
            "ms_per_step": 1e3 * dl / steps,
            "spin_updates_per_s": steps * 12 * n_total / dl,
            "acceptance": float(a.sum() / max(p.sum(), 1)),
            "learned_phase_avg_launch_ms": learned_ms,
        })
        extra["learned_interleaved"] = info
    g.close()
    return line(steps * 11 * n_total / dt, steps, 0, dt,
                "L=64 (4,194,304 spins) x 32 replicas, T 0.02..0.5, 1 Metropolis + 10 "
                "over-relaxation, swaps every 10 sweeps; then with 1 learned pass interleaved",
                extra)


def main():
    if "--calibrate-copy" in sys.argv:
        sizes = [int(a) for a in sys.argv[sys.argv.index("--calibrate-copy") + 1:] if a.isdigit()]
        calibrate_copy(sweep=sizes)
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--phase-launches", type=int, default=400)
    ap.add_argument("--exec", default=None, metavar="STREAMS,GROUPS",
                    help="c3 and c5: run this execution instead of the autotuner's choice "
                         "(replica groups in flight, cache-blocking groups; c3 on 256-cell tiles), "
                         "e.g. for a profiler run, whose per-launch cost would sway the autotuner")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-proposal leg")
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c5"],
                    help="c2 (default): the L=16 x 64 parallel-tempering line; others print "
                         "their own line")
    ap.add_argument("--train-steps", type=int, default=2000)
    ap.add_argument("--obs-sweeps", type=int, default=2000,
                    help="measured sweeps per observables leg (C/N, |m|^2 agreement and tau_int)")
    ap.add_argument("--therm-sweeps", type=int, default=2000,
                    help="plain sweeps before the tau_int legs (independent of --warmup)")
    ap.add_argument("--series-sweeps", type=int, default=3000,
                    help="timed sweeps per tau_int leg (independent of --steps)")
    ap.add_argument("--phase-only", choices=["learned", "or_pair", "metropolis", "metropolis_pair",
                                             "or_pass"],
                    help="only time N launches of one colour-phase kernel on the L=16 x 64 "
                         "workload, eagerly (no captured sweeps): the command the kernel-trace "
                         "summaries of these kernels are made with")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the replicas of "
              "this workload are all resident on one MI355X.", file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))
    if args.phase_only:
        print(json.dumps(phase_only(args)), flush=True)
        return
    if args.workload != "c2":
        fn = {"c1": workload_c1, "c3": workload_c3, "c5": workload_c5}[args.workload]
        print(json.dumps(fn(args)), flush=True)
        return

    from cudacode_amd import pyro, stats

    temps = pyro.geometric_ladder(T_MIN, T_MAX, R)
    log("create ctx")
    g = pyro.Pyrochlore(L, temps, seed=SEED, n_metro=N_METRO, n_or=N_OR, pt_every=PT_EVERY,
                        cos_cone=0.5)
    log("cone adaptation (untimed)")
    g.CalculateSolidAngle(target=0.5, rounds=10, sweeps_per_round=10)
    log(f"warmup {args.warmup} sweeps")
    g.Mc(args.warmup)
    g.reset_acceptance()
    g.record_start(args.steps)
    g.sync()
    log(f"timed {args.steps} sweeps")
    t0 = time.perf_counter()
    g.Mc(args.steps)
    g.sync()
    dt = time.perf_counter() - t0
    n_total = 16 * L ** 3 * R
    updates = args.steps * (N_METRO + N_OR) * n_total
    value = updates / dt
    ms_per_step = 1e3 * dt / args.steps
    acc = float(np.mean(g.CalculateRatio()))
    timed_streams = g.streams()

    sweeps_per_s = args.steps / dt

    log(f"phase timing: {args.phase_launches} whole-pass, pair and single over-relaxation launches")
    pass_ms, _ = g.time_phase(5, args.phase_launches)
    pair_ms, _ = g.time_phase(3, args.phase_launches)
    or_ms, _ = g.time_phase(1, args.phase_launches)
    met_ms, _ = g.time_phase(0, args.phase_launches // 4)
    met_pair_ms, _ = g.time_phase(4, args.phase_launches // 4)
    g.set_kernel_options(0, False)  # the site-per-thread kernel, for comparison
    or_ms_v0, _ = g.time_phase(1, args.phase_launches // 4)
    g.set_kernel_options(1, False)
    bytes_per_launch = BYTES_PER_UPDATE * n_total // 4
    achieved = bytes_per_launch / (or_ms * 1e-3) / 1e9
    pair_bytes = BYTES_PER_UPDATE_FUSED * n_total // 2
    pair_achieved = pair_bytes / (pair_ms * 1e-3) / 1e9
    pass_bytes = BYTES_PER_UPDATE_PASS * n_total
    pass_achieved = pass_bytes / (pass_ms * 1e-3) / 1e9
    # Algorithmic bytes of the whole sweep as it runs at L=16: 36 B per
    # Metropolis update (two colours per launch, in place), 24 B per
    # over-relaxation update (a whole pass per launch).
    sweep_bytes = (BYTES_PER_UPDATE_FUSED * N_METRO + BYTES_PER_UPDATE_PASS * N_OR) * n_total
    # A caller that hands host buffers over every sweep pays the PCIe copies
    # too (pyro_set_spins re-anchors the energy as well): the host-inclusive
    # rate, reported beside the HBM-resident value.
    log("host transfer timing")
    host = g.get_spins()
    t0 = time.perf_counter()
    for _ in range(5):
        host = g.get_spins()
    get_ms = 1e3 * (time.perf_counter() - t0) / 5
    t0 = time.perf_counter()
    for _ in range(5):
        g.set_spins(host)
    set_ms = 1e3 * (time.perf_counter() - t0) / 5
    host_transfer = {
        "bytes": int(host.nbytes), "get_spins_ms": get_ms, "set_spins_ms": set_ms,
        "host_inclusive_updates_per_s": (N_METRO + N_OR) * n_total / (
            (ms_per_step + get_ms + set_ms) * 1e-3),
    }
    # The timed phase launches above reuse one sweep index's draws: sweep back
    # to equilibrium before anything samples from the chain again.
    g.Mc(200)
    # tau_int of the plain sampler from fixed-length legs after a fixed
    # thermalisation (a series as short as --steps cannot resolve it).
    learned = None
    if not args.no_learned:
        learned, plain = learned_leg(g, args)
    else:
        log(f"plain tau_int: {args.therm_sweeps} thermalisation sweeps, 2 legs of "
            f"{args.series_sweeps}")
        plain = plain_tau_leg(g, args)
    g.close()
    taus, tau_err = plain.taus, plain.tau_err
    samples = np.array([stats.independent_samples_per_s(sweeps_per_s, t) for t in taus])
    traffic = measured_traffic(OR_KERNEL)

    out = {
        "metric": "spin updates/s (plain sweep); tau_int-corrected independent samples/s",
        "value": value,
        "unit": "spin updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "parallel tempering: pyrochlore L=16 (65,536 spins) x 64 replicas, "
                        "geometric T 0.01..1.0, 1 Metropolis + 10 over-relaxation per sweep, "
                        "swaps every 10 sweeps, FP32 spins, FP64 reductions",
            "L": L, "replicas": R, "n_metro": N_METRO, "n_or": N_OR, "pt_every": PT_EVERY,
            "seed": SEED, "parallelism": "replicas resident on one GPU",
            "streams": timed_streams,
        },
        "indep_samples_per_s": {
            # sweeps/s of the timed sweeps / (2 tau_int), tau_int per rung of
            # the energy from the plain sampler's fixed-length legs
            "median_over_rungs": float(np.median(samples)),
            "lowest_T": float(samples[-1]),
            "tau_int_sweeps_median": float(np.median(taus)),
            "tau_int_sweeps_lowest_T": float(taus[-1]),
            "tau_int_sweeps_lowest_T_err": float(tau_err[-1]),
            "series_sweeps": int(plain.series.shape[0]),
            "thermalisation_sweeps": args.therm_sweeps,
        },
        "acceptance": acc,
        "host_transfer": host_transfer,
        "roofline": {
            "bound": "hbm",
            "kernel": OR_PASS_KERNEL + " (a whole over-relaxation pass per launch, one thread "
                                       "per cell, one z-plane of a replica per block, plane "
                                       "z+1's first two colours recomputed where the last two "
                                       "read them: planes read at y offsets staged in LDS, x "
                                       "offsets from the DPP row; 10 of 12 launches per sweep "
                                       "and replica group)",
            "achieved": pass_achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": pass_achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(OR_PASS_KERNEL),
            "avg_launch_ms": pass_ms,
            "bytes_per_launch": pass_bytes,
            "pair": {
                "kernel": OR_PAIR_KERNEL + " (two over-relaxation colours per launch)",
                "achieved": pair_achieved, "frac": pair_achieved / HBM_PEAK_GBS,
                "traffic": measured_traffic(OR_PAIR_KERNEL), "avg_launch_ms": pair_ms,
                "bytes_per_launch": pair_bytes},
            "single_phase": {
                "kernel": OR_KERNEL + " (one over-relaxation colour per launch)",
                "achieved": achieved, "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "avg_launch_ms": or_ms, "bytes_per_launch": bytes_per_launch},
            "metropolis_avg_launch_ms": met_ms,
            # the sweep's Metropolis: two colours per launch (k_metro_pair_plane)
            "metropolis_pair_avg_launch_ms": met_pair_ms,
            "metropolis_pair_achieved": (BYTES_PER_UPDATE_FUSED * n_total // 2)
                                        / (met_pair_ms * 1e-3) / 1e9,
            "site_per_thread_or_avg_launch_ms": or_ms_v0,
            # The whole timed sweep at its algorithmic bytes (exchanges,
            # bookkeeping and the replica groups' overlap included).
            "sweep_achieved": sweep_bytes / (ms_per_step * 1e-3) / 1e9,
            "sweep_frac": sweep_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        },
    }
    if learned is not None:
        out["learned_vs_plain"] = learned
    if not args.no_cpu and rank == 0:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline()
    if rank == 0:
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
