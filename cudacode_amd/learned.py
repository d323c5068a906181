"""The learned local proposal as a workflow: train it on environments the plain
sampler generates, judge it against the plain sampler, and fall back to the
plain sampler when it does not pay.

- `train`: gathers (beta h, s) environments from plain sweeps across the
  temperature ladder and trains the network on the GPU (the library's own
  FP32 forward/backward kernels and Adam, pyro_mlp_train).
- `timed_leg`: runs and times sweeps with an energy series recorded, and
  returns sweeps/s, per-rung tau_int with its Madras-Sokal error and
  tau_int-corrected samples/s.
- `paired_legs`: alternating plain and learned legs of a fixed length from
  one equilibrated chain (plain, learned, plain, learned, ...), so that both
  samplers are measured at the same point of the chain and neither carries
  the other's equilibration; `pool` joins the legs of one sampler.
- `judge_paired`: the samples/s ratio per rung with a bootstrap interval of
  its median over rungs, the agreement of the mean energy and C/N per rung,
  and the plain-vs-plain and learned-vs-learned controls of the same lengths.
- `choose`: keeps the learned proposal only if it is faster in samples/s and
  agrees with the plain sampler; otherwise restores the plain schedule.

The proposal is exact whatever the network does (the Metropolis-Hastings
correction), so the plain sampler is the judge of speed, not of validity; a
disagreement beyond the control's spread means a defect, and the workflow
then falls back too.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from . import measure, pyro, stats


def train(g, snapshots: int = 16, per_replica: int = 1024, sweeps_between: int = 5,
          steps: int = 2000, batch: int = 4096, lr: float = 3e-3, lr_final: float = 3e-4,
          init_seed: int = 0) -> dict:
    """Gather snapshots x per_replica x R environments from the ctx's current
    (plain) schedule and train the network from a fresh initialisation."""
    g.clear_environments()
    cap = snapshots * per_replica * g.R
    for k in range(snapshots):
        g.Mc(sweeps_between)
        g.gather_environments(per_replica, 1000 + k, cap)
    g.mlp_set_weights(pyro.mlp_init(init_seed))
    g.sync()
    t0 = time.perf_counter()
    hist = g.mlp_train(steps, batch=batch, lr=lr, lr_final=lr_final)
    g.sync()
    return {"environments": cap, "train_steps": steps, "train_s": time.perf_counter() - t0,
            "loss_first": float(hist[: min(10, hist.size)].mean()),
            "loss_last": float(hist[-min(50, hist.size):].mean())}


@dataclass
class Leg:
    series: np.ndarray        # (sweeps, R) energy per sweep, rungs by descending T
    sweeps_per_s: float
    taus: np.ndarray          # tau_int per rung, in sweeps
    samples: np.ndarray       # tau_int-corrected independent samples/s per rung
    acceptance: float
    tau_err: np.ndarray = None  # Madras-Sokal error of tau_int per rung
    seconds: float = 0.0      # wall time of the timed sweeps


def _taus(series: np.ndarray):
    t = [stats.tau_int(series[:, k]) for k in range(series.shape[1])]
    return np.array([v[0] for v in t]), np.array([v[2] for v in t])


def timed_leg(g, sweeps: int, warmup: int = 0, learned: bool = False) -> Leg:
    """warmup untimed sweeps, then `sweeps` timed ones with the energy series."""
    if warmup:
        g.Mc(warmup)
    g.reset_acceptance()
    g.record_start(sweeps)
    g.sync()
    t0 = time.perf_counter()
    g.Mc(sweeps)
    g.sync()
    dt = time.perf_counter() - t0
    series = g.record_fetch()
    a, p = g.acceptance_learned() if learned else g.acceptance()
    sps = sweeps / dt
    taus, err = _taus(series)
    samples = np.array([stats.independent_samples_per_s(sps, t) for t in taus])
    return Leg(series, sps, taus, samples, float(a.sum() / max(p.sum(), 1)), err, dt)


def paired_legs(g, sweeps: int, rounds: int = 2, warmup: int = 20, n_learned: int = 1):
    """`rounds` rounds of a plain leg then a learned leg, each `sweeps` timed
    sweeps (after `warmup` untimed ones of the same sampler, which also
    capture its graph), continuing one chain. Both samplers leave the
    Boltzmann law stationary, so a chain equilibrated by one is equilibrated
    for the other, and alternating the legs puts both at the same point of
    the chain. Returns (plain legs, learned legs); the ctx ends on its
    previous (plain) schedule."""
    plain, learned = [], []
    prev = (g.n_metro, g.n_or)
    for _ in range(rounds):
        plain.append(timed_leg(g, sweeps, warmup=warmup))
        use_learned(g, n_learned)
        learned.append(timed_leg(g, sweeps, warmup=warmup, learned=True))
        use_plain(g, prev)
    return plain, learned


def pool(legs) -> Leg:
    """One sampler's legs together: sweeps/s over all of them, tau_int the
    mean of the legs' estimates (error: their Madras-Sokal errors combined),
    and the series concatenated (for means and binning errors only)."""
    n = sum(leg.series.shape[0] for leg in legs)
    secs = sum(leg.seconds for leg in legs)
    sps = n / secs
    taus = np.mean([leg.taus for leg in legs], axis=0)
    err = np.sqrt(np.sum([leg.tau_err ** 2 for leg in legs], axis=0)) / len(legs)
    samples = sps / (2.0 * taus)
    acc = float(np.mean([leg.acceptance for leg in legs]))
    return Leg(np.concatenate([leg.series for leg in legs]), sps, taus, samples, acc, err, secs)


def ratio_interval(plain: Leg, learned: Leg, n_boot: int = 4000, seed: int = 0):
    """Percentiles (2.5, 16, 50, 84, 97.5) of the median over rungs of the
    samples/s ratio learned / plain, from a parametric bootstrap that draws
    each rung's tau_int of both samplers from a normal law with its
    Madras-Sokal error (clipped at 1/2, the uncorrelated value); the
    sweeps/s ratio is taken as exact (its run-to-run spread is a few %)."""
    rng = np.random.default_rng(seed)
    R = plain.taus.size
    tp = np.maximum(plain.taus + rng.standard_normal((n_boot, R)) * plain.tau_err, 0.5)
    tl = np.maximum(learned.taus + rng.standard_normal((n_boot, R)) * learned.tau_err, 0.5)
    med = np.median((learned.sweeps_per_s / plain.sweeps_per_s) * tp / tl, axis=1)
    return np.percentile(med, [2.5, 16, 50, 84, 97.5])


def observables_leg(g, sweeps: int, n_bins: int = 20) -> measure.Summary:
    """`sweeps` sweeps measured after every sweep (device rows): per-rung E/N,
    C/N, order-parameter moduli with errors, and tau_int of E and |m_A2|^2 in
    sweeps."""
    g.measure_start(1, sweeps)
    g.Mc(sweeps)
    rows = g.measure_fetch()
    g.measure_stop()
    return measure.summarise(rows, measure.rung_temperatures(g), measure.rung_spins(g), n_bins)


def zdiff(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """|a0 - b0| / hypot(a1, b1) for (value, error) rows."""
    return np.abs(a[..., 0] - b[..., 0]) / np.maximum(np.hypot(a[..., 1], b[..., 1]), 1e-300)


def zscores(a: np.ndarray, b: np.ndarray, n_spins, n_bins: int = 20) -> np.ndarray:
    """|mean difference| / combined binning error of the per-spin energy, per rung."""
    N = np.broadcast_to(np.asarray(n_spins, np.float64), (a.shape[1],))
    z = []
    for k in range(a.shape[1]):
        e1, e2 = a[:, k] / N[k], b[:, k] / N[k]
        s = np.hypot(stats.binning_error(e1, n_bins), stats.binning_error(e2, n_bins))
        z.append(abs(e1.mean() - e2.mean()) / max(s, 1e-300))
    return np.array(z)


def use_learned(g, n_learned: int = 1) -> tuple:
    """Switch the ctx to n_learned learned passes in place of Metropolis;
    returns the previous (n_metro, n_or) for `use_plain`."""
    prev = (g.n_metro, g.n_or)
    g.set_passes(0, g.n_or)
    g.set_learned_passes(n_learned)
    return prev


def use_plain(g, prev: tuple) -> None:
    g.set_learned_passes(0)
    g.set_passes(*prev)


def specific_heats(series: np.ndarray, temps_desc, n_spins, n_bins: int = 20) -> np.ndarray:
    """C/N per rung with its jackknife error from an energy series (sweeps, R)."""
    N = np.broadcast_to(np.asarray(n_spins, np.float64), (series.shape[1],))
    return np.array([stats.specific_heat(series[:, k], float(temps_desc[k]), float(N[k]), n_bins)
                     for k in range(series.shape[1])])


def judge_paired(plain_legs, learned_legs, n_spins, temps_desc, plain_obs=None,
                 learned_obs=None, min_length_taus: float = 50.0) -> dict:
    """Speed and agreement of the learned sampler against the plain one from
    the legs of `paired_legs` (at least two of each): the samples/s ratio per
    rung and the bootstrap interval of its median, the agreement per rung of
    the mean energy and of C/N between the pooled samplers, and the same
    statistics between the first two legs of each sampler (plain vs plain,
    learned vs learned) as controls of what chance gives at these lengths.
    `series_long_enough` is False when a leg is shorter than min_length_taus
    times a rung's tau_int (the window estimate is then biased low)."""
    P, Q = pool(plain_legs), pool(learned_legs)
    ratio = Q.samples / P.samples
    rel = np.hypot(P.tau_err / P.taus, Q.tau_err / Q.taus)
    lo95, lo68, med, hi68, hi95 = ratio_interval(P, Q)
    n_leg = min(leg.series.shape[0] for leg in plain_legs + learned_legs)
    long_enough = bool(n_leg >= min_length_taus * max(P.taus.max(), Q.taus.max()))
    z = zscores(P.series, Q.series, n_spins)
    z_pp = zscores(plain_legs[0].series, plain_legs[1].series, n_spins)
    z_ll = zscores(learned_legs[0].series, learned_legs[1].series, n_spins)
    cP = specific_heats(P.series, temps_desc, n_spins)
    cQ = specific_heats(Q.series, temps_desc, n_spins)
    zc = zdiff(cP, cQ)
    zc_pp = zdiff(specific_heats(plain_legs[0].series, temps_desc, n_spins),
                  specific_heats(plain_legs[1].series, temps_desc, n_spins))
    out = {
        "legs": f"{len(plain_legs)} plain and {len(learned_legs)} learned legs of {n_leg} "
                "timed sweeps, alternating on one equilibrated chain",
        "series_sweeps_per_leg": int(n_leg), "series_long_enough": long_enough,
        "sweeps_per_s": Q.sweeps_per_s, "plain_sweeps_per_s": P.sweeps_per_s,
        "acceptance": Q.acceptance, "plain_acceptance": P.acceptance,
        "tau_int_sweeps_median": float(np.median(Q.taus)),
        "tau_int_sweeps_lowest_T": float(Q.taus[-1]),
        "plain_tau_int_sweeps_median": float(np.median(P.taus)),
        "plain_tau_int_sweeps_lowest_T": float(P.taus[-1]),
        "indep_samples_per_s_median": float(np.median(Q.samples)),
        "indep_samples_per_s_lowest_T": float(Q.samples[-1]),
        "plain_indep_samples_per_s_median": float(np.median(P.samples)),
        "samples_ratio_median": float(np.median(ratio)),
        "samples_ratio_median_68": [float(lo68), float(hi68)],
        "samples_ratio_median_95": [float(lo95), float(hi95)],
        "samples_ratio_lowest_T": float(ratio[-1]),
        "samples_ratio_lowest_T_err": float(ratio[-1] * rel[-1]),
        "samples_ratio_min": float(ratio.min()),
        "energy_agreement_max_z": float(z.max()),
        "energy_agreement_rungs_within_2sigma": int((z < 2).sum()),
        "control_plain_vs_plain_max_z": float(z_pp.max()),
        "control_plain_vs_plain_within_2sigma": int((z_pp < 2).sum()),
        "control_learned_vs_learned_max_z": float(z_ll.max()),
        "control_learned_vs_learned_within_2sigma": int((z_ll < 2).sum()),
        "specific_heat_series_max_z": float(np.nanmax(zc)),
        "specific_heat_series_within_2sigma": int((zc < 2).sum()),
        "control_specific_heat_plain_vs_plain_max_z": float(np.nanmax(zc_pp)),
        "control_specific_heat_plain_vs_plain_within_2sigma": int((zc_pp < 2).sum()),
        "worst_rungs": [int(k) for k in np.argsort(-z)[:3]],
        "tau_int_E_per_rung": [float(t) for t in Q.taus],
        "tau_int_E_err_per_rung": [float(t) for t in Q.tau_err],
        "plain_tau_int_E_per_rung": [float(t) for t in P.taus],
        "plain_tau_int_E_err_per_rung": [float(t) for t in P.tau_err],
        "samples_ratio_per_rung": [float(r) for r in ratio],
    }
    if plain_obs is not None:
        out.update(_obs_verdict(plain_obs, learned_obs))
    return out


def _obs_verdict(p: measure.Summary, q: measure.Summary) -> dict:
    zc = zdiff(p.c, q.c)
    zm = zdiff(p.m2[:, 0], q.m2[:, 0])
    return {
        "specific_heat_max_z": float(np.nanmax(zc)),
        "specific_heat_rungs_within_2sigma": int((zc < 2).sum()),
        "m_A2_sq_max_z": float(np.nanmax(zm)),
        "m_A2_sq_rungs_within_2sigma": int((zm < 2).sum()),
        "tau_int_m_A2_sq_per_rung": [float(t) for t in q.tau_m2],
        "plain_tau_int_m_A2_sq_per_rung": [float(t) for t in p.tau_m2],
        "tau_int_m_A2_sq_median": float(np.nanmedian(q.tau_m2)),
        "plain_tau_int_m_A2_sq_median": float(np.nanmedian(p.tau_m2)),
        # The lowest-temperature rung's estimates, (value, error) per sampler:
        # E/N, C/N (-> 3/4 as T -> 0 for this model) and <|m_A2|^2>.
        "lowest_T": {
            "T": float(q.T[-1]),
            "energy_per_spin": [float(q.e[-1, 0]), float(q.e[-1, 1])],
            "plain_energy_per_spin": [float(p.e[-1, 0]), float(p.e[-1, 1])],
            "specific_heat": [float(q.c[-1, 0]), float(q.c[-1, 1])],
            "plain_specific_heat": [float(p.c[-1, 0]), float(p.c[-1, 1])],
            "m_A2_sq": [float(q.m2[-1, 0, 0]), float(q.m2[-1, 0, 1])],
            "plain_m_A2_sq": [float(p.m2[-1, 0, 0]), float(p.m2[-1, 0, 1])],
        },
    }


def choose(g, verdict: dict, prev: tuple, n_learned: int = 1, min_ratio: float = 1.0,
           z_margin: float = 1.5) -> str:
    """Keep the learned proposal ('learned') if its median samples/s ratio is
    above min_ratio and its worst energy z-score is within z_margin times the
    plain control's; otherwise fall back to the plain schedule ('plain')."""
    ok_speed = verdict["samples_ratio_median"] > min_ratio
    ok_agree = verdict["energy_agreement_max_z"] <= z_margin * max(
        verdict["control_plain_vs_plain_max_z"], 2.0)
    if ok_speed and ok_agree:
        g.set_passes(0, prev[1])
        g.set_learned_passes(n_learned)
        return "learned"
    use_plain(g, prev)
    return "plain"
