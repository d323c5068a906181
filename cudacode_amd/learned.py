"""The learned local proposal as a workflow: train it on environments the plain
sampler generates, judge it against the plain sampler, and fall back to the
plain sampler when it does not pay.

- `train`: gathers (beta h, s) environments from plain sweeps across the
  temperature ladder and trains the network on the GPU (the library's own
  FP32 forward/backward kernels and Adam, pyro_mlp_train).
- `timed_leg`: runs and times sweeps with an energy series recorded, and
  returns sweeps/s, per-rung tau_int and tau_int-corrected samples/s.
- `judge`: compares a learned leg with a plain leg and a plain control leg:
  the samples/s ratio per rung and the agreement of the mean energy per rung
  (z-scores, read against the plain-vs-plain spread of the same run lengths).
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
    taus = np.array([stats.tau_int(series[:, k])[0] for k in range(series.shape[1])])
    samples = np.array([stats.independent_samples_per_s(sps, t) for t in taus])
    return Leg(series, sps, taus, samples, float(a.sum() / max(p.sum(), 1)))


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


def judge(plain: Leg, learned: Leg, control: Leg, n_spins, plain_obs=None,
          learned_obs=None) -> dict:
    """Speed and agreement of a learned leg against the plain sampler; with
    observables legs, also the agreement of C/N and <|m_A2|^2> per rung and
    tau_int per rung for E and |m_A2|^2."""
    ratio = learned.samples / plain.samples
    z = zscores(plain.series, learned.series, n_spins)
    z_ctl = zscores(plain.series, control.series, n_spins)
    z_lc = zscores(learned.series, control.series, n_spins)
    return {
        "sweeps_per_s": learned.sweeps_per_s, "plain_sweeps_per_s": plain.sweeps_per_s,
        "acceptance": learned.acceptance,
        "tau_int_sweeps_median": float(np.median(learned.taus)),
        "tau_int_sweeps_lowest_T": float(learned.taus[-1]),
        "plain_tau_int_sweeps_median": float(np.median(plain.taus)),
        "indep_samples_per_s_median": float(np.median(learned.samples)),
        "indep_samples_per_s_lowest_T": float(learned.samples[-1]),
        "samples_ratio_median": float(np.median(ratio)),
        "samples_ratio_lowest_T": float(ratio[-1]),
        "samples_ratio_min": float(ratio.min()),
        "energy_agreement_max_z": float(z.max()),
        "energy_agreement_rungs_within_2sigma": int((z < 2).sum()),
        "control_plain_vs_plain_max_z": float(z_ctl.max()),
        "control_plain_vs_plain_within_2sigma": int((z_ctl < 2).sum()),
        "learned_vs_control_max_z": float(z_lc.max()),
        "learned_vs_control_within_2sigma": int((z_lc < 2).sum()),
        "worst_rungs": [int(k) for k in np.argsort(-z)[:3]],
        "tau_int_E_per_rung": [float(t) for t in learned.taus],
        "plain_tau_int_E_per_rung": [float(t) for t in plain.taus],
    } | (_obs_verdict(plain_obs, learned_obs) if plain_obs is not None else {})


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
