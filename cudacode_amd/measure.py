"""Measurement pipeline: rung-resolved observables with error bars and the
reference's per-repetition output files.

The reference's measurement loop (Pyrochlore.cu:1604-1803) accumulates, per
temperature and per measurement, the energy and its square, the moduli of the
five order-parameter groups and their squares, and |M|, in `Measure` objects
(class not in the published sources) that it saves every 500 measurements and
at the end to Energies<rep>.dat, OrderParameters<rep>.dat and
Magnetization<rep>.dat. Here the device appends one row per measurement
(Pyrochlore.measure_start / measure_fetch: E, |m_A2|, |m_E|, |m_T1|, |m_F|,
|m_T2| per rung), and this module turns the rows into means with binning
errors, the specific heat with a jackknife error and tau_int, and writes the
three files in the text format documented in `write_*` below (the reference's
format cannot be recovered, so the files state their columns in a header).

`run_protocol` is the reference's main flow for one repetition
(Pyrochlore.cu:1461-1803): simulated annealing from T = 20, parallel
tempering, thermalisation, InitialConf<rep>.dat, then the measurements with
periodic file updates and optional Conf<rep>.dat snapshots.
"""
from __future__ import annotations

import json
import os
import time
from dataclasses import dataclass

import numpy as np

from . import stats

GROUPS = ("A2", "E", "T1", "F", "T2")  # order of the 5 order-parameter columns


@dataclass
class Summary:
    """Per-rung estimates, rungs by descending temperature."""
    T: np.ndarray                 # (R,)
    n: int                        # measurements
    n_spins: np.ndarray           # (R,) spins of the replica at each rung
    e: np.ndarray                 # <E>/N and error, (R, 2)
    e2: np.ndarray                # <(E/N)^2> and error, (R, 2)
    c: np.ndarray                 # C/N = (<E^2> - <E>^2)/(N T^2) and jackknife error, (R, 2)
    tau_e: np.ndarray             # tau_int of E in measurements, (R,)
    m: np.ndarray                 # <|m_k|> and error, (R, 5, 2)
    m2: np.ndarray                # <|m_k|^2> and error, (R, 5, 2)
    tau_m2: np.ndarray            # tau_int of |m_A2|^2 in measurements, (R,)


def _bins(n: int, n_bins: int) -> int:
    return max(2, min(n_bins, n // 2)) if n >= 4 else 0


def summarise(rows: np.ndarray, temps_desc, n_spins, n_bins: int = 32) -> Summary:
    """Estimates from measurement rows (n, R, 6) as returned by measure_fetch.

    temps_desc: the rung temperatures in descending order; n_spins: spins per
    replica (scalar, or per rung for a mixed batch)."""
    rows = np.asarray(rows, np.float64)
    n, R, q = rows.shape
    if q != 6:
        raise ValueError(f"measurement rows have {q} columns, expected 6")
    T = np.asarray(temps_desc, np.float64)
    N = np.broadcast_to(np.asarray(n_spins, np.float64), (R,)).copy()
    nb = _bins(n, n_bins)
    e = np.zeros((R, 2))
    e2 = np.zeros((R, 2))
    c = np.full((R, 2), np.nan)
    tau = np.full(R, np.nan)
    m = np.zeros((R, 5, 2))
    m2 = np.zeros((R, 5, 2))
    tau_m2 = np.full(R, np.nan)
    for k in range(R):
        E = rows[:, k, 0]
        en = E / N[k]
        e[k] = en.mean(), stats.binning_error(en, nb) if nb else np.nan
        e2[k] = (en * en).mean(), stats.binning_error(en * en, nb) if nb else np.nan
        if nb:
            c[k] = stats.specific_heat(E, T[k], N[k], nb)
            tau[k] = stats.tau_int(E)[0]
            tau_m2[k] = stats.tau_int(rows[:, k, 1] ** 2)[0]
        for g in range(5):
            x = rows[:, k, 1 + g]
            m[k, g] = x.mean(), stats.binning_error(x, nb) if nb else np.nan
            m2[k, g] = (x * x).mean(), stats.binning_error(x * x, nb) if nb else np.nan
    return Summary(T=T, n=n, n_spins=N, e=e, e2=e2, c=c, tau_e=tau, m=m, m2=m2, tau_m2=tau_m2)


def _fmt(v) -> str:
    return f"{v:.12e}"


def write_energies(path: str, s: Summary) -> None:
    """Energies<rep>.dat: one line per rung, descending T. Columns:
    T, <E>/N, err, <(E/N)^2>, err, C/N, err, tau_int(E) [measurements], n."""
    with open(path, "w") as f:
        f.write("# T  <E>/N  err  <(E/N)^2>  err  C/N  err  tau_int_E  n_measurements\n")
        for k in range(len(s.T)):
            vals = [s.T[k], *s.e[k], *s.e2[k], *s.c[k], s.tau_e[k]]
            f.write(" ".join(_fmt(v) for v in vals) + f" {s.n}\n")


def write_order_parameters(path: str, s: Summary) -> None:
    """OrderParameters<rep>.dat: one line per rung, descending T. Columns: T,
    then <|m_k|> and its error for k = A2, E, T1, F, T2, then <|m_k|^2> and
    its error in the same order (the reference's 10 accumulated quantities,
    Pyrochlore.cu:1656-1666)."""
    with open(path, "w") as f:
        head = ["T"] + [f"|m_{g}| err" for g in GROUPS] + [f"|m_{g}|^2 err" for g in GROUPS]
        f.write("# " + "  ".join(head) + "\n")
        for k in range(len(s.T)):
            vals = [s.T[k]] + list(s.m[k].ravel()) + list(s.m2[k].ravel())
            f.write(" ".join(_fmt(v) for v in vals) + "\n")


def write_magnetization(path: str, s: Summary) -> None:
    """Magnetization<rep>.dat: one line per rung, descending T. Columns: T,
    <|M|>/N, err, <(|M|/N)^2>, err (|M|/N is |m_F|)."""
    with open(path, "w") as f:
        f.write("# T  <|M|>/N  err  <(|M|/N)^2>  err\n")
        for k in range(len(s.T)):
            vals = [s.T[k], *s.m[k, 3], *s.m2[k, 3]]
            f.write(" ".join(_fmt(v) for v in vals) + "\n")


def read_table(path: str) -> np.ndarray:
    """Any of the three files back as an array (header skipped)."""
    return np.loadtxt(path, comments="#", ndmin=2)


def write_all(out_dir: str, rep: int, s: Summary) -> dict:
    os.makedirs(out_dir, exist_ok=True)
    paths = {
        "energies": os.path.join(out_dir, f"Energies{rep}.dat"),
        "order_parameters": os.path.join(out_dir, f"OrderParameters{rep}.dat"),
        "magnetization": os.path.join(out_dir, f"Magnetization{rep}.dat"),
    }
    write_energies(paths["energies"], s)
    write_order_parameters(paths["order_parameters"], s)
    write_magnetization(paths["magnetization"], s)
    return paths


def log_line(path: str, s: Summary, **extra) -> None:
    """Append one JSON line to the run log: per-rung E/N, C/N, <|m_A2|^2>,
    tau_int(E), plus any extra fields (sweeps, updates/s, acceptance)."""
    rec = {
        "time": time.time(), "n_measurements": s.n, "T": s.T.tolist(),
        "E_per_spin": s.e[:, 0].tolist(), "E_err": s.e[:, 1].tolist(),
        "C_per_spin": s.c[:, 0].tolist(), "C_err": s.c[:, 1].tolist(),
        "m_A2_sq": s.m2[:, 0, 0].tolist(), "tau_int_E": s.tau_e.tolist(),
    }
    rec.update(extra)
    with open(path, "a") as f:
        f.write(json.dumps(rec, allow_nan=True) + "\n")


def rung_temperatures(g) -> np.ndarray:
    """Rung temperatures in rung order (descending T; grouped by L for a batch)."""
    T, at = g.temperatures()
    return np.asarray(T)[np.asarray(at)]


def rung_spins(g) -> np.ndarray:
    """Spins of the replica at each rung (constant per rung: exchanges stay
    within equal-size groups)."""
    _, at = g.temperatures()
    return np.asarray(g.N, np.float64)[np.asarray(at)] if np.ndim(g.N) else np.full(len(at), float(g.N))


def run_protocol(g, out_dir: str, rep: int = 0, *, anneal_steps: int = 100,
                 anneal_sweeps: int = 10, T_start: float = 20.0, pt_rounds: int = 100,
                 thermalisation: int = 1000, n_measures: int = 1000, sweeps_per_measure: int = 10,
                 save_every: int = 500, store_confs: int = 0, confs_every: int = 100,
                 n_bins: int = 32) -> Summary:
    """One repetition of the reference's flow (Pyrochlore.cu:1461-1803) on a
    ctx whose temperatures are the targets: annealing (anneal_steps x
    anneal_sweeps sweeps from T_start), pt_rounds exchange intervals,
    thermalisation sweeps, InitialConf<rep>.dat, then n_measures
    measurements sweeps_per_measure sweeps apart, writing the three result
    files every save_every measurements and at the end, and up to
    store_confs Conf<rep>.dat records every confs_every measurements (both
    files only for a single lattice size, as in the reference). Each
    file update first runs the spin guard (an error if any spin is not
    finite or off the unit sphere) and appends a line to run<rep>.jsonl."""
    os.makedirs(out_dir, exist_ok=True)
    # The reference's configuration files hold one lattice size; a batch of
    # mixed sizes keeps its configurations in checkpoints instead.
    files = len(np.unique(np.atleast_1d(g.N))) == 1
    g.SimulatedAnnealing(anneal_steps, anneal_sweeps, T_start)
    g.Mc(pt_rounds * max(g.pt_every, 1))
    g.Mc(thermalisation)
    # Repetition rep appends its record (the reference passes rep as the
    # append flag, Pyrochlore.cu:1517-1522), so record rep of InitialConf<k>.dat
    # is this repetition's starting point (ReadConfigurations(..., rep)).
    if files:
        g.SaveConfigurations(os.path.join(out_dir, "InitialConf"), append=rep > 0)
    T = rung_temperatures(g)
    N = rung_spins(g)
    # Measurements fall on sweeps whose index + 1 is a multiple of the
    # cadence: align so that the first comes sweeps_per_measure sweeps on.
    sweep, _ = g.sweep_counters()
    if sweep % sweeps_per_measure:
        g.Mc(sweeps_per_measure - sweep % sweeps_per_measure)
    g.measure_start(sweeps_per_measure, n_measures)
    done, confs = 0, 0
    summary = None
    conf_prefix = os.path.join(out_dir, "Conf")
    run_log = os.path.join(out_dir, f"run{rep}.jsonl")
    spins = float(np.sum(N))
    passes = g.n_metro + g.n_or
    t0 = time.perf_counter()
    while done < n_measures:
        # Advance to the next file update or configuration snapshot.
        step = min(save_every - done % save_every, n_measures - done)
        if store_confs:
            step = min(step, confs_every - done % confs_every)
        g.Mc(step * sweeps_per_measure)
        done += step
        if files and store_confs and confs < store_confs and done % confs_every == 0:
            g.SaveConfigurations(conf_prefix, append=True)
            confs += 1
        if done % save_every == 0 or done == n_measures:
            n_bad, first = g.check_spins(1e-4)
            if n_bad:
                raise RuntimeError(f"{n_bad} spins not finite or not unit (first index {first})")
            summary = summarise(g.measure_fetch(), T, N, n_bins)
            write_all(out_dir, rep, summary)
            g.sync()
            dt = time.perf_counter() - t0
            acc, prop = g.acceptance()
            log_line(run_log, summary, sweeps=done * sweeps_per_measure,
                     spin_updates_per_s=done * sweeps_per_measure * passes * spins / dt,
                     acceptance=(acc / np.maximum(prop, 1)).tolist())
    g.measure_stop()
    return summary
