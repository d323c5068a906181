"""Host-side interface over the C ABI of libpyro.so (include/pyro.h).

The class mirrors the reference's host entry points for the Monte Carlo step
(taillefu/cudacode): ``Mc`` (Pyrochlore.cu:74), ``MCMultipleTemperaturesPerSystem``
(Pyrochlore.cu:829), ``OverRelaxationByPiStep`` (Pyrochlore.cu:975),
``TotalEnergyv2`` (Math.cu:309), ``Magnetization`` (Pyrochlore.cu:453),
``DoMeasurement`` (Pyrochlore.cu:1081), ``TrySwapingTemperatures``
(MonteCarlo.cu:105), ``GenerateRandomVectorsOnDevice`` (MonteCarlo.cu:275),
``CalculateRatio`` (MonteCarlo.cu:346) and ``CalculateSolidAngle``
(MonteCarlo.cu:371). Errors raise :class:`PyroError` carrying the library's
message instead of the reference's printf-and-continue.

All compute runs in the gfx950 library; this module only moves arguments.
There is no CPU fallback: without the built library the import fails.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpyro.so")

# Entry points of include/pyro.h, in declaration order.
EXPORTS = [
    "pyro_version", "pyro_last_error", "pyro_create", "pyro_destroy", "pyro_links16",
    "pyro_independent_sets", "pyro_philox4x32_10", "pyro_init_random_spins", "pyro_set_spins",
    "pyro_get_spins", "pyro_spins_device_ptr", "pyro_set_temperatures", "pyro_get_temperatures",
    "pyro_set_cos_cone", "pyro_get_cos_cone", "pyro_metropolis_phase", "pyro_overrelax_phase",
    "pyro_sweep", "pyro_get_sweep", "pyro_get_graph_block", "pyro_set_sweep", "pyro_pt_swap", "pyro_energy",
    "pyro_magnetization", "pyro_order_params", "pyro_sublattice_sums", "pyro_fold_order_params",
    "pyro_energy_incremental",
    "pyro_acceptance", "pyro_reset_acceptance", "pyro_adapt_cone", "pyro_record_start",
    "pyro_record_fetch", "pyro_set_couplings", "pyro_get_couplings", "pyro_exchange_matrices",
    "pyro_zeeman_fields", "pyro_magnetization_g", "pyro_fold_magnetization_g", "pyro_set_streams", "pyro_get_streams", "pyro_set_group_blocking", "pyro_get_group_blocking", "pyro_set_or_fusion", "pyro_check_spins", "pyro_set_post_rotate", "pyro_set_update_schedule", "pyro_anneal", "pyro_measure_start", "pyro_measure_stop", "pyro_measure_fetch",
    "pyro_set_passes", "pyro_set_kernel_options", "pyro_create_batch",
    "pyro_get_layout", "pyro_set_exec_mode", "pyro_get_exec_mode", "pyro_autotune",
    "pyro_checkpoint_save", "pyro_checkpoint_load", "pyro_export_configurations",
    "pyro_import_configurations",
    "pyro_mlp_set_weights", "pyro_mlp_get_weights",
    "pyro_mlp_eval", "pyro_vmf_draw", "pyro_mlp_loss_grad", "pyro_gather_environments",
    "pyro_get_environments", "pyro_clear_environments", "pyro_mlp_train", "pyro_learned_phase",
    "pyro_set_learned_passes", "pyro_acceptance_learned", "pyro_mlp_adam_step", "pyro_time_phase",
    "pyro_observables_noalloc", "pyro_sync",
]

MLP_PARAMS = 4868
MEAS_Q = 6  # per rung and measurement: E, |m_A2|, |m_E|, |m_T1|, |m_F|, |m_T2|


class PyroError(RuntimeError):
    pass


class Config(C.Structure):
    _fields_ = [
        ("L", C.c_int),
        ("n_replicas", C.c_int),
        ("temperatures", C.POINTER(C.c_double)),
        ("seed", C.c_uint64),
        ("J", C.c_double),
        ("spin_len2", C.c_double),
        ("n_metro", C.c_int),
        ("n_or", C.c_int),
        ("pt_every", C.c_int),
        ("cos_cone", C.c_float),
        ("device", C.c_int),
    ]


class BatchConfig(C.Structure):
    _fields_ = [
        ("n_replicas", C.c_int),
        ("L", C.POINTER(C.c_int)),
        ("temperatures", C.POINTER(C.c_double)),
        ("seed", C.c_uint64),
        ("J", C.c_double),
        ("spin_len2", C.c_double),
        ("n_metro", C.c_int),
        ("n_or", C.c_int),
        ("pt_every", C.c_int),
        ("cos_cone", C.c_float),
        ("device", C.c_int),
    ]


_lib = None


def lib():
    """Load libpyro.so. Raises if it has not been built: there is no fallback."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing; build it with `python -m cudacode_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.pyro_version.restype = C.c_char_p
        L.pyro_last_error.restype = C.c_char_p
        L.pyro_create.argtypes = [C.POINTER(Config), C.POINTER(vp)]
        L.pyro_destroy.argtypes = [vp]
        L.pyro_links16.argtypes = [vp]
        L.pyro_independent_sets.argtypes = [vp]
        L.pyro_philox4x32_10.argtypes = [vp, vp, vp]
        L.pyro_init_random_spins.argtypes = [vp]
        L.pyro_set_spins.argtypes = [vp, vp]
        L.pyro_get_spins.argtypes = [vp, vp]
        L.pyro_spins_device_ptr.argtypes = [vp, C.POINTER(vp), C.POINTER(C.c_size_t)]
        L.pyro_set_temperatures.argtypes = [vp, vp]
        L.pyro_get_temperatures.argtypes = [vp, vp, vp]
        L.pyro_set_cos_cone.argtypes = [vp, vp]
        L.pyro_get_cos_cone.argtypes = [vp, vp]
        L.pyro_metropolis_phase.argtypes = [vp, C.c_int, C.c_uint32, C.c_int]
        L.pyro_overrelax_phase.argtypes = [vp, C.c_int]
        L.pyro_sweep.argtypes = [vp, C.c_int, C.c_int]
        L.pyro_get_sweep.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.pyro_get_graph_block.argtypes = [vp, C.POINTER(C.c_int)]
        L.pyro_set_sweep.argtypes = [vp, C.c_uint32, C.c_uint32]
        L.pyro_pt_swap.argtypes = [vp, C.c_int, C.c_uint32, C.POINTER(C.c_int)]
        for name in ("pyro_energy", "pyro_magnetization", "pyro_order_params",
                     "pyro_sublattice_sums", "pyro_energy_incremental"):
            getattr(L, name).argtypes = [vp, vp]
        L.pyro_fold_order_params.argtypes = [vp, C.c_int, C.c_double, vp]
        L.pyro_acceptance.argtypes = [vp, vp, vp]
        L.pyro_reset_acceptance.argtypes = [vp]
        L.pyro_adapt_cone.argtypes = [vp, C.c_double, C.c_int, C.c_int]
        L.pyro_record_start.argtypes = [vp, C.c_int]
        L.pyro_record_fetch.argtypes = [vp, vp, C.POINTER(C.c_int)]
        L.pyro_anneal.argtypes = [vp, C.c_double, C.c_int, C.c_int]
        L.pyro_set_post_rotate.argtypes = [vp, C.c_int]
        L.pyro_set_update_schedule.argtypes = [vp, C.c_int, C.c_int]
        L.pyro_set_streams.argtypes = [vp, C.c_int]
        L.pyro_get_streams.argtypes = [vp, C.POINTER(C.c_int)]
        L.pyro_set_or_fusion.argtypes = [vp, C.c_int]
        L.pyro_set_group_blocking.argtypes = [vp, C.c_int]
        L.pyro_get_group_blocking.argtypes = [vp, C.POINTER(C.c_int)]
        L.pyro_check_spins.argtypes = [vp, C.c_double, C.POINTER(C.c_longlong),
                                       C.POINTER(C.c_longlong)]
        L.pyro_set_couplings.argtypes = [vp, vp, vp]
        L.pyro_get_couplings.argtypes = [vp, C.POINTER(C.c_int), vp]
        L.pyro_exchange_matrices.argtypes = [C.c_double] * 4 + [vp]
        L.pyro_zeeman_fields.argtypes = [vp, C.c_double, C.c_double, vp]
        L.pyro_magnetization_g.argtypes = [vp, C.c_double, C.c_double, vp]
        L.pyro_fold_magnetization_g.argtypes = [vp, C.c_int, C.c_double, C.c_double, vp]
        L.pyro_measure_start.argtypes = [vp, C.c_int, C.c_int]
        L.pyro_measure_stop.argtypes = [vp]
        L.pyro_measure_fetch.argtypes = [vp, vp, C.POINTER(C.c_int)]
        L.pyro_time_phase.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double)]
        L.pyro_set_passes.argtypes = [vp, C.c_int, C.c_int]
        L.pyro_set_kernel_options.argtypes = [vp, C.c_int, C.c_int]
        L.pyro_create_batch.argtypes = [C.POINTER(BatchConfig), C.POINTER(vp)]
        L.pyro_get_layout.argtypes = [vp, vp, vp]
        L.pyro_set_exec_mode.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.pyro_get_exec_mode.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                         C.POINTER(C.c_int)]
        L.pyro_autotune.argtypes = [vp, C.c_int, vp, C.POINTER(C.c_int)]
        L.pyro_checkpoint_save.argtypes = [vp, C.c_char_p]
        L.pyro_checkpoint_load.argtypes = [vp, C.c_char_p]
        L.pyro_export_configurations.argtypes = [vp, C.c_char_p, C.c_int, C.c_int]
        L.pyro_import_configurations.argtypes = [vp, C.c_char_p, C.c_int, C.c_int]
        L.pyro_mlp_set_weights.argtypes = [vp, vp]
        L.pyro_mlp_get_weights.argtypes = [vp, vp]
        L.pyro_mlp_eval.argtypes = [vp, vp, C.c_int, vp]
        L.pyro_vmf_draw.argtypes = [vp, vp, C.c_int, vp]
        L.pyro_mlp_loss_grad.argtypes = [vp, vp, C.c_int, C.POINTER(C.c_double), vp]
        L.pyro_gather_environments.argtypes = [vp, C.c_int, C.c_uint32, C.c_int,
                                               C.POINTER(C.c_int)]
        L.pyro_get_environments.argtypes = [vp, vp, C.POINTER(C.c_int)]
        L.pyro_clear_environments.argtypes = [vp]
        L.pyro_mlp_train.argtypes = [vp, C.c_int, C.c_int, C.c_float, C.c_float, vp]
        L.pyro_mlp_adam_step.argtypes = [vp, vp, C.c_float]
        L.pyro_learned_phase.argtypes = [vp, C.c_int, C.c_uint32, C.c_int]
        L.pyro_set_learned_passes.argtypes = [vp, C.c_int]
        L.pyro_acceptance_learned.argtypes = [vp, vp, vp]
        L.pyro_observables_noalloc.argtypes = [C.c_int, vp, C.c_int, C.c_int, C.c_float, vp, vp]
        L.pyro_sync.argtypes = [vp]
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        msg = lib().pyro_last_error().decode(errors="replace")
        raise PyroError(f"pyro error {rc}: {msg}")


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def links16() -> np.ndarray:
    out = np.zeros((16, 6, 4), np.int8)
    _check(lib().pyro_links16(_p(out)))
    return out


def independent_sets() -> np.ndarray:
    out = np.zeros((24, 4), np.int8)
    _check(lib().pyro_independent_sets(_p(out)))
    return out


def philox4x32_10(ctr, key) -> np.ndarray:
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    _check(lib().pyro_philox4x32_10(_p(c), _p(k), _p(out)))
    return out


def fold_order_params(sub12, n_spins: float) -> np.ndarray:
    """Order parameters (R, 12) from sublattice sums (R, 12); host arithmetic only."""
    s = np.ascontiguousarray(sub12, np.float64).reshape(-1, 12)
    out = np.empty_like(s)
    _check(lib().pyro_fold_order_params(_p(s), s.shape[0], float(n_spins), _p(out)))
    return out


def fold_magnetization_g(sub12, gxy: float, gz: float) -> np.ndarray:
    """g-tensor-weighted magnetisation (R, 3) from sublattice sums (R, 12);
    host arithmetic only (MagnetizationGPU, Pyrochlore.cu:277-342)."""
    s = np.ascontiguousarray(sub12, np.float64).reshape(-1, 12)
    out = np.empty((s.shape[0], 3), np.float64)
    _check(lib().pyro_fold_magnetization_g(_p(s), s.shape[0], float(gxy), float(gz), _p(out)))
    return out


def observables_noalloc(device: int, spins_ptr: int, L: int, R: int, jeff: float, out_ptr: int,
                        stream_ptr: int = 0) -> None:
    """Enqueue the energies and sublattice sums (R x 13 doubles at device
    address out_ptr) of R replicas of spins at device address spins_ptr on the
    HIP stream stream_ptr (0: the null stream); allocates nothing."""
    _check(lib().pyro_observables_noalloc(int(device), C.c_void_p(spins_ptr), int(L), int(R),
                                          float(jeff), C.c_void_p(out_ptr),
                                          C.c_void_p(stream_ptr or None)))


# Exchange constants of the reference's defaults (Pyrochlore.cu:139-142).
YB2TI2O7_J = (-0.09 / 0.29, -0.22 / 0.29, -1.0, 0.0)


def exchange_matrices(J1: float, J2: float, J3: float, J4: float) -> np.ndarray:
    """Exchange matrices C[b, b'] (4, 4, 3, 3) of the J1..J4 parameterisation."""
    out = np.zeros(144, np.float64)
    _check(lib().pyro_exchange_matrices(float(J1), float(J2), float(J3), float(J4), _p(out)))
    return out.reshape(4, 4, 3, 3)


def zeeman_fields(B, gxy: float = 1.0, gz: float = 1.0) -> np.ndarray:
    """Zeeman field per basis site (4, 3) in meV for a field B (3,) in tesla."""
    b = np.ascontiguousarray(B, np.float64).reshape(3)
    out = np.zeros(12, np.float64)
    _check(lib().pyro_zeeman_fields(_p(b), float(gxy), float(gz), _p(out)))
    return out.reshape(4, 3)


def mlp_init(seed: int = 0) -> np.ndarray:
    """Initial weights of the learned proposal in the packed layout
    W1[64][6] b1[64] W2[64][64] b2[64] W3[4][64] b3[4]: He-normal hidden layers,
    a small output layer, zero biases except b3[3] = log(kappa) = 1."""
    r = np.random.default_rng(seed & 0xFFFFFFFF)
    w = np.concatenate([
        r.normal(0, np.sqrt(2 / 6), 64 * 6), np.zeros(64),
        r.normal(0, np.sqrt(2 / 64), 64 * 64), np.zeros(64),
        r.normal(0, 0.1 / np.sqrt(64), 4 * 64), np.array([0.0, 0.0, 0.0, 1.0]),
    ]).astype(np.float32)
    assert w.size == MLP_PARAMS
    return w


def geometric_ladder(t_min: float, t_max: float, n: int) -> np.ndarray:
    """n temperatures T_k = t_min (t_max/t_min)^(k/(n-1))."""
    if n == 1:
        return np.array([t_min], np.float64)
    k = np.arange(n, dtype=np.float64)
    return t_min * (t_max / t_min) ** (k / (n - 1))


class Pyrochlore:
    """One GPU's resident ensemble of R replicas of an L^3-cell pyrochlore lattice."""

    def __init__(self, L: int, temperatures, seed: int = 56081376, J: float = 1.0,
                 spin_len2: float = 1.0, n_metro: int = 1, n_or: int = 0, pt_every: int = 0,
                 cos_cone: float = -1.0, device: int = 0):
        self._ctx = C.c_void_p()
        self._temps_init = np.ascontiguousarray(temperatures, np.float64)
        self.L, self.R = int(L), len(self._temps_init)
        self.ncell = self.L ** 3
        self.N = 16 * self.ncell
        self.seed, self.J, self.spin_len2 = int(seed), float(J), float(spin_len2)
        self.n_metro, self.n_or, self.pt_every = int(n_metro), int(n_or), int(pt_every)
        cfg = Config(self.L, self.R,
                     self._temps_init.ctypes.data_as(C.POINTER(C.c_double)),
                     self.seed & 0xFFFFFFFFFFFFFFFF, self.J, self.spin_len2, self.n_metro,
                     self.n_or, self.pt_every, float(cos_cone), int(device))
        _check(lib().pyro_create(C.byref(cfg), C.byref(self._ctx)))

    # -- lifetime -----------------------------------------------------------
    def close(self) -> None:
        if self._ctx:
            lib().pyro_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- spins --------------------------------------------------------------
    @property
    def n_floats(self) -> int:
        return self.R * 48 * self.ncell

    def get_spins(self) -> np.ndarray:
        out = np.empty(self.n_floats, np.float32)
        _check(lib().pyro_get_spins(self._ctx, _p(out)))
        return out

    def set_spins(self, soa) -> None:
        a = np.ascontiguousarray(soa, np.float32).ravel()
        if a.size != self.n_floats:
            raise PyroError(f"set_spins: expected {self.n_floats} floats, got {a.size}")
        _check(lib().pyro_set_spins(self._ctx, _p(a)))

    def GenerateRandomVectorsOnDevice(self) -> None:
        _check(lib().pyro_init_random_spins(self._ctx))

    def device_ptr(self):
        p, n = C.c_void_p(), C.c_size_t()
        _check(lib().pyro_spins_device_ptr(self._ctx, C.byref(p), C.byref(n)))
        return p.value, n.value

    # -- temperatures -------------------------------------------------------
    def temperatures(self):
        T = np.empty(self.R, np.float64)
        at = np.empty(self.R, np.int32)
        _check(lib().pyro_get_temperatures(self._ctx, _p(T), _p(at)))
        return T, at

    def set_temperatures(self, T) -> None:
        a = np.ascontiguousarray(T, np.float64)
        _check(lib().pyro_set_temperatures(self._ctx, _p(a)))

    def cos_cone(self) -> np.ndarray:
        c = np.empty(self.R, np.float32)
        _check(lib().pyro_get_cos_cone(self._ctx, _p(c)))
        return c

    def set_cos_cone(self, c) -> None:
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(c, np.float32), (self.R,)))
        _check(lib().pyro_set_cos_cone(self._ctx, _p(a)))

    # -- the step -----------------------------------------------------------
    def MCMultipleTemperaturesPerSystem(self, colour: int, sweep: int, pass_: int = 0) -> None:
        _check(lib().pyro_metropolis_phase(self._ctx, colour, sweep, pass_))

    def OverRelaxationByPiStep(self, colour: int) -> None:
        _check(lib().pyro_overrelax_phase(self._ctx, colour))

    def Mc(self, steps: int, use_graph: bool = True) -> None:
        _check(lib().pyro_sweep(self._ctx, int(steps), 1 if use_graph else 0))

    def sweep_counters(self):
        s, r = C.c_uint32(), C.c_uint32()
        _check(lib().pyro_get_sweep(self._ctx, C.byref(s), C.byref(r)))
        return s.value, r.value

    def graph_block(self) -> int:
        """Sweeps per captured graph: Mc replays one graph per whole block that
        starts at a multiple of it."""
        n = C.c_int()
        _check(lib().pyro_get_graph_block(self._ctx, C.byref(n)))
        return n.value

    def set_sweep_counters(self, sweep: int, pt_round: int) -> None:
        _check(lib().pyro_set_sweep(self._ctx, sweep, pt_round))

    def TrySwapingTemperatures(self, parity: int, round_: int) -> int:
        n = C.c_int()
        _check(lib().pyro_pt_swap(self._ctx, parity, round_, C.byref(n)))
        return n.value

    # -- observables --------------------------------------------------------
    def _vec(self, fn, width):
        out = np.empty(self.R * width, np.float64)
        _check(fn(self._ctx, _p(out)))
        return out.reshape(self.R, width) if width > 1 else out

    def TotalEnergyv2(self) -> np.ndarray:
        return self._vec(lib().pyro_energy, 1)

    def energy_incremental(self) -> np.ndarray:
        return self._vec(lib().pyro_energy_incremental, 1)

    def Magnetization(self) -> np.ndarray:
        return self._vec(lib().pyro_magnetization, 3)

    def MagnetizationG(self, gxy: float, gz: float) -> np.ndarray:
        """Sum over spins of g_b s_i, g_b the g-tensor of the spin's basis
        site (MagnetizationGPU, Pyrochlore.cu:277-342), (R, 3)."""
        out = np.empty(self.R * 3, np.float64)
        _check(lib().pyro_magnetization_g(self._ctx, float(gxy), float(gz), _p(out)))
        return out.reshape(self.R, 3)

    def DoMeasurement(self) -> np.ndarray:
        return self._vec(lib().pyro_order_params, 12)

    def sublattice_sums(self) -> np.ndarray:
        return self._vec(lib().pyro_sublattice_sums, 12)

    def acceptance(self):
        a = np.empty(self.R, np.uint64)
        p = np.empty(self.R, np.uint64)
        _check(lib().pyro_acceptance(self._ctx, _p(a), _p(p)))
        return a, p

    def reset_acceptance(self) -> None:
        _check(lib().pyro_reset_acceptance(self._ctx))

    def CalculateRatio(self) -> np.ndarray:
        a, p = self.acceptance()
        return a / np.maximum(p, 1)

    def CalculateSolidAngle(self, target: float = 0.5, rounds: int = 20,
                            sweeps_per_round: int = 10) -> np.ndarray:
        _check(lib().pyro_adapt_cone(self._ctx, target, rounds, sweeps_per_round))
        return self.cos_cone()

    # -- energy series ------------------------------------------------------
    def record_start(self, capacity: int) -> None:
        _check(lib().pyro_record_start(self._ctx, int(capacity)))

    def record_fetch(self) -> np.ndarray:
        n = C.c_int()
        _check(lib().pyro_record_fetch(self._ctx, None, C.byref(n)))
        out = np.empty(max(n.value, 1) * self.R, np.float64)
        _check(lib().pyro_record_fetch(self._ctx, _p(out), C.byref(n)))
        return out[: n.value * self.R].reshape(n.value, self.R)

    def set_couplings(self, C=None, B_site=None) -> None:
        """Matrix field model: C (4, 4, 3, 3) exchange matrices, B_site (4, 3)
        Zeeman fields (or None); C=None returns to the scalar J model."""
        if C is None:
            _check(lib().pyro_set_couplings(self._ctx, None, None))
            return
        c = np.ascontiguousarray(C, np.float64).reshape(144)
        b = None if B_site is None else np.ascontiguousarray(B_site, np.float64).reshape(12)
        _check(lib().pyro_set_couplings(self._ctx, _p(c), None if b is None else _p(b)))

    def set_streams(self, n: int) -> None:
        """1 to 4 streams, one replica group each, for the phases of a sweep
        (identical results)."""
        _check(lib().pyro_set_streams(self._ctx, int(n)))

    def set_or_fusion(self, level) -> None:
        """Over-relaxation colours per launch, identical results: 0 (or False)
        one, 1 (or True) two, 2 (default) a whole pass where a kernel for it
        exists (L = 16 or 64, untracked over-relaxation), else two."""
        _check(lib().pyro_set_or_fusion(self._ctx, int(level)))

    def streams(self) -> int:
        n = C.c_int()
        _check(lib().pyro_get_streams(self._ctx, C.byref(n)))
        return n.value

    def set_group_blocking(self, groups: int) -> None:
        """Cache-blocked sweeps: each exchange interval runs replica group after
        group (0 or 1: off). Identical results."""
        _check(lib().pyro_set_group_blocking(self._ctx, int(groups)))

    def group_blocking(self) -> int:
        n = C.c_int()
        _check(lib().pyro_get_group_blocking(self._ctx, C.byref(n)))
        return n.value

    def check_spins(self, tol: float = 1e-5):
        """(number of non-finite or non-unit spins, first such index or -1)."""
        n, first = C.c_longlong(), C.c_longlong()
        _check(lib().pyro_check_spins(self._ctx, float(tol), C.byref(n), C.byref(first)))
        return n.value, first.value

    def set_post_rotate(self, on: bool) -> None:
        """Rotate each spin about its field after the Metropolis test (Math.cu:168-197)."""
        _check(lib().pyro_set_post_rotate(self._ctx, 1 if on else 0))

    def set_update_schedule(self, independent_sets=0, or_random_angle: bool = False) -> None:
        """independent_sets 1 (or True): phases on independent sets drawn from
        the 24 (Pyrochlore.cu:848, IndependentSubLattices[Update[0] % 24]); 2:
        the colour classes of each pass in a drawn order; 0 (or False): the
        colour classes in order. or_random_angle: over-relaxation by a random
        angle about the field (OverRelaxationStep, Pyrochlore.cu:1009-1045)
        instead of by pi."""
        _check(lib().pyro_set_update_schedule(self._ctx, int(independent_sets),
                                              1 if or_random_angle else 0))

    def couplings(self):
        """(is_matrix, the 156 FP32 values the kernels use)."""
        m = C.c_int()
        out = np.zeros(156, np.float32)
        _check(lib().pyro_get_couplings(self._ctx, C.byref(m), _p(out)))
        return bool(m.value), out

    def SimulatedAnnealing(self, steps: int, sweeps_per_step: int,
                           T_start: float = 20.0) -> None:
        """Linear schedule from T_start toward each replica's temperature
        (MonteCarlo.cu:68-104,295-344), no exchanges meanwhile."""
        _check(lib().pyro_anneal(self._ctx, float(T_start), int(steps), int(sweeps_per_step)))

    def measure_start(self, every: int, capacity: int) -> None:
        """Append a measurement row per rung every `every` sweeps, on the device."""
        _check(lib().pyro_measure_start(self._ctx, int(every), int(capacity)))

    def measure_stop(self) -> None:
        _check(lib().pyro_measure_stop(self._ctx))

    def measure_fetch(self) -> np.ndarray:
        """Rows so far, shape (n, R, MEAS_Q), rungs by descending T."""
        n = C.c_int()
        _check(lib().pyro_measure_fetch(self._ctx, None, C.byref(n)))
        out = np.empty(max(n.value, 1) * self.R * MEAS_Q, np.float64)
        _check(lib().pyro_measure_fetch(self._ctx, _p(out), C.byref(n)))
        return out[: n.value * self.R * MEAS_Q].reshape(n.value, self.R, MEAS_Q)

    def set_kernel_options(self, phase_variant: int = 1, track_or_energy: bool = False) -> None:
        _check(lib().pyro_set_kernel_options(self._ctx, phase_variant, 1 if track_or_energy else 0))

    def set_exec_mode(self, mode: int, block: int = None, n_bins: int = 0) -> None:
        """0: per-phase kernels; 1: one persistent kernel per sweep (block:
        256/512/1024 threads, default 1024); 2: tile-list phases (block: cells
        per tile, 128/256/512, default 256)."""
        if block is None:
            block = 1024 if mode == 1 else (256 if mode == 2 else 0)
        _check(lib().pyro_set_exec_mode(self._ctx, mode, block, n_bins))

    def exec_mode(self):
        m, b, n = C.c_int(), C.c_int(), C.c_int()
        _check(lib().pyro_get_exec_mode(self._ctx, C.byref(m), C.byref(b), C.byref(n)))
        return m.value, b.value, n.value

    AUTOTUNE_MAX = 64  # PYRO_AUTOTUNE_MAX

    def autotune(self, sweeps: int = 5) -> np.ndarray:
        """Try every execution, keep the fastest; rows (mode, block, variant,
        streams, cache-blocking groups, ms/sweep)."""
        res = np.zeros(6 * self.AUTOTUNE_MAX, np.float64)
        n = C.c_int()
        _check(lib().pyro_autotune(self._ctx, sweeps, _p(res), C.byref(n)))
        return res[: 6 * n.value].reshape(n.value, 6)

    # -- checkpoint / configuration files ----------------------------------
    def save_checkpoint(self, path: str) -> None:
        _check(lib().pyro_checkpoint_save(self._ctx, os.fsencode(path)))

    def load_checkpoint(self, path: str) -> None:
        _check(lib().pyro_checkpoint_load(self._ctx, os.fsencode(path)))

    def SaveConfigurations(self, prefix: str, append: bool = True, bytes_per_value: int = 8):
        """Reference files <prefix><k>.dat (k = temperature rank), AoS records (Math.cu:618)."""
        _check(lib().pyro_export_configurations(self._ctx, os.fsencode(prefix), int(append),
                                                bytes_per_value))

    def ReadConfigurations(self, prefix: str, rep: int = 0, bytes_per_value: int = 8):
        """Load record `rep` of the reference files (Math.cu:662)."""
        _check(lib().pyro_import_configurations(self._ctx, os.fsencode(prefix), rep,
                                                bytes_per_value))

    def layout(self):
        """(L per replica, float offset of each replica's spins in get_spins())."""
        L = np.empty(self.R, np.int32)
        off = np.empty(self.R, np.int64)
        _check(lib().pyro_get_layout(self._ctx, _p(L), _p(off)))
        return L, off

    def set_passes(self, n_metro: int, n_or: int) -> None:
        _check(lib().pyro_set_passes(self._ctx, n_metro, n_or))
        self.n_metro, self.n_or = n_metro, n_or

    # -- learned proposal ---------------------------------------------------
    def mlp_set_weights(self, w) -> None:
        a = np.ascontiguousarray(w, np.float32)
        if a.size != MLP_PARAMS:
            raise PyroError(f"mlp_set_weights: expected {MLP_PARAMS} floats, got {a.size}")
        _check(lib().pyro_mlp_set_weights(self._ctx, _p(a)))

    def mlp_get_weights(self) -> np.ndarray:
        w = np.empty(MLP_PARAMS, np.float32)
        _check(lib().pyro_mlp_get_weights(self._ctx, _p(w)))
        return w

    def mlp_eval(self, rows9) -> np.ndarray:
        a = np.ascontiguousarray(rows9, np.float32).reshape(-1, 9)
        out = np.empty_like(a)
        _check(lib().pyro_mlp_eval(self._ctx, _p(a), a.shape[0], _p(out)))
        return out

    def vmf_draw(self, rows4) -> np.ndarray:
        a = np.ascontiguousarray(rows4, np.float32).reshape(-1, 4)
        out = np.empty((a.shape[0], 3), np.float32)
        _check(lib().pyro_vmf_draw(self._ctx, _p(a), a.shape[0], _p(out)))
        return out

    def mlp_loss_grad(self, env6):
        a = np.ascontiguousarray(env6, np.float32).reshape(-1, 6)
        g = np.empty(MLP_PARAMS, np.float32)
        loss = C.c_double()
        _check(lib().pyro_mlp_loss_grad(self._ctx, _p(a), a.shape[0], C.byref(loss), _p(g)))
        return loss.value, g

    def gather_environments(self, per_replica: int, draw: int, capacity: int) -> int:
        n = C.c_int()
        _check(lib().pyro_gather_environments(self._ctx, per_replica, draw, capacity, C.byref(n)))
        return n.value

    def environments(self) -> np.ndarray:
        n = C.c_int()
        _check(lib().pyro_get_environments(self._ctx, None, C.byref(n)))
        out = np.empty((max(n.value, 1), 6), np.float32)
        _check(lib().pyro_get_environments(self._ctx, _p(out), C.byref(n)))
        return out[: n.value]

    def clear_environments(self) -> None:
        _check(lib().pyro_clear_environments(self._ctx))

    def mlp_adam_step(self, grad, lr: float) -> None:
        g = np.ascontiguousarray(grad, np.float32)
        _check(lib().pyro_mlp_adam_step(self._ctx, _p(g), float(lr)))

    def mlp_train(self, steps: int, batch: int = 4096, lr: float = 3e-3,
                  lr_final: float = 3e-4) -> np.ndarray:
        hist = np.zeros(max(steps, 1), np.float32)
        _check(lib().pyro_mlp_train(self._ctx, steps, batch, lr, lr_final, _p(hist)))
        return hist[:steps]

    def learned_phase(self, colour: int, sweep: int, pass_: int = 0) -> None:
        _check(lib().pyro_learned_phase(self._ctx, colour, sweep, pass_))

    def set_learned_passes(self, n: int) -> None:
        _check(lib().pyro_set_learned_passes(self._ctx, n))

    def acceptance_learned(self):
        a = np.empty(self.R, np.uint64)
        p = np.empty(self.R, np.uint64)
        _check(lib().pyro_acceptance_learned(self._ctx, _p(a), _p(p)))
        return a, p

    def time_phase(self, kind: int, n_launches: int):
        """(mean ms per launch, first-to-last ms) of n back-to-back phase launches
        (kind 0 Metropolis, 1 over-relaxation, 2 learned, 3 fused over-relaxation
        pair, 4 fused Metropolis pair, 5 whole over-relaxation pass), timed with
        HIP events on the ctx stream."""
        avg, span = C.c_double(), C.c_double()
        _check(lib().pyro_time_phase(self._ctx, kind, n_launches, C.byref(avg), C.byref(span)))
        return avg.value, span.value

    def sync(self) -> None:
        _check(lib().pyro_sync(self._ctx))


class Batch(Pyrochlore):
    """Replicas of different lattice sizes in one ctx (sizes in {4,8,12,16,24,32}).

    Replica r's spins are the 48*L[r]^3 floats at layout()[1][r] of get_spins();
    exchanges happen within groups of equal size. Sweeps run as one persistent
    kernel per sweep (mode 1) or as tile-list phases (mode 2, the default);
    autotune() picks the faster.
    """

    def __init__(self, L, temperatures, seed: int = 56081376, J: float = 1.0,
                 spin_len2: float = 1.0, n_metro: int = 1, n_or: int = 0, pt_every: int = 0,
                 cos_cone: float = -1.0, device: int = 0):
        self._ctx = C.c_void_p()
        self._L = np.ascontiguousarray(L, np.int32)
        self._temps_init = np.ascontiguousarray(temperatures, np.float64)
        if self._L.size != self._temps_init.size:
            raise PyroError("Batch: one temperature per replica")
        self.R = int(self._L.size)
        self.L = None
        self.N = 16 * self._L.astype(np.int64) ** 3
        self.seed, self.J, self.spin_len2 = int(seed), float(J), float(spin_len2)
        self.n_metro, self.n_or, self.pt_every = int(n_metro), int(n_or), int(pt_every)
        cfg = BatchConfig(self.R, self._L.ctypes.data_as(C.POINTER(C.c_int)),
                          self._temps_init.ctypes.data_as(C.POINTER(C.c_double)),
                          self.seed & 0xFFFFFFFFFFFFFFFF, self.J, self.spin_len2, self.n_metro,
                          self.n_or, self.pt_every, float(cos_cone), int(device))
        _check(lib().pyro_create_batch(C.byref(cfg), C.byref(self._ctx)))

    @property
    def n_floats(self) -> int:
        return int(3 * self.N.sum())
