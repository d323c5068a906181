"""ctypes binding of the CPU oracle (oracle/libpyro_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker. The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpyro_oracle.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_links16.argtypes = [_i8p]
        L.orc_independent_sets.argtypes = [_i8p]
        L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.orc_exp_neg.argtypes = [C.c_float]
        L.orc_exp_neg.restype = C.c_float
        L.orc_sincospi.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.orc_cone_proposal.argtypes = [_f32p, C.c_float, C.c_float, C.c_float, _f32p]
        L.orc_overrelax.argtypes = [_f32p, _f32p, _f32p]
        L.orc_renormalise.argtypes = [_f32p, _f32p]
        L.orc_metropolis_accept.argtypes = [_f32p, _f32p, _f32p, C.c_float, C.c_float]
        L.orc_metropolis_accept.restype = C.c_int
        L.orc_init_spins.argtypes = [_f32p, C.c_int, C.c_int, C.c_uint64]
        L.orc_metropolis_phase.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_int,
                                           _f32p, _f32p, C.c_float, C.c_uint64, C.c_void_p,
                                           C.c_void_p]
        L.orc_overrelax_phase.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_void_p]
        L.orc_observables.argtypes = [_f32p, C.c_int, C.c_int, C.c_float, _f64p]
        L.orc_pt_swap.argtypes = [C.c_int, C.c_int, C.c_uint32, C.c_uint64, _f64p, _f64p, _f32p,
                                  _f32p, _i32p]
        L.orc_pt_swap.restype = C.c_int
        L.orc_run.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_int, C.c_int,
                              C.c_int, C.POINTER(C.c_uint32), _f64p, _f32p, _f32p, _i32p,
                              C.c_float, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p,
                              C.POINTER(C.c_int)]
        L.orc_set_update_order.argtypes = [C.c_void_p]
        L.orc_set_post_rotate.argtypes = [C.c_int]
        L.orc_set_update_schedule.argtypes = [C.c_int, C.c_int]
        L.orc_drawn_set.argtypes = [C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_uint64]
        L.orc_drawn_set.restype = C.c_int
        L.orc_drawn_colour.argtypes = [C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_uint64]
        L.orc_drawn_colour.restype = C.c_int
        L.orc_overrelax_phase_ex.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_int,
                                             C.c_float, C.c_uint64, C.c_void_p]
        L.orc_rotate_about.argtypes = [_f32p, _f32p, C.c_float, _f32p]
        L.orc_set_couplings.argtypes = [C.c_void_p]
        L.orc_magnetization_g.argtypes = [_f64p, C.c_double, C.c_double, _f64p]
        L.orc_adapt_cone_step.argtypes = [C.c_int, C.c_double, _u64p, _u64p, _f32p]
        L.orc_num_threads.restype = C.c_int
        L.orc_set_num_threads.argtypes = [C.c_int]
        L.orc_set_replica_base.argtypes = [C.c_int]
        L.orc_set_swap_group.argtypes = [C.c_int]
        L.orc_mlp_nparams.restype = C.c_int
        L.orc_mlp_features.argtypes = [_f64p, _f64p, _f64p]
        L.orc_mlp_forward.argtypes = [_f64p, _f64p, _f64p]
        L.orc_mlp_decode.argtypes = [_f64p, _f64p, C.POINTER(C.c_double)]
        L.orc_mlp_loss_grad.argtypes = [_f64p, _f64p, C.c_int, C.c_void_p]
        L.orc_mlp_loss_grad.restype = C.c_double
        L.orc_vmf_log_norm.argtypes = [C.c_double]
        L.orc_vmf_log_norm.restype = C.c_double
        L.orc_vmf_log_density.argtypes = [_f64p, C.c_double, _f64p]
        L.orc_vmf_log_density.restype = C.c_double
        L.orc_vmf_sample.argtypes = [_f64p, C.c_double, C.c_double, C.c_double, _f64p]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def links16() -> np.ndarray:
    out = np.zeros((16, 6, 4), np.int8)
    lib().orc_links16(out)
    return out


def independent_sets() -> np.ndarray:
    out = np.zeros((24, 4), np.int8)
    lib().orc_independent_sets(out)
    return out


def philox(ctr, key) -> np.ndarray:
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def exp_neg(x: float) -> float:
    return float(lib().orc_exp_neg(x))


def sincospi(a: float):
    s, c = C.c_float(), C.c_float()
    lib().orc_sincospi(a, C.byref(s), C.byref(c))
    return s.value, c.value


def cone_proposal(s, cos_cone, u_cos, u_phi) -> np.ndarray:
    out = np.zeros(3, np.float32)
    lib().orc_cone_proposal(np.asarray(s, np.float32), cos_cone, u_cos, u_phi, out)
    return out


def overrelax(h, s) -> np.ndarray:
    out = np.zeros(3, np.float32)
    lib().orc_overrelax(np.asarray(h, np.float32), np.asarray(s, np.float32), out)
    return out


def renormalise(v) -> np.ndarray:
    out = np.zeros(3, np.float32)
    lib().orc_renormalise(np.ascontiguousarray(v, np.float32), out)
    return out


def metropolis_accept(h, s, sp, beta, u) -> bool:
    return bool(lib().orc_metropolis_accept(np.asarray(h, np.float32), np.asarray(s, np.float32),
                                            np.asarray(sp, np.float32), beta, u))


def init_spins(L: int, R: int, seed: int) -> np.ndarray:
    s = np.zeros(R * 48 * L ** 3, np.float32)
    lib().orc_init_spins(s, L, R, seed)
    return s


def metropolis_phase(spins, L, R, colour, sweep, pass_, beta, cosc, jeff, seed, acc=None, dE=None):
    lib().orc_metropolis_phase(spins, L, R, colour, sweep, pass_, np.asarray(beta, np.float32),
                               np.asarray(cosc, np.float32), jeff, seed, _ptr(acc), _ptr(dE))


def overrelax_phase(spins, L, R, colour, jeff, dE=None):
    lib().orc_overrelax_phase(spins, L, R, colour, jeff, _ptr(dE))


def observables(spins, L, R, jeff) -> np.ndarray:
    out = np.zeros(R * 13, np.float64)
    lib().orc_observables(spins, L, R, jeff, out)
    return out.reshape(R, 13)


def pt_swap(R, parity, rnd, seed, energy, temps, beta, cosc, replica_at) -> int:
    return lib().orc_pt_swap(R, parity, rnd, seed, np.ascontiguousarray(energy, np.float64), temps,
                             beta, cosc, replica_at)


class OracleRun:
    """Whole-run state mirrored from the GPU ctx (spins, temperatures, rungs)."""

    def __init__(self, L, temps, seed, J=1.0, spin_len2=1.0, n_metro=1, n_or=0, pt_every=0,
                 cos_cone=-1.0):
        self.L, self.R = L, len(temps)
        self.seed = seed
        self.jeff = np.float32(J * spin_len2)
        self.n_metro, self.n_or, self.pt_every = n_metro, n_or, pt_every
        self.temps = np.array(temps, np.float64)
        self.beta = (1.0 / self.temps).astype(np.float32)
        self.cosc = np.full(self.R, cos_cone, np.float32)
        self.replica_at = np.argsort(-self.temps, kind="stable").astype(np.int32)
        self.spins = init_spins(L, self.R, seed)
        self.sweep = 0
        self.pt_round = C.c_uint32(0)
        self.accepted = np.zeros(self.R, np.uint64)
        self.dE = np.zeros(self.R, np.float64)
        self.swaps = C.c_int(0)
        self.anchor = observables(self.spins, L, self.R, self.jeff)[:, 0].copy()

    def run(self, n):
        lib().orc_run(self.spins, self.L, self.R, n, self.sweep, self.n_metro, self.n_or,
                      self.pt_every, C.byref(self.pt_round), self.temps, self.beta, self.cosc,
                      self.replica_at, self.jeff, self.seed, _ptr(self.accepted), _ptr(self.dE),
                      _ptr(self.anchor), C.byref(self.swaps))
        self.sweep += n

    def observables(self):
        return observables(self.spins, self.L, self.R, self.jeff)

    def adapt_cone(self, target, rounds, sweeps_per_round):
        """Cone adaptation: rounds of sweeps_per_round sweeps, each followed by
        one adapt_cone_step on the acceptance of that round."""
        per_sweep = np.uint64(self.n_metro * 16 * self.L ** 3)
        for _ in range(rounds):
            self.accepted[:] = 0
            self.run(sweeps_per_round)
            proposed = np.full(self.R, per_sweep * np.uint64(sweeps_per_round), np.uint64)
            adapt_cone_step(target, self.accepted, proposed, self.cosc)
        self.accepted[:] = 0


_order_keep = None


def set_update_order(perm) -> None:
    """Visit the updates of every colour phase in this order (cell * 4 + k), or
    naturally for None."""
    global _order_keep
    if perm is None:
        _order_keep = None
        lib().orc_set_update_order(None)
        return
    _order_keep = np.ascontiguousarray(perm, np.int64).copy()
    lib().orc_set_update_order(_order_keep.ctypes.data_as(C.c_void_p))


def set_post_rotate(on: bool) -> None:
    lib().orc_set_post_rotate(1 if on else 0)


def set_update_schedule(independent_sets=0, or_random_angle: bool = False) -> None:
    """Every following phase: independent_sets 1 (or True) draws each phase's
    set from the 24, 2 runs the colour classes in a drawn order per pass, 0
    (or False) the colour classes in order; over-relaxation by a random angle
    about h (else by pi)."""
    lib().orc_set_update_schedule(int(independent_sets), 1 if or_random_angle else 0)


def drawn_colour(sweep, stage, pass_, phase, seed) -> int:
    """Colour of a phase under the drawn colour order (independent_sets=2)."""
    return lib().orc_drawn_colour(sweep, stage, pass_, phase, seed)


def drawn_set(sweep, stage, pass_, phase, seed) -> int:
    """Index (0..23) of the independent set of a phase under the drawn-set
    schedule; stage 0 = Metropolis, 1 = over-relaxation."""
    return lib().orc_drawn_set(sweep, stage, pass_, phase, seed)


def overrelax_phase_ex(spins, L, R, phase, sweep, pass_, jeff, seed, dE=None) -> None:
    lib().orc_overrelax_phase_ex(spins, L, R, phase, sweep, pass_, jeff, seed, _ptr(dE))


def rotate_about(h, s, u) -> np.ndarray:
    out = np.zeros(3, np.float32)
    lib().orc_rotate_about(np.ascontiguousarray(h, np.float32), np.ascontiguousarray(s, np.float32),
                           float(u), out)
    return out


_coup_keep = None


def set_couplings(coup156) -> None:
    """Matrix field model (156 float32: spin_len2 C, then 2 spin_len2 B) for
    the following oracle calls, or None for the scalar model."""
    global _coup_keep
    if coup156 is None:
        _coup_keep = None
        lib().orc_set_couplings(None)
        return
    _coup_keep = np.ascontiguousarray(coup156, np.float32).copy()
    assert _coup_keep.size == 156
    lib().orc_set_couplings(_coup_keep.ctypes.data_as(C.c_void_p))


def magnetization_g(sub12, gxy: float, gz: float) -> np.ndarray:
    """g-tensor-weighted magnetisation (3,) from one replica's 12 sublattice sums."""
    out = np.zeros(3, np.float64)
    lib().orc_magnetization_g(np.ascontiguousarray(sub12, np.float64).reshape(12), float(gxy),
                              float(gz), out)
    return out


def adapt_cone_step(target: float, accepted, proposed, cos_cone) -> None:
    """One round of the cone rule on cos_cone (float32, updated in place)."""
    assert cos_cone.dtype == np.float32 and cos_cone.flags.c_contiguous
    lib().orc_adapt_cone_step(cos_cone.size, float(target),
                              np.ascontiguousarray(accepted, np.uint64),
                              np.ascontiguousarray(proposed, np.uint64), cos_cone)


def set_replica_base(base: int) -> None:
    lib().orc_set_replica_base(base)


def set_swap_group(group: int) -> None:
    lib().orc_set_swap_group(group)


def num_threads() -> int:
    return lib().orc_num_threads()


def set_num_threads(n: int) -> None:
    lib().orc_set_num_threads(n)


# -- learned proposal -------------------------------------------------------
def mlp_nparams() -> int:
    return lib().orc_mlp_nparams()


def mlp_features(u, s) -> np.ndarray:
    x = np.zeros(6)
    lib().orc_mlp_features(np.ascontiguousarray(u, np.float64), np.ascontiguousarray(s, np.float64), x)
    return x


def mlp_forward(w, x) -> np.ndarray:
    o = np.zeros(4)
    lib().orc_mlp_forward(np.ascontiguousarray(w, np.float64), np.ascontiguousarray(x, np.float64), o)
    return o


def mlp_decode(o):
    mu = np.zeros(3)
    k = C.c_double()
    lib().orc_mlp_decode(np.ascontiguousarray(o, np.float64), mu, C.byref(k))
    return mu, k.value


def mlp_loss_grad(w, env, with_grad=True):
    """Mean reverse KL over env rows (u, s) and its gradient (FP64)."""
    env = np.ascontiguousarray(env, np.float64).reshape(-1, 6)
    g = np.zeros(mlp_nparams()) if with_grad else None
    loss = lib().orc_mlp_loss_grad(np.ascontiguousarray(w, np.float64), env, env.shape[0],
                                   None if g is None else g.ctypes.data_as(C.c_void_p))
    return loss, g


def vmf_log_norm(k) -> float:
    return lib().orc_vmf_log_norm(k)


def vmf_log_density(mu, k, y) -> float:
    return lib().orc_vmf_log_density(np.ascontiguousarray(mu, np.float64), k,
                                     np.ascontiguousarray(y, np.float64))


def vmf_sample(mu, k, u1, u2) -> np.ndarray:
    out = np.zeros(3)
    lib().orc_vmf_sample(np.ascontiguousarray(mu, np.float64), k, u1, u2, out)
    return out


def mlp_init(seed: int = 0) -> np.ndarray:
    """He-style initial weights in the packed layout (W1 b1 W2 b2 W3 b3)."""
    r = np.random.default_rng(seed)
    w = np.concatenate([
        r.normal(0, np.sqrt(2 / 6), 64 * 6), np.zeros(64),
        r.normal(0, np.sqrt(2 / 64), 64 * 64), np.zeros(64),
        r.normal(0, np.sqrt(1 / 64), 4 * 64), np.zeros(4),
    ])
    assert w.size == mlp_nparams()
    return w
