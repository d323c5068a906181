// gfx950 kernels of the pyrochlore Heisenberg Monte Carlo step: the scalar
// (Heisenberg J) colour phases and observables from pyro_phase.h, spin
// initialisation, the observables' second stage, parallel-tempering
// exchanges, the per-sweep bookkeeping and the measurement rows.
#include <mutex>

#include "pyro_phase.h"

namespace pyro {

// Uniform initial spins (Generate3DVectorList, Math.cu:132-143, written SoA).
// grid = (ceil(ncell/256), R), block = 256, each thread one cell x 16 sites.
__global__ __launch_bounds__(256) void k_init_spins(float *__restrict__ spins, int ncell,
                                                    uint32_t key0, uint32_t key1) {
  const int r = blockIdx.y;
  const int cell = blockIdx.x * 256 + threadIdx.x;
  if (cell >= ncell) return;
  const size_t plane = (size_t)16 * ncell;
  float *Sx = spins + (size_t)r * 3 * plane;
  for (int site = 0; site < 16; ++site) {
    const u32x4 w = philox4x32_10(
        rng_counter((uint32_t)cell, (uint32_t)site, 0, kKindInit, (uint32_t)r, 0), key0, key1);
    const f3 v = uniform_sphere(u01(w.x), u01(w.y));
    const size_t i = (size_t)site * ncell + cell;
    Sx[i] = v.x;
    Sx[plane + i] = v.y;
    Sx[2 * plane + i] = v.z;
  }
}

// Fixed-order second stage: obs[r][0] = escale * sum over b and blocks of the
// energy partials (escale = jeff/2 for the scalar model, each bond having been
// visited from both ends; -1 for the matrix model); obs[r][1 + 3b + c] = sum
// over blocks of component c of sublattice b. grid = R, block = 256.
__global__ __launch_bounds__(256) void k_observables_final(const double *__restrict__ partials,
                                                           int nblk, double escale,
                                                           double *__restrict__ obs,
                                                           double *__restrict__ anchor,
                                                           double *__restrict__ dE) {
  const int r = blockIdx.x;
  const double *p = partials + (size_t)r * 4 * nblk * 4;
  double v[kObs];
#pragma unroll
  for (int q = 0; q < kObs; ++q) v[q] = 0.0;
  // Each thread sums its strided share of every sublattice's block partials
  // (rows b * nblk + k), then the block reduces all 13 sums at once: wave
  // butterflies and the 4 waves in order, a fixed order throughout.
#pragma unroll
  for (int b = 0; b < 4; ++b)
    for (int k = threadIdx.x; k < nblk; k += 256) {
      const double *row = p + ((size_t)b * nblk + k) * 4;
      v[0] += row[0];
      v[1 + 3 * b] += row[1];
      v[2 + 3 * b] += row[2];
      v[3 + 3 * b] += row[3];
    }
  __shared__ double s_v[4][kObs];
  const int wv = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < kObs; ++q) {
    const double t = wave_sum(v[q]);
    if ((threadIdx.x & 63) == 0) s_v[wv][q] = t;
  }
  __syncthreads();
  if (threadIdx.x < kObs) {
    const int q = threadIdx.x;
    const double t = (s_v[0][q] + s_v[1][q]) + (s_v[2][q] + s_v[3][q]);
    const double val = (q == 0) ? escale * t : t;
    obs[(size_t)r * kObs + q] = val;
    if (q == 0 && anchor != nullptr) {  // re-anchor the incremental energy here
      anchor[r] = val;
      dE[r] = 0.0;
    }
  }
}

// End of n sweeps (n = 1, or the sweeps of an interval run without a
// per-sweep booking): when recording, the row of the last of them
// (incremental energy per rung) is written here if record_last (the rows of
// the earlier ones were written by the phase kernels) and the row count
// advances by n up to the capacity; the sweep counter advances by n. Called by
// every thread of one block.
__device__ void end_interval(int R, uint32_t *__restrict__ ctl, const double *__restrict__ anchor,
                             const double *__restrict__ dE, const int *__restrict__ replica_at,
                             double *__restrict__ series, uint32_t n, int record_last) {
  const uint32_t on = ctl[kCtlRecordOn];
  const uint32_t row = ctl[kCtlRecordRow];
  const uint32_t cap = ctl[kCtlRecordCap];
  const uint32_t last = row + n - 1u;
  const bool rec = on && series != nullptr && record_last && last < cap;
  if (rec) {
    for (int k = threadIdx.x; k < R; k += blockDim.x) {
      const int rep = replica_at[k];
      series[(size_t)last * R + k] = anchor[rep] + dE[rep];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (on && series != nullptr) ctl[kCtlRecordRow] = row + n < cap ? row + n : (row < cap ? cap : row);
    ctl[kCtlSweep] = ctl[kCtlSweep] + n;
  }
}

// Parallel-tempering exchange (TrySwapingTemperatures, MonteCarlo.cu:105-172)
// on the energies in obs. Rungs are ordered by descending temperature; pairs
// (c, c+1), c = parity, parity+2, ... are disjoint, so each thread decides its
// own pairs (the reference visits them in order; no pair reads another's
// result, so the outcome is the same). A batch of mixed lattice sizes
// exchanges within groups of equal size: R is the group's rung count,
// replica_at its rungs, `group` keys its draws, and the exchange-round
// counter advances after the last group. With end_n > 0 the last group's
// launch also ends the interval (end_interval over all R_all rungs, after the
// exchange). grid = 1, block = 256.
__global__ __launch_bounds__(256) void k_pt_swap(
    int R, const double *__restrict__ obs, double *__restrict__ temps, float *__restrict__ beta,
    float *__restrict__ cosc, int *__restrict__ replica_at, uint32_t *__restrict__ ctl,
    uint32_t key0, uint32_t key1, int parity, uint32_t round, int from_ctl, int group,
    int last_group, uint32_t end_n, int R_all, const int *__restrict__ replica_at_all,
    const double *__restrict__ anchor, const double *__restrict__ dE,
    double *__restrict__ series) {
  __shared__ uint32_t s_round;
  __shared__ int s_nacc;
  if (threadIdx.x == 0) {
    s_round = from_ctl ? ctl[kCtlPtRound] : round;
    s_nacc = 0;
  }
  __syncthreads();
  const uint32_t rnd = s_round;
  const int par = from_ctl ? (int)(rnd & 1u) : parity;
  int nacc = 0;
  for (int c = par + 2 * (int)threadIdx.x; c + 1 < R; c += 2 * (int)blockDim.x) {
    const int t1 = replica_at[c], t2 = replica_at[c + 1];
    const double T1 = temps[t1], T2 = temps[t2];
    const double arg = (1.0 / T1 - 1.0 / T2) * (obs[(size_t)t1 * kObs] - obs[(size_t)t2 * kObs]);
    const u32x4 w = philox4x32_10(
        rng_counter((uint32_t)c, 0, 0, kKindSwap, (uint32_t)group, rnd), key0, key1);
    const double u = u01_53(w.x, w.y);
    if (arg >= 0.0 || u < exp(arg)) {
      temps[t1] = T2;
      temps[t2] = T1;
      const float b = beta[t1];
      beta[t1] = beta[t2];
      beta[t2] = b;
      const float cc = cosc[t1];
      cosc[t1] = cosc[t2];
      cosc[t2] = cc;
      replica_at[c] = t2;
      replica_at[c + 1] = t1;
      ++nacc;
    }
  }
  if (nacc) atomicAdd(&s_nacc, nacc);
  __syncthreads();  // also makes every thread's replica_at writes visible to the block
  if (threadIdx.x == 0) {
    ctl[kCtlSwapsAcc] += (uint32_t)s_nacc;
    if (from_ctl && last_group) ctl[kCtlPtRound] = rnd + 1u;
  }
  if (last_group && end_n > 0u)
    end_interval(R_all, ctl, anchor, dE, replica_at_all, series, end_n, 1);
}

// End of n sweeps without an exchange (see end_interval). grid = 1, block = 256.
__global__ __launch_bounds__(256) void k_end_sweep(int R, uint32_t *__restrict__ ctl,
                                                   const double *__restrict__ anchor,
                                                   const double *__restrict__ dE,
                                                   const int *__restrict__ replica_at,
                                                   double *__restrict__ series, uint32_t n,
                                                   int record_last) {
  end_interval(R, ctl, anchor, dE, replica_at, series, n, record_last);
}

// Series entries of one replica group after sweep ctl[kCtlSweep] + delta of
// a split run (the row count advances at the join): the rungs whose replica
// lies in [r0, r0 + nrep). Rungs do not move between joins. grid = 1,
// block = 256, on the group's own stream.
__global__ __launch_bounds__(256) void k_record_part(int R, int r0, int nrep, uint32_t delta,
                                                     const uint32_t *__restrict__ ctl,
                                                     const double *__restrict__ anchor,
                                                     const double *__restrict__ dE,
                                                     const int *__restrict__ replica_at,
                                                     double *__restrict__ series) {
  const uint32_t row = ctl[kCtlRecordRow] + delta;
  if (!ctl[kCtlRecordOn] || row >= ctl[kCtlRecordCap]) return;  // uniform over the block
  for (int k = threadIdx.x; k < R; k += blockDim.x) {
    const int rep = replica_at[k];
    if (rep >= r0 && rep < r0 + nrep) series[(size_t)row * R + k] = anchor[rep] + dE[rep];
  }
}

// Measurement row (the per-measurement quantities of Pyrochlore.cu:1641-1672):
// for each rung k (descending T) the full energy and the moduli of the 5
// order-parameter groups of the replica at that rung, appended at row
// ctl[kCtlMeasRow] while it is below ctl[kCtlMeasCap]. dL gives per-replica
// sizes in a mixed batch (else L). grid = 1, block = 256.
__global__ __launch_bounds__(256) void k_measure_row(int R, const double *__restrict__ obs,
                                                     const int *__restrict__ replica_at,
                                                     const int *__restrict__ dL, int L,
                                                     uint32_t *__restrict__ ctl,
                                                     double *__restrict__ rows) {
  const uint32_t row = ctl[kCtlMeasRow];
  if (row >= ctl[kCtlMeasCap]) return;  // uniform over the block
  for (int k = threadIdx.x; k < R; k += blockDim.x) {
    const int rep = replica_at[k];
    const double Lr = dL ? (double)dL[rep] : (double)L;
    const double *o = obs + (size_t)rep * kObs;
    double m[12];
    fold_order_params(o + 1, 16.0 * Lr * Lr * Lr, m);
    double *out = rows + ((size_t)row * R + k) * kMeasQ;
    out[0] = o[0];
    out[1] = fabs(m[0]);
    out[2] = sqrt(m[1] * m[1] + m[2] * m[2]);
    out[3] = sqrt(m[3] * m[3] + m[4] * m[4] + m[5] * m[5]);
    out[4] = sqrt(m[6] * m[6] + m[7] * m[7] + m[8] * m[8]);
    out[5] = sqrt(m[9] * m[9] + m[10] * m[10] + m[11] * m[11]);
  }
  __syncthreads();
  if (threadIdx.x == 0) ctl[kCtlMeasRow] = row + 1u;
}

// Annealing step m of `steps` (SimulatedAnnealingThread, MonteCarlo.cu:88-92):
// replica r runs at T0 - (T0 - T_r) * m * dT with dT = 1 / steps, evaluated in
// that order in FP64; m = steps restores beta = 1 / T_r exactly.
__global__ void k_anneal_beta(int R, const double *__restrict__ temps, double T0, int m, double dT,
                              int restore, float *__restrict__ beta) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  const double T = restore ? temps[r] : T0 - (T0 - temps[r]) * m * dT;
  beta[r] = (float)(1.0 / T);
}

// Spin guard: counts spins that are not finite or whose squared norm is off 1
// by more than tol, and the smallest such global spin index (replica r's
// spins are numbered from off[r] / 3). grid = (blocks, R), block = 256.
__global__ __launch_bounds__(256) void k_check_spins(const float *__restrict__ spins,
                                                     const long long *__restrict__ off,
                                                     const int *__restrict__ dL, double tol,
                                                     unsigned long long *__restrict__ out) {
  const int r = blockIdx.y;
  const long long L = dL[r];
  const long long n = 16 * L * L * L;
  const float *S = spins + off[r];
  unsigned long long bad = 0, first = ~0ull;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float x = S[i], y = S[n + i], z = S[2 * n + i];
    const double q = (double)x * x + (double)y * y + (double)z * z;
    if (!(fabs(q - 1.0) <= tol)) {  // also catches NaN and infinities
      ++bad;
      const unsigned long long g = (unsigned long long)(off[r] / 3 + i);
      first = first < g ? first : g;
    }
  }
  if (bad) {
    atomicAdd(&out[0], bad);
    atomicMin(&out[1], first);
  }
}

__global__ void k_set_anchor(int R, const double *__restrict__ obs, double *__restrict__ anchor,
                             double *__restrict__ dE) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < R) {
    anchor[r] = obs[(size_t)r * kObs];
    dE[r] = 0.0;
  }
}

__global__ void k_set_sweep(uint32_t *__restrict__ ctl, uint32_t sweep, uint32_t round) {
  if (threadIdx.x == 0) {
    ctl[kCtlSweep] = sweep;
    ctl[kCtlPtRound] = round;
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

int launch_init_spins(pyro_ctx *ctx) {
  if (ctx->hetero) return launch_batch_init_spins(ctx);
  dim3 grid((ctx->ncell + 255) / 256, ctx->R);
  k_init_spins<<<grid, 256, 0, ctx->stream>>>(ctx->d_spins, ctx->ncell, key_lo(ctx), key_hi(ctx));
  PYRO_HIP(hipGetLastError());
  return 0;
}

int launch_metropolis_phase(pyro_ctx *ctx, int colour, int pass, const uint32_t *) {
  if (ctx->hetero) return set_error(PYRO_ESTATE, "single colour phases need a uniform lattice size");
  if (ctx->mat) return launch_phase_mat(ctx, 0, colour, pass);
  launch_phase<0, false>(ctx, colour, pass);
  PYRO_HIP(hipGetLastError());
  return 0;
}

bool or_fusable(const pyro_ctx *ctx) {
  if (!ctx->fuse_or || !ctx->d_spins_alt || ctx->hetero || ctx->mat || ctx->exec_mode != 0 ||
      ctx->phase_variant != 1 || ctx->set_schedule || ctx->or_angle)
    return false;
  switch (ctx->L) {
    case 4: case 8: case 12: case 16: case 24: case 32: case 64: return true;
    default: return false;
  }
}

bool metro_fusable(const pyro_ctx *ctx) {
  return ctx->fuse_or && !ctx->hetero && !ctx->mat && ctx->exec_mode == 0 &&
         ctx->phase_variant == 1 && !ctx->post_rotate && !ctx->set_schedule && ctx->L == 16;
}

int launch_metropolis_pair(pyro_ctx *ctx, int ca, int pass) {
  if (!metro_fusable(ctx)) return set_error(PYRO_ESTATE, "fused Metropolis pair: not for this ctx");
  if (ca == 0)
    launch_metro_pair_t<0>(ctx, pass);
  else
    launch_metro_pair_t<2>(ctx, pass);
  PYRO_HIP(hipGetLastError());
  return 0;
}

int launch_overrelax_pair(pyro_ctx *ctx, int ca, const float *cur, float *oth) {
  if (!launch_or_pair(ctx, ca, cur, oth))
    return set_error(PYRO_ESTATE, "fused over-relaxation: no compiled copy for L=%d", ctx->L);
  PYRO_HIP(hipGetLastError());
  return 0;
}

bool or_whole_pass(const pyro_ctx *ctx) {
  return ctx->fuse_or == 2 && or_fusable(ctx) && (ctx->L == 16 || ctx->L == 64) && !ctx->track_or;
}

int launch_overrelax_pass_whole(pyro_ctx *ctx, const float *cur, float *oth) {
  if (!or_whole_pass(ctx)) return set_error(PYRO_ESTATE, "whole over-relaxation pass: not for this ctx");
  launch_or_pass(ctx, cur, oth);
  PYRO_HIP(hipGetLastError());
  return 0;
}

int launch_overrelax_pass_fused(pyro_ctx *ctx, int parity) {
  const float *cur = parity ? ctx->d_spins_alt : ctx->d_spins;
  float *oth = parity ? ctx->d_spins : ctx->d_spins_alt;
  if (or_whole_pass(ctx)) return launch_overrelax_pass_whole(ctx, cur, oth);
  if (!launch_or_pair(ctx, 0, cur, oth) || !launch_or_pair(ctx, 2, cur, oth))
    return set_error(PYRO_ESTATE, "fused over-relaxation: no compiled copy for L=%d", ctx->L);
  PYRO_HIP(hipGetLastError());
  return 0;
}

int launch_overrelax_phase(pyro_ctx *ctx, int colour, int pass) {
  if (ctx->hetero) return set_error(PYRO_ESTATE, "single colour phases need a uniform lattice size");
  if (ctx->mat) return launch_phase_mat(ctx, ctx->or_angle ? 2 : 1, colour, pass);
  if (ctx->or_angle)
    launch_phase<2, false>(ctx, colour, pass);
  else
    launch_phase<1, false>(ctx, colour, pass);
  PYRO_HIP(hipGetLastError());
  return 0;
}

int launch_observables(pyro_ctx *ctx, bool anchor) {
  if (ctx->hetero) {
    PYRO_TRY(launch_batch_observables(ctx));
    return anchor ? launch_set_anchor(ctx) : 0;
  }
  if (ctx->mat) {
    PYRO_TRY(launch_observables_partials_mat(ctx));
  } else {
    launch_observables_partials<false>(ctx);
    PYRO_HIP(hipGetLastError());
  }
  const double escale = ctx->mat ? -1.0 : 0.5 * (double)ctx->jeff;
  k_observables_final<<<ctx->R, 256, 0, ctx->stream>>>(
      ctx->d_partials, ctx->obs_blocks, escale, ctx->d_obs, anchor ? ctx->d_anchor : nullptr,
      anchor ? ctx->d_dE : nullptr);
  PYRO_HIP(hipGetLastError());
  return 0;
}

int launch_pt_swap(pyro_ctx *ctx, int parity, uint32_t round, int from_ctl, uint32_t end_n) {
  const int ng = (int)ctx->h_group_off.size();
  for (int g = 0; g < ng; ++g) {
    k_pt_swap<<<1, 256, 0, ctx->stream>>>(
        ctx->h_group_len[g], ctx->d_obs, ctx->d_temps, ctx->d_beta, ctx->d_cosc,
        ctx->d_replica_at + ctx->h_group_off[g], ctx->d_ctl, key_lo(ctx), key_hi(ctx), parity,
        round, from_ctl, g, g == ng - 1, end_n, ctx->R, ctx->d_replica_at, ctx->d_anchor,
        ctx->d_dE, ctx->d_series);
    PYRO_HIP(hipGetLastError());
  }
  return 0;
}

int launch_end_sweep(pyro_ctx *ctx, uint32_t n, int record_last) {
  k_end_sweep<<<1, 256, 0, ctx->stream>>>(ctx->R, ctx->d_ctl, ctx->d_anchor, ctx->d_dE,
                                          ctx->d_replica_at, ctx->d_series, n, record_last);
  PYRO_HIP(hipGetLastError());
  return 0;
}

int launch_record_part(pyro_ctx *ctx) {
  if (!ctx->d_series) return 0;
  k_record_part<<<1, 256, 0, launch_stream(ctx)>>>(ctx->R, launch_r0(ctx), launch_nrep(ctx),
                                                   ctx->launch_delta, ctx->d_ctl, ctx->d_anchor,
                                                   ctx->d_dE, ctx->d_replica_at, ctx->d_series);
  PYRO_HIP(hipGetLastError());
  return 0;
}

int launch_measure_row(pyro_ctx *ctx) {
  k_measure_row<<<1, 256, 0, ctx->stream>>>(ctx->R, ctx->d_obs, ctx->d_replica_at,
                                            ctx->hetero ? ctx->d_L : nullptr, ctx->L, ctx->d_ctl,
                                            ctx->d_meas);
  PYRO_HIP(hipGetLastError());
  return 0;
}

int launch_anneal_beta(pyro_ctx *ctx, double T0, int m, int steps, int restore) {
  k_anneal_beta<<<(ctx->R + 255) / 256, 256, 0, ctx->stream>>>(ctx->R, ctx->d_temps, T0, m,
                                                               1.0 / (double)steps, restore,
                                                               ctx->d_beta);
  PYRO_HIP(hipGetLastError());
  return 0;
}

int check_spins(pyro_ctx *ctx, double tol, unsigned long long *d_out) {
  int Lmax = 0;
  for (int L : ctx->h_L) Lmax = L > Lmax ? L : Lmax;
  const long long n = 16ll * Lmax * Lmax * Lmax;
  const int blocks = (int)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024);
  k_check_spins<<<dim3(blocks, ctx->R), 256, 0, ctx->stream>>>(ctx->d_spins, ctx->d_off, ctx->d_L,
                                                                 tol, d_out);
  PYRO_HIP(hipGetLastError());
  return 0;
}

int launch_set_anchor(pyro_ctx *ctx) {
  k_set_anchor<<<(ctx->R + 255) / 256, 256, 0, ctx->stream>>>(ctx->R, ctx->d_obs, ctx->d_anchor,
                                                               ctx->d_dE);
  PYRO_HIP(hipGetLastError());
  return 0;
}

int launch_set_sweep(pyro_ctx *ctx, uint32_t sweep, uint32_t round) {
  k_set_sweep<<<1, 64, 0, ctx->stream>>>(ctx->d_ctl, sweep, round);
  PYRO_HIP(hipGetLastError());
  return 0;
}

// Scratch of pyro_observables_noalloc: the first-stage partial sums of one
// call, R * 4 * ceil(L^3 / 256) * 4 doubles, at most kNoAllocScratchDoubles
// (4 MiB: L = 16 x 64 replicas needs 128 KiB, L = 64 x 32 all of it). Each GPU
// holds its own copy until the process ends; every call on a GPU shares it.
// Host code uses the per-device addresses from hipGetSymbolAddress, never the
// array's name (a host-side stand-in).
constexpr size_t kNoAllocScratchDoubles = size_t(1) << 19;
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

constexpr int kMaxDevices = 64;
static double *g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mutex;

// The device address of `scratch` on `device`, which must be the current
// device; resolved once per device.
int noalloc_scratch(int device, double **ptr) {
  if (device < 0 || device >= kMaxDevices)
    return set_error(PYRO_EINVAL, "no-allocation scratch: device %d outside [0, %d)", device, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(PYRO_EHIP, "no-allocation scratch: hipGetSymbolAddress failed on device %d: %s",
                       device, hipGetErrorString(e));
    g_scratch_ptr[device] = (double *)p;
  }
  *ptr = g_scratch_ptr[device];
  return 0;
}

}  // namespace pyro

using namespace pyro;

extern "C" int pyro_observables_noalloc(int device, const float *spins, int L, int R, float jeff,
                                        double *obs13, void *stream) {
  if (!spins || !obs13) return set_error(PYRO_EINVAL, "pyro_observables_noalloc: null pointer");
  if (L < 2 || L > 256 || (size_t)L * L * L * 16 > (size_t)INT32_MAX)
    return set_error(PYRO_EINVAL, "pyro_observables_noalloc: L=%d outside [2, 256]", L);
  if (R < 1 || R > 65535)
    return set_error(PYRO_EINVAL, "pyro_observables_noalloc: R=%d outside [1, 65535]", R);
  const int ncell = L * L * L;
  const int nblk = (ncell + kObsBlock - 1) / kObsBlock;
  const size_t need = (size_t)R * 4 * nblk * 4;
  if (need > kNoAllocScratchDoubles)
    return set_error(PYRO_EINVAL,
                     "pyro_observables_noalloc: L=%d x %d replicas need %zu bytes of scratch, above "
                     "the %zu-byte maximum",
                     L, R, need * sizeof(double), kNoAllocScratchDoubles * sizeof(double));
  int ndev = 0, prev = 0;
  PYRO_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev)
    return set_error(PYRO_EINVAL, "pyro_observables_noalloc: device %d of %d", device, ndev);
  PYRO_HIP(hipGetDevice(&prev));
  if (prev != device) PYRO_HIP(hipSetDevice(device));
  double *part = nullptr;
  int rc = noalloc_scratch(device, &part);
  if (rc == 0) {
    hipStream_t st = (hipStream_t)stream;
    // Every partial is written by its block, so nothing needs clearing.
    launch_obs_partials<false>(spins, L, R, part, nullptr, st);
    k_observables_final<<<R, 256, 0, st>>>(part, nblk, 0.5 * (double)jeff, obs13, nullptr, nullptr);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
      rc = set_error(PYRO_EHIP, "pyro_observables_noalloc: launch failed: %s", hipGetErrorString(e));
  }
  if (prev != device) (void)hipSetDevice(prev);
  return rc;
}
