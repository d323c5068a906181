// C ABI and host glue: context lifetime, the sweep driver, parallel
// tempering cadence, cone adaptation, observables and the energy series.
// Host C++; every HIP status is checked and turned into a PYRO_E* code.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>

#include "pyro_internal.h"

namespace pyro {

static thread_local std::string g_error;

int set_error(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_error = buf;
  return code;
}

static int device_guard(pyro_ctx *ctx) {
  PYRO_HIP(hipSetDevice(ctx->device));
  return 0;
}

static void drop_graph(pyro_ctx *ctx) {
  if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
  if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
  ctx->graph_exec = nullptr;
  ctx->graph = nullptr;
  ctx->graph_sweeps = 0;
}

// Recompute the full energy and make it the anchor of the incremental one.
static int reanchor(pyro_ctx *ctx) { return launch_observables(ctx, true); }

// After a change of the energy model: the captured sweeps hold the old
// kernels, and the incremental energy restarts from a full recomputation.
int reset_model_state(pyro_ctx *ctx) {
  PYRO_TRY(device_guard(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  drop_graph(ctx);
  PYRO_TRY(reanchor(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

// The colour phases of one sweep for the current launch range. With record,
// the range's energy-series entries after the sweep are written too: by the
// first over-relaxation launch when its kernel can take them (untracked
// over-relaxation leaves dE final after the Metropolis and learned phases),
// else by a separate launch at the end.
static int enqueue_phases(pyro_ctx *ctx, bool record = false) {
  if (ctx->exec_mode == 2) {  // tile lists: no series hook in these kernels
    PYRO_TRY(launch_batch_sweep_body(ctx));
    return record ? launch_record_part(ctx) : 0;
  }
  for (int p = 0; p < ctx->n_metro; ++p) {
    if (metro_fusable(ctx)) {
      PYRO_TRY(launch_metropolis_pair(ctx, 0, p));
      PYRO_TRY(launch_metropolis_pair(ctx, 2, p));
    } else {
      for (int c = 0; c < kColours; ++c) PYRO_TRY(launch_metropolis_phase(ctx, c, p, nullptr));
    }
  }
  for (int p = 0; p < ctx->n_learned; ++p)
    for (int c = 0; c < kColours; ++c) PYRO_TRY(launch_learned_phase(ctx, c, p, p > 0));
  if (ctx->n_learned > 0) PYRO_TRY(launch_learned_fold(ctx));
  ctx->launch_series_done = 0;
  int o = 0;
  if (or_fusable(ctx)) {
    // An even number of fused passes (two colours per launch, alternating
    // between the spin buffers) ends in d_spins; an odd last pass runs in place.
    for (; o < (ctx->n_or & ~1); ++o) {
      ctx->launch_series = record && o == 0 ? ctx->d_series : nullptr;
      const int rc = launch_overrelax_pass_fused(ctx, o & 1);
      ctx->launch_series = nullptr;
      if (rc) return rc;
    }
  }
  for (; o < ctx->n_or; ++o)
    for (int c = 0; c < kColours; ++c) {
      ctx->launch_series = record && o == 0 && c == 0 ? ctx->d_series : nullptr;
      const int rc = launch_overrelax_phase(ctx, c, o);
      ctx->launch_series = nullptr;
      if (rc) return rc;
    }
  const bool recorded = ctx->launch_series_done != 0;
  ctx->launch_series_done = 0;
  if (record && !recorded) PYRO_TRY(launch_record_part(ctx));
  return 0;
}

// Whether the sweeps run cache-blocked (block_groups): execution mode 0 on a
// uniform lattice with up to R groups, or the tile lists (mode 2) with up to
// kMaxGroups groups (one prebuilt tile list each).
static bool group_blocked(const pyro_ctx *ctx) {
  const int G = ctx->block_groups;
  if (G < 2 || G > ctx->R) return false;
  return (ctx->exec_mode == 0 && !ctx->hetero) || (ctx->exec_mode == 2 && G <= kMaxGroups);
}

static bool split_sweep(const pyro_ctx *ctx) {
  return !group_blocked(ctx) && ctx->n_parts > 1 && ctx->n_parts <= ctx->n_streams &&
         ctx->n_parts <= ctx->R && (ctx->exec_mode == 2 || (ctx->exec_mode == 0 && !ctx->hetero));
}

// The n sweeps of one interval, cache-blocked: group p runs all n before
// the next group of its stream starts. With S = n_parts streams (1 when
// n_parts does not name a usable split), stream s takes groups s, s + S, ...
// in turn, so S groups are in flight (and in the cache) at a time; the
// streams fork from and join the ctx stream around the interval. The draws
// of sweep d of the interval are keyed on ctl[kCtlSweep] + d (launch_delta),
// as in a split sweep, so spins equal an unblocked run's bit for bit. Every
// sweep's series entries are written by the phase launches, except the last
// one's when an exchange or measurement ends the interval (the interval end
// writes it).
static int enqueue_blocked_interval(pyro_ctx *ctx, uint32_t n, bool joined_end) {
  const int G = ctx->block_groups, parts0 = ctx->n_parts;
  const int S = parts0 >= 2 && parts0 <= ctx->n_streams ? std::min(parts0, G) : 1;
  if (S > 1) {
    PYRO_HIP(hipEventRecord(ctx->ev_fork, ctx->stream));
    for (int s = 1; s < S; ++s) PYRO_HIP(hipStreamWaitEvent(ctx->part_stream[s], ctx->ev_fork, 0));
  }
  if (ctx->exec_mode == 2) ctx->n_parts = G;  // the tile list of group p of G
  int rc = 0;
  for (int p = 0; p < G && rc == 0; ++p) {
    const int r0 = p * ctx->R / G, r1 = (p + 1) * ctx->R / G;
    ctx->launch_stream = ctx->part_stream[p % S];
    ctx->launch_part = p;
    ctx->launch_r0 = r0;
    ctx->launch_nrep = r1 - r0;
    for (uint32_t d = 0; d < n && rc == 0; ++d) {
      ctx->launch_delta = d;
      rc = enqueue_phases(ctx, ctx->d_series != nullptr && !(joined_end && d + 1 == n));
    }
  }
  ctx->n_parts = parts0;
  ctx->launch_stream = nullptr;
  ctx->launch_nrep = 0;
  ctx->launch_part = 0;
  ctx->launch_delta = 0;
  for (int s = 1; s < S; ++s) {  // joined even after an error: no stream left running ahead
    if (hipEventRecord(ctx->ev_join[s], ctx->part_stream[s]) != hipSuccess ||
        hipStreamWaitEvent(ctx->stream, ctx->ev_join[s], 0) != hipSuccess)
      rc = rc ? rc : set_error(PYRO_EHIP, "blocked interval: stream join failed");
  }
  return rc;
}

// The other groups' streams join the ctx stream (no booking).
static int split_join_streams(pyro_ctx *ctx) {
  for (int p = 1; p < ctx->n_parts; ++p) {
    PYRO_HIP(hipEventRecord(ctx->ev_join[p], ctx->part_stream[p]));
    PYRO_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_join[p], 0));
  }
  ctx->forked = 0;
  return 0;
}

// Ends a split run: the other streams join the ctx stream, and the sweeps
// made since the fork are booked (sweep counter, series rows). With
// record_last the series row of the last of them is written here (after an
// exchange moved the rungs); otherwise the phase launches of each group
// wrote every row (enqueue_phases with record).
static int split_join(pyro_ctx *ctx, int record_last) {
  if (!ctx->forked) return 0;
  PYRO_TRY(split_join_streams(ctx));
  const uint32_t n = ctx->unjoined;
  ctx->unjoined = 0;
  return launch_end_sweep(ctx, n, record_last);
}

// The phases of one sweep on n_parts streams, one replica group each (the
// highest group first, group 0 on the ctx stream last), plus each group's
// series entries when the sweep is not followed by a join.
static int enqueue_split_phases(pyro_ctx *ctx, bool join_after) {
  const int np = ctx->n_parts;
  if (!ctx->forked) {
    PYRO_HIP(hipEventRecord(ctx->ev_fork, ctx->stream));
    for (int p = 1; p < np; ++p) PYRO_HIP(hipStreamWaitEvent(ctx->part_stream[p], ctx->ev_fork, 0));
    ctx->forked = 1;
    ctx->unjoined = 0;
  }
  ctx->launch_delta = ctx->unjoined;
  int rc = 0;
  for (int p = np - 1; p >= 0 && rc == 0; --p) {
    const int r0 = p * ctx->R / np, r1 = (p + 1) * ctx->R / np;
    ctx->launch_stream = ctx->part_stream[p];
    ctx->launch_part = p;
    ctx->launch_r0 = r0;
    ctx->launch_nrep = r1 - r0;
    rc = enqueue_phases(ctx, !join_after && ctx->d_series != nullptr);
  }
  ctx->launch_stream = nullptr;
  ctx->launch_nrep = 0;
  ctx->launch_part = 0;
  ctx->launch_delta = 0;
  ++ctx->unjoined;
  return rc;
}

// Whether an exchange / a measurement follows sweep `sweep_index`.
static bool exchange_after(const pyro_ctx *ctx, uint32_t sweep_index) {
  return ctx->pt_every > 0 && (sweep_index + 1u) % (uint32_t)ctx->pt_every == 0u;
}
static bool measure_after(const pyro_ctx *ctx, uint32_t sweep_index) {
  return ctx->measure_every > 0 && (sweep_index + 1u) % (uint32_t)ctx->measure_every == 0u;
}

// What follows the last sweep of an interval: full energies (which also
// re-anchor the incremental energy, so FP32 rounding in the untracked
// over-relaxation passes cannot accumulate beyond one exchange or measurement
// interval), the measurement row, the exchange, and the booking of the
// interval's n sweeps with the series row of the last one.
static int enqueue_interval_end(pyro_ctx *ctx, bool exchange, bool measure, uint32_t n) {
  if (exchange || measure) PYRO_TRY(launch_observables(ctx, true));
  // The measurement sees the configuration of this sweep at the rungs it was
  // simulated at, before the exchange moves temperatures.
  if (measure) PYRO_TRY(launch_measure_row(ctx));
  if (exchange) return launch_pt_swap(ctx, 0, 0, 1, n);  // books the interval too
  return launch_end_sweep(ctx, n, 1);
}

// One sweep as kernel launches (also the body that gets captured). Replicas
// are independent between exchanges, so the phases of n_parts replica groups
// can run on as many streams: one group's kernels fill the GPU while
// another's drain, which hides most of the per-launch ramp of these short
// kernels. The streams run apart until the next exchange or measurement
// (which need every replica) or the end of the call (split_join).
static int enqueue_sweep(pyro_ctx *ctx, uint32_t sweep_index) {
  const bool exchange = exchange_after(ctx, sweep_index);
  const bool measure = measure_after(ctx, sweep_index);
  const bool split = split_sweep(ctx);
  if (split) {
    PYRO_TRY(enqueue_split_phases(ctx, exchange || measure));
    if (!(exchange || measure)) return 0;
    // Join without booking yet: the counters advance after the exchange, as
    // for an unsplit sweep.
    PYRO_TRY(split_join_streams(ctx));
  } else if (ctx->exec_mode != 0) {
    PYRO_TRY(launch_batch_sweep_body(ctx));
  } else {
    PYRO_TRY(enqueue_phases(ctx));
  }
  const uint32_t n = split ? ctx->unjoined : 1u;
  ctx->unjoined = 0;
  return enqueue_interval_end(ctx, exchange, measure, n);
}

// Sweeps first .. first + count - 1 (first = the sweep index that sets the
// exchange and measurement cadence). The persistent execution runs each
// interval up to an exchange or measurement (or to the end) in one launch;
// the others go sweep by sweep.
static int enqueue_sweeps(pyro_ctx *ctx, uint32_t first, uint32_t count) {
  if (group_blocked(ctx)) {
    uint32_t i = 0;
    while (i < count) {
      uint32_t j = i;
      while (j + 1 < count && !exchange_after(ctx, first + j) && !measure_after(ctx, first + j)) ++j;
      const bool exchange = exchange_after(ctx, first + j), measure = measure_after(ctx, first + j);
      const uint32_t n = j - i + 1;
      PYRO_TRY(enqueue_blocked_interval(ctx, n, exchange || measure));
      if (exchange || measure) {
        PYRO_TRY(enqueue_interval_end(ctx, exchange, measure, n));
      } else {
        PYRO_TRY(launch_end_sweep(ctx, n, 0));
      }
      i = j + 1;
    }
    return 0;
  }
  if (ctx->exec_mode != 1) {
    for (uint32_t i = 0; i < count; ++i) PYRO_TRY(enqueue_sweep(ctx, first + i));
    return 0;
  }
  uint32_t i = 0;
  while (i < count) {
    uint32_t j = i;
    while (j + 1 < count && !exchange_after(ctx, first + j) && !measure_after(ctx, first + j)) ++j;
    const bool exchange = exchange_after(ctx, first + j), measure = measure_after(ctx, first + j);
    const uint32_t n = j - i + 1;
    PYRO_TRY(launch_batch_sweeps(ctx, (int)n, exchange || measure ? 0 : 1));
    if (exchange || measure) {
      PYRO_TRY(enqueue_interval_end(ctx, exchange, measure, n));
    } else {
      PYRO_TRY(launch_end_sweep(ctx, n, 0));
    }
    i = j + 1;
  }
  return 0;
}

// Sweeps per captured graph: a whole number of exchange and measurement
// intervals, so that every replay starts aligned with both cadences, and at
// least kMinGraphSweeps so that small lattices (a few microseconds of work per
// sweep) do not pay a graph launch per sweep.
constexpr uint32_t kMinGraphSweeps = 16;
static uint32_t graph_block(const pyro_ctx *ctx) {
  const uint32_t a = ctx->pt_every > 0 ? (uint32_t)ctx->pt_every : 1u;
  const uint32_t b = ctx->measure_every > 0 ? (uint32_t)ctx->measure_every : 1u;
  const uint32_t l = (uint32_t)std::lcm(a, b);
  return l * ((kMinGraphSweeps + l - 1) / l);
}

static uint64_t replica_cells(const pyro_ctx *ctx, int r) {
  const uint64_t L = (uint64_t)ctx->h_L[r];
  return L * L * L;
}

static void account_sweeps(pyro_ctx *ctx, uint32_t n) {
  for (int r = 0; r < ctx->R; ++r)
    ctx->h_proposed[r] += (uint64_t)ctx->n_metro * 16ull * replica_cells(ctx, r) * n;
  for (size_t r = 0; r < ctx->h_proposed_learned.size(); ++r)
    ctx->h_proposed_learned[r] += (uint64_t)ctx->n_learned * 16ull * replica_cells(ctx, (int)r) * n;
  for (uint32_t i = 0; i < n; ++i) {
    if (ctx->pt_every > 0 && (ctx->h_sweep + 1u) % (uint32_t)ctx->pt_every == 0u) ++ctx->h_pt_round;
    ++ctx->h_sweep;
  }
}

}  // namespace pyro

using namespace pyro;

extern "C" {

const char *pyro_version(void) { return "pyro-mi355x 0.1 (gfx950)"; }

const char *pyro_last_error(void) { return g_error.c_str(); }

int pyro_links16(int8_t out[16][6][4]) {
  if (!out) return set_error(PYRO_EINVAL, "pyro_links16: null output");
  static const Link t[16][6] = PYRO_LINKS16_INIT;
  for (int s = 0; s < 16; ++s)
    for (int n = 0; n < 6; ++n) {
      out[s][n][0] = t[s][n].dx;
      out[s][n][1] = t[s][n].dy;
      out[s][n][2] = t[s][n].dz;
      out[s][n][3] = t[s][n].site;
    }
  return 0;
}

int pyro_independent_sets(int8_t out[24][4]) {
  if (!out) return set_error(PYRO_EINVAL, "pyro_independent_sets: null output");
  static const int8_t t[24][4] = PYRO_INDEPENDENT_SETS_INIT;
  std::memcpy(out, t, sizeof t);
  return 0;
}

int pyro_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  if (!ctr || !key || !out) return set_error(PYRO_EINVAL, "pyro_philox4x32_10: null argument");
  u32x4 c = {ctr[0], ctr[1], ctr[2], ctr[3]};
  u32x4 o = philox4x32_10(c, key[0], key[1]);
  out[0] = o.x;
  out[1] = o.y;
  out[2] = o.z;
  out[3] = o.w;
  return 0;
}

// Shared construction of uniform and batch ctxs: Ls[r] is replica r's size.
static int create_ctx(const std::vector<int> &Ls, const double *temps, uint64_t seed, double J,
                      double spin_len2, int n_metro, int n_or, int pt_every, float cos_cone,
                      int device, pyro_ctx **out) {
  const int R = (int)Ls.size();
  if (R < 1 || R > 65535) return set_error(PYRO_EINVAL, "pyro_create: n_replicas=%d outside [1, 65535]", R);
  if (!temps) return set_error(PYRO_EINVAL, "pyro_create: null temperatures");
  for (int r = 0; r < R; ++r)
    if (!(temps[r] > 0.0) || !std::isfinite(temps[r]))
      return set_error(PYRO_EINVAL, "pyro_create: temperature[%d]=%g not positive", r, temps[r]);
  if (n_metro < 0 || n_or < 0 || pt_every < 0 || n_metro > 255)
    return set_error(PYRO_EINVAL, "pyro_create: negative pass counts");
  if (!(cos_cone >= -1.0f && cos_cone < 1.0f))
    return set_error(PYRO_EINVAL, "pyro_create: cos_cone=%g outside [-1, 1)", cos_cone);
  int ndev = 0;
  PYRO_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev)
    return set_error(PYRO_EINVAL, "pyro_create: device %d of %d", device, ndev);
  PYRO_HIP(hipSetDevice(device));

  pyro_ctx *ctx = new pyro_ctx();
  ctx->R = R;
  ctx->h_L = Ls;
  ctx->hetero = false;
  for (int r = 1; r < R; ++r)
    if (Ls[r] != Ls[0]) ctx->hetero = true;
  ctx->L = ctx->hetero ? 0 : Ls[0];
  ctx->ncell = ctx->hetero ? 0 : Ls[0] * Ls[0] * Ls[0];
  ctx->h_off.resize(R);
  size_t total = 0;
  for (int r = 0; r < R; ++r) {
    ctx->h_off[r] = (long long)total;
    total += (size_t)48 * Ls[r] * Ls[r] * Ls[r];
  }
  ctx->total_floats = total;
  // Exchange groups: one per lattice size, ascending.
  std::vector<int> sizes(Ls);
  std::sort(sizes.begin(), sizes.end());
  sizes.erase(std::unique(sizes.begin(), sizes.end()), sizes.end());
  int go = 0;
  for (int L : sizes) {
    const int n = (int)std::count(Ls.begin(), Ls.end(), L);
    ctx->h_group_off.push_back(go);
    ctx->h_group_len.push_back(n);
    go += n;
  }
  ctx->exec_mode = ctx->hetero ? 2 : 0;
  ctx->device = device;
  ctx->seed = seed;
  ctx->jeff = (float)(J * spin_len2);
  ctx->spin_len2 = spin_len2;
  ctx->n_metro = n_metro;
  ctx->n_or = n_or;
  ctx->pt_every = pt_every;
  ctx->obs_blocks = ctx->hetero ? 1 : (ctx->ncell + kObsBlock - 1) / kObsBlock;
  ctx->h_proposed.assign(R, 0);

  auto fail = [&](int rc) {
    pyro_destroy(ctx);
    return rc;
  };
#define ALLOC(ptr, count)                                                                  \
  do {                                                                                     \
    if (hipMalloc((void **)&(ptr), sizeof(*(ptr)) * (size_t)(count)) != hipSuccess)        \
      return fail(set_error(PYRO_ENOMEM, "pyro_create: hipMalloc of %zu bytes for %s failed", \
                            sizeof(*(ptr)) * (size_t)(count), #ptr));                      \
  } while (0)
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(set_error(PYRO_EHIP, "pyro_create: stream creation failed"));
  ctx->part_stream[0] = ctx->stream;
  if (R >= 2) {  // streams for split sweeps (default priority)
    ctx->n_streams = std::min(R, kMaxParts);
    if (hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess)
      return fail(set_error(PYRO_EHIP, "pyro_create: event creation failed"));
    for (int p = 1; p < ctx->n_streams; ++p)
      if (hipStreamCreateWithFlags(&ctx->part_stream[p], hipStreamNonBlocking) != hipSuccess ||
          hipEventCreateWithFlags(&ctx->ev_join[p], hipEventDisableTiming) != hipSuccess)
        return fail(set_error(PYRO_EHIP, "pyro_create: stream creation failed"));
    // Splitting pays when each group's launch still spans a few hundred
    // workgroups; below that, the fork and join cost more than the overlap
    // hides (measured: L=16 x 64 replicas 0.396 ms per sweep on one stream,
    // 0.342 in two groups, 0.335 in three; L=4 x 64 2.2x slower in two).
    // Four groups ran 0.328-0.363 ms from one captured graph to the next
    // (four streams of this ctx on the default four hardware queues); with
    // the plane-staged fused pair, 0.3115-0.3125 ms on one MI355X and
    // 0.43-0.45 ms on another, against 0.319-0.325 for three on both. So the
    // default is the larger of 3 and 2 groups whose smallest group spans
    // kSplitMinBlocks; pyro_set_streams allows 4. With the whole
    // over-relaxation pass at L=16 x 64, four groups ran 0.2041-0.2054 ms per
    // sweep against 0.2122-0.2164 for three on three GPUs, and 0.2375-0.2440
    // against 0.2188-0.2249 (20 sweeps after the cone adaptation) on a
    // fourth, so the default stays 3 and pyro_autotune tries 4 there.
    for (int n : {3, 2}) {
      if (n > ctx->n_streams) continue;
      long long blocks = 0;  // group 0: the first R / n replicas (the fewest)
      for (int r = 0; r < R / n; ++r) blocks += ((long long)Ls[r] * Ls[r] * Ls[r] + 255) / 256;
      if (blocks >= kSplitMinBlocks) {
        ctx->n_parts = n;
        break;
      }
    }
  }
  ALLOC(ctx->d_spins, total);
  // The second buffer of the fused over-relaxation passes (without it, if it
  // does not fit, the passes run one colour per launch).
  if (hipMalloc((void **)&ctx->d_spins_alt, sizeof(float) * total) != hipSuccess) {
    ctx->d_spins_alt = nullptr;
    (void)hipGetLastError();
  }
  ALLOC(ctx->d_temps, R);
  ALLOC(ctx->d_beta, R);
  ALLOC(ctx->d_cosc, R);
  ALLOC(ctx->d_replica_at, R);
  ALLOC(ctx->d_acc, R);
  ALLOC(ctx->d_dE, R);
  ALLOC(ctx->d_anchor, R);
  ALLOC(ctx->d_partials, (size_t)R * ctx->obs_blocks * kObsPartials);
  ALLOC(ctx->d_obs, (size_t)R * kObs);
  ALLOC(ctx->d_ctl, kCtlWords);
#undef ALLOC

  std::vector<uint32_t> ctl(kCtlWords, 0u);
  std::vector<float> cosc(R, cos_cone);
  if (hipMemcpy(ctx->d_ctl, ctl.data(), sizeof(uint32_t) * kCtlWords, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemset(ctx->d_acc, 0, sizeof(unsigned long long) * R) != hipSuccess ||
      hipMemset(ctx->d_dE, 0, sizeof(double) * R) != hipSuccess ||
      hipMemset(ctx->d_anchor, 0, sizeof(double) * R) != hipSuccess ||
      hipMemcpy(ctx->d_cosc, cosc.data(), sizeof(float) * R, hipMemcpyHostToDevice) != hipSuccess)
    return fail(set_error(PYRO_EHIP, "pyro_create: initial upload failed"));
  // Resolve this device's no-allocation scratch now, outside any capture
  // (pyro_observables_noalloc resolves it on its first call otherwise, and
  // reports a failure itself: the ctx does not need it).
  double *scratch_ptr = nullptr;
  (void)noalloc_scratch(device, &scratch_ptr);
  int rc = batch_setup(ctx, 0);
  if (rc) return fail(rc);
  rc = pyro_set_temperatures(ctx, temps);
  if (rc) return fail(rc);
  rc = pyro_init_random_spins(ctx);
  if (rc) return fail(rc);
  *out = ctx;
  return 0;
}

int pyro_create(const pyro_config *cfg, pyro_ctx **out) {
  if (!cfg || !out) return set_error(PYRO_EINVAL, "pyro_create: null argument");
  *out = nullptr;
  if (cfg->L < 2 || cfg->L > 256)
    return set_error(PYRO_EINVAL, "pyro_create: L=%d outside [2, 256]", cfg->L);
  if ((size_t)cfg->L * cfg->L * cfg->L * 16 > (size_t)INT32_MAX)
    return set_error(PYRO_EINVAL, "pyro_create: L=%d too large for 32-bit site indices", cfg->L);
  if (cfg->n_replicas < 1 || cfg->n_replicas > 65535)
    return set_error(PYRO_EINVAL, "pyro_create: n_replicas=%d outside [1, 65535]", cfg->n_replicas);
  std::vector<int> Ls(cfg->n_replicas, cfg->L);
  return create_ctx(Ls, cfg->temperatures, cfg->seed, cfg->J, cfg->spin_len2, cfg->n_metro,
                    cfg->n_or, cfg->pt_every, cfg->cos_cone, cfg->device, out);
}

int pyro_create_batch(const pyro_batch_config *cfg, pyro_ctx **out) {
  if (!cfg || !out || !cfg->L) return set_error(PYRO_EINVAL, "pyro_create_batch: null argument");
  *out = nullptr;
  if (cfg->n_replicas < 1 || cfg->n_replicas > 65535)
    return set_error(PYRO_EINVAL, "pyro_create_batch: n_replicas=%d outside [1, 65535]",
                     cfg->n_replicas);
  std::vector<int> Ls(cfg->L, cfg->L + cfg->n_replicas);
  for (int r = 0; r < cfg->n_replicas; ++r)
    if (!batch_size_supported(Ls[r]))
      return set_error(PYRO_EINVAL, "pyro_create_batch: L[%d]=%d not in {4,8,12,16,24,32}", r, Ls[r]);
  return create_ctx(Ls, cfg->temperatures, cfg->seed, cfg->J, cfg->spin_len2, cfg->n_metro,
                    cfg->n_or, cfg->pt_every, cfg->cos_cone, cfg->device, out);
}

int pyro_destroy(pyro_ctx *ctx) {
  if (!ctx) return 0;
  (void)hipSetDevice(ctx->device);
  for (hipStream_t st : ctx->part_stream)
    if (st) (void)hipStreamSynchronize(st);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  drop_graph(ctx);
  mlp_free(ctx);
  batch_free(ctx);
  void *ptrs[] = {ctx->d_spins, ctx->d_temps, ctx->d_beta,     ctx->d_cosc, ctx->d_replica_at,
                  ctx->d_acc,   ctx->d_dE,    ctx->d_anchor,   ctx->d_partials, ctx->d_obs,
                  ctx->d_ctl,   ctx->d_series, ctx->d_meas, ctx->d_coup, ctx->d_spins_alt};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
  for (int p = 1; p < kMaxParts; ++p) {
    if (ctx->ev_join[p]) (void)hipEventDestroy(ctx->ev_join[p]);
    if (ctx->part_stream[p]) (void)hipStreamDestroy(ctx->part_stream[p]);
  }
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return 0;
}

int pyro_init_random_spins(pyro_ctx *ctx) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  PYRO_TRY(device_guard(ctx));
  PYRO_TRY(launch_init_spins(ctx));
  PYRO_TRY(reanchor(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

int pyro_set_spins(pyro_ctx *ctx, const float *soa) {
  if (!ctx || !soa) return set_error(PYRO_EINVAL, "pyro_set_spins: null argument");
  PYRO_TRY(device_guard(ctx));
  const size_t n = ctx->total_floats;
  PYRO_HIP(hipMemcpyAsync(ctx->d_spins, soa, n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  PYRO_TRY(reanchor(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

int pyro_get_spins(pyro_ctx *ctx, float *soa) {
  if (!ctx || !soa) return set_error(PYRO_EINVAL, "pyro_get_spins: null argument");
  PYRO_TRY(device_guard(ctx));
  const size_t n = ctx->total_floats;
  PYRO_HIP(hipMemcpyAsync(soa, ctx->d_spins, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

int pyro_spins_device_ptr(pyro_ctx *ctx, float **dev_ptr, size_t *n_floats) {
  if (!ctx || !dev_ptr || !n_floats) return set_error(PYRO_EINVAL, "pyro_spins_device_ptr: null");
  *dev_ptr = ctx->d_spins;
  *n_floats = ctx->total_floats;
  return 0;
}

int pyro_set_temperatures(pyro_ctx *ctx, const double *T) {
  if (!ctx || !T) return set_error(PYRO_EINVAL, "pyro_set_temperatures: null argument");
  PYRO_TRY(device_guard(ctx));
  const int R = ctx->R;
  std::vector<float> beta(R);
  std::vector<int> order(R);
  for (int r = 0; r < R; ++r) {
    if (!(T[r] > 0.0) || !std::isfinite(T[r]))
      return set_error(PYRO_EINVAL, "pyro_set_temperatures: T[%d]=%g not positive", r, T[r]);
    beta[r] = (float)(1.0 / T[r]);
  }
  // Rungs grouped by lattice size (ascending), each group by descending
  // temperature, ties by replica index (gsl_sort_largest_index order).
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    if (ctx->h_L[a] != ctx->h_L[b]) return ctx->h_L[a] < ctx->h_L[b];
    return T[a] > T[b];
  });
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  PYRO_HIP(hipMemcpy(ctx->d_temps, T, sizeof(double) * R, hipMemcpyHostToDevice));
  PYRO_HIP(hipMemcpy(ctx->d_beta, beta.data(), sizeof(float) * R, hipMemcpyHostToDevice));
  PYRO_HIP(hipMemcpy(ctx->d_replica_at, order.data(), sizeof(int) * R, hipMemcpyHostToDevice));
  return 0;
}

int pyro_get_temperatures(pyro_ctx *ctx, double *T, int *replica_at) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  PYRO_TRY(device_guard(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  if (T) PYRO_HIP(hipMemcpy(T, ctx->d_temps, sizeof(double) * ctx->R, hipMemcpyDeviceToHost));
  if (replica_at)
    PYRO_HIP(hipMemcpy(replica_at, ctx->d_replica_at, sizeof(int) * ctx->R, hipMemcpyDeviceToHost));
  return 0;
}

int pyro_set_cos_cone(pyro_ctx *ctx, const float *cos_cone) {
  if (!ctx || !cos_cone) return set_error(PYRO_EINVAL, "pyro_set_cos_cone: null argument");
  for (int r = 0; r < ctx->R; ++r)
    if (!(cos_cone[r] >= -1.0f && cos_cone[r] < 1.0f))
      return set_error(PYRO_EINVAL, "pyro_set_cos_cone: cos_cone[%d]=%g outside [-1, 1)", r,
                       cos_cone[r]);
  PYRO_TRY(device_guard(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  PYRO_HIP(hipMemcpy(ctx->d_cosc, cos_cone, sizeof(float) * ctx->R, hipMemcpyHostToDevice));
  return 0;
}

int pyro_get_cos_cone(pyro_ctx *ctx, float *cos_cone) {
  if (!ctx || !cos_cone) return set_error(PYRO_EINVAL, "pyro_get_cos_cone: null argument");
  PYRO_TRY(device_guard(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  PYRO_HIP(hipMemcpy(cos_cone, ctx->d_cosc, sizeof(float) * ctx->R, hipMemcpyDeviceToHost));
  return 0;
}

int pyro_metropolis_phase(pyro_ctx *ctx, int colour, uint32_t sweep, int pass) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  if (colour < 0 || colour > 3) return set_error(PYRO_EINVAL, "colour %d outside 0..3", colour);
  if (pass < 0 || pass > 255) return set_error(PYRO_EINVAL, "pass %d outside 0..255", pass);
  PYRO_TRY(device_guard(ctx));
  // The kernel reads the sweep index from the control words.
  PYRO_TRY(launch_set_sweep(ctx, sweep, ctx->h_pt_round));
  PYRO_TRY(launch_metropolis_phase(ctx, colour, pass, nullptr));
  PYRO_TRY(launch_set_sweep(ctx, ctx->h_sweep, ctx->h_pt_round));
  for (auto &p : ctx->h_proposed) p += 4ull * (uint64_t)ctx->ncell;
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

int pyro_overrelax_phase(pyro_ctx *ctx, int colour) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  if (colour < 0 || colour > 3) return set_error(PYRO_EINVAL, "colour %d outside 0..3", colour);
  PYRO_TRY(device_guard(ctx));
  PYRO_TRY(launch_overrelax_phase(ctx, colour));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

static int run_sweeps(pyro_ctx *ctx, int n_sweeps, int use_graph);

int pyro_sweep(pyro_ctx *ctx, int n_sweeps, int use_graph) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  if (n_sweeps < 0) return set_error(PYRO_EINVAL, "n_sweeps %d < 0", n_sweeps);
  PYRO_TRY(device_guard(ctx));
  const int rc = run_sweeps(ctx, n_sweeps, use_graph);
  if (rc && ctx->forked) (void)split_join(ctx, 0);  // never leave the other streams running ahead
  return rc;
}

static int run_sweeps(pyro_ctx *ctx, int n_sweeps, int use_graph) {
  const uint32_t block = graph_block(ctx);
  uint32_t left = (uint32_t)n_sweeps;
  if (!use_graph && left > 0) {
    PYRO_TRY(enqueue_sweeps(ctx, ctx->h_sweep, left));
    account_sweeps(ctx, left);
    left = 0;
  }
  // A captured block always starts on a multiple of the exchange and
  // measurement cadences, so both land after the same sweeps as in eager mode.
  if (left > 0 && ctx->h_sweep % block != 0u) {
    const uint32_t k = std::min(left, block - ctx->h_sweep % block);
    PYRO_TRY(enqueue_sweeps(ctx, ctx->h_sweep, k));
    account_sweeps(ctx, k);
    left -= k;
  }
  if (left >= block) {
    PYRO_TRY(split_join(ctx, 0));  // a captured block starts and ends joined
    if (!ctx->graph_exec || ctx->graph_sweeps != (int)block) {
      drop_graph(ctx);
      PYRO_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      int rc = enqueue_sweeps(ctx, 0, block);
      if (rc == 0) rc = split_join(ctx, 0);
      ctx->forked = 0;
      ctx->unjoined = 0;
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(ctx->stream, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return set_error(PYRO_EHIP, "graph capture failed: %s", hipGetErrorString(e));
      ctx->graph = g;
      PYRO_HIP(hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0));
      ctx->graph_sweeps = (int)block;
    }
    while (left >= block) {
      PYRO_HIP(hipGraphLaunch(ctx->graph_exec, ctx->stream));
      account_sweeps(ctx, block);
      left -= block;
    }
  }
  if (left > 0) {
    PYRO_TRY(enqueue_sweeps(ctx, ctx->h_sweep, left));
    account_sweeps(ctx, left);
  }
  PYRO_TRY(split_join(ctx, 0));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

int pyro_get_graph_block(pyro_ctx *ctx, int *sweeps) {
  if (!ctx || !sweeps) return set_error(PYRO_EINVAL, "null argument");
  *sweeps = (int)graph_block(ctx);
  return 0;
}

int pyro_get_sweep(pyro_ctx *ctx, uint32_t *sweep, uint32_t *pt_round) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  PYRO_TRY(device_guard(ctx));
  uint32_t ctl[kCtlWords];
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  PYRO_HIP(hipMemcpy(ctl, ctx->d_ctl, sizeof ctl, hipMemcpyDeviceToHost));
  if (ctl[kCtlSweep] != ctx->h_sweep || ctl[kCtlPtRound] != ctx->h_pt_round)
    return set_error(PYRO_ESTATE, "sweep counters diverged: device %u/%u host %u/%u", ctl[kCtlSweep],
                     ctl[kCtlPtRound], ctx->h_sweep, ctx->h_pt_round);
  if (sweep) *sweep = ctl[kCtlSweep];
  if (pt_round) *pt_round = ctl[kCtlPtRound];
  return 0;
}

int pyro_set_sweep(pyro_ctx *ctx, uint32_t sweep, uint32_t pt_round) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  PYRO_TRY(device_guard(ctx));
  PYRO_TRY(launch_set_sweep(ctx, sweep, pt_round));
  ctx->h_sweep = sweep;
  ctx->h_pt_round = pt_round;
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

int pyro_pt_swap(pyro_ctx *ctx, int parity, uint32_t round, int *n_accepted) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  if (parity != 0 && parity != 1) return set_error(PYRO_EINVAL, "parity %d not 0/1", parity);
  PYRO_TRY(device_guard(ctx));
  uint32_t before = 0, after = 0;
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  PYRO_HIP(hipMemcpy(&before, ctx->d_ctl + kCtlSwapsAcc, sizeof before, hipMemcpyDeviceToHost));
  PYRO_TRY(launch_observables(ctx));
  PYRO_TRY(launch_pt_swap(ctx, parity, round, 0));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  PYRO_HIP(hipMemcpy(&after, ctx->d_ctl + kCtlSwapsAcc, sizeof after, hipMemcpyDeviceToHost));
  if (n_accepted) *n_accepted = (int)(after - before);
  return 0;
}

static int fetch_obs(pyro_ctx *ctx, std::vector<double> &obs) {
  PYRO_TRY(device_guard(ctx));
  PYRO_TRY(launch_observables(ctx));
  obs.resize((size_t)ctx->R * kObs);
  PYRO_HIP(hipMemcpyAsync(obs.data(), ctx->d_obs, sizeof(double) * obs.size(), hipMemcpyDeviceToHost,
                          ctx->stream));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

int pyro_energy(pyro_ctx *ctx, double *E) {
  if (!ctx || !E) return set_error(PYRO_EINVAL, "pyro_energy: null argument");
  std::vector<double> obs;
  PYRO_TRY(fetch_obs(ctx, obs));
  for (int r = 0; r < ctx->R; ++r) E[r] = obs[(size_t)r * kObs];
  return 0;
}

int pyro_sublattice_sums(pyro_ctx *ctx, double *sub12) {
  if (!ctx || !sub12) return set_error(PYRO_EINVAL, "pyro_sublattice_sums: null argument");
  std::vector<double> obs;
  PYRO_TRY(fetch_obs(ctx, obs));
  for (int r = 0; r < ctx->R; ++r)
    for (int q = 0; q < 12; ++q) sub12[(size_t)r * 12 + q] = obs[(size_t)r * kObs + 1 + q];
  return 0;
}

int pyro_magnetization(pyro_ctx *ctx, double *M3) {
  if (!ctx || !M3) return set_error(PYRO_EINVAL, "pyro_magnetization: null argument");
  std::vector<double> obs;
  PYRO_TRY(fetch_obs(ctx, obs));
  for (int r = 0; r < ctx->R; ++r)
    for (int k = 0; k < 3; ++k) {
      const double *s = &obs[(size_t)r * kObs + 1];
      M3[(size_t)r * 3 + k] = ((s[k] + s[3 + k]) + s[6 + k]) + s[9 + k];
    }
  return 0;
}

// q = 0 order parameters from the sublattice sums s_b: the twelve rows of
// Measurement (Pyrochlore.cu:580-624) -- A2 (1), E (2), T1 (3), ferro (3),
// T2 (3) -- divided by N so the ideal ordered state has modulus 1.
int pyro_fold_order_params(const double *sub12, int R, double n_spins, double *op12) {
  if (!sub12 || !op12 || R < 0 || !(n_spins > 0.0))
    return set_error(PYRO_EINVAL, "pyro_fold_order_params: bad argument");
  for (int r = 0; r < R; ++r) fold_order_params(sub12 + (size_t)r * 12, n_spins, op12 + (size_t)r * 12);
  return 0;
}

int pyro_order_params(pyro_ctx *ctx, double *op12) {
  if (!ctx || !op12) return set_error(PYRO_EINVAL, "pyro_order_params: null argument");
  std::vector<double> sub((size_t)ctx->R * 12);
  PYRO_TRY(pyro_sublattice_sums(ctx, sub.data()));
  for (int r = 0; r < ctx->R; ++r)
    PYRO_TRY(pyro_fold_order_params(sub.data() + 12 * (size_t)r, 1,
                                    16.0 * (double)replica_cells(ctx, r), op12 + 12 * (size_t)r));
  return 0;
}

int pyro_energy_incremental(pyro_ctx *ctx, double *E) {
  if (!ctx || !E) return set_error(PYRO_EINVAL, "pyro_energy_incremental: null argument");
  PYRO_TRY(device_guard(ctx));
  std::vector<double> a(ctx->R), d(ctx->R);
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  PYRO_HIP(hipMemcpy(a.data(), ctx->d_anchor, sizeof(double) * ctx->R, hipMemcpyDeviceToHost));
  PYRO_HIP(hipMemcpy(d.data(), ctx->d_dE, sizeof(double) * ctx->R, hipMemcpyDeviceToHost));
  for (int r = 0; r < ctx->R; ++r) E[r] = a[r] + d[r];
  return 0;
}

int pyro_acceptance(pyro_ctx *ctx, uint64_t *accepted, uint64_t *proposed) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  PYRO_TRY(device_guard(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  if (accepted)
    PYRO_HIP(hipMemcpy(accepted, ctx->d_acc, sizeof(uint64_t) * ctx->R, hipMemcpyDeviceToHost));
  if (proposed) std::memcpy(proposed, ctx->h_proposed.data(), sizeof(uint64_t) * ctx->R);
  return 0;
}

int pyro_reset_acceptance(pyro_ctx *ctx) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  PYRO_TRY(device_guard(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  PYRO_HIP(hipMemset(ctx->d_acc, 0, sizeof(unsigned long long) * ctx->R));
  std::fill(ctx->h_proposed.begin(), ctx->h_proposed.end(), 0ull);
  if (ctx->d_acc_learned)
    PYRO_HIP(hipMemset(ctx->d_acc_learned, 0, sizeof(unsigned long long) * ctx->R));
  std::fill(ctx->h_proposed_learned.begin(), ctx->h_proposed_learned.end(), 0ull);
  return 0;
}

int pyro_set_learned_passes(pyro_ctx *ctx, int n_learned) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  if (n_learned < 0 || n_learned > 255) return set_error(PYRO_EINVAL, "n_learned %d", n_learned);
  if (n_learned > 0 && !ctx->mlp_ready)
    return set_error(PYRO_ESTATE, "pyro_set_learned_passes: set or train the network first");
  if (n_learned > 0 && ctx->hetero)
    return set_error(PYRO_ESTATE, "pyro_set_learned_passes: needs a uniform lattice size");
  if (n_learned > 0 && ctx->mat)
    return set_error(PYRO_ESTATE, "pyro_set_learned_passes: the learned proposal is for the scalar model");
  PYRO_TRY(device_guard(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  drop_graph(ctx);  // the captured sweep has the old schedule
  ctx->n_learned = n_learned;
  if (n_learned > 0) ctx->exec_mode = 0;  // learned passes run as per-phase kernels
  return 0;
}

int pyro_set_kernel_options(pyro_ctx *ctx, int phase_variant, int track_or_energy) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  if (phase_variant < 0 || phase_variant > 1)
    return set_error(PYRO_EINVAL, "phase_variant %d not 0/1", phase_variant);
  PYRO_TRY(device_guard(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  drop_graph(ctx);
  ctx->phase_variant = phase_variant;
  ctx->track_or = track_or_energy != 0;
  return 0;
}

int pyro_set_exec_mode(pyro_ctx *ctx, int mode, int block, int n_bins) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  if (mode < 0 || mode > 2) return set_error(PYRO_EINVAL, "exec mode %d not 0/1/2", mode);
  if (mode == 0 && ctx->hetero)
    return set_error(PYRO_EINVAL, "exec mode 0 needs a uniform lattice size");
  if (mode != 0 && ctx->mat)
    return set_error(PYRO_ESTATE, "exec mode %d: exchange matrices run in execution mode 0 only", mode);
  if (mode != 0 && ctx->post_rotate)
    return set_error(PYRO_ESTATE, "exec mode %d: the post-Metropolis rotation runs in execution mode 0 only", mode);
  if (mode != 0 && (ctx->set_schedule || ctx->or_angle))
    return set_error(PYRO_ESTATE, "exec mode %d: the update schedule options run in execution mode 0 only", mode);
  if (mode != 0 && !batch_size_supported(ctx->hetero ? 4 : ctx->L))
    return set_error(PYRO_EINVAL, "exec mode %d needs L in {4,8,12,16,24,32}", mode);
  if (mode != 0 && ctx->n_learned > 0)
    return set_error(PYRO_ESTATE, "learned passes run only in exec mode 0");
  if (mode == 1 && block != 256 && block != 512 && block != 1024)
    return set_error(PYRO_EINVAL, "persistent block size %d not 256/512/1024", block);
  if (mode == 2 && block != 0 && block != 128 && block != 256 && block != 512)
    return set_error(PYRO_EINVAL, "tile size %d not 128/256/512 (0: 256)", block);
  PYRO_TRY(device_guard(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  drop_graph(ctx);
  if (mode == 1 && n_bins != ctx->n_bins) PYRO_TRY(batch_setup(ctx, n_bins));
  ctx->exec_mode = mode;
  if (mode == 1) ctx->batch_block = block;
  if (mode == 2) ctx->tile_cells = block ? block : 256;
  return 0;
}

int pyro_get_exec_mode(pyro_ctx *ctx, int *mode, int *block, int *n_bins) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  if (mode) *mode = ctx->exec_mode;
  if (block) *block = ctx->exec_mode == 2 ? ctx->tile_cells : ctx->batch_block;
  if (n_bins) *n_bins = ctx->n_bins;
  return 0;
}

int pyro_get_layout(pyro_ctx *ctx, int *L, long long *offset) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  for (int r = 0; r < ctx->R; ++r) {
    if (L) L[r] = ctx->h_L[r];
    if (offset) offset[r] = ctx->h_off[r];
  }
  return 0;
}

// Time `sweeps` sweeps (eager launches, HIP events on the ctx stream) in each
// candidate execution, restoring the full state after each trial so tuning
// leaves the chain untouched; keep the fastest. Candidates: per-phase kernels
// (uniform ctx: cell-per-thread and site-per-thread), the persistent sweep
// kernel with 256/512/1024 threads and one replica per block, and the tile
// list. results (optional) receives {mode, block, variant, ms per sweep} rows.
int pyro_autotune(pyro_ctx *ctx, int sweeps, double *results, int *n_results) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  if (sweeps < 1 || sweeps > 1000) return set_error(PYRO_EINVAL, "pyro_autotune: sweeps %d", sweeps);
  if (ctx->n_learned > 0) return set_error(PYRO_ESTATE, "pyro_autotune: learned passes are per-phase only");
  if (ctx->mat) return set_error(PYRO_ESTATE, "pyro_autotune: exchange matrices run in execution mode 0 only");
  if (ctx->post_rotate)
    return set_error(PYRO_ESTATE, "pyro_autotune: the post-Metropolis rotation runs in execution mode 0 only");
  if (ctx->set_schedule || ctx->or_angle)
    return set_error(PYRO_ESTATE, "pyro_autotune: the update schedule options run in execution mode 0 only");
  PYRO_TRY(device_guard(ctx));
  struct Cand { int mode, block, variant, streams, groups; };
  std::vector<Cand> cands;
  if (!ctx->hetero) {
    for (int v : {1, 0})
      for (int n : {1, 2, 3})
        if (n <= ctx->n_streams) cands.push_back({0, 0, v, n, 0});
    if (or_whole_pass(ctx) && ctx->n_streams >= 4) cands.push_back({0, 0, 1, 4, 0});
    // cache-blocked: G groups, S of them in flight at a time (S = G is the
    // stream split above)
    for (int S : {1, 2, 3})
      for (int G : {2, 4, 8, 16, 32})
        if (G > S && G <= ctx->R && S <= ctx->n_streams) cands.push_back({0, 0, 1, S, G});
  }
  if (batch_size_supported(ctx->hetero ? 4 : ctx->L)) {
    if (persistent_candidate(ctx)) {
      cands.push_back({1, 256, 1, 1, 0});
      cands.push_back({1, 512, 1, 1, 0});
      cands.push_back({1, 1024, 1, 1, 0});
    }
    for (int tile : {256, 128, 512})
      for (int n : {1, 2, 3})
        if (n <= ctx->n_streams) cands.push_back({2, tile, 1, n, 0});
    for (int S : {1, 2, 3})
      for (int G : {2, 3, 4, 6, 8})
        if (G > S && G <= ctx->R && S <= ctx->n_streams) cands.push_back({2, 256, 1, S, G});
  }
  if (cands.empty()) return 0;
  if ((int)cands.size() > PYRO_AUTOTUNE_MAX)
    return set_error(PYRO_EINVAL, "pyro_autotune: %d candidates", (int)cands.size());
  // Snapshot of everything a sweep changes (an exchange or measurement
  // inside a trial also re-anchors the incremental energy: anchor and dE).
  const int R = ctx->R;
  float *s_spins = nullptr;
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  PYRO_HIP(hipMalloc((void **)&s_spins, sizeof(float) * ctx->total_floats));
  std::vector<double> temps(R), dE(R), anchor(R);
  std::vector<float> beta(R), cosc(R);
  std::vector<int> at(R);
  std::vector<unsigned long long> acc(R);
  std::vector<uint32_t> ctl(kCtlWords);
  int rc = 0;
  auto snap = [&](bool save) -> int {
    const hipMemcpyKind k = save ? hipMemcpyDeviceToHost : hipMemcpyHostToDevice;
    if (hipMemcpy(save ? (void *)s_spins : (void *)ctx->d_spins,
                  save ? (void *)ctx->d_spins : (void *)s_spins, sizeof(float) * ctx->total_floats,
                  hipMemcpyDeviceToDevice) != hipSuccess ||
        hipMemcpy(save ? (void *)temps.data() : (void *)ctx->d_temps, save ? (void *)ctx->d_temps : (void *)temps.data(), sizeof(double) * R, k) != hipSuccess ||
        hipMemcpy(save ? (void *)dE.data() : (void *)ctx->d_dE, save ? (void *)ctx->d_dE : (void *)dE.data(), sizeof(double) * R, k) != hipSuccess ||
        hipMemcpy(save ? (void *)anchor.data() : (void *)ctx->d_anchor, save ? (void *)ctx->d_anchor : (void *)anchor.data(), sizeof(double) * R, k) != hipSuccess ||
        hipMemcpy(save ? (void *)beta.data() : (void *)ctx->d_beta, save ? (void *)ctx->d_beta : (void *)beta.data(), sizeof(float) * R, k) != hipSuccess ||
        hipMemcpy(save ? (void *)cosc.data() : (void *)ctx->d_cosc, save ? (void *)ctx->d_cosc : (void *)cosc.data(), sizeof(float) * R, k) != hipSuccess ||
        hipMemcpy(save ? (void *)at.data() : (void *)ctx->d_replica_at, save ? (void *)ctx->d_replica_at : (void *)at.data(), sizeof(int) * R, k) != hipSuccess ||
        hipMemcpy(save ? (void *)acc.data() : (void *)ctx->d_acc, save ? (void *)ctx->d_acc : (void *)acc.data(), sizeof(unsigned long long) * R, k) != hipSuccess ||
        hipMemcpy(save ? (void *)ctl.data() : (void *)ctx->d_ctl, save ? (void *)ctx->d_ctl : (void *)ctl.data(), sizeof(uint32_t) * kCtlWords, k) != hipSuccess)
      return set_error(PYRO_EHIP, "pyro_autotune: state snapshot failed");
    return 0;
  };
  rc = snap(true);
  const uint32_t h_sweep = ctx->h_sweep, h_round = ctx->h_pt_round;
  const std::vector<uint64_t> h_prop = ctx->h_proposed;
  const int mode0 = ctx->exec_mode, block0 = ctx->batch_block, variant0 = ctx->phase_variant;
  const int parts0 = ctx->n_parts, tile0 = ctx->tile_cells, groups0 = ctx->block_groups;
  double best = 1e30;
  Cand best_c = {mode0, mode0 == 2 ? tile0 : block0, variant0, parts0, groups0};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = set_error(PYRO_EHIP, "pyro_autotune: event creation failed");
  int nres = 0;
  for (size_t i = 0; i < cands.size() && !rc; ++i) {
    const Cand c = cands[i];
    drop_graph(ctx);
    if (c.mode == 1 && ctx->n_bins != R) rc = batch_setup(ctx, R);
    if (rc) break;
    ctx->exec_mode = c.mode;
    ctx->batch_block = c.mode == 1 ? c.block : block0;
    ctx->tile_cells = c.mode == 2 ? c.block : tile0;
    ctx->phase_variant = c.variant;
    ctx->n_parts = c.streams;
    ctx->block_groups = c.groups;
    // Time the candidate as pyro_sweep runs it: `sweeps` sweeps captured in
    // one graph, replayed once to warm up (code load, caches) and once timed.
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      rc = set_error(PYRO_EHIP, "pyro_autotune: capture failed");
      break;
    }
    rc = enqueue_sweeps(ctx, ctx->h_sweep, (uint32_t)sweeps);
    if (!rc) rc = split_join(ctx, 0);
    ctx->forked = 0;
    ctx->unjoined = 0;
    const hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
    if (!rc && ce != hipSuccess) rc = set_error(PYRO_EHIP, "pyro_autotune: capture failed");
    if (!rc && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
      rc = set_error(PYRO_EHIP, "pyro_autotune: graph instantiation failed");
    float ms = 0.0f;
    if (!rc && (hipGraphLaunch(exec, ctx->stream) != hipSuccess ||
                hipEventRecord(e0, ctx->stream) != hipSuccess ||
                hipGraphLaunch(exec, ctx->stream) != hipSuccess ||
                hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
      rc = set_error(PYRO_EHIP, "pyro_autotune: timing failed");
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (rc) break;
    const double per = ms / sweeps;
    if (results) {
      results[6 * nres + 0] = c.mode;
      results[6 * nres + 1] = c.block;
      results[6 * nres + 2] = c.variant;
      results[6 * nres + 3] = c.streams;
      results[6 * nres + 4] = c.groups;
      results[6 * nres + 5] = per;
    }
    ++nres;
    if (per < best) {
      best = per;
      best_c = c;
    }
    rc = snap(false);
    ctx->h_sweep = h_sweep;
    ctx->h_pt_round = h_round;
    ctx->h_proposed = h_prop;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(s_spins);
  drop_graph(ctx);
  ctx->exec_mode = best_c.mode;
  ctx->batch_block = best_c.mode == 1 ? best_c.block : block0;
  ctx->tile_cells = best_c.mode == 2 ? best_c.block : tile0;
  ctx->phase_variant = best_c.variant;
  ctx->n_parts = best_c.streams;
  ctx->block_groups = best_c.groups;
  if (n_results) *n_results = nres;
  if (rc) {
    ctx->exec_mode = mode0;
    ctx->batch_block = block0;
    ctx->tile_cells = tile0;
    ctx->phase_variant = variant0;
    ctx->n_parts = parts0;
    ctx->block_groups = groups0;
  }
  return rc;
}

int pyro_set_passes(pyro_ctx *ctx, int n_metro, int n_or) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  if (n_metro < 0 || n_or < 0 || n_metro > 255) return set_error(PYRO_EINVAL, "pass counts");
  PYRO_TRY(device_guard(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  drop_graph(ctx);
  ctx->n_metro = n_metro;
  ctx->n_or = n_or;
  return 0;
}

int pyro_adapt_cone(pyro_ctx *ctx, double target, int rounds, int sweeps_per_round) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  if (!(target > 0.0 && target < 1.0) || rounds < 0 || sweeps_per_round < 1)
    return set_error(PYRO_EINVAL, "pyro_adapt_cone: bad target/rounds/sweeps");
  if (ctx->n_metro < 1) return set_error(PYRO_ESTATE, "pyro_adapt_cone: n_metro = 0");
  const int R = ctx->R;
  std::vector<float> cosc(R);
  std::vector<uint64_t> acc(R), prop(R);
  for (int it = 0; it < rounds; ++it) {
    PYRO_TRY(pyro_reset_acceptance(ctx));
    PYRO_TRY(pyro_sweep(ctx, sweeps_per_round, 1));
    PYRO_TRY(pyro_acceptance(ctx, acc.data(), prop.data()));
    PYRO_TRY(pyro_get_cos_cone(ctx, cosc.data()));
    for (int r = 0; r < R; ++r) {
      const double ratio = prop[r] ? (double)acc[r] / (double)prop[r] : 0.0;
      double theta = std::acos(std::max(-1.0, std::min(1.0, (double)cosc[r])));
      theta *= (ratio < target) ? 0.9 : 1.0 / 0.9;
      theta = std::min(theta, M_PI);
      theta = std::max(theta, 1e-4);
      float c = (float)std::cos(theta);
      cosc[r] = std::max(-1.0f, std::min(c, 0.99999994f));
    }
    PYRO_TRY(pyro_set_cos_cone(ctx, cosc.data()));
  }
  PYRO_TRY(pyro_reset_acceptance(ctx));
  return 0;
}

int pyro_record_start(pyro_ctx *ctx, int capacity) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  if (capacity < 0) return set_error(PYRO_EINVAL, "capacity %d < 0", capacity);
  PYRO_TRY(device_guard(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  if (capacity > ctx->record_cap) {
    drop_graph(ctx);  // the captured end-of-sweep kernel holds the old pointer
    if (ctx->d_series) PYRO_HIP(hipFree(ctx->d_series));
    ctx->d_series = nullptr;
    ctx->record_cap = 0;
    PYRO_HIP(hipMalloc((void **)&ctx->d_series, sizeof(double) * (size_t)capacity * ctx->R));
    ctx->record_cap = capacity;
  }
  uint32_t w[3] = {0u, capacity > 0 ? 1u : 0u, (uint32_t)capacity};
  PYRO_HIP(hipMemcpy(ctx->d_ctl + kCtlRecordRow, w, sizeof w, hipMemcpyHostToDevice));
  return 0;
}

int pyro_record_fetch(pyro_ctx *ctx, double *rows, int *n_rows) {
  if (!ctx || !n_rows) return set_error(PYRO_EINVAL, "pyro_record_fetch: null argument");
  PYRO_TRY(device_guard(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  uint32_t row = 0;
  PYRO_HIP(hipMemcpy(&row, ctx->d_ctl + kCtlRecordRow, sizeof row, hipMemcpyDeviceToHost));
  *n_rows = (int)row;
  if (rows && row > 0)
    PYRO_HIP(hipMemcpy(rows, ctx->d_series, sizeof(double) * (size_t)row * ctx->R,
                       hipMemcpyDeviceToHost));
  return 0;
}

int pyro_check_spins(pyro_ctx *ctx, double tol, long long *n_bad, long long *first_bad) {
  if (!ctx || !n_bad) return set_error(PYRO_EINVAL, "pyro_check_spins: null argument");
  if (!(tol >= 0.0)) return set_error(PYRO_EINVAL, "pyro_check_spins: tol %g < 0", tol);
  PYRO_TRY(device_guard(ctx));
  unsigned long long *d = nullptr;
  const unsigned long long init[2] = {0ull, ~0ull};
  PYRO_HIP(hipMalloc((void **)&d, sizeof init));
  unsigned long long out[2] = {0ull, ~0ull};
  int rc = 0;
  if (hipMemcpyAsync(d, init, sizeof init, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
    rc = set_error(PYRO_EHIP, "pyro_check_spins: upload failed");
  if (!rc) rc = check_spins(ctx, tol, d);
  if (!rc && (hipMemcpyAsync(out, d, sizeof out, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
              hipStreamSynchronize(ctx->stream) != hipSuccess))
    rc = set_error(PYRO_EHIP, "pyro_check_spins: download failed");
  (void)hipFree(d);
  if (rc) return rc;
  *n_bad = (long long)out[0];
  if (first_bad) *first_bad = out[0] ? (long long)out[1] : -1;
  return 0;
}

int pyro_set_streams(pyro_ctx *ctx, int n) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  if (n < 1 || n > kMaxParts)
    return set_error(PYRO_EINVAL, "pyro_set_streams: %d streams (1 to %d)", n, kMaxParts);
  if (n > ctx->n_streams)
    return set_error(PYRO_ESTATE, "pyro_set_streams: %d streams need >= %d replicas", n, n);
  PYRO_TRY(device_guard(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  drop_graph(ctx);
  ctx->n_parts = n;
  return 0;
}

int pyro_set_or_fusion(pyro_ctx *ctx, int on) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  if (on < 0 || on > 2) return set_error(PYRO_EINVAL, "pyro_set_or_fusion: level %d not 0/1/2", on);
  PYRO_TRY(device_guard(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  drop_graph(ctx);
  ctx->fuse_or = on;
  return 0;
}

int pyro_set_group_blocking(pyro_ctx *ctx, int groups) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  if (groups < 0 || groups > ctx->R)
    return set_error(PYRO_EINVAL, "pyro_set_group_blocking: %d groups (0 to %d)", groups, ctx->R);
  if (groups >= 2 && ctx->exec_mode == 1)
    return set_error(PYRO_ESTATE, "pyro_set_group_blocking: not in the persistent execution");
  if (groups > kMaxGroups && (ctx->exec_mode == 2 || ctx->hetero))
    return set_error(PYRO_EINVAL, "pyro_set_group_blocking: at most %d groups of tile lists", kMaxGroups);
  PYRO_TRY(device_guard(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  drop_graph(ctx);
  ctx->block_groups = groups >= 2 ? groups : 0;
  return 0;
}

int pyro_get_group_blocking(pyro_ctx *ctx, int *groups) {
  if (!ctx || !groups) return set_error(PYRO_EINVAL, "null argument");
  *groups = ctx->block_groups;
  return 0;
}

int pyro_get_streams(pyro_ctx *ctx, int *n) {
  if (!ctx || !n) return set_error(PYRO_EINVAL, "null argument");
  *n = ctx->n_parts;
  return 0;
}

int pyro_set_post_rotate(pyro_ctx *ctx, int on) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  if (on && (ctx->hetero || ctx->exec_mode != 0))
    return set_error(PYRO_ESTATE, "pyro_set_post_rotate: execution mode 0 on a uniform lattice only");
  PYRO_TRY(device_guard(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  drop_graph(ctx);  // the captured sweeps carry the old flag
  ctx->post_rotate = on != 0;
  return 0;
}

int pyro_set_update_schedule(pyro_ctx *ctx, int independent_sets, int or_random_angle) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  if (independent_sets < 0 || independent_sets > 2)
    return set_error(PYRO_EINVAL, "pyro_set_update_schedule: independent_sets %d not 0/1/2",
                     independent_sets);
  if ((independent_sets || or_random_angle) && (ctx->hetero || ctx->exec_mode != 0))
    return set_error(PYRO_ESTATE, "pyro_set_update_schedule: execution mode 0 on a uniform lattice only");
  PYRO_TRY(device_guard(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  drop_graph(ctx);  // the captured sweeps carry the old schedule
  ctx->set_schedule = independent_sets;
  ctx->or_angle = or_random_angle != 0;
  return 0;
}

int pyro_anneal(pyro_ctx *ctx, double T_start, int steps, int sweeps_per_step) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  if (!(T_start > 0.0) || !std::isfinite(T_start) || steps < 1 || sweeps_per_step < 1)
    return set_error(PYRO_EINVAL, "pyro_anneal: T_start %g, steps %d, sweeps_per_step %d", T_start,
                     steps, sweeps_per_step);
  PYRO_TRY(device_guard(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  // Annealing sweeps have no exchanges and no measurements (the reference
  // runs Mc alone at each step); both cadences come back afterwards.
  const int pt = ctx->pt_every, me = ctx->measure_every;
  ctx->pt_every = 0;
  ctx->measure_every = 0;
  drop_graph(ctx);
  int rc = 0;
  for (int m = 0; m < steps && rc == 0; ++m) {
    rc = launch_anneal_beta(ctx, T_start, m, steps, 0);
    if (rc == 0) rc = pyro_sweep(ctx, sweeps_per_step, 1);
  }
  const int rc2 = launch_anneal_beta(ctx, T_start, steps, steps, 1);
  drop_graph(ctx);
  ctx->pt_every = pt;
  ctx->measure_every = me;
  if (rc) return rc;
  if (rc2) return rc2;
  PYRO_TRY(reanchor(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

int pyro_measure_start(pyro_ctx *ctx, int every, int capacity) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  if (every < 1 || capacity < 1)
    return set_error(PYRO_EINVAL, "pyro_measure_start: every %d / capacity %d must be >= 1", every,
                     capacity);
  const uint32_t a = ctx->pt_every > 0 ? (uint32_t)ctx->pt_every : 1u;
  if (std::lcm(a, (uint32_t)every) > 100000u)
    return set_error(PYRO_EINVAL,
                     "pyro_measure_start: every %d and pt_every %d give a capture block above 1e5 sweeps",
                     every, ctx->pt_every);
  PYRO_TRY(device_guard(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  drop_graph(ctx);  // the captured sweeps have the old cadence and pointers
  if (capacity > ctx->meas_cap) {
    if (ctx->d_meas) PYRO_HIP(hipFree(ctx->d_meas));
    ctx->d_meas = nullptr;
    ctx->meas_cap = 0;
    if (hipMalloc((void **)&ctx->d_meas, sizeof(double) * kMeasQ * (size_t)capacity * ctx->R) !=
        hipSuccess)
      return set_error(PYRO_ENOMEM, "pyro_measure_start: %d rows do not fit", capacity);
    ctx->meas_cap = capacity;
  }
  ctx->measure_every = every;
  const uint32_t w[2] = {0u, (uint32_t)capacity};
  PYRO_HIP(hipMemcpy(ctx->d_ctl + kCtlMeasRow, w, sizeof w, hipMemcpyHostToDevice));
  return 0;
}

int pyro_measure_stop(pyro_ctx *ctx) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  PYRO_TRY(device_guard(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->measure_every != 0) drop_graph(ctx);
  ctx->measure_every = 0;
  return 0;
}

int pyro_measure_fetch(pyro_ctx *ctx, double *rows, int *n_rows) {
  if (!ctx || !n_rows) return set_error(PYRO_EINVAL, "pyro_measure_fetch: null argument");
  PYRO_TRY(device_guard(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  uint32_t row = 0;
  PYRO_HIP(hipMemcpy(&row, ctx->d_ctl + kCtlMeasRow, sizeof row, hipMemcpyDeviceToHost));
  *n_rows = (int)row;
  if (rows && row > 0)
    PYRO_HIP(hipMemcpy(rows, ctx->d_meas, sizeof(double) * kMeasQ * (size_t)row * ctx->R,
                       hipMemcpyDeviceToHost));
  return 0;
}

int pyro_time_phase(pyro_ctx *ctx, int kind, int n_launches, double *avg_ms, double *span_ms) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  if (kind < 0 || kind > 5 || n_launches < 1 || n_launches > 100000)
    return set_error(PYRO_EINVAL, "pyro_time_phase: kind %d / n_launches %d", kind, n_launches);
  if (kind == 4 && !metro_fusable(ctx))
    return set_error(PYRO_ESTATE, "pyro_time_phase: no fused Metropolis pair for this ctx");
  if (kind == 3 && !or_fusable(ctx))
    return set_error(PYRO_ESTATE, "pyro_time_phase: no fused over-relaxation for this ctx");
  if (kind == 5 && !or_whole_pass(ctx))
    return set_error(PYRO_ESTATE, "pyro_time_phase: no whole over-relaxation pass for this ctx");
  PYRO_TRY(device_guard(ctx));
  std::vector<hipEvent_t> ev(2 * (size_t)n_launches, nullptr);
  int rc = 0;
  for (auto &e : ev)
    if (hipEventCreate(&e) != hipSuccess) {
      rc = set_error(PYRO_EHIP, "pyro_time_phase: event creation failed");
      break;
    }
  // The events ride in the kernel's own dispatch (hipExtLaunchKernelGGL), so
  // each pair brackets the kernel's execution and nothing else.
  for (int i = 0; i < n_launches && rc == 0; ++i) {
    ctx->time_ev0 = ev[2 * i];
    ctx->time_ev1 = ev[2 * i + 1];
    if (kind == 3) {  // fused pairs (0,1) and (2,3) from d_spins into the second buffer:
                      // d_spins, the state, is left as it was
      rc = launch_overrelax_pair(ctx, 2 * (i & 1), ctx->d_spins, ctx->d_spins_alt);
    } else if (kind == 5) {  // whole passes from d_spins into the second buffer
      rc = launch_overrelax_pass_whole(ctx, ctx->d_spins, ctx->d_spins_alt);
    } else if (kind == 4) {  // fused Metropolis pairs (0,1), (2,3), in place
      rc = launch_metropolis_pair(ctx, 2 * (i & 1), 0);
    } else {
      rc = kind == 0   ? launch_metropolis_phase(ctx, i & 3, 0, nullptr)
           : kind == 1 ? launch_overrelax_phase(ctx, i & 3)
                       : launch_learned_phase(ctx, i & 3, 0, i >= kColours);
    }
    ctx->time_ev0 = ctx->time_ev1 = nullptr;
  }
  if (rc == 0 && hipStreamSynchronize(ctx->stream) != hipSuccess)
    rc = set_error(PYRO_EHIP, "pyro_time_phase: synchronise failed");
  double sum = 0.0;
  float span = 0.0f;
  for (int i = 0; i < n_launches && rc == 0; ++i) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]) != hipSuccess)
      rc = set_error(PYRO_EHIP, "pyro_time_phase: elapsed time failed");
    sum += ms;
  }
  if (rc == 0 && hipEventElapsedTime(&span, ev[0], ev[2 * (size_t)n_launches - 1]) != hipSuccess)
    rc = set_error(PYRO_EHIP, "pyro_time_phase: elapsed time failed");
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
  if (rc) return rc;
  if (kind == 2) {  // the learned launches' slots into the books (after the timed region)
    PYRO_TRY(launch_learned_fold(ctx));
    PYRO_HIP(hipStreamSynchronize(ctx->stream));
  }
  if (kind == 0)
    for (auto &p : ctx->h_proposed) p += 4ull * (uint64_t)ctx->ncell * (uint64_t)n_launches;
  if (kind == 2)
    for (auto &p : ctx->h_proposed_learned) p += 4ull * (uint64_t)ctx->ncell * (uint64_t)n_launches;
  if (avg_ms) *avg_ms = sum / n_launches;
  if (span_ms) *span_ms = span;
  return 0;
}

int pyro_sync(pyro_ctx *ctx) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  PYRO_TRY(device_guard(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

}  // extern "C"
