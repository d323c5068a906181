// Internal state shared by the translation units of libpyro.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pyro.h"
#include "pyro_math.h"

namespace pyro {

// Control words in device memory, so that captured graphs can be replayed:
// the kernels read the sweep index and the exchange round from here.
enum CtlWord : int {
  kCtlSweep = 0,      // current sweep index (keys the Metropolis draws)
  kCtlPtRound = 1,    // parallel-tempering round counter (parity = round & 1)
  kCtlSwapsAcc = 2,   // accepted swaps, cumulative
  kCtlRecordRow = 3,  // rows written to the energy series
  kCtlRecordOn = 4,   // 1 while recording
  kCtlRecordCap = 5,  // series capacity in rows
  kCtlMeasRow = 6,    // rows written to the measurement series
  kCtlMeasCap = 7,    // measurement series capacity in rows (0 = off)
  kCtlWords = 8,
};

constexpr int kObs = 13;  // energy + 4 sublattices x 3 components
constexpr int kObsPartials = 16;
constexpr int kCoupFloats = 156;  // matrix field model: 16 x 9 exchange + 4 x 3 Zeeman
// Replica groups of a split sweep (one stream each), and the workgroups each
// group's launch must span for the split to be chosen by default.
constexpr int kMaxParts = 4;
constexpr long long kSplitMinBlocks = 256;
// Replica groups of a cache-blocked sweep in the tile-list execution (one
// prebuilt tile list per group; any number up to R in execution mode 0).
constexpr int kMaxGroups = 8;
// Measurement row, per rung: E, |m_A2|, |m_E|, |m_T1|, |m_F| (= |M| / N), |m_T2|.
constexpr int kMeasQ = 6;  // per observables block: 4 bases x (energy, mx, my, mz)
constexpr int kPhaseBlock = 256;   // 4 waves: one per site of the colour
constexpr int kPhaseCells = 64;    // cells per block (one per lane)
constexpr int kObsBlock = 256;     // cells per block in the observable pass

// Learned proposal network 6 -> 64 -> 64 -> 4 (ReLU), packed FP32 parameters
// W1[64][6] b1[64] W2[64][64] b2[64] W3[4][64] b3[4].
constexpr int kMlpIn = 6, kMlpH = 64, kMlpOut = 4;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpH * kMlpIn;
constexpr int kOffW2 = kOffB1 + kMlpH;
constexpr int kOffB2 = kOffW2 + kMlpH * kMlpH;
constexpr int kOffW3 = kOffB2 + kMlpH;
constexpr int kOffB3 = kOffW3 + kMlpOut * kMlpH;
constexpr int kMlpParams = kOffB3 + kMlpOut;  // 4868
// bf16 operand images: W1|b1 as [64][8] (column 6 = bias, 7 = 0), W2 [64][64],
// W3 padded to [32][64] (rows 4..31 zero); then b2 as 64 floats (two entries
// each, 16-byte aligned), already scaled as the forward pass uses it.
constexpr int kPackP1 = 0;
constexpr int kPackP2 = kPackP1 + 64 * 8;
constexpr int kPackP3 = kPackP2 + 64 * 64;
constexpr int kPackB2 = kPackP3 + 32 * 64;
constexpr int kPackSize = kPackB2 + 2 * 64;
static_assert(kPackB2 % 8 == 0, "b2 floats 16-byte aligned");
// Per-sample training buffers, [unit][batch]: x(6) a1(64) a2(64) d3(4) d2(64) d1(64) loss(1)
constexpr float kLogKappaMin = -4.60517019f;  // ln 0.01
constexpr float kLogKappaMax = 8.31776617f;   // ln 4096


// Access at a 32-bit byte offset from a base that is uniform over the wave
// (a replica's component plane): the base stays in scalar registers and the
// global load or store takes the offset as one 32-bit vector register, with
// no 64-bit address arithmetic per access. Offsets stay below 2^31: a
// replica of L <= 64 spans 48 L^3 floats = 50 MB.
__device__ __forceinline__ float ld_at(const float *__restrict__ base, uint32_t off) {
  return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + off);
}
__device__ __forceinline__ void st_at(float *__restrict__ base, uint32_t off, float v) {
  *reinterpret_cast<float *>(reinterpret_cast<char *>(base) + off) = v;
}

// One block's partial sums of a learned phase (see k_learned_phase).
struct LearnedSlot {
  double de;
  unsigned long long acc;
};

}  // namespace pyro

struct pyro_ctx {
  // configuration
  int L = 0, R = 0;
  int ncell = 0;          // L^3
  int device = 0;
  uint64_t seed = 0;
  float jeff = 1.0f;      // J * spin_len2
  int n_metro = 1, n_or = 0, pt_every = 0;
  int phase_variant = 1;   // 0: wave per site (any L); 1: cell per thread (compile-time L)
  bool track_or = false;   // FP64 energy bookkeeping also in over-relaxation phases
  // When set, the next phase launch records these events as part of its
  // dispatch (hipExtLaunchKernelGGL), timing the kernel itself.
  hipEvent_t time_ev0 = nullptr, time_ev1 = nullptr;
  hipStream_t stream = nullptr;

  // device state
  float *d_spins = nullptr;        // R * 48 * ncell
  double *d_temps = nullptr;       // R, temperature held by replica r
  float *d_beta = nullptr;         // R, 1/T as float (what the updates use)
  float *d_cosc = nullptr;         // R, cosine of the proposal cone
  int *d_replica_at = nullptr;     // R, replica holding rung k (descending T)
  unsigned long long *d_acc = nullptr;  // R, accepted Metropolis proposals
  double *d_dE = nullptr;          // R, accumulated energy change since the anchor
  double *d_anchor = nullptr;      // R, full energy at the anchor
  double *d_partials = nullptr;    // R * obs_blocks * kObsPartials
  double *d_obs = nullptr;         // R * kObs
  uint32_t *d_ctl = nullptr;       // kCtlWords
  double *d_series = nullptr;      // record_cap * R
  int obs_blocks = 0;
  int record_cap = 0;

  // Replica geometry. Uniform ctx: every replica has L; a batch ctx mixes
  // sizes, replica r's spins start at float offset h_off[r].
  bool hetero = false;
  std::vector<int> h_L;
  std::vector<long long> h_off;
  size_t total_floats = 0;
  int *d_L = nullptr;
  long long *d_off = nullptr;
  // Exchange groups (equal L): rungs h_group_off[g] .. +h_group_len[g] of d_replica_at.
  std::vector<int> h_group_off, h_group_len;
  // Execution mode of a sweep: 0 per-phase kernels (uniform only), 1 one
  // persistent kernel per sweep (a block owns whole replicas), 2 per-phase
  // kernels over a work list of 256-cell tiles of all replicas.
  int exec_mode = 0;
  int batch_block = 256;            // block size of the persistent kernel
  int tile_cells = 256;             // cells per tile (= threads per block) of the tile lists
  int *d_order = nullptr;           // replicas in bin order
  int *d_bin_start = nullptr;       // n_bins + 1 offsets into d_order
  int n_bins = 0;
  int *d_tiles = nullptr;           // 2 ints per tile: replica, tile index (-1: padding)
  int n_tiles = 0;
  int *d_part_off = nullptr;        // R + 1: first observables partial of each replica
  double *d_tile_part = nullptr;    // kObs doubles per tile: observables partials
  // The same lists for each replica group of a split or cache-blocked sweep:
  // d_tiles_part[n][p] covers group p of n (replicas [p R / n, (p + 1) R / n)).
  int *d_tiles_part[pyro::kMaxGroups + 1][pyro::kMaxGroups] = {};
  int n_tiles_part[pyro::kMaxGroups + 1][pyro::kMaxGroups] = {};
  // The same lists with 128-cell (index 0) and 512-cell (index 1) tiles, for
  // the phases of the tile-list execution (the observables keep 256).
  int *d_tiles_x[2] = {};
  int n_tiles_x[2] = {};
  int *d_tiles_part_x[2][pyro::kMaxGroups + 1][pyro::kMaxGroups] = {};
  int n_tiles_part_x[2][pyro::kMaxGroups + 1][pyro::kMaxGroups] = {};

  // learned proposal
  int n_learned = 0;                    // learned passes per sweep (after Metropolis)
  float *d_mlp_w = nullptr;              // packed FP32 master weights (kMlpParams)
  float *d_mlp_m = nullptr, *d_mlp_v = nullptr;  // Adam moments
  uint16_t *d_mlp_pack = nullptr;        // bf16 operand images for the MFMA forward
  float *d_grad = nullptr;               // kMlpParams gradient + 1 loss slot
  float *d_env = nullptr;                // environments (u = beta h, s), 6 floats each
  int env_cap = 0, env_n = 0;
  float *d_train = nullptr;              // per-block partial gradients of a batch
  int train_cap = 0;
  int adam_t = 0;
  bool mlp_ready = false;
  unsigned long long *d_acc_learned = nullptr;  // R, accepted learned proposals
  // Per-block partial sums of the learned phases (energy change, accepted
  // count): learned_nslots per replica and colour, moved into d_dE /
  // d_acc_learned by launch_learned_fold after the learned passes of a sweep.
  pyro::LearnedSlot *d_learned_slots = nullptr;
  int learned_nslots = 0;
  std::vector<uint64_t> h_proposed_learned;

  // local-field model: scalar jeff, or (mat) exchange matrices and Zeeman
  // fields per basis site, kCoupFloats floats scaled by spin_len2 (d_coup;
  // h_coup the same values on the host)
  double spin_len2 = 1.0;
  bool mat = false;
  bool post_rotate = false;  // rotate about h after each Metropolis test
  // Update schedule: set_schedule 1 draws each phase's independent set from
  // the 24, 2 runs the colour classes of each pass in a drawn order (0: the
  // colour classes in order); or_angle makes over-relaxation a rotation about
  // h by a random angle (else by pi).
  int set_schedule = 0;
  bool or_angle = false;

  // Sweeps split into n_parts replica groups, each on its own stream (one
  // group's kernel ramps up while another's drains): group p holds replicas
  // [p R / n_parts, (p + 1) R / n_parts) and runs on part_stream[p], where
  // part_stream[0] is the ctx stream and the others exist for p < n_streams.
  // The launch range below (and launch_part) is set by the sweep driver
  // around each group and is empty otherwise.
  int n_parts = 1;
  int n_streams = 1;
  // Cache blocking: with block_groups >= 2 the sweeps of each exchange or
  // measurement interval run group after group on the ctx stream, group p
  // (replicas [p R / G, (p + 1) R / G)) through all of the interval's sweeps
  // before group p + 1 starts, so that one group's two spin buffers stay in
  // the Infinity Cache across the interval. 0: off (n_parts decides).
  int block_groups = 0;
  hipStream_t part_stream[pyro::kMaxParts] = {};
  hipEvent_t ev_fork = nullptr;
  hipEvent_t ev_join[pyro::kMaxParts] = {};
  hipStream_t launch_stream = nullptr;
  int launch_r0 = 0, launch_nrep = 0, launch_part = 0;
  // Split sweeps join the other streams only at exchanges, measurements and
  // the end of a call: `forked` while they run ahead, `unjoined` sweeps since
  // the fork. Phase kernels key their draws on ctl[kCtlSweep] + launch_delta,
  // because the counter itself advances only at the join.
  int forked = 0;
  uint32_t unjoined = 0;
  uint32_t launch_delta = 0;
  // Series hook offered to the next over-relaxation launch of a split sweep;
  // launch_series_done is set if the launched kernel took it.
  double *launch_series = nullptr;
  int launch_series_done = 0;
  // Second spin buffer for the fused over-relaxation passes (k_or_pair): a
  // pass reads one buffer and writes the other, so passes alternate and an
  // even number of them ends back in d_spins. nullptr: no fusion.
  // fuse_or: 0 one colour per launch, 1 two, 2 a whole pass where a kernel
  // for it exists (L = 16, untracked over-relaxation), else two.
  float *d_spins_alt = nullptr;
  int fuse_or = 2;
  float *d_coup = nullptr;
  float h_coup[pyro::kCoupFloats] = {};

  // measurement series: every measure_every sweeps one row of kMeasQ values
  // per rung (descending T) is appended on the device
  int measure_every = 0;
  double *d_meas = nullptr;
  int meas_cap = 0;

  // host mirrors
  uint32_t h_sweep = 0, h_pt_round = 0;
  std::vector<uint64_t> h_proposed;  // per replica, Metropolis proposals since reset

  // captured sweep blocks
  hipGraphExec_t graph_exec = nullptr;
  hipGraph_t graph = nullptr;
  int graph_sweeps = 0;
};

namespace pyro {

// The replicas and stream of the next phase launch: all replicas on the ctx
// stream, unless a sweep split into replica groups on several streams set a
// range.
static inline hipStream_t launch_stream(const pyro_ctx *c) {
  return c->launch_stream ? c->launch_stream : c->stream;
}
static inline int launch_r0(const pyro_ctx *c) { return c->launch_nrep ? c->launch_r0 : 0; }
static inline int launch_nrep(const pyro_ctx *c) { return c->launch_nrep ? c->launch_nrep : c->R; }

int set_error(int code, const char *fmt, ...);

// Kernel launchers (pyro_kernels.hip). All enqueue on ctx->stream, except
// that the phase launchers and launch_record_part follow the launch range
// (launch_stream / launch_r0 / launch_nrep) while a split sweep sets one.
int launch_init_spins(pyro_ctx *ctx);
int launch_metropolis_phase(pyro_ctx *ctx, int colour, int pass, const uint32_t *sweep_override);
int launch_overrelax_phase(pyro_ctx *ctx, int colour, int pass = 0);
// Whether the over-relaxation passes of a sweep run fused (two colours per
// launch between the spin buffers), and one fused pass: parity 0 reads
// d_spins and writes d_spins_alt, parity 1 the reverse.
bool or_fusable(const pyro_ctx *ctx);
// Metropolis colour pairs (0,1), (2,3) in one launch each, in place: L = 16,
// scalar model, cell kernels, no post-Metropolis rotation, fusion on.
bool metro_fusable(const pyro_ctx *ctx);
int launch_metropolis_pair(pyro_ctx *ctx, int ca, int pass);
int launch_overrelax_pass_fused(pyro_ctx *ctx, int parity);
// A whole over-relaxation pass per launch (k_or_pass_plane): fusion level 2
// on an or_fusable ctx at L = 16 without over-relaxation energy tracking.
bool or_whole_pass(const pyro_ctx *ctx);
int launch_overrelax_pass_whole(pyro_ctx *ctx, const float *cur, float *oth);
int launch_overrelax_pair(pyro_ctx *ctx, int ca, const float *cur, float *oth);
// Fills d_obs; with anchor, also re-anchors the incremental energy (anchor = E, dE = 0).
int launch_observables(pyro_ctx *ctx, bool anchor = false);
// end_n > 0: the exchange also ends an interval of end_n sweeps (series row
// of the last, sweep counter), as launch_end_sweep(ctx, end_n, 1) after it.
int launch_pt_swap(pyro_ctx *ctx, int parity, uint32_t round, int from_ctl, uint32_t end_n = 0);
int launch_end_sweep(pyro_ctx *ctx, uint32_t n, int record_last);  // series row, sweep counter
int launch_record_part(pyro_ctx *ctx);           // series entries of the launch range
int launch_measure_row(pyro_ctx *ctx);           // append a measurement row from d_obs
// Matrix field model; kind 0 Metropolis, 1 over-relaxation by pi, 2 by a random angle.
int launch_phase_mat(pyro_ctx *ctx, int kind, int colour, int pass);
int reset_model_state(pyro_ctx *ctx);  // drop captured sweeps, re-anchor the energy
int check_spins(pyro_ctx *ctx, double tol, unsigned long long *d_out);  // d_out: {count, first}
int launch_observables_partials_mat(pyro_ctx *ctx);
int launch_anneal_beta(pyro_ctx *ctx, double T0, int m, int steps, int restore);
int launch_set_anchor(pyro_ctx *ctx);            // d_anchor = d_obs energy, d_dE = 0
int launch_set_sweep(pyro_ctx *ctx, uint32_t sweep, uint32_t round);
// Address of the no-allocation scratch array on `device` (the current device).
int noalloc_scratch(int device, double **ptr);

// Batches and alternative sweep executions (pyro_batch.hip).
bool batch_size_supported(int L);
bool persistent_candidate(const pyro_ctx *ctx);  // autotuner tries the persistent execution
int batch_setup(pyro_ctx *ctx, int n_bins);       // bins, tiles, device geometry arrays
int launch_batch_init_spins(pyro_ctx *ctx);
int launch_batch_observables(pyro_ctx *ctx);      // fills d_obs
int launch_batch_sweep_body(pyro_ctx *ctx);       // Metropolis + OR passes of one sweep
// Persistent execution: n_sweeps sweeps in one launch, with the series rows of
// all but the last (and the last too if record_last).
int launch_batch_sweeps(pyro_ctx *ctx, int n_sweeps, int record_last);
void batch_free(pyro_ctx *ctx);

// Learned proposal (pyro_learned.hip).
// accumulate = 0: the launch's per-block slots of this colour are written
// (the first learned pass since the last fold); 1: added to.
int launch_learned_phase(pyro_ctx *ctx, int colour, int pass, int accumulate);
// Adds the learned phases' per-block slots to d_dE and d_acc_learned (launch range).
int launch_learned_fold(pyro_ctx *ctx);
int mlp_alloc(pyro_ctx *ctx);
int mlp_pack(pyro_ctx *ctx);
void mlp_free(pyro_ctx *ctx);

}  // namespace pyro

#define PYRO_HIP(call)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return ::pyro::set_error(PYRO_EHIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

#define PYRO_TRY(expr)       \
  do {                       \
    int rc_ = (expr);        \
    if (rc_ != 0) return rc_; \
  } while (0)
