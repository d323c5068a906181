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
  kCtlWords = 8,
};

constexpr int kObs = 13;  // energy + 4 sublattices x 3 components
constexpr int kPhaseBlock = 256;   // 4 waves: one per site of the colour
constexpr int kPhaseCells = 64;    // cells per block (one per lane)
constexpr int kObsBlock = 256;     // cells per block in the observable pass

}  // namespace pyro

struct pyro_ctx {
  // configuration
  int L = 0, R = 0;
  int ncell = 0;          // L^3
  int device = 0;
  uint64_t seed = 0;
  float jeff = 1.0f;      // J * spin_len2
  int n_metro = 1, n_or = 0, pt_every = 0;
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
  double *d_partials = nullptr;    // R * obs_blocks * kObs
  double *d_obs = nullptr;         // R * kObs
  uint32_t *d_ctl = nullptr;       // kCtlWords
  double *d_series = nullptr;      // record_cap * R
  int obs_blocks = 0;
  int record_cap = 0;

  // host mirrors
  uint32_t h_sweep = 0, h_pt_round = 0;
  std::vector<uint64_t> h_proposed;  // per replica, Metropolis proposals since reset

  // captured sweep blocks
  hipGraphExec_t graph_exec = nullptr;
  hipGraph_t graph = nullptr;
  int graph_sweeps = 0;
};

namespace pyro {

int set_error(int code, const char *fmt, ...);

// Kernel launchers (pyro_kernels.hip). All enqueue on ctx->stream.
int launch_init_spins(pyro_ctx *ctx);
int launch_metropolis_phase(pyro_ctx *ctx, int colour, int pass, const uint32_t *sweep_override);
int launch_overrelax_phase(pyro_ctx *ctx, int colour);
int launch_observables(pyro_ctx *ctx);           // fills d_obs
int launch_pt_swap(pyro_ctx *ctx, int parity, uint32_t round, int from_ctl);
int launch_end_sweep(pyro_ctx *ctx);             // record row, advance sweep counter
int launch_set_anchor(pyro_ctx *ctx);            // d_anchor = d_obs energy, d_dE = 0
int launch_set_sweep(pyro_ctx *ctx, uint32_t sweep, uint32_t round);

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
