// gfx950 kernels of the pyrochlore Heisenberg Monte Carlo step.
//
// Colour-phase kernels. A colour c is the independent set {c, c+4, c+8, c+12}
// of every cell (Pyrochlore.cu:36-43: no two sites of a class are bonded), so
// all N/4 updates of a phase are independent. A block is 4 waves of 64 lanes;
// wave k updates site c + 4k of 64 consecutive cells, so every wave reads and
// writes one contiguous 256-byte run per component plane, and the 6 neighbour
// gathers of a wave hit the same or adjacent runs (cell +-1, +-L, +-L^2 of the
// other 12 site planes), which the 4 waves of a block share through L1/L2.
// Random numbers are drawn in registers (Philox4x32-10, one block function
// per update), so the reference's pre-filled uniform buffer and its HBM
// round trip (MonteCarlo.cu:3-13, 16 B per spin per sweep) are gone.
#include <hip/hip_ext.h>

#include "pyro_cell.h"
#include "pyro_internal.h"

namespace pyro {

__constant__ Link c_links[16][6] = PYRO_LINKS16_INIT;

__device__ __forceinline__ int wrap(int v, int L) {
  return v < 0 ? v + L : (v >= L ? v - L : v);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// One colour phase on every replica. KIND 0: Metropolis with the cone
// proposal (MCMultipleTemperaturesPerSystem, Pyrochlore.cu:829-885).
// KIND 1: over-relaxation by pi (OverRelaxationByPiStep,
// Pyrochlore.cu:975-1007). grid = (ceil(ncell/64), R), block = 256.
// LT > 0 fixes the lattice size at compile time (index arithmetic by
// constants); LT = 0 reads it from L.
template <int KIND, int LT>
__global__ __launch_bounds__(kPhaseBlock) void k_phase(
    float *__restrict__ spins, int L_rt, int ncell_rt, int colour, int pass,
    const float *__restrict__ beta, const float *__restrict__ cosc,
    const uint32_t *__restrict__ ctl, uint32_t key0, uint32_t key1, float jeff,
    unsigned long long *__restrict__ acc, double *__restrict__ dE) {
  const int L = LT > 0 ? LT : L_rt;
  const int ncell = LT > 0 ? LT * LT * LT : ncell_rt;
  const int r = blockIdx.y;
  const int k = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int cell = blockIdx.x * kPhaseCells + lane;
  const int site = colour + 4 * k;
  const size_t plane = (size_t)16 * ncell;
  float *Sx = spins + (size_t)r * 3 * plane;
  float *Sy = Sx + plane;
  float *Sz = Sy + plane;

  double de = 0.0;
  int accepted = 0;
  if (cell < ncell) {
    const int LL = L * L;
    const int z = cell / LL;
    const int rem = cell - z * LL;
    const int y = rem / L;
    const int x = rem - y * L;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
    for (int n = 0; n < 6; ++n) {
      const Link lk = c_links[site][n];
      const int x1 = wrap(x + lk.dx, L);
      const int y1 = wrap(y + lk.dy, L);
      const int z1 = wrap(z + lk.dz, L);
      const int j = lk.site * ncell + (z1 * L + y1) * L + x1;
      const float a = Sx[j], b = Sy[j], c = Sz[j];
      fx = fx + a;
      fy = fy + b;
      fz = fz + c;
      gx += (double)a;
      gy += (double)b;
      gz += (double)c;
    }
    const int i = site * ncell + cell;
    const f3 s = {Sx[i], Sy[i], Sz[i]};
    f3 sp;
    bool write;
    if (KIND == 0) {
      const u32x4 w = philox4x32_10(
          rng_counter((uint32_t)cell, (uint32_t)site, (uint32_t)pass, kKindMetropolis,
                      (uint32_t)r, ctl[kCtlSweep]),
          key0, key1);
      const f3 h = {-jeff * fx, -jeff * fy, -jeff * fz};
      sp = cone_proposal(s, cosc[r], u01(w.x), u01(w.y));
      write = metropolis_accept(h, s, sp, beta[r], u01(w.z));
      accepted = write ? 1 : 0;
    } else {
      const f3 f = {fx, fy, fz};
      sp = overrelax(f, s);
      write = true;
    }
    if (write) {
      de = (double)jeff * (((double)sp.x - (double)s.x) * gx + ((double)sp.y - (double)s.y) * gy +
                           ((double)sp.z - (double)s.z) * gz);
      Sx[i] = sp.x;
      Sy[i] = sp.y;
      Sz[i] = sp.z;
    }
  }

  __shared__ double s_de[4];
  __shared__ unsigned s_acc[4];
  de = wave_sum(de);
  unsigned nacc = 0;
  if (KIND == 0) nacc = (unsigned)__popcll(__ballot(accepted));
  if (lane == 0) {
    s_de[k] = de;
    s_acc[k] = nacc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double t = (s_de[0] + s_de[1]) + (s_de[2] + s_de[3]);
    if (t != 0.0) atomicAdd(&dE[r], t);
    if (KIND == 0) {
      const unsigned a = s_acc[0] + s_acc[1] + s_acc[2] + s_acc[3];
      if (a) atomicAdd(&acc[r], (unsigned long long)a);
    }
  }
}

// ---------------------------------------------------------------------------
// Cell-per-thread colour phase (pyro_cell.h) with the colour and the lattice
// size fixed at compile time. Same Philox counters and float expressions as
// k_phase, hence the same results bit for bit.
//
// Blocks are remapped so the hardware's round-robin dispatch over the 8 XCDs
// (block b runs on XCD b % 8) gives each XCD a contiguous run of cells: the
// z-1 / z+1 neighbour planes of a block then sit in the same XCD's L2.
// grid = 1-D, n_blocks = R * ceil(ncell / 256), block = 256; a block covers
// cells of one replica only, so one atomic per block.
// ---------------------------------------------------------------------------
template <int KIND, int COLOUR, int LT, bool TRACK>
__global__ __launch_bounds__(256) void k_phase_cells(
    float *__restrict__ spins, int blocks_per_replica, int pass, const float *__restrict__ beta,
    const float *__restrict__ cosc, const uint32_t *__restrict__ ctl, uint32_t key0,
    uint32_t key1, float jeff, unsigned long long *__restrict__ acc, double *__restrict__ dE) {
  constexpr int NC = LT * LT * LT;
  const int nb = gridDim.x;
  const int b = blockIdx.x;
  const int lb = (nb & 7) == 0 ? (b & 7) * (nb >> 3) + (b >> 3) : b;
  const int r = lb / blocks_per_replica;
  const int cell = (lb - r * blocks_per_replica) * 256 + threadIdx.x;
  double de = 0.0;
  unsigned nacc = 0;
  if (cell < NC)
    cell_update<KIND, COLOUR, LT, TRACK>(spins + (size_t)r * 48 * NC, cell, r, pass,
                                         KIND == 0 ? beta[r] : 0.0f, KIND == 0 ? cosc[r] : 0.0f,
                                         KIND == 0 ? ctl[kCtlSweep] : 0u, key0, key1, jeff, de,
                                         nacc);
  if (KIND == 0 || TRACK) {
    block_sum<256>(de, nacc);
    if (threadIdx.x == 0) {
      if (de != 0.0) atomicAdd(&dE[r], de);
      if (KIND == 0 && nacc) atomicAdd(&acc[r], (unsigned long long)nacc);
    }
  }
}

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

// Per-block FP64 partial sums of the energy and the sublattice-b magnetisation
// (TotalEnergyGPU, Pyrochlore.cu:344-389; MagnetizationGPU, :277-342;
// Measurement, :526-578). grid = (obs_blocks, 4, R): blockIdx.y is the basis
// index b, and a thread sums the 4 sites 4t + b of its cell with all 72
// neighbour loads independent, so a launch has 4x the waves of a cell-per-thread
// layout and no dependent load chains. partials[((r * 4 + b) * nblk + blk) * 4 + q],
// q = (energy, mx, my, mz).
template <int B>
__device__ __forceinline__ void obs_cell(const float *__restrict__ Sx, const float *__restrict__ Sy,
                                         const float *__restrict__ Sz, int L, int ncell, int cell,
                                         double v[4]) {
  const int LL = L * L;
  const int z = cell / LL;
  const int rem = cell - z * LL;
  const int y = rem / L;
  const int x = rem - y * L;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int site = 4 * t + B;
    double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
    for (int n = 0; n < 6; ++n) {
      const Link lk = kLinks[site][n];
      const int x1 = wrap(x + lk.dx, L);
      const int y1 = wrap(y + lk.dy, L);
      const int z1 = wrap(z + lk.dz, L);
      const size_t j = (size_t)lk.site * ncell + (size_t)((z1 * L + y1) * L + x1);
      gx += (double)Sx[j];
      gy += (double)Sy[j];
      gz += (double)Sz[j];
    }
    const size_t i = (size_t)site * ncell + cell;
    const double sx = Sx[i], sy = Sy[i], sz = Sz[i];
    v[0] += sx * gx + sy * gy + sz * gz;
    v[1] += sx;
    v[2] += sy;
    v[3] += sz;
  }
}

__global__ __launch_bounds__(kObsBlock) void k_observables(const float *__restrict__ spins,
                                                           int L, int ncell,
                                                           double *__restrict__ partials) {
  const int b = blockIdx.y;
  const int r = blockIdx.z;
  const int cell = blockIdx.x * kObsBlock + threadIdx.x;
  const size_t plane = (size_t)16 * ncell;
  const float *Sx = spins + (size_t)r * 3 * plane;
  const float *Sy = Sx + plane;
  const float *Sz = Sy + plane;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (cell < ncell) {
    switch (b) {
      case 0: obs_cell<0>(Sx, Sy, Sz, L, ncell, cell, v); break;
      case 1: obs_cell<1>(Sx, Sy, Sz, L, ncell, cell, v); break;
      case 2: obs_cell<2>(Sx, Sy, Sz, L, ncell, cell, v); break;
      default: obs_cell<3>(Sx, Sy, Sz, L, ncell, cell, v); break;
    }
  }
  __shared__ double s_v[4][4];
  const int wv = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double t = wave_sum(v[q]);
    if (lane == 0) s_v[wv][q] = t;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int q = threadIdx.x;
    const double t = (s_v[0][q] + s_v[1][q]) + (s_v[2][q] + s_v[3][q]);
    partials[(((size_t)r * 4 + b) * gridDim.x + blockIdx.x) * 4 + q] = t;
  }
}

// Fixed-order second stage: obs[r][0] = jeff/2 * sum over b and blocks of the
// energy partials (each bond was visited from both ends); obs[r][1 + 3b + c] =
// sum over blocks of component c of sublattice b. grid = R, block = 256.
__global__ __launch_bounds__(256) void k_observables_final(const double *__restrict__ partials,
                                                           int nblk, float jeff,
                                                           double *__restrict__ obs) {
  const int r = blockIdx.x;
  __shared__ double s_v[256];
  for (int q = 0; q < kObs; ++q) {
    double t = 0.0;
    if (q == 0) {
      for (int k = threadIdx.x; k < 4 * nblk; k += 256)
        t += partials[((size_t)r * 4 * nblk + k) * 4];
    } else {
      const int b = (q - 1) / 3, c = (q - 1) % 3;
      for (int k = threadIdx.x; k < nblk; k += 256)
        t += partials[(((size_t)r * 4 + b) * nblk + k) * 4 + 1 + c];
    }
    s_v[threadIdx.x] = t;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) s_v[threadIdx.x] += s_v[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) obs[(size_t)r * kObs + q] = (q == 0) ? 0.5 * (double)jeff * s_v[0] : s_v[0];
    __syncthreads();
  }
}

// Parallel-tempering exchange (TrySwapingTemperatures, MonteCarlo.cu:105-172)
// on the energies in obs. Rungs are ordered by descending temperature;
// pairs (c, c+1), c = parity, parity+2, ... One thread: R is small and the
// pairs are visited in order, as in the reference.
// A batch of mixed lattice sizes exchanges within groups of equal size: R is
// the group's rung count, replica_at its rungs, `group` keys its draws, and
// the exchange-round counter advances after the last group.
__global__ void k_pt_swap(int R, const double *__restrict__ obs, double *__restrict__ temps,
                          float *__restrict__ beta, float *__restrict__ cosc,
                          int *__restrict__ replica_at, uint32_t *__restrict__ ctl,
                          uint32_t key0, uint32_t key1, int parity, uint32_t round,
                          int from_ctl, int group, int last_group) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (from_ctl) {
    round = ctl[kCtlPtRound];
    parity = (int)(round & 1u);
  }
  int nacc = 0;
  for (int c = parity ? 1 : 0; c + 1 < R; c += 2) {
    const int t1 = replica_at[c], t2 = replica_at[c + 1];
    const double T1 = temps[t1], T2 = temps[t2];
    const double arg = (1.0 / T1 - 1.0 / T2) * (obs[(size_t)t1 * kObs] - obs[(size_t)t2 * kObs]);
    const u32x4 w = philox4x32_10(
        rng_counter((uint32_t)c, 0, 0, kKindSwap, (uint32_t)group, round), key0, key1);
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
  ctl[kCtlSwapsAcc] += (uint32_t)nacc;
  if (from_ctl && last_group) ctl[kCtlPtRound] = round + 1u;
}

// End of a sweep: append one row (incremental energy per rung) to the series
// when recording, then advance the sweep counter. grid = 1, block = 256.
__global__ __launch_bounds__(256) void k_end_sweep(int R, uint32_t *__restrict__ ctl,
                                                   const double *__restrict__ anchor,
                                                   const double *__restrict__ dE,
                                                   const int *__restrict__ replica_at,
                                                   double *__restrict__ series) {
  const uint32_t on = ctl[kCtlRecordOn];
  const uint32_t row = ctl[kCtlRecordRow];
  const uint32_t cap = ctl[kCtlRecordCap];
  const bool rec = on && series != nullptr && row < cap;
  if (rec) {
    for (int k = threadIdx.x; k < R; k += blockDim.x) {
      const int rep = replica_at[k];
      series[(size_t)row * R + k] = anchor[rep] + dE[rep];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (rec) ctl[kCtlRecordRow] = row + 1u;
    ctl[kCtlSweep] = ctl[kCtlSweep] + 1u;
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
static inline uint32_t key_lo(const pyro_ctx *c) { return (uint32_t)c->seed; }
static inline uint32_t key_hi(const pyro_ctx *c) { return (uint32_t)(c->seed >> 32); }

int launch_init_spins(pyro_ctx *ctx) {
  if (ctx->hetero) return launch_batch_init_spins(ctx);
  dim3 grid((ctx->ncell + 255) / 256, ctx->R);
  k_init_spins<<<grid, 256, 0, ctx->stream>>>(ctx->d_spins, ctx->ncell, key_lo(ctx), key_hi(ctx));
  PYRO_HIP(hipGetLastError());
  return 0;
}

template <int KIND, int LT>
static void launch_phase_t(pyro_ctx *ctx, int colour, int pass) {
  dim3 grid((ctx->ncell + kPhaseCells - 1) / kPhaseCells, ctx->R);
  if (ctx->time_ev0) {
    hipExtLaunchKernelGGL(k_phase<KIND, LT>, grid, dim3(kPhaseBlock), 0, ctx->stream, ctx->time_ev0,
                          ctx->time_ev1, 0, ctx->d_spins, ctx->L, ctx->ncell, colour, pass,
                          (const float *)ctx->d_beta, (const float *)ctx->d_cosc,
                          (const uint32_t *)ctx->d_ctl, key_lo(ctx), key_hi(ctx), ctx->jeff,
                          ctx->d_acc, ctx->d_dE);
    return;
  }
  k_phase<KIND, LT><<<grid, kPhaseBlock, 0, ctx->stream>>>(
      ctx->d_spins, ctx->L, ctx->ncell, colour, pass, ctx->d_beta, ctx->d_cosc, ctx->d_ctl,
      key_lo(ctx), key_hi(ctx), ctx->jeff, ctx->d_acc, ctx->d_dE);
}

template <int KIND, int COLOUR, int LT, bool TRACK>
static void launch_cells_t(pyro_ctx *ctx, int pass) {
  const int bpr = (LT * LT * LT + 255) / 256;
  if (ctx->time_ev0) {
    hipExtLaunchKernelGGL(k_phase_cells<KIND, COLOUR, LT, TRACK>, dim3(bpr * ctx->R), dim3(256), 0,
                          ctx->stream, ctx->time_ev0, ctx->time_ev1, 0, ctx->d_spins, bpr, pass,
                          (const float *)ctx->d_beta, (const float *)ctx->d_cosc,
                          (const uint32_t *)ctx->d_ctl, key_lo(ctx), key_hi(ctx), ctx->jeff,
                          ctx->d_acc, ctx->d_dE);
    return;
  }
  k_phase_cells<KIND, COLOUR, LT, TRACK><<<bpr * ctx->R, 256, 0, ctx->stream>>>(
      ctx->d_spins, bpr, pass, ctx->d_beta, ctx->d_cosc, ctx->d_ctl, key_lo(ctx), key_hi(ctx),
      ctx->jeff, ctx->d_acc, ctx->d_dE);
}

template <int KIND, int LT>
static void launch_cells_l(pyro_ctx *ctx, int colour, int pass) {
  const bool track = KIND == 1 && ctx->track_or;
#define PYRO_CELLS_CASE(C)                                            \
  case C:                                                             \
    if (track) launch_cells_t<KIND, C, LT, true>(ctx, pass);          \
    else launch_cells_t<KIND, C, LT, false>(ctx, pass);               \
    break;
  switch (colour) {
    PYRO_CELLS_CASE(0)
    PYRO_CELLS_CASE(1)
    PYRO_CELLS_CASE(2)
    PYRO_CELLS_CASE(3)
  }
#undef PYRO_CELLS_CASE
}

// Cell-per-thread kernels for the lattice sizes of the benchmark
// configurations; returns false when L has no compile-time copy.
template <int KIND>
static bool launch_cells(pyro_ctx *ctx, int colour, int pass) {
  switch (ctx->L) {
    case 4: launch_cells_l<KIND, 4>(ctx, colour, pass); return true;
    case 8: launch_cells_l<KIND, 8>(ctx, colour, pass); return true;
    case 12: launch_cells_l<KIND, 12>(ctx, colour, pass); return true;
    case 16: launch_cells_l<KIND, 16>(ctx, colour, pass); return true;
    case 24: launch_cells_l<KIND, 24>(ctx, colour, pass); return true;
    case 32: launch_cells_l<KIND, 32>(ctx, colour, pass); return true;
    case 64: launch_cells_l<KIND, 64>(ctx, colour, pass); return true;
    default: return false;
  }
}

// Lattice sizes of the benchmark configurations get a compile-time copy.
template <int KIND>
static void launch_phase(pyro_ctx *ctx, int colour, int pass) {
  if (ctx->phase_variant == 1 && launch_cells<KIND>(ctx, colour, pass)) return;
  // The wave-per-site kernel always tracks the FP64 energy change.
  switch (ctx->L) {
    case 4: launch_phase_t<KIND, 4>(ctx, colour, pass); break;
    case 8: launch_phase_t<KIND, 8>(ctx, colour, pass); break;
    case 12: launch_phase_t<KIND, 12>(ctx, colour, pass); break;
    case 16: launch_phase_t<KIND, 16>(ctx, colour, pass); break;
    case 24: launch_phase_t<KIND, 24>(ctx, colour, pass); break;
    case 32: launch_phase_t<KIND, 32>(ctx, colour, pass); break;
    case 64: launch_phase_t<KIND, 64>(ctx, colour, pass); break;
    default: launch_phase_t<KIND, 0>(ctx, colour, pass); break;
  }
}

int launch_metropolis_phase(pyro_ctx *ctx, int colour, int pass, const uint32_t *) {
  if (ctx->hetero) return set_error(PYRO_ESTATE, "single colour phases need a uniform lattice size");
  launch_phase<0>(ctx, colour, pass);
  PYRO_HIP(hipGetLastError());
  return 0;
}

int launch_overrelax_phase(pyro_ctx *ctx, int colour) {
  if (ctx->hetero) return set_error(PYRO_ESTATE, "single colour phases need a uniform lattice size");
  launch_phase<1>(ctx, colour, 0);
  PYRO_HIP(hipGetLastError());
  return 0;
}

int launch_observables(pyro_ctx *ctx) {
  if (ctx->hetero) return launch_batch_observables(ctx);
  dim3 grid(ctx->obs_blocks, 4, ctx->R);
  k_observables<<<grid, kObsBlock, 0, ctx->stream>>>(ctx->d_spins, ctx->L, ctx->ncell,
                                                     ctx->d_partials);
  PYRO_HIP(hipGetLastError());
  k_observables_final<<<ctx->R, 256, 0, ctx->stream>>>(ctx->d_partials, ctx->obs_blocks, ctx->jeff,
                                                       ctx->d_obs);
  PYRO_HIP(hipGetLastError());
  return 0;
}

int launch_pt_swap(pyro_ctx *ctx, int parity, uint32_t round, int from_ctl) {
  const int ng = (int)ctx->h_group_off.size();
  for (int g = 0; g < ng; ++g) {
    k_pt_swap<<<1, 64, 0, ctx->stream>>>(ctx->h_group_len[g], ctx->d_obs, ctx->d_temps,
                                         ctx->d_beta, ctx->d_cosc,
                                         ctx->d_replica_at + ctx->h_group_off[g], ctx->d_ctl,
                                         key_lo(ctx), key_hi(ctx), parity, round, from_ctl, g,
                                         g == ng - 1);
    PYRO_HIP(hipGetLastError());
  }
  return 0;
}

int launch_end_sweep(pyro_ctx *ctx) {
  k_end_sweep<<<1, 256, 0, ctx->stream>>>(ctx->R, ctx->d_ctl, ctx->d_anchor, ctx->d_dE,
                                          ctx->d_replica_at, ctx->d_series);
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

}  // namespace pyro
