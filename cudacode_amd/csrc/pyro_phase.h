// Colour-phase and observables kernels of uniform lattices, templated on the
// local-field model (MAT: false = scalar Heisenberg J, true = exchange
// matrices and Zeeman field). Included by pyro_kernels.hip, which instantiates
// the scalar model, and by pyro_aniso.hip, which instantiates the matrix
// model, so the two sets compile in parallel translation units.
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
#pragma once

#include <hip/hip_ext.h>

#include "pyro_cell.h"
#include "pyro_internal.h"

namespace pyro {

static inline uint32_t key_lo(const pyro_ctx *c) { return (uint32_t)c->seed; }
static inline uint32_t key_hi(const pyro_ctx *c) { return (uint32_t)(c->seed >> 32); }

static __constant__ Link c_links[16][6] = PYRO_LINKS16_INIT;

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
// constants); LT = 0 reads it from L. MAT: the matrix field model (coup).
template <int KIND, int LT, bool MAT>
__global__ __launch_bounds__(kPhaseBlock) void k_phase(
    float *__restrict__ spins, int L_rt, int ncell_rt, int colour, int pass,
    const float *__restrict__ beta, const float *__restrict__ cosc,
    const uint32_t *__restrict__ ctl, uint32_t key0, uint32_t key1, float jeff,
    unsigned long long *__restrict__ acc, double *__restrict__ dE, const float *__restrict__ coup,
    int post_rotate, int r0, uint32_t sweep_delta) {
  const int L = LT > 0 ? LT : L_rt;
  const int ncell = LT > 0 ? LT * LT * LT : ncell_rt;
  const int r = r0 + (int)blockIdx.y;
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
    float fx, fy, fz;
    field_start<MAT>(coup, colour, fx, fy, fz);
    double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
    for (int n = 0; n < 6; ++n) {
      const Link lk = c_links[site][n];
      const int x1 = wrap(x + lk.dx, L);
      const int y1 = wrap(y + lk.dy, L);
      const int z1 = wrap(z + lk.dz, L);
      const int j = lk.site * ncell + (z1 * L + y1) * L + x1;
      const float a = Sx[j], b = Sy[j], c = Sz[j];
      field_add<MAT, true>(coup, colour, lk.site & 3, a, b, c, fx, fy, fz, gx, gy, gz);
    }
    const int i = site * ncell + cell;
    const f3 s = {Sx[i], Sy[i], Sz[i]};
    f3 sp;
    bool write;
    if (KIND == 0) {
      const u32x4 w = philox4x32_10(
          rng_counter((uint32_t)cell, (uint32_t)site, (uint32_t)pass, kKindMetropolis,
                      (uint32_t)r, ctl[kCtlSweep] + sweep_delta),
          key0, key1);
      const f3 h = field_h<MAT>(jeff, fx, fy, fz);
      sp = cone_proposal(s, cosc[r], u01(w.x), u01(w.y));
      write = metropolis_accept(h, s, sp, beta[r], u01(w.z));
      accepted = write ? 1 : 0;
      if (post_rotate) {
        sp = rotate_about(h, write ? sp : s, u01(w.w));
        write = true;
      }
    } else {
      const f3 f = {fx, fy, fz};
      sp = overrelax(f, s);
      write = true;
    }
    if (write) {
      de = field_de<MAT>(jeff, coup, colour, s, sp, gx, gy, gz);
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
//
// Series hook (over-relaxation without FP64 tracking only, which leaves dE
// alone): with series != nullptr, block 0 also writes the energy-series
// entries (anchor + dE) of the rungs whose replica is in this launch's range,
// at row ctl[kCtlRecordRow] + sweep_delta. A split sweep passes it to its
// first over-relaxation launch, whose blocks start after the sweep's
// Metropolis or learned phases have finished changing dE, so no separate
// recording launch breaks the stream.
// ---------------------------------------------------------------------------
template <int KIND, int COLOUR, int LT, bool TRACK, bool MAT>
__global__ __launch_bounds__(256) void k_phase_cells(
    float *__restrict__ spins, int blocks_per_replica, int pass, const float *__restrict__ beta,
    const float *__restrict__ cosc, const uint32_t *__restrict__ ctl, uint32_t key0,
    uint32_t key1, float jeff, unsigned long long *__restrict__ acc, double *__restrict__ dE,
    const float *__restrict__ coup, int post_rotate, int r0, uint32_t sweep_delta,
    double *__restrict__ series, int R_total, const int *__restrict__ replica_at,
    const double *__restrict__ anchor) {
  constexpr int NC = LT * LT * LT;
  const int nb = gridDim.x;
  const int b = blockIdx.x;
  if (KIND == 1 && !TRACK && series != nullptr && b == 0) {
    const uint32_t row = ctl[kCtlRecordRow] + sweep_delta;
    if (ctl[kCtlRecordOn] && row < ctl[kCtlRecordCap]) {
      const int r1 = r0 + nb / blocks_per_replica;
      for (int k = threadIdx.x; k < R_total; k += 256) {
        const int rep = replica_at[k];
        if (rep >= r0 && rep < r1) series[(size_t)row * R_total + k] = anchor[rep] + dE[rep];
      }
    }
  }
  const int lb = (nb & 7) == 0 ? (b & 7) * (nb >> 3) + (b >> 3) : b;
  const int q = lb / blocks_per_replica;  // replica within the launch range
  const int r = r0 + q;
  const int cell = (lb - q * blocks_per_replica) * 256 + threadIdx.x;
  double de = 0.0;
  unsigned nacc = 0;
  if (cell < NC)
    cell_update<KIND, COLOUR, LT, TRACK, MAT>(spins + (size_t)r * 48 * NC, cell, r, pass,
                                              KIND == 0 ? beta[r] : 0.0f,
                                              KIND == 0 ? cosc[r] : 0.0f,
                                              KIND == 0 ? ctl[kCtlSweep] + sweep_delta : 0u, key0, key1, jeff,
                                              de, nacc, coup, KIND == 0 && post_rotate != 0);
  if (KIND == 0 || TRACK) {
    block_sum<256>(de, nacc);
    if (threadIdx.x == 0) {
      if (de != 0.0) atomicAdd(&dE[r], de);
      if (KIND == 0 && nacc) atomicAdd(&acc[r], (unsigned long long)nacc);
    }
  }
}

// Two over-relaxation colour phases (CA, CA + 1) in one launch between the two
// spin buffers (or_pair_cell), scalar model. Same block layout, XCD remap and
// series hook as k_phase_cells; a pass is k_or_pair<0> then k_or_pair<2>,
// after which `oth` holds every colour and the buffers swap roles.
// Algorithmic bytes: 12 N read + 6 N written for N/2 updates = 36 B/update.
// Occupancy: the first pair fits 128 VGPRs (4 waves per SIMD) without spills;
// the second, reading two buffers, would spill one register in the middle of
// its load burst at that bound (a wait for every outstanding load), so it
// runs at 3 waves per SIMD (measured: L=16 x 64 sweep 0.322 ms against 0.337
// with both at 4 and 0.335 with both at 3).
template <int CA, int LT, bool TRACK>
__global__ __launch_bounds__(256, TRACK ? 1 : (CA == 0 ? 4 : 3)) void k_or_pair(
    const float *__restrict__ cur, float *__restrict__ oth, int blocks_per_replica, float jeff,
    double *__restrict__ dE, int r0, const uint32_t *__restrict__ ctl, uint32_t sweep_delta,
    double *__restrict__ series, int R_total, const int *__restrict__ replica_at,
    const double *__restrict__ anchor) {
  constexpr int NC = LT * LT * LT;
  const int nb = gridDim.x;
  const int b = blockIdx.x;
  const int lb = (nb & 7) == 0 ? (b & 7) * (nb >> 3) + (b >> 3) : b;
  const int q = lb / blocks_per_replica;
  const int r = r0 + q;
  const int cell = (lb - q * blocks_per_replica) * 256 + threadIdx.x;
  if (!TRACK && series != nullptr && b == 0) {
    const uint32_t row = ctl[kCtlRecordRow] + sweep_delta;
    if (ctl[kCtlRecordOn] && row < ctl[kCtlRecordCap]) {
      const int r1 = r0 + nb / blocks_per_replica;
      for (int k = threadIdx.x; k < R_total; k += 256) {
        const int rep = replica_at[k];
        if (rep >= r0 && rep < r1) series[(size_t)row * R_total + k] = anchor[rep] + dE[rep];
      }
    }
  }
  double de = 0.0;
  if (cell < NC)
    or_pair_cell<CA, LT, TRACK>(cur + (size_t)r * 48 * NC, oth + (size_t)r * 48 * NC, cell, jeff,
                                de);
  if (TRACK) {
    unsigned nacc = 0;
    block_sum<256>(de, nacc);
    if (threadIdx.x == 0 && de != 0.0) atomicAdd(&dE[r], de);
  }
}

// Per-block FP64 partial sums of the energy and the sublattice-b magnetisation
// (TotalEnergyGPU, Pyrochlore.cu:344-389; MagnetizationGPU, :277-342;
// Measurement, :526-578). grid = (obs_blocks, 4, R): blockIdx.y is the basis
// index b, and a thread sums the 4 sites 4t + b of its cell with all 72
// neighbour loads independent, so a launch has 4x the waves of a cell-per-thread
// layout and no dependent load chains. partials[((r * 4 + b) * nblk + blk) * 4 + q],
// q = (energy, mx, my, mz). The energy partial is sum_i s_i . g_i (scalar
// model: g = neighbour sum) or sum_i s_i . (B_b + g_i / 2) (matrix model:
// g = sum_n C s_n); the second stage scales it by jeff / 2 or -1.
template <int B, bool MAT>
__device__ __forceinline__ void obs_cell(const float *__restrict__ Sx, const float *__restrict__ Sy,
                                         const float *__restrict__ Sz, int L, int ncell, int cell,
                                         const float *__restrict__ coup, double v[4]) {
  const int LL = L * L;
  const int z = cell / LL;
  const int rem = cell - z * LL;
  const int y = rem / L;
  const int x = rem - y * L;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int site = 4 * t + B;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    float fx, fy, fz;  // unused: the FP32 field
    field_start<false>(coup, B, fx, fy, fz);
#pragma unroll
    for (int n = 0; n < 6; ++n) {
      const Link lk = kLinks[site][n];
      const int x1 = wrap(x + lk.dx, L);
      const int y1 = wrap(y + lk.dy, L);
      const int z1 = wrap(z + lk.dz, L);
      const size_t j = (size_t)lk.site * ncell + (size_t)((z1 * L + y1) * L + x1);
      field_add<MAT, true>(coup, B, lk.site & 3, Sx[j], Sy[j], Sz[j], fx, fy, fz, gx, gy, gz);
    }
    const size_t i = (size_t)site * ncell + cell;
    const double sx = Sx[i], sy = Sy[i], sz = Sz[i];
    if constexpr (MAT)
      v[0] += sx * ((double)coup[144 + 3 * B] + 0.5 * gx) +
              sy * ((double)coup[145 + 3 * B] + 0.5 * gy) +
              sz * ((double)coup[146 + 3 * B] + 0.5 * gz);
    else
      v[0] += sx * gx + sy * gy + sz * gz;
    v[1] += sx;
    v[2] += sy;
    v[3] += sz;
  }
}

template <bool MAT>
__global__ __launch_bounds__(kObsBlock) void k_observables(const float *__restrict__ spins,
                                                           int L, int ncell,
                                                           double *__restrict__ partials,
                                                           const float *__restrict__ coup) {
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
      case 0: obs_cell<0, MAT>(Sx, Sy, Sz, L, ncell, cell, coup, v); break;
      case 1: obs_cell<1, MAT>(Sx, Sy, Sz, L, ncell, cell, coup, v); break;
      case 2: obs_cell<2, MAT>(Sx, Sy, Sz, L, ncell, cell, coup, v); break;
      default: obs_cell<3, MAT>(Sx, Sy, Sz, L, ncell, cell, coup, v); break;
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

template <int KIND, int LT, bool MAT>
static void launch_phase_t(pyro_ctx *ctx, int colour, int pass) {
  dim3 grid((ctx->ncell + kPhaseCells - 1) / kPhaseCells, launch_nrep(ctx));
  if (ctx->time_ev0) {
    hipExtLaunchKernelGGL(k_phase<KIND, LT, MAT>, grid, dim3(kPhaseBlock), 0, launch_stream(ctx),
                          ctx->time_ev0, ctx->time_ev1, 0, ctx->d_spins, ctx->L, ctx->ncell, colour,
                          pass, (const float *)ctx->d_beta, (const float *)ctx->d_cosc,
                          (const uint32_t *)ctx->d_ctl, key_lo(ctx), key_hi(ctx), ctx->jeff,
                          ctx->d_acc, ctx->d_dE, (const float *)ctx->d_coup, (int)ctx->post_rotate,
                          launch_r0(ctx), ctx->launch_delta);
    return;
  }
  k_phase<KIND, LT, MAT><<<grid, kPhaseBlock, 0, launch_stream(ctx)>>>(
      ctx->d_spins, ctx->L, ctx->ncell, colour, pass, ctx->d_beta, ctx->d_cosc, ctx->d_ctl,
      key_lo(ctx), key_hi(ctx), ctx->jeff, ctx->d_acc, ctx->d_dE, ctx->d_coup,
      (int)ctx->post_rotate, launch_r0(ctx), ctx->launch_delta);
}

template <int KIND, int COLOUR, int LT, bool TRACK, bool MAT>
static void launch_cells_t(pyro_ctx *ctx, int pass) {
  const int bpr = (LT * LT * LT + 255) / 256;
  const int nrep = launch_nrep(ctx);
  double *series = nullptr;
  if (KIND == 1 && !TRACK && ctx->launch_series) {
    series = ctx->launch_series;
    ctx->launch_series_done = 1;
  }
  if (ctx->time_ev0) {
    hipExtLaunchKernelGGL(k_phase_cells<KIND, COLOUR, LT, TRACK, MAT>, dim3(bpr * nrep),
                          dim3(256), 0, launch_stream(ctx), ctx->time_ev0, ctx->time_ev1, 0,
                          ctx->d_spins, bpr, pass, (const float *)ctx->d_beta,
                          (const float *)ctx->d_cosc, (const uint32_t *)ctx->d_ctl, key_lo(ctx),
                          key_hi(ctx), ctx->jeff, ctx->d_acc, ctx->d_dE,
                          (const float *)ctx->d_coup, (int)ctx->post_rotate, launch_r0(ctx),
                          ctx->launch_delta, series, ctx->R, (const int *)ctx->d_replica_at,
                          (const double *)ctx->d_anchor);
    return;
  }
  k_phase_cells<KIND, COLOUR, LT, TRACK, MAT><<<bpr * nrep, 256, 0, launch_stream(ctx)>>>(
      ctx->d_spins, bpr, pass, ctx->d_beta, ctx->d_cosc, ctx->d_ctl, key_lo(ctx), key_hi(ctx),
      ctx->jeff, ctx->d_acc, ctx->d_dE, ctx->d_coup, (int)ctx->post_rotate, launch_r0(ctx),
      ctx->launch_delta, series, ctx->R, ctx->d_replica_at, ctx->d_anchor);
}

template <int KIND, int LT, bool MAT>
static void launch_cells_l(pyro_ctx *ctx, int colour, int pass) {
  const bool track = KIND == 1 && ctx->track_or;
#define PYRO_CELLS_CASE(C)                                            \
  case C:                                                             \
    if (track) launch_cells_t<KIND, C, LT, true, MAT>(ctx, pass);     \
    else launch_cells_t<KIND, C, LT, false, MAT>(ctx, pass);          \
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
template <int KIND, bool MAT>
static bool launch_cells(pyro_ctx *ctx, int colour, int pass) {
  switch (ctx->L) {
    case 4: launch_cells_l<KIND, 4, MAT>(ctx, colour, pass); return true;
    case 8: launch_cells_l<KIND, 8, MAT>(ctx, colour, pass); return true;
    case 12: launch_cells_l<KIND, 12, MAT>(ctx, colour, pass); return true;
    case 16: launch_cells_l<KIND, 16, MAT>(ctx, colour, pass); return true;
    case 24: launch_cells_l<KIND, 24, MAT>(ctx, colour, pass); return true;
    case 32: launch_cells_l<KIND, 32, MAT>(ctx, colour, pass); return true;
    case 64: launch_cells_l<KIND, 64, MAT>(ctx, colour, pass); return true;
    default: return false;
  }
}

template <int CA, int LT, bool TRACK>
static void launch_or_pair_t(pyro_ctx *ctx, const float *cur, float *oth) {
  const int bpr = (LT * LT * LT + 255) / 256;
  const int nrep = launch_nrep(ctx);
  double *series = nullptr;
  if (CA == 0 && !TRACK && ctx->launch_series) {
    series = ctx->launch_series;
    ctx->launch_series_done = 1;
  }
  if (ctx->time_ev0) {
    hipExtLaunchKernelGGL(k_or_pair<CA, LT, TRACK>, dim3(bpr * nrep), dim3(256), 0,
                          launch_stream(ctx), ctx->time_ev0, ctx->time_ev1, 0, cur, oth, bpr,
                          ctx->jeff, ctx->d_dE, launch_r0(ctx), (const uint32_t *)ctx->d_ctl,
                          ctx->launch_delta, series, ctx->R, (const int *)ctx->d_replica_at,
                          (const double *)ctx->d_anchor);
    return;
  }
  k_or_pair<CA, LT, TRACK><<<bpr * nrep, 256, 0, launch_stream(ctx)>>>(
      cur, oth, bpr, ctx->jeff, ctx->d_dE, launch_r0(ctx), ctx->d_ctl, ctx->launch_delta, series,
      ctx->R, ctx->d_replica_at, ctx->d_anchor);
}

template <int LT>
static void launch_or_pair_l(pyro_ctx *ctx, int ca, const float *cur, float *oth) {
  if (ctx->track_or) {
    if (ca == 0) launch_or_pair_t<0, LT, true>(ctx, cur, oth);
    else launch_or_pair_t<2, LT, true>(ctx, cur, oth);
  } else {
    if (ca == 0) launch_or_pair_t<0, LT, false>(ctx, cur, oth);
    else launch_or_pair_t<2, LT, false>(ctx, cur, oth);
  }
}

// The fused over-relaxation colour pair ca (0 or 2) from cur into oth, for
// the compile-time lattice sizes; false when L has no copy.
static inline bool launch_or_pair(pyro_ctx *ctx, int ca, const float *cur, float *oth) {
  switch (ctx->L) {
    case 4: launch_or_pair_l<4>(ctx, ca, cur, oth); return true;
    case 8: launch_or_pair_l<8>(ctx, ca, cur, oth); return true;
    case 12: launch_or_pair_l<12>(ctx, ca, cur, oth); return true;
    case 16: launch_or_pair_l<16>(ctx, ca, cur, oth); return true;
    case 24: launch_or_pair_l<24>(ctx, ca, cur, oth); return true;
    case 32: launch_or_pair_l<32>(ctx, ca, cur, oth); return true;
    case 64: launch_or_pair_l<64>(ctx, ca, cur, oth); return true;
    default: return false;
  }
}

// Lattice sizes of the benchmark configurations get a compile-time copy.
template <int KIND, bool MAT>
static void launch_phase(pyro_ctx *ctx, int colour, int pass) {
  if (ctx->phase_variant == 1 && launch_cells<KIND, MAT>(ctx, colour, pass)) return;
  // The wave-per-site kernel always tracks the FP64 energy change.
  switch (ctx->L) {
    case 4: launch_phase_t<KIND, 4, MAT>(ctx, colour, pass); break;
    case 8: launch_phase_t<KIND, 8, MAT>(ctx, colour, pass); break;
    case 12: launch_phase_t<KIND, 12, MAT>(ctx, colour, pass); break;
    case 16: launch_phase_t<KIND, 16, MAT>(ctx, colour, pass); break;
    case 24: launch_phase_t<KIND, 24, MAT>(ctx, colour, pass); break;
    case 32: launch_phase_t<KIND, 32, MAT>(ctx, colour, pass); break;
    case 64: launch_phase_t<KIND, 64, MAT>(ctx, colour, pass); break;
    default: launch_phase_t<KIND, 0, MAT>(ctx, colour, pass); break;
  }
}

// Partial sums of the observables (first stage) on the ctx stream.
template <bool MAT>
static void launch_observables_partials(pyro_ctx *ctx) {
  dim3 grid(ctx->obs_blocks, 4, ctx->R);
  k_observables<MAT><<<grid, kObsBlock, 0, ctx->stream>>>(ctx->d_spins, ctx->L, ctx->ncell,
                                                          ctx->d_partials, ctx->d_coup);
}

}  // namespace pyro
