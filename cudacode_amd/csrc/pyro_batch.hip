// Batches of replicas with different lattice sizes, and the two sweep
// executions that serve them (no reference counterpart: the reference runs
// one lattice size per process, Pyrochlore.cu:1251-1257).
//
//  exec_mode 1 -- one persistent kernel per run of sweeps up to the next
//    exchange or measurement. Replicas are independent between exchanges, so
//    a block that owns whole replicas runs every colour phase of every pass of
//    those sweeps with block barriers only: no grid-wide synchronisation, one
//    launch per interval. Replicas are packed into bins of balanced work
//    (longest-processing-time first); bins = blocks. A replica of L <= 8
//    (at most 96 KiB of spins) is staged in LDS for the interval, so its
//    neighbour gathers are LDS reads; larger ones work in global memory
//    (L2-resident per block).
//  exec_mode 2 -- per-phase kernels over a work list of 256-cell tiles of all
//    replicas (perfect balance, one launch per colour phase), tiles ordered so
//    each XCD gets a contiguous run of one replica's cells.
//
// Both call the same cell_update as the uniform kernels, with the replica's
// global index in the Philox counters: every replica's trajectory is the one
// a single-replica run of its size would produce, bit for bit.
#include <algorithm>
#include <numeric>

#include "pyro_cell.h"
#include "pyro_internal.h"

namespace pyro {

__constant__ Link c_links_batch[16][6] = PYRO_LINKS16_INIT;

bool batch_size_supported(int L) {
  return L == 4 || L == 8 || L == 12 || L == 16 || L == 24 || L == 32;
}

// --- persistent sweep ------------------------------------------------------
template <int KIND, int LT, int BLOCK, bool TRACK>
__device__ __forceinline__ void replica_pass(float *Sx, int r, int pass, float beta_r, float cosc_r,
                                             uint32_t sweep, uint32_t key0, uint32_t key1, float jeff,
                                             double &de, unsigned &nacc) {
  constexpr int NC = LT * LT * LT;
  // Item u = k * NC + cell updates site C + 4k of that cell: consecutive
  // threads take consecutive cells of one site plane (coalesced), one site
  // per iteration (the register budget of a 1024-thread block).
#define PYRO_COLOUR(C)                                                                     \
  _Pragma("unroll 1") for (int u = threadIdx.x; u < 4 * NC; u += BLOCK) {                  \
    const int k = u / NC;                                                                  \
    site_update<KIND, LT, TRACK>(Sx, c_links_batch, u - k * NC, C + 4 * k, r, pass, beta_r, \
                                 cosc_r, sweep, key0, key1, jeff, de, nacc);               \
  }                                                                                        \
  __syncthreads();
  PYRO_COLOUR(0)
  PYRO_COLOUR(1)
  PYRO_COLOUR(2)
  PYRO_COLOUR(3)
#undef PYRO_COLOUR
}

// Replicas up to this size are staged in LDS (48 L^3 floats: 12 KiB at L=4,
// 96 KiB at L=8). The array is referenced by name in the staged variant, so
// the compiler knows those spins live in LDS and emits LDS instructions.
constexpr int kLdsMaxL = 8;
extern __shared__ float lds_spins[];

// One sweep of one replica (spins at Sg, or their LDS copy when STAGED). One
// out-of-line instantiation per variant keeps each within the register
// budget of a 1024-thread block.
template <int LT, int BLOCK, bool TRACK, bool STAGED>
__device__ __attribute__((noinline)) void replica_sweep(float *Sg, int r, int n_metro, int n_or,
                                                        float beta_r, float cosc_r, uint32_t sweep,
                                                        uint32_t key0, uint32_t key1, float jeff,
                                                        double &de, unsigned &nacc) {
  float *Sx = STAGED ? lds_spins : Sg;
  for (int p = 0; p < n_metro; ++p)
    replica_pass<0, LT, BLOCK, TRACK>(Sx, r, p, beta_r, cosc_r, sweep, key0, key1, jeff, de, nacc);
  for (int o = 0; o < n_or; ++o)
    replica_pass<1, LT, BLOCK, TRACK>(Sx, r, 0, beta_r, cosc_r, sweep, key0, key1, jeff, de, nacc);
}

// The sweeps [0, n_sweeps) of one replica: after each, the block's FP64
// energy change and acceptance count go to dE[r] and acc[r] (this block owns
// the replica: no atomics), and, when recording, the series entry anchor + dE
// of the replica's rung at row ctl[kCtlRecordRow] + s; the last sweep's entry
// only if record_last (else the end-of-sweep kernel writes it after an
// exchange). A replica of L <= kLdsMaxL is copied to LDS for the run when
// staged.
template <int LT, int BLOCK, bool TRACK>
__device__ __forceinline__ void replica_run(float *Sg, bool staged, int r, int n_metro, int n_or,
                                            float beta_r, float cosc_r,
                                            const uint32_t *__restrict__ ctl, int n_sweeps,
                                            int record_last, int rung, int R,
                                            double *__restrict__ series,
                                            const double *__restrict__ anchor, uint32_t key0,
                                            uint32_t key1, float jeff,
                                            unsigned long long *__restrict__ acc,
                                            double *__restrict__ dE) {
  constexpr int NF = 48 * LT * LT * LT;
  bool lds = false;
  if constexpr (LT <= kLdsMaxL) lds = staged;
  if (lds) {
    for (int i = threadIdx.x; i < NF; i += BLOCK) lds_spins[i] = Sg[i];
    __syncthreads();
  }
  const uint32_t sweep0 = ctl[kCtlSweep];
  const uint32_t row0 = ctl[kCtlRecordRow];
  const bool rec = series != nullptr && rung >= 0 && ctl[kCtlRecordOn] != 0u;
  const uint32_t cap = ctl[kCtlRecordCap];
  for (int s = 0; s < n_sweeps; ++s) {
    double de = 0.0;
    unsigned nacc = 0;
    const uint32_t sweep = sweep0 + (uint32_t)s;
    if constexpr (LT <= kLdsMaxL) {
      if (lds)
        replica_sweep<LT, BLOCK, TRACK, true>(Sg, r, n_metro, n_or, beta_r, cosc_r, sweep, key0,
                                              key1, jeff, de, nacc);
      else
        replica_sweep<LT, BLOCK, TRACK, false>(Sg, r, n_metro, n_or, beta_r, cosc_r, sweep, key0,
                                               key1, jeff, de, nacc);
    } else {
      replica_sweep<LT, BLOCK, TRACK, false>(Sg, r, n_metro, n_or, beta_r, cosc_r, sweep, key0,
                                             key1, jeff, de, nacc);
    }
    block_sum<BLOCK>(de, nacc);
    if (threadIdx.x == 0) {
      dE[r] += de;
      acc[r] += nacc;
      const uint32_t row = row0 + (uint32_t)s;
      if (rec && (s + 1 < n_sweeps || record_last) && row < cap)
        series[(size_t)row * R + rung] = anchor[r] + dE[r];
    }
  }
  if (lds) {
    __syncthreads();
    for (int i = threadIdx.x; i < NF; i += BLOCK) Sg[i] = lds_spins[i];
    __syncthreads();  // the next replica of the bin reuses the LDS
  }
}

// grid = n_bins, block = BLOCK, dynamic LDS = 48 L^3 floats of the largest
// staged size in the batch (0: nothing staged). Bin b owns replicas
// order[bin_start[b] .. bin_start[b+1]); each replica is swept through all
// n_sweeps sweeps by this block only, so its counters are updated without
// atomics. rung_of[r] is the replica's rung (for the series), or nullptr.
template <int BLOCK, bool TRACK>
__global__ __launch_bounds__(BLOCK) void k_batch_sweep(
    float *__restrict__ spins, const long long *__restrict__ off, const int *__restrict__ Ls,
    const int *__restrict__ order, const int *__restrict__ bin_start, int n_metro, int n_or,
    const float *__restrict__ beta, const float *__restrict__ cosc,
    const uint32_t *__restrict__ ctl, uint32_t key0, uint32_t key1, float jeff,
    unsigned long long *__restrict__ acc, double *__restrict__ dE, int n_sweeps, int record_last,
    int lds_max_L, int R, const int *__restrict__ replica_at, double *__restrict__ series,
    const double *__restrict__ anchor) {
  __shared__ int s_rung;
  for (int q = bin_start[blockIdx.x]; q < bin_start[blockIdx.x + 1]; ++q) {
    const int r = order[q];
    float *Sx = spins + off[r];
    const float b = beta[r], c = cosc[r];
    const int L = Ls[r];
    const bool staged = L <= lds_max_L;
    int rung = -1;
    if (series != nullptr) {
      if (threadIdx.x == 0) s_rung = -1;
      __syncthreads();
      for (int k = threadIdx.x; k < R; k += BLOCK)
        if (replica_at[k] == r) s_rung = k;
      __syncthreads();
      rung = s_rung;
      __syncthreads();
    }
#define PYRO_RUN(LV)                                                                          \
  case LV:                                                                                    \
    replica_run<LV, BLOCK, TRACK>(Sx, staged, r, n_metro, n_or, b, c, ctl, n_sweeps, record_last, \
                                  rung, R, series, anchor, key0, key1, jeff, acc, dE);         \
    break;
    switch (L) {
      PYRO_RUN(4)
      PYRO_RUN(8)
      PYRO_RUN(12)
      PYRO_RUN(16)
      PYRO_RUN(24)
      PYRO_RUN(32)
      default: break;
    }
#undef PYRO_RUN
  }
}

// --- tile-list phases --------------------------------------------------------
// grid = n_tiles (padded to a multiple of 8 with r = -1), block = TILE, the
// cells of a tile (128, 256 or 512).
template <int KIND, int COLOUR, bool TRACK, int TILE>
__global__ __launch_bounds__(TILE) void k_batch_phase(
    float *__restrict__ spins, const long long *__restrict__ off, const int *__restrict__ Ls,
    const int *__restrict__ tiles, int pass, const float *__restrict__ beta,
    const float *__restrict__ cosc, const uint32_t *__restrict__ ctl, uint32_t key0, uint32_t key1,
    float jeff, unsigned long long *__restrict__ acc, double *__restrict__ dE, uint32_t sweep_delta) {
  const int r = tiles[2 * blockIdx.x];
  if (r < 0) return;  // padding tile: the whole block leaves together
  const int cell = tiles[2 * blockIdx.x + 1] * TILE + threadIdx.x;
  const int L = Ls[r];
  float *Sx = spins + off[r];
  const float b = KIND == 0 ? beta[r] : 0.0f, c = KIND == 0 ? cosc[r] : 0.0f;
  const uint32_t sweep = KIND == 0 ? ctl[kCtlSweep] + sweep_delta : 0u;
  double de = 0.0;
  unsigned nacc = 0;
  if (cell < L * L * L) {
    switch (L) {
      case 4: cell_update<KIND, COLOUR, 4, TRACK>(Sx, cell, r, pass, b, c, sweep, key0, key1, jeff, de, nacc); break;
      case 8: cell_update<KIND, COLOUR, 8, TRACK>(Sx, cell, r, pass, b, c, sweep, key0, key1, jeff, de, nacc); break;
      case 12: cell_update<KIND, COLOUR, 12, TRACK>(Sx, cell, r, pass, b, c, sweep, key0, key1, jeff, de, nacc); break;
      case 16: cell_update<KIND, COLOUR, 16, TRACK>(Sx, cell, r, pass, b, c, sweep, key0, key1, jeff, de, nacc); break;
      case 24: cell_update<KIND, COLOUR, 24, TRACK>(Sx, cell, r, pass, b, c, sweep, key0, key1, jeff, de, nacc); break;
      case 32: cell_update<KIND, COLOUR, 32, TRACK>(Sx, cell, r, pass, b, c, sweep, key0, key1, jeff, de, nacc); break;
      default: break;
    }
  }
  if (KIND == 0 || TRACK) {
    block_sum<TILE>(de, nacc);
    if (threadIdx.x == 0) {
      if (de != 0.0) atomicAdd(&dE[r], de);
      if (KIND == 0 && nacc) atomicAdd(&acc[r], (unsigned long long)nacc);
    }
  }
}

// Fused over-relaxation pair (colours CA, CA + 1, or_pair_cell) over the tile
// list, between the two spin buffers. grid = n_tiles, block = TILE.
template <int CA, bool TRACK, int TILE>
__global__ __launch_bounds__(TILE) void k_batch_or_pair(
    const float *__restrict__ cur, float *__restrict__ oth, const long long *__restrict__ off,
    const int *__restrict__ Ls, const int *__restrict__ tiles, float jeff,
    double *__restrict__ dE) {
  const int r = tiles[2 * blockIdx.x];
  if (r < 0) return;  // padding tile: the whole block leaves together
  const int cell = tiles[2 * blockIdx.x + 1] * TILE + threadIdx.x;
  const int L = Ls[r];
  const float *c = cur + off[r];
  float *o = oth + off[r];
  double de = 0.0;
  if (cell < L * L * L) {
    switch (L) {
      case 4: or_pair_cell<CA, 4, TRACK>(c, o, cell, jeff, de); break;
      case 8: or_pair_cell<CA, 8, TRACK>(c, o, cell, jeff, de); break;
      case 12: or_pair_cell<CA, 12, TRACK>(c, o, cell, jeff, de); break;
      case 16: or_pair_cell<CA, 16, TRACK>(c, o, cell, jeff, de); break;
      case 24: or_pair_cell<CA, 24, TRACK>(c, o, cell, jeff, de); break;
      case 32: or_pair_cell<CA, 32, TRACK>(c, o, cell, jeff, de); break;
      default: break;
    }
  }
  if (TRACK) {
    unsigned nacc = 0;
    block_sum<TILE>(de, nacc);
    if (threadIdx.x == 0 && de != 0.0) atomicAdd(&dE[r], de);
  }
}

// --- initial spins and observables ------------------------------------------
// grid = (ceil(max ncell / 256), R), block = 256; the same draws as k_init_spins.
__global__ __launch_bounds__(256) void k_batch_init(float *__restrict__ spins,
                                                    const long long *__restrict__ off,
                                                    const int *__restrict__ Ls, uint32_t key0,
                                                    uint32_t key1) {
  const int r = blockIdx.y;
  const int L = Ls[r];
  const int nc = L * L * L;
  const int cell = blockIdx.x * 256 + threadIdx.x;
  if (cell >= nc) return;
  float *Sx = spins + off[r];
  const size_t plane = (size_t)16 * nc;
  for (int site = 0; site < 16; ++site) {
    const u32x4 w = philox4x32_10(
        rng_counter((uint32_t)cell, (uint32_t)site, 0, kKindInit, (uint32_t)r, 0), key0, key1);
    const f3 v = uniform_sphere(u01(w.x), u01(w.y));
    const size_t i = (size_t)site * nc + cell;
    Sx[i] = v.x;
    Sx[plane + i] = v.y;
    Sx[2 * plane + i] = v.z;
  }
}

__device__ __forceinline__ int wrap_b(int v, int L) { return v < 0 ? v + L : (v >= L ? v - L : v); }

// Exchange and measurement observables of a batch, first stage: the FP64
// energy partial (each bond from both ends) and the 12 sublattice sums of one
// 256-cell tile, one block per tile of the tile list, written to the tile's
// slot in its replica's run of partials, part[(part_off[r] + t) * kObs + q].
// A block per tile spreads every replica over many CUs (one block per
// replica took 1.26 ms per call for 256 mixed replicas of L up to 24).
// grid = the padded tile list, block = 256.
__global__ __launch_bounds__(256) void k_batch_obs_tiles(const float *__restrict__ spins,
                                                         const long long *__restrict__ off,
                                                         const int *__restrict__ Ls,
                                                         const int *__restrict__ tiles,
                                                         const int *__restrict__ part_off,
                                                         double *__restrict__ part) {
  const int r = tiles[2 * blockIdx.x];
  if (r < 0) return;  // padding tile: the whole block leaves together
  const int t = tiles[2 * blockIdx.x + 1];
  const int L = Ls[r], LL = L * L, nc = L * L * L;
  const int cell = t * 256 + threadIdx.x;
  const float *Sx = spins + off[r];
  const float *Sy = Sx + 16 * nc;
  const float *Sz = Sy + 16 * nc;
  double v[kObs];
#pragma unroll
  for (int q = 0; q < kObs; ++q) v[q] = 0.0;
  if (cell < nc) {
    const int z = cell / LL, rem = cell - z * LL, y = rem / L, x = rem - y * L;
    for (int tt = 0; tt < 4; ++tt) {
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int site = 4 * tt + b;
        double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
        for (int n = 0; n < 6; ++n) {
          const Link lk = kLinks[site][n];
          const int j = lk.site * nc + (wrap_b(z + lk.dz, L) * L + wrap_b(y + lk.dy, L)) * L +
                        wrap_b(x + lk.dx, L);
          gx += (double)Sx[j];
          gy += (double)Sy[j];
          gz += (double)Sz[j];
        }
        const int i = site * nc + cell;
        const double sx = Sx[i], sy = Sy[i], sz = Sz[i];
        v[0] += sx * gx + sy * gy + sz * gz;
        v[1 + 3 * b] += sx;
        v[2 + 3 * b] += sy;
        v[3 + 3 * b] += sz;
      }
    }
  }
  __shared__ double s_v[4][kObs];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < kObs; ++q) {
    const double w = wave_sum_d(v[q]);
    if (lane == 0) s_v[wv][q] = w;
  }
  __syncthreads();
  if (threadIdx.x < kObs) {
    const int q = threadIdx.x;
    part[((size_t)part_off[r] + t) * kObs + q] = (s_v[0][q] + s_v[1][q]) + (s_v[2][q] + s_v[3][q]);
  }
}

// Second stage: per replica, the tiles' partials in tile order (fixed, so the
// result does not depend on which tile finished first); the energy scaled by
// jeff / 2. grid = R, block = 256.
__global__ __launch_bounds__(256) void k_batch_obs_final(const double *__restrict__ part,
                                                         const int *__restrict__ part_off,
                                                         float jeff, double *__restrict__ obs) {
  const int r = blockIdx.x;
  const int k0 = part_off[r], n = part_off[r + 1] - k0;
  __shared__ double s_v[256];
  for (int q = 0; q < kObs; ++q) {
    double t = 0.0;
    for (int k = threadIdx.x; k < n; k += 256) t += part[((size_t)k0 + k) * kObs + q];
    s_v[threadIdx.x] = t;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) s_v[threadIdx.x] += s_v[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) obs[(size_t)r * kObs + q] = q == 0 ? 0.5 * (double)jeff * s_v[0] : s_v[0];
    __syncthreads();
  }
}

// --- host side ----------------------------------------------------------------
static inline uint32_t bk0(const pyro_ctx *c) { return (uint32_t)c->seed; }
static inline uint32_t bk1(const pyro_ctx *c) { return (uint32_t)(c->seed >> 32); }

void batch_free(pyro_ctx *ctx) {
  void *ptrs[] = {ctx->d_L,     ctx->d_off,      ctx->d_order,   ctx->d_bin_start,
                  ctx->d_tiles, ctx->d_part_off, ctx->d_tile_part};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (auto &row : ctx->d_tiles_part)
    for (int *&p : row) {
      if (p) (void)hipFree(p);
      p = nullptr;
    }
  for (auto &row : ctx->n_tiles_part)
    for (int &n : row) n = 0;
  for (int x = 0; x < 2; ++x) {
    if (ctx->d_tiles_x[x]) (void)hipFree(ctx->d_tiles_x[x]);
    ctx->d_tiles_x[x] = nullptr;
    ctx->n_tiles_x[x] = 0;
    for (int n = 0; n <= kMaxGroups; ++n)
      for (int p = 0; p < kMaxGroups; ++p) {
        if (ctx->d_tiles_part_x[x][n][p]) (void)hipFree(ctx->d_tiles_part_x[x][n][p]);
        ctx->d_tiles_part_x[x][n][p] = nullptr;
        ctx->n_tiles_part_x[x][n][p] = 0;
      }
  }
  ctx->d_L = nullptr;
  ctx->d_off = nullptr;
  ctx->d_order = ctx->d_bin_start = ctx->d_tiles = ctx->d_part_off = nullptr;
  ctx->d_tile_part = nullptr;
}

// Tiles of `tile` cells of replicas [r_begin, r_end) in replica order, dealt
// to the 8 XCDs in contiguous chunks: hardware block b runs on XCD b % 8, so
// list slot b takes logical tile (b % 8) * chunk + b / 8. Padded to a
// multiple of 8 with r = -1.
static std::vector<int> tile_list(const pyro_ctx *ctx, int r_begin, int r_end, int tile = 256) {
  std::vector<int> logical;
  for (int r = r_begin; r < r_end; ++r) {
    const int nc = ctx->h_L[r] * ctx->h_L[r] * ctx->h_L[r];
    for (int t = 0; t < (nc + tile - 1) / tile; ++t) {
      logical.push_back(r);
      logical.push_back(t);
    }
  }
  const int nt = (int)logical.size() / 2;
  const int chunk = (nt + 7) / 8;
  const int padded = chunk * 8;
  std::vector<int> tiles(2 * (size_t)padded, -1);
  for (int b = 0; b < padded; ++b) {
    const int lt = (b % 8) * chunk + b / 8;
    if (lt < nt) {
      tiles[2 * b] = logical[2 * lt];
      tiles[2 * b + 1] = logical[2 * lt + 1];
    }
  }
  return tiles;
}

// Geometry tables, persistent-kernel bins (LPT packing of work ~ L^3 into
// n_bins bins; n_bins <= 0 means one bin per replica) and the tile list.
int batch_setup(pyro_ctx *ctx, int n_bins) {
  const int R = ctx->R;
  if (n_bins <= 0 || n_bins > R) n_bins = R;
  std::vector<int> order(R);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ctx->h_L[a] > ctx->h_L[b]; });
  std::vector<std::vector<int>> bins(n_bins);
  std::vector<long long> load(n_bins, 0);
  for (int r : order) {
    const int b = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    bins[b].push_back(r);
    load[b] += (long long)ctx->h_L[r] * ctx->h_L[r] * ctx->h_L[r];
  }
  std::vector<int> flat, start(1, 0);
  for (auto &bv : bins) {
    flat.insert(flat.end(), bv.begin(), bv.end());
    start.push_back((int)flat.size());
  }
  const std::vector<int> tiles = tile_list(ctx, 0, R);
  // Each replica's run of tile partials for the observables.
  std::vector<int> part_off(R + 1, 0);
  for (int r = 0; r < R; ++r)
    part_off[r + 1] = part_off[r] + (ctx->h_L[r] * ctx->h_L[r] * ctx->h_L[r] + 255) / 256;
  // parts[n][p]: group p of a split into n groups (n = 2 .. kMaxGroups, n <= R).
  std::vector<int> parts[kMaxGroups + 1][kMaxGroups];
  for (int n = 2; n <= kMaxGroups && n <= R; ++n)
    for (int p = 0; p < n; ++p) parts[n][p] = tile_list(ctx, p * R / n, (p + 1) * R / n);
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  batch_free(ctx);
  PYRO_HIP(hipMalloc((void **)&ctx->d_L, sizeof(int) * R));
  PYRO_HIP(hipMalloc((void **)&ctx->d_off, sizeof(long long) * R));
  PYRO_HIP(hipMalloc((void **)&ctx->d_order, sizeof(int) * R));
  PYRO_HIP(hipMalloc((void **)&ctx->d_bin_start, sizeof(int) * (n_bins + 1)));
  PYRO_HIP(hipMalloc((void **)&ctx->d_tiles, sizeof(int) * tiles.size()));
  PYRO_HIP(hipMemcpy(ctx->d_L, ctx->h_L.data(), sizeof(int) * R, hipMemcpyHostToDevice));
  PYRO_HIP(hipMemcpy(ctx->d_off, ctx->h_off.data(), sizeof(long long) * R, hipMemcpyHostToDevice));
  PYRO_HIP(hipMemcpy(ctx->d_order, flat.data(), sizeof(int) * R, hipMemcpyHostToDevice));
  PYRO_HIP(hipMemcpy(ctx->d_bin_start, start.data(), sizeof(int) * (n_bins + 1), hipMemcpyHostToDevice));
  PYRO_HIP(hipMemcpy(ctx->d_tiles, tiles.data(), sizeof(int) * tiles.size(), hipMemcpyHostToDevice));
  PYRO_HIP(hipMalloc((void **)&ctx->d_part_off, sizeof(int) * (R + 1)));
  PYRO_HIP(hipMemcpy(ctx->d_part_off, part_off.data(), sizeof(int) * (R + 1), hipMemcpyHostToDevice));
  PYRO_HIP(hipMalloc((void **)&ctx->d_tile_part, sizeof(double) * kObs * (size_t)part_off[R]));
  for (int n = 2; n <= kMaxGroups; ++n)
    for (int p = 0; p < n; ++p) {
      const std::vector<int> &t = parts[n][p];
      if (t.empty()) continue;
      PYRO_HIP(hipMalloc((void **)&ctx->d_tiles_part[n][p], sizeof(int) * t.size()));
      PYRO_HIP(hipMemcpy(ctx->d_tiles_part[n][p], t.data(), sizeof(int) * t.size(),
                         hipMemcpyHostToDevice));
      ctx->n_tiles_part[n][p] = (int)t.size() / 2;
    }
  // The lists of 128- and 512-cell tiles, whole and per replica group.
  for (int x = 0; x < 2; ++x) {
    const int tile = x == 0 ? 128 : 512;
    const std::vector<int> all = tile_list(ctx, 0, R, tile);
    PYRO_HIP(hipMalloc((void **)&ctx->d_tiles_x[x], sizeof(int) * all.size()));
    PYRO_HIP(hipMemcpy(ctx->d_tiles_x[x], all.data(), sizeof(int) * all.size(), hipMemcpyHostToDevice));
    ctx->n_tiles_x[x] = (int)all.size() / 2;
    for (int n = 2; n <= kMaxGroups && n <= R; ++n)
      for (int p = 0; p < n; ++p) {
        const std::vector<int> t = tile_list(ctx, p * R / n, (p + 1) * R / n, tile);
        if (t.empty()) continue;
        PYRO_HIP(hipMalloc((void **)&ctx->d_tiles_part_x[x][n][p], sizeof(int) * t.size()));
        PYRO_HIP(hipMemcpy(ctx->d_tiles_part_x[x][n][p], t.data(), sizeof(int) * t.size(),
                           hipMemcpyHostToDevice));
        ctx->n_tiles_part_x[x][n][p] = (int)t.size() / 2;
      }
  }
  ctx->n_bins = n_bins;
  ctx->n_tiles = (int)tiles.size() / 2;
  return 0;
}

int launch_batch_init_spins(pyro_ctx *ctx) {
  int maxL = *std::max_element(ctx->h_L.begin(), ctx->h_L.end());
  dim3 grid((maxL * maxL * maxL + 255) / 256, ctx->R);
  k_batch_init<<<grid, 256, 0, ctx->stream>>>(ctx->d_spins, ctx->d_off, ctx->d_L, bk0(ctx), bk1(ctx));
  PYRO_HIP(hipGetLastError());
  return 0;
}

int launch_batch_observables(pyro_ctx *ctx) {
  k_batch_obs_tiles<<<ctx->n_tiles, 256, 0, ctx->stream>>>(ctx->d_spins, ctx->d_off, ctx->d_L,
                                                           ctx->d_tiles, ctx->d_part_off,
                                                           ctx->d_tile_part);
  k_batch_obs_final<<<ctx->R, 256, 0, ctx->stream>>>(ctx->d_tile_part, ctx->d_part_off, ctx->jeff,
                                                     ctx->d_obs);
  PYRO_HIP(hipGetLastError());
  return 0;
}

// Whether the autotuner tries the persistent execution: only when every
// replica is staged in LDS (L <= kLdsMaxL). A block owns whole replicas, so a
// larger replica serialises on one CU: 256 mixed replicas of L in
// {8, 12, 16, 24} took 5.9-12 ms per sweep persistent against 1.58 ms as
// tile-list phases, while one L=4 replica took 21.6 us persistent against
// 70 us as per-phase kernels.
bool persistent_candidate(const pyro_ctx *ctx) {
  for (int L : ctx->h_L)
    if (L > kLdsMaxL) return false;
  return true;
}

// Largest replica size of the batch that the persistent kernel stages in LDS
// (0: none), and the dynamic LDS bytes that takes.
static int lds_staged_L(const pyro_ctx *ctx) {
  int m = 0;
  for (int L : ctx->h_L)
    if (L <= kLdsMaxL && L > m) m = L;
  return m;
}

template <int BLOCK, bool TRACK>
static int launch_persistent_t(pyro_ctx *ctx, int n_sweeps, int record_last) {
  const int lmax = lds_staged_L(ctx);
  const size_t lds = sizeof(float) * 48 * (size_t)lmax * lmax * lmax;
  if (lds > 65536)
    PYRO_HIP(hipFuncSetAttribute((const void *)k_batch_sweep<BLOCK, TRACK>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  k_batch_sweep<BLOCK, TRACK><<<ctx->n_bins, BLOCK, lds, ctx->stream>>>(
      ctx->d_spins, ctx->d_off, ctx->d_L, ctx->d_order, ctx->d_bin_start, ctx->n_metro, ctx->n_or,
      ctx->d_beta, ctx->d_cosc, ctx->d_ctl, bk0(ctx), bk1(ctx), ctx->jeff, ctx->d_acc, ctx->d_dE,
      n_sweeps, record_last, lmax, ctx->R, ctx->d_replica_at, ctx->d_series, ctx->d_anchor);
  return 0;
}

// All tiles of the current tile size, or the list of one replica group while
// a split sweep has set a launch range (always group launch_part of n_parts).
static void launch_tile_range(const pyro_ctx *ctx, const int *&tiles, int &n) {
  const int x = ctx->tile_cells == 128 ? 0 : (ctx->tile_cells == 512 ? 1 : -1);
  if (x < 0) {
    tiles = ctx->d_tiles;
    n = ctx->n_tiles;
    if (ctx->launch_nrep) {
      tiles = ctx->d_tiles_part[ctx->n_parts][ctx->launch_part];
      n = ctx->n_tiles_part[ctx->n_parts][ctx->launch_part];
    }
    return;
  }
  tiles = ctx->d_tiles_x[x];
  n = ctx->n_tiles_x[x];
  if (ctx->launch_nrep) {
    tiles = ctx->d_tiles_part_x[x][ctx->n_parts][ctx->launch_part];
    n = ctx->n_tiles_part_x[x][ctx->n_parts][ctx->launch_part];
  }
}

template <int KIND, int COLOUR, bool TRACK, int TILE>
static void launch_tiles_tt(pyro_ctx *ctx, int pass, const int *tiles, int n) {
  k_batch_phase<KIND, COLOUR, TRACK, TILE><<<n, TILE, 0, launch_stream(ctx)>>>(
      ctx->d_spins, ctx->d_off, ctx->d_L, tiles, pass, ctx->d_beta, ctx->d_cosc, ctx->d_ctl,
      bk0(ctx), bk1(ctx), ctx->jeff, ctx->d_acc, ctx->d_dE, ctx->launch_delta);
}

template <int KIND, int COLOUR, bool TRACK>
static void launch_tiles_t(pyro_ctx *ctx, int pass) {
  const int *tiles;
  int n;
  launch_tile_range(ctx, tiles, n);
  if (n == 0) return;
  if (ctx->tile_cells == 128)
    launch_tiles_tt<KIND, COLOUR, TRACK, 128>(ctx, pass, tiles, n);
  else if (ctx->tile_cells == 512)
    launch_tiles_tt<KIND, COLOUR, TRACK, 512>(ctx, pass, tiles, n);
  else
    launch_tiles_tt<KIND, COLOUR, TRACK, 256>(ctx, pass, tiles, n);
}

template <bool TRACK, int TILE>
static void launch_tiles_or_tt(pyro_ctx *ctx, const float *cur, float *oth, const int *tiles, int n) {
  k_batch_or_pair<0, TRACK, TILE><<<n, TILE, 0, launch_stream(ctx)>>>(cur, oth, ctx->d_off, ctx->d_L,
                                                                     tiles, ctx->jeff, ctx->d_dE);
  k_batch_or_pair<2, TRACK, TILE><<<n, TILE, 0, launch_stream(ctx)>>>(cur, oth, ctx->d_off, ctx->d_L,
                                                                     tiles, ctx->jeff, ctx->d_dE);
}

// The fused over-relaxation pass `parity` (0: d_spins -> second buffer,
// 1: back) over the tile list of the current launch range.
template <bool TRACK>
static void launch_tiles_or_pass(pyro_ctx *ctx, int parity) {
  const int *tiles;
  int n;
  launch_tile_range(ctx, tiles, n);
  if (n == 0) return;
  const float *cur = parity ? ctx->d_spins_alt : ctx->d_spins;
  float *oth = parity ? ctx->d_spins : ctx->d_spins_alt;
  if (ctx->tile_cells == 128)
    launch_tiles_or_tt<TRACK, 128>(ctx, cur, oth, tiles, n);
  else if (ctx->tile_cells == 512)
    launch_tiles_or_tt<TRACK, 512>(ctx, cur, oth, tiles, n);
  else
    launch_tiles_or_tt<TRACK, 256>(ctx, cur, oth, tiles, n);
}

template <int KIND, bool TRACK>
static void launch_tiles_pass(pyro_ctx *ctx, int pass) {
  launch_tiles_t<KIND, 0, TRACK>(ctx, pass);
  launch_tiles_t<KIND, 1, TRACK>(ctx, pass);
  launch_tiles_t<KIND, 2, TRACK>(ctx, pass);
  launch_tiles_t<KIND, 3, TRACK>(ctx, pass);
}

int launch_batch_sweeps(pyro_ctx *ctx, int n_sweeps, int record_last) {
  const bool tr = ctx->track_or;
  int rc;
  if (ctx->batch_block == 1024)
    rc = tr ? launch_persistent_t<1024, true>(ctx, n_sweeps, record_last)
            : launch_persistent_t<1024, false>(ctx, n_sweeps, record_last);
  else if (ctx->batch_block == 512)
    rc = tr ? launch_persistent_t<512, true>(ctx, n_sweeps, record_last)
            : launch_persistent_t<512, false>(ctx, n_sweeps, record_last);
  else
    rc = tr ? launch_persistent_t<256, true>(ctx, n_sweeps, record_last)
            : launch_persistent_t<256, false>(ctx, n_sweeps, record_last);
  if (rc) return rc;
  PYRO_HIP(hipGetLastError());
  return 0;
}

int launch_batch_sweep_body(pyro_ctx *ctx) {
  const bool tr = ctx->track_or;
  if (ctx->exec_mode == 1) {
    PYRO_TRY(launch_batch_sweeps(ctx, 1, 0));
  } else {
    for (int p = 0; p < ctx->n_metro; ++p) launch_tiles_pass<0, false>(ctx, p);
    int o = 0;
    if (ctx->fuse_or && ctx->d_spins_alt && !ctx->mat) {
      // Two colours per launch, alternating between the spin buffers; an even
      // number of fused passes ends in d_spins, an odd last pass runs in place.
      for (; o < (ctx->n_or & ~1); ++o) {
        if (tr) launch_tiles_or_pass<true>(ctx, o & 1);
        else launch_tiles_or_pass<false>(ctx, o & 1);
      }
    }
    for (; o < ctx->n_or; ++o) {
      if (tr) launch_tiles_pass<1, true>(ctx, 0);
      else launch_tiles_pass<1, false>(ctx, 0);
    }
  }
  PYRO_HIP(hipGetLastError());
  return 0;
}

}  // namespace pyro
