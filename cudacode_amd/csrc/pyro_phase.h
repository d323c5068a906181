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
static __constant__ int8_t c_sets[24][4] = PYRO_INDEPENDENT_SETS_INIT;

__device__ __forceinline__ int wrap(int v, int L) {
  return v < 0 ? v + L : (v >= L ? v - L : v);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// One phase on every replica. KIND 0: Metropolis with the cone proposal
// (MCMultipleTemperaturesPerSystem, Pyrochlore.cu:829-885). KIND 1:
// over-relaxation by pi (OverRelaxationByPiStep, Pyrochlore.cu:975-1007).
// KIND 2: rotation about the local field by a uniform random angle
// (OverRelaxationStep, Pyrochlore.cu:1009-1045). grid = (ceil(ncell/64), R),
// block = 256. LT > 0 fixes the lattice size at compile time (index
// arithmetic by constants); LT = 0 reads it from L. MAT: the matrix field
// model (coup).
//
// Phase `colour` updates the colour class {c, c+4, c+8, c+12}. set_mode 1:
// it updates one of the 24 independent sets instead, drawn from Philox per
// (sweep, stage, pass, phase) and the same for every replica and block (the
// reference's IndependentSubLattices[Update[0] % 24], Pyrochlore.cu:848); a
// site may then be updated by two phases of one pass, so the move draws of a
// drawn set carry the phase in their site field. set_mode 2: it updates the
// colour class perm[colour] of a permutation of the four drawn per (sweep,
// stage, pass) by a Fisher-Yates shuffle on one Philox word (every site
// still moves once a pass; the draws keep their counters).
// `pass` numbers the Metropolis or over-relaxation pass within the sweep.
template <int KIND, int LT, bool MAT>
__global__ __launch_bounds__(kPhaseBlock) void k_phase(
    float *__restrict__ spins, int L_rt, int ncell_rt, int colour, int pass,
    const float *__restrict__ beta, const float *__restrict__ cosc,
    const uint32_t *__restrict__ ctl, uint32_t key0, uint32_t key1, float jeff,
    unsigned long long *__restrict__ acc, double *__restrict__ dE, const float *__restrict__ coup,
    int post_rotate, int r0, uint32_t sweep_delta, int set_mode) {
  const int L = LT > 0 ? LT : L_rt;
  const int ncell = LT > 0 ? LT * LT * LT : ncell_rt;
  const int r = r0 + (int)blockIdx.y;
  const int k = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int cell = blockIdx.x * kPhaseCells + lane;
  const uint32_t sweep = ctl[kCtlSweep] + sweep_delta;
  int site = colour + 4 * k;
  uint32_t tag = (uint32_t)site;  // site field of this update's draws
  if (set_mode == 1) {
    const u32x4 q = philox4x32_10(
        rng_counter((uint32_t)colour, KIND == 0 ? 0u : 1u, (uint32_t)pass, kKindSets, 0u, sweep),
        key0, key1);
    site = __builtin_amdgcn_readfirstlane((int)c_sets[q.x % 24u][k]);
    tag = (uint32_t)site | ((uint32_t)colour << 4);
  } else if (set_mode == 2) {
    const u32x4 q = philox4x32_10(
        rng_counter(0u, KIND == 0 ? 0u : 1u, (uint32_t)pass, kKindSets, 1u, sweep), key0, key1);
    int perm[4] = {0, 1, 2, 3};
    uint32_t v = q.x;
#pragma unroll
    for (int i = 3; i > 0; --i) {
      const int j = (int)(v % (uint32_t)(i + 1));
      v /= (uint32_t)(i + 1);
      const int t = perm[i];
      perm[i] = perm[j];
      perm[j] = t;
    }
    int c = perm[0];
#pragma unroll
    for (int p = 1; p < 4; ++p)
      if (colour == p) c = perm[p];
    site = __builtin_amdgcn_readfirstlane(c + 4 * k);
    tag = (uint32_t)site;
  }
  const int b = site & 3;  // basis site (the matrix model's couplings)
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
    field_start<MAT>(coup, b, fx, fy, fz);
    double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
    for (int n = 0; n < 6; ++n) {
      const Link lk = c_links[site][n];
      const int x1 = wrap(x + lk.dx, L);
      const int y1 = wrap(y + lk.dy, L);
      const int z1 = wrap(z + lk.dz, L);
      const int j = lk.site * ncell + (z1 * L + y1) * L + x1;
      const float sa = Sx[j], sb = Sy[j], sc = Sz[j];
      field_add<MAT, true>(coup, b, lk.site & 3, sa, sb, sc, fx, fy, fz, gx, gy, gz);
    }
    const int i = site * ncell + cell;
    const f3 s = {Sx[i], Sy[i], Sz[i]};
    f3 sp;
    bool write;
    if (KIND == 0) {
      const u32x4 w = philox4x32_10(
          rng_counter((uint32_t)cell, tag, (uint32_t)pass, kKindMetropolis, (uint32_t)r, sweep),
          key0, key1);
      const f3 h = field_h<MAT>(jeff, fx, fy, fz);
      sp = cone_proposal(s, cosc[r], u01(w.x), u01(w.y));
      write = metropolis_accept(h, s, sp, beta[r], u01(w.z));
      accepted = write ? 1 : 0;
      if (post_rotate) {
        sp = rotate_about(h, write ? sp : s, u01(w.w));
        write = true;
      }
    } else if (KIND == 2) {
      const u32x4 w = philox4x32_10(
          rng_counter((uint32_t)cell, tag, (uint32_t)pass, kKindOrAngle, (uint32_t)r, sweep),
          key0, key1);
      sp = rotate_about(field_h<MAT>(jeff, fx, fy, fz), s, u01(w.x));
      write = true;
    } else {
      const f3 f = {fx, fy, fz};
      sp = overrelax(f, s);
      write = true;
    }
    if (write) {
      de = field_de<MAT>(jeff, coup, b, s, sp, gx, gy, gz);
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

// The site-planes (site, dz) that the fused pair of a z-plane of cells reads:
// the 16 sites of the plane itself and the 4 sites of one adjacent plane that
// the pair's links reach (z - 1 for CA = 0, z + 1 for CA = 2), 20 in all;
// slot[site][dz + 1] is a plane's position in that list (or -1). MODE 3
// takes the 3 foreign colour-CA values from their owner lanes (all three lie
// in the block's plane) instead of recomputing them, which leaves 29 values
// a cell reads. A plane is held in one of three ways (MODE 0: all staged in
// LDS; MODE 1: LDS or own; MODE 2 and 3: all three):
//  - kPlaneLds: staged in LDS, read at in-plane offsets that change y;
//  - kPlaneOwn: read only at the cell's own position: loaded into registers;
//  - kPlaneRow: read only at x offsets -1, 0, 1: loaded into registers at the
//    own position and rotated along x within the 16-lane DPP row, which at
//    L = 16 is exactly one periodic row of cells.
enum : int8_t { kPlaneLds = 0, kPlaneOwn = 1, kPlaneRow = 2 };

template <int CA, int MODE>
struct PairPlanes {
  int8_t site[24] = {}, dz[24] = {}, kind[24] = {}, lds[24] = {};
  int8_t slot[16][3] = {};
  // LDS row k = 3 * lds + component: the row's float offset in a replica's
  // buffer without its z term (component * 16 N + site * N, N = 4096),
  // (dz + 1) << 20 and (read from the second buffer) << 22.
  uint32_t row[72] = {};
  int n = 0, n_lds = 0;
  constexpr PairPlanes() {
    constexpr PairNeeds<CA, (MODE < 3)> nd{};
    for (int s = 0; s < 16; ++s)
      for (int d = 0; d < 3; ++d) slot[s][d] = -1;
    bool dx_used[24] = {}, dy_used[24] = {};
    for (int i = 0; i < nd.n; ++i) {
      const int s = nd.v[i].site, d = nd.v[i].dz + 1;
      if (slot[s][d] < 0) {
        slot[s][d] = (int8_t)n;
        site[n] = (int8_t)s;
        dz[n] = (int8_t)(d - 1);
        ++n;
      }
      dx_used[slot[s][d]] |= nd.v[i].dx != 0;
      dy_used[slot[s][d]] |= nd.v[i].dy != 0;
    }
    for (int sl = 0; sl < n; ++sl) {
      kind[sl] = kPlaneLds;
      if (MODE >= 1 && !dy_used[sl] && !dx_used[sl]) kind[sl] = kPlaneOwn;
      if (MODE >= 2 && !dy_used[sl] && dx_used[sl]) kind[sl] = kPlaneRow;
      lds[sl] = -1;
      if (kind[sl] == kPlaneLds) {
        for (int comp = 0; comp < 3; ++comp)
          row[3 * n_lds + comp] = (uint32_t)(comp * 16 * 4096 + site[sl] * 4096) |
                                  (uint32_t)(dz[sl] + 1) << 20 | (uint32_t)((site[sl] & 3) < CA) << 22;
        lds[sl] = (int8_t)n_lds++;
      }
    }
  }
};

// The LDS staging rows as a device constant (a wave indexes it with a uniform
// value: one scalar load per row).
template <int CA, int MODE>
__constant__ constexpr uint32_t kPairRows[72] = {
#define PYRO_ROW(k) PairPlanes<CA, MODE>{}.row[k]
    PYRO_ROW(0),  PYRO_ROW(1),  PYRO_ROW(2),  PYRO_ROW(3),  PYRO_ROW(4),  PYRO_ROW(5),
    PYRO_ROW(6),  PYRO_ROW(7),  PYRO_ROW(8),  PYRO_ROW(9),  PYRO_ROW(10), PYRO_ROW(11),
    PYRO_ROW(12), PYRO_ROW(13), PYRO_ROW(14), PYRO_ROW(15), PYRO_ROW(16), PYRO_ROW(17),
    PYRO_ROW(18), PYRO_ROW(19), PYRO_ROW(20), PYRO_ROW(21), PYRO_ROW(22), PYRO_ROW(23),
    PYRO_ROW(24), PYRO_ROW(25), PYRO_ROW(26), PYRO_ROW(27), PYRO_ROW(28), PYRO_ROW(29),
    PYRO_ROW(30), PYRO_ROW(31), PYRO_ROW(32), PYRO_ROW(33), PYRO_ROW(34), PYRO_ROW(35),
    PYRO_ROW(36), PYRO_ROW(37), PYRO_ROW(38), PYRO_ROW(39), PYRO_ROW(40), PYRO_ROW(41),
    PYRO_ROW(42), PYRO_ROW(43), PYRO_ROW(44), PYRO_ROW(45), PYRO_ROW(46), PYRO_ROW(47),
    PYRO_ROW(48), PYRO_ROW(49), PYRO_ROW(50), PYRO_ROW(51), PYRO_ROW(52), PYRO_ROW(53),
    PYRO_ROW(54), PYRO_ROW(55), PYRO_ROW(56), PYRO_ROW(57), PYRO_ROW(58), PYRO_ROW(59),
#undef PYRO_ROW
};

// Value of the cell x + DX (DX = -1 or 1, periodic within the row of 16) from
// the lanes of the same 16-lane DPP row: row_ror:n gives lane i the value of
// lane (i - n) mod 16.
template <int DX>
__device__ __forceinline__ float row_neighbour(float v) {
  constexpr int ctrl = 0x120 + (DX < 0 ? 1 : 15);  // row_ror:1 or row_ror:15
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), ctrl, 0xF, 0xF, false));
}

// The colour-CA sites of a cell whose new values its neighbours in other rows
// of the plane need (foreign neighbours at dy != 0): k[e] = the site's index
// k (site CA + 4k) in the exchange, slot[k] = its exchange row (or -1).
template <int CA>
struct PairExchange {
  int k[4] = {}, slot[4] = {-1, -1, -1, -1};
  int n = 0;
  constexpr PairExchange() {
    constexpr PairNeeds<CA, false> nd{};
    for (int f = 0; f < nd.nf; ++f) {
      const int kk = nd.fv[f].site >> 2;
      if (nd.fv[f].dy != 0 && slot[kk] < 0) {
        slot[kk] = n;
        k[n++] = kk;
      }
    }
  }
};
template <int CA>
constexpr PairExchange<CA> kPairExchange{};

// Blocks per CU that the staged pair's LDS allows (160 KB per CU).
constexpr int or_plane_blocks(int mode) { return mode == 0 ? 2 : mode == 1 ? 3 : mode == 2 ? 4 : 5; }

// The fused pair at L = 16, one z-plane of a replica per block (256 cells, a
// cell per thread), with the planes it reads at neighbouring positions staged
// in LDS: each wave copies whole rows of 256 floats (one site-plane component)
// with 16-byte loads, all issued in one burst with the thread's own register
// loads, and after a barrier every thread takes its cell's 41 values from LDS
// (in-plane neighbours wrap within the staged rows), from its registers or from
// its DPP row. Same arithmetic (or_pair_phase_a / _b) and stores as k_or_pair,
// so the same results bit for bit. k_or_pair issues 123 four-byte loads per
// thread, twice what a wave may have in flight, i.e. two memory round trips.
// LDS per block: 20 / 17 / 13 / 6 planes x 3 KB for MODE 0 / 1 / 2 / 3 (and
// MODE 3's 6 KB exchange rows), or_plane_blocks(MODE) blocks per CU. Block order: replica-major z-planes behind the XCD remap, so
// the planes of one replica run on one XCD and share its L2 for the halo.
template <int CA, int MODE>
__global__ __launch_bounds__(256, or_plane_blocks(MODE)) void k_or_pair_plane(
    const float *__restrict__ cur, float *__restrict__ oth, float jeff, int r0,
    const uint32_t *__restrict__ ctl, uint32_t sweep_delta, double *__restrict__ series,
    int R_total, const int *__restrict__ replica_at, const double *__restrict__ anchor,
    const double *__restrict__ dE) {
  constexpr int LT = 16, NC = LT * LT * LT, P = 16 * NC;
  constexpr bool RECOMP = MODE < 3;
  constexpr PairPlanes<CA, MODE> pp{};
  constexpr PairNeeds<CA, RECOMP> nd{};
  constexpr int NS = pp.n, NN = nd.n, ROWS = 3 * pp.n_lds;
  static_assert(NS == 20 && ROWS <= 72, "20 site-planes");
  __shared__ float4 sm[ROWS * 64];
  const int nb = gridDim.x;
  const int b = blockIdx.x;
  const int lb = (nb & 7) == 0 ? (b & 7) * (nb >> 3) + (b >> 3) : b;
  const int q = lb >> 4;  // 16 planes per replica
  const int r = r0 + q;
  const int z = lb & 15;
  if (series != nullptr && b == 0) {
    const uint32_t row = ctl[kCtlRecordRow] + sweep_delta;
    if (ctl[kCtlRecordOn] && row < ctl[kCtlRecordCap]) {
      const int r1 = r0 + (nb >> 4);
      for (int k = threadIdx.x; k < R_total; k += 256) {
        const int rep = replica_at[k];
        if (rep >= r0 && rep < r1) series[(size_t)row * R_total + k] = anchor[rep] + dE[rep];
      }
    }
  }
  const float *rc = cur + (size_t)r * 48 * NC;
  float *ro = oth + (size_t)r * 48 * NC;
  const int t = threadIdx.x, x = t & 15, y = t >> 4;
  const int lane = t & 63;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  // The planes held in registers: this cell's value.
  float ox[NS], oy[NS], oz[NS];
#pragma unroll
  for (int sl = 0; sl < NS; ++sl) {
    if (pp.kind[sl] == kPlaneLds) continue;
    const float *src = (pp.site[sl] & 3) < CA ? ro : rc;
    const uint32_t j = 4u * (uint32_t)(pp.site[sl] * NC + ((z + pp.dz[sl]) & 15) * 256 + t);
    ox[sl] = ld_at(src, j);
    oy[sl] = ld_at(src + P, j);
    oz[sl] = ld_at(src + 2 * P, j);
  }
  // LDS row k = 3 * lds + component; wave w copies rows w, w + 4, ... (uniform).
  // Every load is issued before the first LDS write, and no write sits in a
  // branch (a write that waited for its own load inside one would serialise
  // the round trips): a wave with no row left in the last round copies the
  // last row again, writing the same bytes to the same place as its owner.
  constexpr int NI = (ROWS + 3) / 4;
  float4 st[NI];
#pragma unroll
  for (int i = 0; i < NI; ++i) {
    const int k = min(4 * i + w, ROWS - 1);
    const uint32_t e = kPairRows<CA, MODE>[k];
    const float *src = (e >> 22) ? ro : rc;
    const int zz = (z + (int)((e >> 20) & 3) - 1) & 15;
    const uint32_t off = 4u * ((e & 0xFFFFFu) + (uint32_t)zz * 256u) + 16u * (uint32_t)lane;
    st[i] = *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(src) + off);
  }
#pragma unroll
  for (int i = 0; i < NI; ++i) sm[min(4 * i + w, ROWS - 1) * 64 + lane] = st[i];
  __syncthreads();
  const float *sf = reinterpret_cast<const float *>(sm);
  const int X[3] = {(x + 15) & 15, x, (x + 1) & 15};
  const int Y[3] = {((y + 15) & 15) * 16, y * 16, ((y + 1) & 15) * 16};
  float vx[NN], vy[NN], vz[NN];
#pragma unroll
  for (int i = 0; i < NN; ++i) {
    const PairNeed qn = nd.v[i];
    const int sl = pp.slot[qn.site][qn.dz + 1];
    if (pp.kind[sl] == kPlaneLds) {
      const int at = Y[qn.dy + 1] + X[qn.dx + 1];
      const int l = pp.lds[sl];
      vx[i] = sf[(3 * l + 0) * 256 + at];
      vy[i] = sf[(3 * l + 1) * 256 + at];
      vz[i] = sf[(3 * l + 2) * 256 + at];
    } else if (qn.dx == 0) {
      vx[i] = ox[sl];
      vy[i] = oy[sl];
      vz[i] = oz[sl];
    } else if (qn.dx < 0) {
      vx[i] = row_neighbour<-1>(ox[sl]);
      vy[i] = row_neighbour<-1>(oy[sl]);
      vz[i] = row_neighbour<-1>(oz[sl]);
    } else {
      vx[i] = row_neighbour<1>(ox[sl]);
      vy[i] = row_neighbour<1>(oy[sl]);
      vz[i] = row_neighbour<1>(oz[sl]);
    }
  }
  f3 na[4], nbv[4];
  double de = 0.0;
  if constexpr (RECOMP) {
    or_pair_math<CA, false>(vx, vy, vz, jeff, de, na, nbv);
  } else {
    // The foreign colour-CA neighbours' new values from their owners: along
    // the row by DPP, across rows through LDS (a row per component of each
    // such site), written after phase CA and read after a barrier.
    static_assert(nd.nf == 3, "3 foreign neighbours per pair");
    constexpr int NX = kPairExchange<CA>.n;
    __shared__ float xs[(NX > 0 ? NX : 1) * 3 * 256];
    or_pair_phase_a<CA, false, false>(vx, vy, vz, jeff, de, na);
#pragma unroll
    for (int e = 0; e < NX; ++e) {
      const f3 v = na[kPairExchange<CA>.k[e]];
      xs[(3 * e + 0) * 256 + t] = v.x;
      xs[(3 * e + 1) * 256 + t] = v.y;
      xs[(3 * e + 2) * 256 + t] = v.z;
    }
    __syncthreads();
    f3 fo[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      const PairNeed q = nd.fv[f];
      static_assert(nd.fv[0].dz == 0 && nd.fv[1].dz == 0 && nd.fv[2].dz == 0,
                    "foreign neighbours in the block's plane");
      const f3 own = na[q.site >> 2];
      if (q.dy == 0 && q.dx < 0) {
        fo[f] = {row_neighbour<-1>(own.x), row_neighbour<-1>(own.y), row_neighbour<-1>(own.z)};
      } else if (q.dy == 0) {
        fo[f] = {row_neighbour<1>(own.x), row_neighbour<1>(own.y), row_neighbour<1>(own.z)};
      } else {
        const int e = kPairExchange<CA>.slot[q.site >> 2];
        const int at = Y[q.dy + 1] + X[q.dx + 1];
        fo[f] = {xs[(3 * e + 0) * 256 + at], xs[(3 * e + 1) * 256 + at],
                 xs[(3 * e + 2) * 256 + at]};
      }
    }
    or_pair_phase_b<CA, false, false>(vx, vy, vz, jeff, de, na, fo, nbv);
  }
  or_pair_store<CA, NC>(ro, z * 256 + t, na, nbv);
}

// The staging the L = 16 fused pair uses. Measured over L = 16 x 64 replicas
// (one launch over all replicas; sweep on three streams), against 15.65 us and
// 0.326 ms for k_or_pair: MODE 0 16.3 us, 0.329 ms; MODE 1 14.5 us, 0.321 ms;
// MODE 2 13.2 us (5.7 TB/s at 36 B per update), 0.322 ms; MODE 3 (the foreign
// values exchanged: 29 values a cell, 6 planes in LDS, ~600 instead of ~780
// VALU per wave, 94 VGPRs, 5 blocks per CU) 12.4 us (6.1 TB/s), 0.308 ms.
constexpr int kOrPlaneMode = 3;

// A whole over-relaxation pass (colours 0, 1, 2, 3 in turn) of one z-plane of
// an L = 16 replica in one launch, from cur into oth (all 16 sites written):
// 24 instead of 36 B per update. The colour-2 and -3 sites of plane z link to
// colour-0 and -1 sites of plane z + 1 (10, 11, 14, 15 to 0, 1, 4, 5), whose
// new values the block recomputes from planes z + 1 and z with their owners'
// float expressions in their owners' order (over-relaxation is deterministic),
// so they equal what the block of plane z + 1 computes bit for bit; the
// colour-0 and -1 sites link down to plane z - 1, whose old colour-2 and -3
// values they need: every read is from cur, which no block of the launch
// writes. PassPlan lists the site-planes read (plane index 0, 1, 2 =
// z - 1, z, z + 1) and holds each as the pair kernels do: in LDS when read at
// a y offset (kPlaneLds), in registers when read only at the own position
// (kPlaneOwn), in registers rotated along the DPP row when read only at x
// offsets (kPlaneRow). A new value replaces the old one where it is held: no
// site is read by another thread while its own colour is being updated, and a
// barrier separates a colour's writes from the first read of them elsewhere.
struct PassPlan {
  int8_t kind[16][3] = {}, slot[16][3] = {};
  bool redo[16] = {};  // colour-0/1 sites recomputed on plane z + 1
  // LDS row k = 3 * slot + component: the row's float offset in a replica's
  // buffer without its z term (component * 16 N + site * N, N = 4096) and
  // the plane index << 20.
  uint32_t row[72] = {};
  int n_lds = 0, n_reg = 0;
  constexpr void use(bool (&u)[16][3], bool (&ux)[16][3], bool (&uy)[16][3], int s, int p, int dx,
                     int dy) {
    u[s][p] = true;
    ux[s][p] = ux[s][p] || dx != 0;
    uy[s][p] = uy[s][p] || dy != 0;
  }
  constexpr PassPlan() {
    bool u[16][3] = {}, ux[16][3] = {}, uy[16][3] = {};
    for (int s = 0; s < 16; ++s) {
      use(u, ux, uy, s, 1, 0, 0);
      for (int m = 0; m < 6; ++m)
        use(u, ux, uy, kLinks[s][m].site, 1 + kLinks[s][m].dz, kLinks[s][m].dx, kLinks[s][m].dy);
      if ((s & 3) >= 2)
        for (int m = 0; m < 6; ++m)
          if (kLinks[s][m].dz == 1) redo[kLinks[s][m].site] = true;
    }
    for (int s = 0; s < 16; ++s)
      if (redo[s]) {
        use(u, ux, uy, s, 2, 0, 0);
        for (int m = 0; m < 6; ++m)
          use(u, ux, uy, kLinks[s][m].site, 2 + kLinks[s][m].dz, kLinks[s][m].dx, kLinks[s][m].dy);
      }
    for (int p = 0; p < 3; ++p)
      for (int s = 0; s < 16; ++s) {
        kind[s][p] = -1;
        slot[s][p] = -1;
        if (!u[s][p]) continue;
        if (uy[s][p]) {
          kind[s][p] = kPlaneLds;
          for (int comp = 0; comp < 3; ++comp)
            row[3 * n_lds + comp] = (uint32_t)(comp * 16 * 4096 + s * 4096) | (uint32_t)p << 20;
          slot[s][p] = (int8_t)n_lds++;
        } else {
          kind[s][p] = ux[s][p] ? kPlaneRow : kPlaneOwn;
          slot[s][p] = (int8_t)n_reg++;
        }
      }
  }
};
constexpr PassPlan kPassPlan{};
static_assert(kPassPlan.n_lds == 12 && kPassPlan.n_reg == 16,
              "plane z - 1: 2 LDS + 2 own; z: 8 LDS + 8 register; z + 1: 2 LDS + 6 own");

__constant__ constexpr uint32_t kPassRows[36] = {
#define PYRO_ROW(k) kPassPlan.row[k]
    PYRO_ROW(0),  PYRO_ROW(1),  PYRO_ROW(2),  PYRO_ROW(3),  PYRO_ROW(4),  PYRO_ROW(5),
    PYRO_ROW(6),  PYRO_ROW(7),  PYRO_ROW(8),  PYRO_ROW(9),  PYRO_ROW(10), PYRO_ROW(11),
    PYRO_ROW(12), PYRO_ROW(13), PYRO_ROW(14), PYRO_ROW(15), PYRO_ROW(16), PYRO_ROW(17),
    PYRO_ROW(18), PYRO_ROW(19), PYRO_ROW(20), PYRO_ROW(21), PYRO_ROW(22), PYRO_ROW(23),
    PYRO_ROW(24), PYRO_ROW(25), PYRO_ROW(26), PYRO_ROW(27), PYRO_ROW(28), PYRO_ROW(29),
    PYRO_ROW(30), PYRO_ROW(31), PYRO_ROW(32), PYRO_ROW(33), PYRO_ROW(34), PYRO_ROW(35),
#undef PYRO_ROW
};

// Value of site s on plane p at cell offset (dx, dy) from where PassPlan
// holds it (all arguments compile-time constants after unrolling).
__device__ __forceinline__ f3 pass_get(const float *sf, const float (&rx)[16],
                                       const float (&ry)[16], const float (&rz)[16], int s, int p,
                                       int dx, int dy, const int (&X)[3], const int (&Y)[3]) {
  const int k = kPassPlan.kind[s][p], sl = kPassPlan.slot[s][p];
  if (k == kPlaneLds) {
    const int at = Y[dy + 1] + X[dx + 1];
    return {sf[(3 * sl + 0) * 256 + at], sf[(3 * sl + 1) * 256 + at], sf[(3 * sl + 2) * 256 + at]};
  }
  if (dx < 0) return {row_neighbour<-1>(rx[sl]), row_neighbour<-1>(ry[sl]), row_neighbour<-1>(rz[sl])};
  if (dx > 0) return {row_neighbour<1>(rx[sl]), row_neighbour<1>(ry[sl]), row_neighbour<1>(rz[sl])};
  return {rx[sl], ry[sl], rz[sl]};
}

// Over-relaxation of site s on plane p from its 6 links in link order (the
// sum and the move of every over-relaxation kernel).
__device__ __forceinline__ f3 pass_relax(const float *sf, const float (&rx)[16],
                                         const float (&ry)[16], const float (&rz)[16], int s, int p,
                                         const int (&X)[3], const int (&Y)[3]) {
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    const Link lk = kLinks[s][m];
    const f3 v = pass_get(sf, rx, ry, rz, lk.site, p + lk.dz, lk.dx, lk.dy, X, Y);
    fx = fx + v.x;
    fy = fy + v.y;
    fz = fz + v.z;
  }
  return overrelax({fx, fy, fz}, pass_get(sf, rx, ry, rz, s, p, 0, 0, X, Y));
}

// The new value of site s on plane p replaces the old one where it is held.
__device__ __forceinline__ void pass_put(float *sf, float (&rx)[16], float (&ry)[16],
                                         float (&rz)[16], int s, int p, f3 v, int t) {
  const int sl = kPassPlan.slot[s][p];
  if (kPassPlan.kind[s][p] == kPlaneLds) {
    sf[(3 * sl + 0) * 256 + t] = v.x;
    sf[(3 * sl + 1) * 256 + t] = v.y;
    sf[(3 * sl + 2) * 256 + t] = v.z;
  } else {
    rx[sl] = v.x;
    ry[sl] = v.y;
    rz[sl] = v.z;
  }
}

// grid = 16 planes x replicas (XCD-dealt as k_or_pair_plane), block = 256
// (one cell of the plane per thread). LDS 36 KB: 4 blocks per CU.
template <int LT>
__global__ __launch_bounds__(256, 4) void k_or_pass_plane(
    const float *__restrict__ cur, float *__restrict__ oth, int r0,
    const uint32_t *__restrict__ ctl, uint32_t sweep_delta, double *__restrict__ series,
    int R_total, const int *__restrict__ replica_at, const double *__restrict__ anchor,
    const double *__restrict__ dE) {
  static_assert(LT == 16, "a z-plane is one block of 256 cells, a row one DPP row");
  constexpr int NC = LT * LT * LT, P = 16 * NC;
  constexpr int ROWS = 3 * kPassPlan.n_lds;
  __shared__ float4 sm[ROWS * 64];
  const int nb = gridDim.x;
  const int b = blockIdx.x;
  const int lb = (nb & 7) == 0 ? (b & 7) * (nb >> 3) + (b >> 3) : b;
  const int r = r0 + (lb >> 4);  // 16 planes per replica
  const int z = lb & 15;
  if (series != nullptr && b == 0) {
    const uint32_t row = ctl[kCtlRecordRow] + sweep_delta;
    if (ctl[kCtlRecordOn] && row < ctl[kCtlRecordCap]) {
      const int r1 = r0 + (nb >> 4);
      for (int k = threadIdx.x; k < R_total; k += 256) {
        const int rep = replica_at[k];
        if (rep >= r0 && rep < r1) series[(size_t)row * R_total + k] = anchor[rep] + dE[rep];
      }
    }
  }
  const float *rc = cur + (size_t)r * 48 * NC;
  float *ro = oth + (size_t)r * 48 * NC;
  const int t = threadIdx.x, x = t & 15, y = t >> 4;
  const int lane = t & 63;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  // The site-planes held in registers: this cell's value.
  float rx[16], ry[16], rz[16];
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const int k = kPassPlan.kind[s][p];
      if (k != kPlaneOwn && k != kPlaneRow) continue;
      const int sl = kPassPlan.slot[s][p];
      const uint32_t j = 4u * (uint32_t)(s * NC + ((z + p - 1) & 15) * 256 + t);
      rx[sl] = ld_at(rc, j);
      ry[sl] = ld_at(rc + P, j);
      rz[sl] = ld_at(rc + 2 * P, j);
    }
  // LDS rows: wave w copies rows w, w + 4, ... with 16-byte loads, all issued
  // before the first LDS write (as k_or_pair_plane).
  constexpr int NI = ROWS / 4;
  static_assert(ROWS % 4 == 0, "whole rounds of rows");
  float4 st[NI];
#pragma unroll
  for (int i = 0; i < NI; ++i) {
    const uint32_t e = kPassRows[4 * i + w];
    const int zz = (z + (int)(e >> 20) - 1) & 15;
    const uint32_t off = 4u * ((e & 0xFFFFFu) + (uint32_t)zz * 256u) + 16u * (uint32_t)lane;
    st[i] = *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(rc) + off);
  }
#pragma unroll
  for (int i = 0; i < NI; ++i) sm[(4 * i + w) * 64 + lane] = st[i];
  __syncthreads();
  float *sf = reinterpret_cast<float *>(sm);
  const int X[3] = {(x + 15) & 15, x, (x + 1) & 15};
  const int Y[3] = {((y + 15) & 15) * 16, y * 16, ((y + 1) & 15) * 16};
  // Plane z + 1: its colour-0, then colour-1 sites that plane z's colours 2
  // and 3 read (they read plane z's colour 2 and 3 sites, still old).
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    f3 nv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (kPassPlan.redo[c + 4 * k]) nv[k] = pass_relax(sf, rx, ry, rz, c + 4 * k, 2, X, Y);
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (kPassPlan.redo[c + 4 * k]) pass_put(sf, rx, ry, rz, c + 4 * k, 2, nv[k], t);
    if (c == 0) __syncthreads();
  }
  // Plane z, colours 0 to 3; each colour's new values go to oth at once.
  float *oy = ro + P, *oz = ro + 2 * P;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    f3 nv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) nv[k] = pass_relax(sf, rx, ry, rz, c + 4 * k, 1, X, Y);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t i = 4u * (uint32_t)((c + 4 * k) * NC + z * 256 + t);
      st_at(ro, i, nv[k].x);
      st_at(oy, i, nv[k].y);
      st_at(oz, i, nv[k].z);
      if (c < 3) pass_put(sf, rx, ry, rz, c + 4 * k, 1, nv[k], t);
    }
    if (c < 3) __syncthreads();
  }
}

// ---- A whole over-relaxation pass per launch at L = 64 ----
//
// A block updates a strip of 4 rows of 64 cells (y0 .. y0 + 3) of plane z,
// one row per wave: a row is periodic in x within the wave, x neighbours come
// from the adjacent lanes. Along z the colours are ordered (links up go to
// lower colours), so as at L = 16 the block recomputes the new values of
// plane z + 1 that plane z's later colours read. Along y they are not: links
// reach lower colours in both directions, so the block also recomputes what
// its rows read of the row below (y0 - 1) and the row above (y0 + 4) — a fifth
// wave does those two rows. StripPlan finds the nodes (plane, row, site) to
// compute as the closure of the strip's own updates over reads of lower
// colours. Rows rho of a strip are indexed rho + 2 (-2 .. 5), planes q = 0, 1,
// 2 for z - 1, z, z + 1. A site-plane read at a y offset by any node is staged
// in LDS over the rows read; one read only at the row of the reading node is
// held in the registers of the wave computing that row.
constexpr int kStripMaxRowComp = 3 * 16 * 8;
struct StripRowTable {
  // staged row-component k: (component * 16 + site) | q << 8 | (rho + 2) << 12,
  // and its float offset in LDS
  uint32_t ent[kStripMaxRowComp] = {}, lds[kStripMaxRowComp] = {};
};
struct StripPlan {
  bool comp[3][8][16] = {};
  int8_t lds_slot[3][16] = {}, lds_r0[3][16] = {}, lds_nr[3][16] = {};
  int16_t lds_base[3][16] = {};  // in units of 64 floats
  int8_t reg_s[3][16] = {}, reg_lo[3][16] = {}, reg_hi[3][16] = {};
  StripRowTable rows{};
  int n_lds = 0, n_rowcomp = 0, n_lds_floats = 0, n_reg_s = 0, n_reg_lo = 0, n_reg_hi = 0;
  bool ok = true;
  constexpr StripPlan() {
    for (int r = 2; r < 6; ++r)
      for (int s = 0; s < 16; ++s) comp[1][r][s] = true;
    for (bool changed = true; changed;) {
      changed = false;
      for (int q = 0; q < 3; ++q)
        for (int r = 0; r < 8; ++r)
          for (int s = 0; s < 16; ++s) {
            if (!comp[q][r][s]) continue;
            for (int m = 0; m < 6; ++m) {
              const Link lk = kLinks[s][m];
              if ((lk.site & 3) >= (s & 3)) continue;
              const int q2 = q + lk.dz, r2 = r + lk.dy;
              if (q2 < 0 || q2 > 2 || r2 < 0 || r2 > 7) {
                ok = false;
                continue;
              }
              if (!comp[q2][r2][lk.site]) {
                comp[q2][r2][lk.site] = true;
                changed = true;
              }
            }
          }
    }
    bool dy[3][16] = {}, ns[3][16] = {}, nlo[3][16] = {}, nhi[3][16] = {};
    int lo[3][16] = {}, hi[3][16] = {};
    for (int q = 0; q < 3; ++q)
      for (int s = 0; s < 16; ++s) {
        lo[q][s] = 99;
        hi[q][s] = -99;
      }
    for (int q = 0; q < 3; ++q)
      for (int r = 0; r < 8; ++r)
        for (int s = 0; s < 16; ++s) {
          if (!comp[q][r][s]) continue;
          // computed rows: the halo rows 1 and 6 and the strip rows 2..5,
          // the strip rows all alike
          if (r < 1 || r > 6) ok = false;
          if (r >= 2 && r <= 5 && comp[q][r][s] != comp[q][2][s]) ok = false;
          for (int m = -1; m < 6; ++m) {
            const Link lk = m < 0 ? Link{0, 0, 0, (int8_t)s} : kLinks[s][m];
            const int q2 = q + lk.dz, r2 = r + lk.dy, s2 = lk.site;
            if (q2 < 0 || q2 > 2 || r2 < 0 || r2 > 7) {
              ok = false;
              continue;
            }
            dy[q2][s2] = dy[q2][s2] || lk.dy != 0;
            lo[q2][s2] = r2 < lo[q2][s2] ? r2 : lo[q2][s2];
            hi[q2][s2] = r2 > hi[q2][s2] ? r2 : hi[q2][s2];
            if (lk.dy == 0) {
              if (r >= 2 && r <= 5) ns[q2][s2] = true;
              if (r == 1) nlo[q2][s2] = true;
              if (r == 6) nhi[q2][s2] = true;
            }
          }
        }
    for (int q = 0; q < 3; ++q)
      for (int s = 0; s < 16; ++s) {
        lds_slot[q][s] = reg_s[q][s] = reg_lo[q][s] = reg_hi[q][s] = -1;
        if (dy[q][s]) {
          const int nr = hi[q][s] - lo[q][s] + 1;
          lds_slot[q][s] = (int8_t)n_lds++;
          lds_r0[q][s] = (int8_t)lo[q][s];
          lds_nr[q][s] = (int8_t)nr;
          lds_base[q][s] = (int16_t)(n_lds_floats / 64);
          for (int c = 0; c < 3; ++c)
            for (int rr = 0; rr < nr; ++rr) {
              rows.ent[n_rowcomp] =
                  (uint32_t)(c * 16 + s) | (uint32_t)q << 8 | (uint32_t)(lo[q][s] + rr) << 12;
              rows.lds[n_rowcomp] = (uint32_t)(n_lds_floats + (c * nr + rr) * 64);
              ++n_rowcomp;
            }
          n_lds_floats += 3 * nr * 64;
        } else {
          if (ns[q][s]) reg_s[q][s] = (int8_t)n_reg_s++;
          if (nlo[q][s]) reg_lo[q][s] = (int8_t)n_reg_lo++;
          if (nhi[q][s]) reg_hi[q][s] = (int8_t)n_reg_hi++;
        }
      }
  }
};
constexpr StripPlan kStripPlan{};
static_assert(kStripPlan.ok, "the closure stays within planes z .. z + 1 and rows y0 - 1 .. y0 + 4");
static_assert(kStripPlan.n_lds == 12 && kStripPlan.n_rowcomp == 180 &&
                  kStripPlan.n_reg_s == 16 && kStripPlan.n_reg_lo == 14 && kStripPlan.n_reg_hi == 7,
              "12 site-planes of 5 rows in LDS (45 KB); 16, 14 and 7 register planes");
constexpr bool strip_halo_ok() {
  for (int q = 0; q < 3; ++q)
    for (int s = 0; s < 16; ++s) {
      if ((s & 3) == 3 && (kStripPlan.comp[q][1][s] || kStripPlan.comp[q][6][s])) return false;
      if ((s & 3) == 3 && kStripPlan.comp[2][2][s]) return false;
    }
  return true;
}
static_assert(strip_halo_ok(), "colour 3 runs on the strip's own plane only");
__constant__ constexpr StripRowTable kStripRows = kStripPlan.rows;
constexpr int kStripRegs = kStripPlan.n_reg_s > kStripPlan.n_reg_lo + kStripPlan.n_reg_hi
                               ? kStripPlan.n_reg_s
                               : kStripPlan.n_reg_lo + kStripPlan.n_reg_hi;

// Value of the cell x + DX (DX = -1 or 1) of the wave's 64-cell row:
// wave_ror:1 gives lane i the value of lane i - 1, wave_rol:1 that of lane
// i + 1, periodic over the 64 lanes.
template <int DX>
__device__ __forceinline__ float wave_neighbour(float v) {
  constexpr int ctrl = DX < 0 ? 0x13C : 0x134;
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), ctrl, 0xF, 0xF, false));
}

// Register slot of site-plane (q, s) for the wave's role: 0 a strip row, 1
// the row below the strip, 2 the row above.
template <int ROLE>
__device__ __forceinline__ int strip_reg(int q, int s) {
  if (ROLE == 0) return kStripPlan.reg_s[q][s];
  if (ROLE == 1) return kStripPlan.reg_lo[q][s];
  return kStripPlan.n_reg_lo + kStripPlan.reg_hi[q][s];
}

// Value of site s on plane q at row index rr (rho + 2) and x offset dx, for a
// node of a wave of role ROLE (all but rr compile-time after unrolling).
template <int ROLE>
__device__ __forceinline__ f3 strip_get(const float *sf, const float (&vx)[kStripRegs],
                                        const float (&vy)[kStripRegs],
                                        const float (&vz)[kStripRegs], int q, int s, int rr, int dx,
                                        int x) {
  if (kStripPlan.lds_slot[q][s] >= 0) {
    const int nr = kStripPlan.lds_nr[q][s];
    const int at = (kStripPlan.lds_base[q][s] + rr - kStripPlan.lds_r0[q][s]) * 64 + ((x + dx) & 63);
    return {sf[at], sf[at + nr * 64], sf[at + 2 * nr * 64]};
  }
  const int sl = strip_reg<ROLE>(q, s);
  if (dx < 0) return {wave_neighbour<-1>(vx[sl]), wave_neighbour<-1>(vy[sl]), wave_neighbour<-1>(vz[sl])};
  if (dx > 0) return {wave_neighbour<1>(vx[sl]), wave_neighbour<1>(vy[sl]), wave_neighbour<1>(vz[sl])};
  return {vx[sl], vy[sl], vz[sl]};
}

// Over-relaxation of node (q, rr, s) from its 6 links in link order; the new
// value replaces the old one where it is held (no node reads another of its
// own colour, so in place within a colour step).
template <int ROLE>
__device__ __forceinline__ f3 strip_relax(float *sf, float (&vx)[kStripRegs],
                                          float (&vy)[kStripRegs], float (&vz)[kStripRegs], int q,
                                          int s, int rr, int x) {
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    const Link lk = kLinks[s][m];
    const f3 v = strip_get<ROLE>(sf, vx, vy, vz, q + lk.dz, lk.site, rr + lk.dy, lk.dx, x);
    fx = fx + v.x;
    fy = fy + v.y;
    fz = fz + v.z;
  }
  const f3 nv = overrelax({fx, fy, fz}, strip_get<ROLE>(sf, vx, vy, vz, q, s, rr, 0, x));
  if (kStripPlan.lds_slot[q][s] >= 0) {
    const int nr = kStripPlan.lds_nr[q][s];
    const int at = (kStripPlan.lds_base[q][s] + rr - kStripPlan.lds_r0[q][s]) * 64 + x;
    sf[at] = nv.x;
    sf[at + nr * 64] = nv.y;
    sf[at + 2 * nr * 64] = nv.z;
  } else {
    const int sl = strip_reg<ROLE>(q, s);
    vx[sl] = nv.x;
    vy[sl] = nv.y;
    vz[sl] = nv.z;
  }
  return nv;
}

// Colour C of a strip row (row index rr = wave + 2): plane z's sites, stored
// to the second buffer, then plane z + 1's recomputed ones.
template <int C>
__device__ __forceinline__ void strip_step(float *sf, float (&vx)[kStripRegs],
                                           float (&vy)[kStripRegs], float (&vz)[kStripRegs],
                                           int rr, int x, float *ro, uint32_t cell) {
  constexpr int NC = 64 * 64 * 64, P = 16 * NC;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int s = C + 4 * k;
    const f3 nv = strip_relax<0>(sf, vx, vy, vz, 1, s, rr, x);
    const uint32_t i = 4u * (uint32_t)(s * NC) + cell;
    st_at(ro, i, nv.x);
    st_at(ro + P, i, nv.y);
    st_at(ro + 2 * P, i, nv.z);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (kStripPlan.comp[2][2][C + 4 * k]) (void)strip_relax<0>(sf, vx, vy, vz, 2, C + 4 * k, rr, x);
}

// Colour C of the rows below (rr = 1) and above (rr = 6) the strip.
template <int C>
__device__ __forceinline__ void strip_halo_step(float *sf, float (&vx)[kStripRegs],
                                                float (&vy)[kStripRegs], float (&vz)[kStripRegs],
                                                int x) {
#pragma unroll
  for (int q = 1; q < 3; ++q)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (kStripPlan.comp[q][1][C + 4 * k]) (void)strip_relax<1>(sf, vx, vy, vz, q, C + 4 * k, 1, x);
      if (kStripPlan.comp[q][6][C + 4 * k]) (void)strip_relax<2>(sf, vx, vy, vz, q, C + 4 * k, 6, x);
    }
}

template <int ROLE>
__device__ __forceinline__ void strip_load_regs(const float *__restrict__ rc, int z, int y, int x,
                                                float (&vx)[kStripRegs], float (&vy)[kStripRegs],
                                                float (&vz)[kStripRegs]) {
  constexpr int NC = 64 * 64 * 64, P = 16 * NC;
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const int raw = ROLE == 0 ? kStripPlan.reg_s[q][s]
                                : (ROLE == 1 ? kStripPlan.reg_lo[q][s] : kStripPlan.reg_hi[q][s]);
      if (raw < 0) continue;
      const int sl = strip_reg<ROLE>(q, s);
      const uint32_t j = 4u * (uint32_t)(s * NC + ((z + q - 1) & 63) * 4096 + y * 64 + x);
      vx[sl] = ld_at(rc, j);
      vy[sl] = ld_at(rc + P, j);
      vz[sl] = ld_at(rc + 2 * P, j);
    }
}

// grid = 1,024 strips (16 per plane, 64 planes) x replicas (XCD-dealt), block
// = 320 (waves 0-3 the strip's rows, wave 4 the rows around it). LDS 45 KB,
// held to 128 VGPRs (a few spill) for 3 blocks per CU: at 134 VGPRs, 2 blocks
// per CU ran 990 against 950 us per launch over 32 replicas.
template <int LT>
__global__ __launch_bounds__(320) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_or_pass_strip(
    const float *__restrict__ cur, float *__restrict__ oth, int r0,
    const uint32_t *__restrict__ ctl, uint32_t sweep_delta, double *__restrict__ series,
    int R_total, const int *__restrict__ replica_at, const double *__restrict__ anchor,
    const double *__restrict__ dE) {
  static_assert(LT == 64, "a row is one wave of 64 cells");
  constexpr int NC = LT * LT * LT;
  constexpr int NITEM = kStripPlan.n_rowcomp * 16;  // 16-byte pieces of the staged rows
  constexpr int NI = (NITEM + 319) / 320;
  __shared__ float4 sm[kStripPlan.n_lds_floats / 4];
  const int nb = gridDim.x;
  const int b = blockIdx.x;
  const int lb = (nb & 7) == 0 ? (b & 7) * (nb >> 3) + (b >> 3) : b;
  const int r = r0 + (lb >> 10);
  const int z = (lb >> 4) & 63;
  const int y0 = (lb & 15) * 4;
  if (series != nullptr && b == 0) {
    const uint32_t row = ctl[kCtlRecordRow] + sweep_delta;
    if (ctl[kCtlRecordOn] && row < ctl[kCtlRecordCap]) {
      const int r1 = r0 + (nb >> 10);
      for (int k = threadIdx.x; k < R_total; k += 320) {
        const int rep = replica_at[k];
        if (rep >= r0 && rep < r1) series[(size_t)row * R_total + k] = anchor[rep] + dE[rep];
      }
    }
  }
  const float *rc = cur + (size_t)r * 48 * NC;
  float *ro = oth + (size_t)r * 48 * NC;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // 0..4
  const int x = threadIdx.x & 63;
  float4 st[NI];
#pragma unroll
  for (int i = 0; i < NI; ++i) {
    const int item = min((int)threadIdx.x + 320 * i, NITEM - 1);
    const uint32_t e = kStripRows.ent[item >> 4];
    const int zz = (z + (int)((e >> 8) & 3u) - 1) & 63;
    const int yy = (y0 + (int)((e >> 12) & 7u) - 2) & 63;
    const uint32_t off =
        4u * ((e & 0xFFu) * (uint32_t)NC + (uint32_t)(zz * 4096 + yy * 64)) + 16u * (uint32_t)(item & 15);
    st[i] = *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(rc) + off);
  }
  float vx[kStripRegs], vy[kStripRegs], vz[kStripRegs];
  if (w < 4) {
    strip_load_regs<0>(rc, z, y0 + w, x, vx, vy, vz);
  } else {
    strip_load_regs<1>(rc, z, (y0 + 63) & 63, x, vx, vy, vz);
    strip_load_regs<2>(rc, z, (y0 + 4) & 63, x, vx, vy, vz);
  }
#pragma unroll
  for (int i = 0; i < NI; ++i) {
    const int item = (int)threadIdx.x + 320 * i;
    if (NITEM % 320 == 0 || item < NITEM)
      sm[(kStripRows.lds[item >> 4] >> 2) + (item & 15)] = st[i];
  }
  __syncthreads();
  float *sf = reinterpret_cast<float *>(sm);
  const int rr = w + 2;
  const uint32_t cell = 4u * (uint32_t)(z * 4096 + (y0 + w) * 64 + x);
#define PYRO_STRIP_STEP(C)                                  \
  if (w < 4)                                                \
    strip_step<C>(sf, vx, vy, vz, rr, x, ro, cell);         \
  else                                                      \
    strip_halo_step<C>(sf, vx, vy, vz, x);
  PYRO_STRIP_STEP(0)
  __syncthreads();
  PYRO_STRIP_STEP(1)
  __syncthreads();
  PYRO_STRIP_STEP(2)
  __syncthreads();
  if (w < 4) strip_step<3>(sf, vx, vy, vz, rr, x, ro, cell);
#undef PYRO_STRIP_STEP
}

// Two Metropolis colour phases (CA, CA + 1) at L = 16 in one launch, in
// place, scalar model: one z-plane of a replica per block, staged as
// k_or_pair_plane<CA, 3> stages it (the same 29 values a cell, one buffer),
// then phase CA of the cell's 4 sites, the exchange of the foreign colour-CA
// values between lanes, phase CA + 1, and the stores of the accepted moves.
// Each site's step is metropolis_site with the Philox counters and field sums
// of the one-colour kernel, so spins and acceptance counts equal two
// one-colour launches bit for bit (the FP64 energy changes are summed per
// block over both colours, a different order of the same terms). In place is
// race-free: a block reads its own plane before its first write, and the
// other planes it reads (z - 1 for colours 0/1, z + 1 for 2/3) only at sites
// outside the pair, which no block of the launch writes. 36 instead of 60 B
// per update. 2 waves per SIMD (171-190 VGPRs): at 3 (168) it spills, and
// measured at L = 16 x 64 the pair launch takes 24.4-24.6 us against
// 25.1 us at 3 and 2 x 14.1-14.4 us for two one-colour launches; sweep
// 0.286-0.288 ms against 0.297 and 0.300 ms.
template <int CA>
__global__ __launch_bounds__(256, 2) void k_metro_pair_plane(
    float *__restrict__ spins, int pass, const float *__restrict__ beta,
    const float *__restrict__ cosc, const uint32_t *__restrict__ ctl, uint32_t key0,
    uint32_t key1, float jeff, unsigned long long *__restrict__ acc, double *__restrict__ dE,
    int r0, uint32_t sweep_delta) {
  constexpr int LT = 16, NC = LT * LT * LT, P = 16 * NC, CB = CA + 1;
  constexpr PairPlanes<CA, 3> pp{};
  constexpr PairNeeds<CA, false> nd{};
  constexpr int NS = pp.n, NN = nd.n, ROWS = 3 * pp.n_lds;
  constexpr int NX = kPairExchange<CA>.n;
  static_assert(NS == 20 && ROWS <= 72 && nd.nf == 3, "20 site-planes, 3 foreign neighbours");
  __shared__ float4 sm[ROWS * 64];
  __shared__ float xs[(NX > 0 ? NX : 1) * 3 * 256];
  const int nb = gridDim.x;
  const int b = blockIdx.x;
  const int lb = (nb & 7) == 0 ? (b & 7) * (nb >> 3) + (b >> 3) : b;
  const int r = r0 + (lb >> 4);  // 16 planes per replica
  const int z = lb & 15;
  float *rs = spins + (size_t)r * 48 * NC;
  const int t = threadIdx.x, x = t & 15, y = t >> 4;
  const int lane = t & 63;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int cell = z * 256 + t;
  float ox[NS], oy[NS], oz[NS];
#pragma unroll
  for (int sl = 0; sl < NS; ++sl) {
    if (pp.kind[sl] == kPlaneLds) continue;
    const uint32_t j = 4u * (uint32_t)(pp.site[sl] * NC + ((z + pp.dz[sl]) & 15) * 256 + t);
    ox[sl] = ld_at(rs, j);
    oy[sl] = ld_at(rs + P, j);
    oz[sl] = ld_at(rs + 2 * P, j);
  }
  constexpr int NI = (ROWS + 3) / 4;
  float4 st[NI];
#pragma unroll
  for (int i = 0; i < NI; ++i) {  // every load before the first LDS write
    const uint32_t e = kPairRows<CA, 3>[min(4 * i + w, ROWS - 1)];
    const int zz = (z + (int)((e >> 20) & 3) - 1) & 15;
    const uint32_t off = 4u * ((e & 0xFFFFFu) + (uint32_t)zz * 256u) + 16u * (uint32_t)lane;
    st[i] = *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(rs) + off);
  }
#pragma unroll
  for (int i = 0; i < NI; ++i) sm[min(4 * i + w, ROWS - 1) * 64 + lane] = st[i];
  __syncthreads();
  const float *sf = reinterpret_cast<const float *>(sm);
  const int X[3] = {(x + 15) & 15, x, (x + 1) & 15};
  const int Y[3] = {((y + 15) & 15) * 16, y * 16, ((y + 1) & 15) * 16};
  float vx[NN], vy[NN], vz[NN];
#pragma unroll
  for (int i = 0; i < NN; ++i) {
    const PairNeed qn = nd.v[i];
    const int sl = pp.slot[qn.site][qn.dz + 1];
    if (pp.kind[sl] == kPlaneLds) {
      const int at = Y[qn.dy + 1] + X[qn.dx + 1];
      const int l = pp.lds[sl];
      vx[i] = sf[(3 * l + 0) * 256 + at];
      vy[i] = sf[(3 * l + 1) * 256 + at];
      vz[i] = sf[(3 * l + 2) * 256 + at];
    } else if (qn.dx == 0) {
      vx[i] = ox[sl];
      vy[i] = oy[sl];
      vz[i] = oz[sl];
    } else if (qn.dx < 0) {
      vx[i] = row_neighbour<-1>(ox[sl]);
      vy[i] = row_neighbour<-1>(oy[sl]);
      vz[i] = row_neighbour<-1>(oz[sl]);
    } else {
      vx[i] = row_neighbour<1>(ox[sl]);
      vy[i] = row_neighbour<1>(oy[sl]);
      vz[i] = row_neighbour<1>(oz[sl]);
    }
  }
  const float bt = beta[r], cc = cosc[r];
  const uint32_t sweep = ctl[kCtlSweep] + sweep_delta;
  double de = 0.0;
  unsigned nacc = 0;
  // Phase CA: na[k] = the new (or kept) spin of site CA + 4k.
  f3 na[4];
  bool wa[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    float fx, fy, fz;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    field_start<false>(nullptr, CA, fx, fy, fz);
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      const int i = nd.a_nb[k][m];
      field_add<false, true>(nullptr, CA, kLinks[CA + 4 * k][m].site & 3, vx[i], vy[i], vz[i], fx,
                             fy, fz, gx, gy, gz);
    }
    const int is = nd.a_self[k];
    const f3 s = {vx[is], vy[is], vz[is]};
    const f3 sp = metropolis_site<false>(s, fx, fy, fz, gx, gy, gz, cell, CA + 4 * k, CA, r, pass,
                                         bt, cc, sweep, key0, key1, jeff, nullptr, false, wa[k], de,
                                         nacc);
    na[k] = {wa[k] ? sp.x : s.x, wa[k] ? sp.y : s.y, wa[k] ? sp.z : s.z};  // component-wise: no stack copy
  }
  // The foreign colour-CA values from their owners (as k_or_pair_plane<CA, 3>).
#pragma unroll
  for (int e = 0; e < NX; ++e) {
    const f3 v = na[kPairExchange<CA>.k[e]];
    xs[(3 * e + 0) * 256 + t] = v.x;
    xs[(3 * e + 1) * 256 + t] = v.y;
    xs[(3 * e + 2) * 256 + t] = v.z;
  }
  __syncthreads();
  f3 fo[3];
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    const PairNeed q = nd.fv[f];
    const f3 own = na[q.site >> 2];
    if (q.dy == 0 && q.dx < 0) {
      fo[f] = {row_neighbour<-1>(own.x), row_neighbour<-1>(own.y), row_neighbour<-1>(own.z)};
    } else if (q.dy == 0) {
      fo[f] = {row_neighbour<1>(own.x), row_neighbour<1>(own.y), row_neighbour<1>(own.z)};
    } else {
      const int e = kPairExchange<CA>.slot[q.site >> 2];
      const int at = Y[q.dy + 1] + X[q.dx + 1];
      fo[f] = {xs[(3 * e + 0) * 256 + at], xs[(3 * e + 1) * 256 + at], xs[(3 * e + 2) * 256 + at]};
    }
  }
  // Phase CA + 1 from the old values, na and fo.
  f3 nbv[4];
  bool wb[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int tsite = CB + 4 * k;
    float fx, fy, fz;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    field_start<false>(nullptr, CB, fx, fy, fz);
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      const Link lk = kLinks[tsite][m];
      f3 v;
      if ((lk.site & 3) == CA) {
        v = nd.f_idx[k][m] < 0 ? na[lk.site >> 2] : fo[nd.f_idx[k][m]];
      } else {
        const int i = nd.t_nb[k][m];
        v = {vx[i], vy[i], vz[i]};
      }
      field_add<false, true>(nullptr, CB, lk.site & 3, v.x, v.y, v.z, fx, fy, fz, gx, gy, gz);
    }
    const int is = nd.t_self[k];
    const f3 s = {vx[is], vy[is], vz[is]};
    nbv[k] = metropolis_site<false>(s, fx, fy, fz, gx, gy, gz, cell, tsite, CB, r, pass, bt, cc,
                                    sweep, key0, key1, jeff, nullptr, false, wb[k], de, nacc);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t ia = 4u * (uint32_t)((CA + 4 * k) * NC + cell);
    const uint32_t ib = 4u * (uint32_t)((CB + 4 * k) * NC + cell);
    if (wa[k]) {
      st_at(rs, ia, na[k].x);
      st_at(rs + P, ia, na[k].y);
      st_at(rs + 2 * P, ia, na[k].z);
    }
    if (wb[k]) {
      st_at(rs, ib, nbv[k].x);
      st_at(rs + P, ib, nbv[k].y);
      st_at(rs + 2 * P, ib, nbv[k].z);
    }
  }
  block_sum<256>(de, nacc);
  if (threadIdx.x == 0) {
    if (de != 0.0) atomicAdd(&dE[r], de);
    if (nacc) atomicAdd(&acc[r], (unsigned long long)nacc);
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
// LT > 0: the lattice size at compile time (constant strides, one wrapped
// coordinate triple per axis); LT = 0: L at run time. Spin offsets are 32-bit
// byte offsets from the replica's planes.
template <int B, bool MAT, int LT>
__device__ __forceinline__ void obs_cell(const float *__restrict__ Sx, const float *__restrict__ Sy,
                                         const float *__restrict__ Sz, int L_rt, int ncell_rt,
                                         int cell, const float *__restrict__ coup, double v[4]) {
  const int L = LT > 0 ? LT : L_rt;
  const int ncell = LT > 0 ? LT * LT * LT : ncell_rt;
  const int LL = L * L;
  const int z = cell / LL;
  const int rem = cell - z * LL;
  const int y = rem / L;
  const int x = rem - y * L;
  const int X[3] = {x == 0 ? L - 1 : x - 1, x, x == L - 1 ? 0 : x + 1};
  const int Y[3] = {(y == 0 ? L - 1 : y - 1) * L, y * L, (y == L - 1 ? 0 : y + 1) * L};
  const int Z[3] = {(z == 0 ? L - 1 : z - 1) * LL, z * LL, (z == L - 1 ? 0 : z + 1) * LL};
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int site = 4 * t + B;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    float fx, fy, fz;  // unused: the FP32 field
    field_start<false>(coup, B, fx, fy, fz);
#pragma unroll
    for (int n = 0; n < 6; ++n) {
      const Link lk = kLinks[site][n];
      const uint32_t j = 4u * (uint32_t)(lk.site * ncell + Z[lk.dz + 1] + Y[lk.dy + 1] + X[lk.dx + 1]);
      field_add<MAT, true>(coup, B, lk.site & 3, ld_at(Sx, j), ld_at(Sy, j), ld_at(Sz, j), fx, fy,
                           fz, gx, gy, gz);
    }
    const uint32_t i = 4u * (uint32_t)(site * ncell + cell);
    const double sx = ld_at(Sx, i), sy = ld_at(Sy, i), sz = ld_at(Sz, i);
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

template <bool MAT, int LT>
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
      case 0: obs_cell<0, MAT, LT>(Sx, Sy, Sz, L, ncell, cell, coup, v); break;
      case 1: obs_cell<1, MAT, LT>(Sx, Sy, Sz, L, ncell, cell, coup, v); break;
      case 2: obs_cell<2, MAT, LT>(Sx, Sy, Sz, L, ncell, cell, coup, v); break;
      default: obs_cell<3, MAT, LT>(Sx, Sy, Sz, L, ncell, cell, coup, v); break;
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
                          launch_r0(ctx), ctx->launch_delta, (int)ctx->set_schedule);
    return;
  }
  k_phase<KIND, LT, MAT><<<grid, kPhaseBlock, 0, launch_stream(ctx)>>>(
      ctx->d_spins, ctx->L, ctx->ncell, colour, pass, ctx->d_beta, ctx->d_cosc, ctx->d_ctl,
      key_lo(ctx), key_hi(ctx), ctx->jeff, ctx->d_acc, ctx->d_dE, ctx->d_coup,
      (int)ctx->post_rotate, launch_r0(ctx), ctx->launch_delta, (int)ctx->set_schedule);
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
  if constexpr (LT == 16 && !TRACK) {  // planes staged in LDS
    if (ctx->time_ev0) {
      hipExtLaunchKernelGGL(k_or_pair_plane<CA, kOrPlaneMode>, dim3(16 * nrep), dim3(256), 0, launch_stream(ctx),
                            ctx->time_ev0, ctx->time_ev1, 0, cur, oth, ctx->jeff, launch_r0(ctx),
                            (const uint32_t *)ctx->d_ctl, ctx->launch_delta, series, ctx->R,
                            (const int *)ctx->d_replica_at, (const double *)ctx->d_anchor,
                            (const double *)ctx->d_dE);
      return;
    }
    k_or_pair_plane<CA, kOrPlaneMode><<<16 * nrep, 256, 0, launch_stream(ctx)>>>(
        cur, oth, ctx->jeff, launch_r0(ctx), ctx->d_ctl, ctx->launch_delta, series, ctx->R,
        ctx->d_replica_at, ctx->d_anchor, ctx->d_dE);
    return;
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

// A whole over-relaxation pass at L = 16 from cur into oth (one launch).
static inline void launch_or_pass(pyro_ctx *ctx, const float *cur, float *oth) {
  const int nrep = launch_nrep(ctx);
  double *series = nullptr;
  if (ctx->launch_series) {
    series = ctx->launch_series;
    ctx->launch_series_done = 1;
  }
  if (ctx->L == 64) {
    if (ctx->time_ev0) {
      hipExtLaunchKernelGGL(k_or_pass_strip<64>, dim3(1024 * nrep), dim3(320), 0,
                            launch_stream(ctx), ctx->time_ev0, ctx->time_ev1, 0, cur, oth,
                            launch_r0(ctx), (const uint32_t *)ctx->d_ctl, ctx->launch_delta, series,
                            ctx->R, (const int *)ctx->d_replica_at, (const double *)ctx->d_anchor,
                            (const double *)ctx->d_dE);
      return;
    }
    k_or_pass_strip<64><<<1024 * nrep, 320, 0, launch_stream(ctx)>>>(
        cur, oth, launch_r0(ctx), ctx->d_ctl, ctx->launch_delta, series, ctx->R,
        ctx->d_replica_at, ctx->d_anchor, ctx->d_dE);
    return;
  }
  if (ctx->time_ev0) {
    hipExtLaunchKernelGGL(k_or_pass_plane<16>, dim3(16 * nrep), dim3(256), 0, launch_stream(ctx),
                          ctx->time_ev0, ctx->time_ev1, 0, cur, oth, launch_r0(ctx),
                          (const uint32_t *)ctx->d_ctl, ctx->launch_delta, series, ctx->R,
                          (const int *)ctx->d_replica_at, (const double *)ctx->d_anchor,
                          (const double *)ctx->d_dE);
    return;
  }
  k_or_pass_plane<16><<<16 * nrep, 256, 0, launch_stream(ctx)>>>(
      cur, oth, launch_r0(ctx), ctx->d_ctl, ctx->launch_delta, series, ctx->R, ctx->d_replica_at,
      ctx->d_anchor, ctx->d_dE);
}

// The fused Metropolis colour pair ca (0 or 2) at L = 16, in place on d_spins.
template <int CA>
static void launch_metro_pair_t(pyro_ctx *ctx, int pass) {
  const int nrep = launch_nrep(ctx);
  if (ctx->time_ev0) {
    hipExtLaunchKernelGGL(k_metro_pair_plane<CA>, dim3(16 * nrep), dim3(256), 0,
                          launch_stream(ctx), ctx->time_ev0, ctx->time_ev1, 0, ctx->d_spins, pass,
                          (const float *)ctx->d_beta, (const float *)ctx->d_cosc,
                          (const uint32_t *)ctx->d_ctl, key_lo(ctx), key_hi(ctx), ctx->jeff,
                          ctx->d_acc, ctx->d_dE, launch_r0(ctx), ctx->launch_delta);
    return;
  }
  k_metro_pair_plane<CA><<<16 * nrep, 256, 0, launch_stream(ctx)>>>(
      ctx->d_spins, pass, ctx->d_beta, ctx->d_cosc, ctx->d_ctl, key_lo(ctx), key_hi(ctx),
      ctx->jeff, ctx->d_acc, ctx->d_dE, launch_r0(ctx), ctx->launch_delta);
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
// The cell-per-thread kernels cover the colour classes and the moves by pi;
// the drawn-set schedule and the random-angle over-relaxation (KIND 2) run
// in the wave-per-site kernel.
template <int KIND, bool MAT>
static void launch_phase(pyro_ctx *ctx, int colour, int pass) {
  if constexpr (KIND != 2)
    if (ctx->phase_variant == 1 && !ctx->set_schedule && launch_cells<KIND, MAT>(ctx, colour, pass))
      return;
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

// Partial sums of the observables (first stage) of R replicas of L^3 cells
// into partials (R x 4 x ceil(L^3 / kObsBlock) x 4 doubles), on `stream`.
template <bool MAT>
static void launch_obs_partials(const float *spins, int L, int R, double *partials,
                                const float *coup, hipStream_t stream) {
  const int ncell = L * L * L;
  const dim3 grid((ncell + kObsBlock - 1) / kObsBlock, 4, R);
#define PYRO_OBS_CASE(LV)                                                                    \
  case LV:                                                                                   \
    k_observables<MAT, LV><<<grid, kObsBlock, 0, stream>>>(spins, L, ncell, partials, coup); \
    return;
  switch (L) {
    PYRO_OBS_CASE(4)
    PYRO_OBS_CASE(8)
    PYRO_OBS_CASE(12)
    PYRO_OBS_CASE(16)
    PYRO_OBS_CASE(24)
    PYRO_OBS_CASE(32)
    PYRO_OBS_CASE(64)
    default: k_observables<MAT, 0><<<grid, kObsBlock, 0, stream>>>(spins, L, ncell, partials, coup);
  }
#undef PYRO_OBS_CASE
}

// Partial sums of the observables (first stage) on the ctx stream.
template <bool MAT>
static void launch_observables_partials(pyro_ctx *ctx) {
  launch_obs_partials<MAT>(ctx->d_spins, ctx->L, ctx->R, ctx->d_partials, ctx->d_coup, ctx->stream);
}

}  // namespace pyro
