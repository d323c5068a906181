// One cell's update for one colour phase, shared by every phase kernel (the
// uniform-lattice kernels and the heterogeneous batch kernels), so all of them
// produce the same spins bit for bit.
//
// A colour c is the independent set {c, c+4, c+8, c+12} of every cell
// (Pyrochlore.cu:36-43). With the colour and the lattice size fixed at compile
// time the 24 links of the cell's 4 target sites fold into constant offsets
// from three wrapped coordinates per axis; a neighbour shared by two targets
// is loaded once; the 4 results are stored after all loads (targets are
// mutually non-adjacent, so deferring the stores changes nothing).
#pragma once

#include "pyro_internal.h"

namespace pyro {

constexpr Link kLinks[16][6] = PYRO_LINKS16_INIT;

// Local-field models. Scalar (the Heisenberg J of pyro_create): a site's
// field is -jeff times the plain FP32 neighbour sum f. Matrix (anisotropic
// exchange and Zeeman term, pyro_set_couplings; the reference's
// CalculateLocalExchangeField, Pyrochlore.cu:210-275): f = B_b +
// sum_n C_{b, b_n} s_n is the field itself. coup holds kCoupFloats floats:
// C as 9 row-major floats per ordered basis pair (b, b_n) at 9 (4 b + b_n),
// scaled by spin_len2, then B as 3 floats per basis site at 144 + 3 b, scaled
// by 2 spin_len2. The FP64 sums g feed the energy bookkeeping: g = sum_n s_n
// (scalar) or sum_n C s_n (matrix, without B).
template <bool MAT>
__device__ __forceinline__ void field_start(const float *__restrict__ coup, int b, float &fx,
                                            float &fy, float &fz) {
  if constexpr (MAT) {
    fx = coup[144 + 3 * b];
    fy = coup[145 + 3 * b];
    fz = coup[146 + 3 * b];
  } else {
    fx = 0.0f;
    fy = 0.0f;
    fz = 0.0f;
  }
}

template <bool MAT, bool G>
__device__ __forceinline__ void field_add(const float *__restrict__ coup, int b, int bn, float a,
                                          float bb, float c, float &fx, float &fy, float &fz,
                                          double &gx, double &gy, double &gz) {
  if constexpr (MAT) {
    const float *m = coup + 9 * (4 * b + bn);
    fx = fx + ((m[0] * a + m[1] * bb) + m[2] * c);
    fy = fy + ((m[3] * a + m[4] * bb) + m[5] * c);
    fz = fz + ((m[6] * a + m[7] * bb) + m[8] * c);
    if (G) {
      gx += ((double)m[0] * a + (double)m[1] * bb) + (double)m[2] * c;
      gy += ((double)m[3] * a + (double)m[4] * bb) + (double)m[5] * c;
      gz += ((double)m[6] * a + (double)m[7] * bb) + (double)m[8] * c;
    }
  } else {
    fx = fx + a;
    fy = fy + bb;
    fz = fz + c;
    if (G) {
      gx += (double)a;
      gy += (double)bb;
      gz += (double)c;
    }
  }
}

// The field of the Metropolis test.
template <bool MAT>
__device__ __forceinline__ f3 field_h(float jeff, float fx, float fy, float fz) {
  if constexpr (MAT) return {fx, fy, fz};
  else return {-jeff * fx, -jeff * fy, -jeff * fz};
}

// FP64 energy change of the move s -> sp: -(sp - s).h with h in FP64.
template <bool MAT>
__device__ __forceinline__ double field_de(float jeff, const float *__restrict__ coup, int b, f3 s,
                                           f3 sp, double gx, double gy, double gz) {
  if constexpr (MAT)
    return -(((double)sp.x - (double)s.x) * (gx + (double)coup[144 + 3 * b]) +
             ((double)sp.y - (double)s.y) * (gy + (double)coup[145 + 3 * b]) +
             ((double)sp.z - (double)s.z) * (gz + (double)coup[146 + 3 * b]));
  else
    return (double)jeff * (((double)sp.x - (double)s.x) * gx + ((double)sp.y - (double)s.y) * gy +
                           ((double)sp.z - (double)s.z) * gz);
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// One site's Metropolis step from its local-field sums (fx: FP32 sum in link
// order, as the move uses it; gx: FP64 sum, for the energy books): Philox
// draw keyed by (cell, site, pass, replica, sweep), cone proposal, test, and
// optionally the rotation about the field. Returns the proposal (or the
// rotated spin); wr tells whether the site is to be written, de gets the FP64
// energy change of a written move, nacc counts the accepted ones.
template <bool MAT>
__device__ __forceinline__ f3 metropolis_site(f3 s, float fx, float fy, float fz, double gx,
                                              double gy, double gz, int cell, int site, int colour,
                                              int r, int pass, float beta_r, float cosc_r,
                                              uint32_t sweep, uint32_t key0, uint32_t key1,
                                              float jeff, const float *__restrict__ coup,
                                              bool post_rotate, bool &wr, double &de,
                                              unsigned &nacc) {
  const u32x4 w = philox4x32_10(
      rng_counter((uint32_t)cell, (uint32_t)site, (uint32_t)pass, kKindMetropolis, (uint32_t)r,
                  sweep),
      key0, key1);
  const f3 h = field_h<MAT>(jeff, fx, fy, fz);
  f3 sp = cone_proposal(s, cosc_r, u01(w.x), u01(w.y));
  wr = metropolis_accept(h, s, sp, beta_r, u01(w.z));
  nacc += wr ? 1u : 0u;
  if (post_rotate) {  // uniform over the block: a kernel argument
    sp = rotate_about(h, wr ? sp : s, u01(w.w));
    wr = true;
  }
  if (wr) de += field_de<MAT>(jeff, coup, colour, s, sp, gx, gy, gz);
  return sp;
}

// KIND 0: Metropolis with the cone proposal (MCMultipleTemperaturesPerSystem,
// Pyrochlore.cu:829-885). KIND 1: over-relaxation by pi (OverRelaxationByPiStep,
// Pyrochlore.cu:975-1007). TRACK: FP64 energy change also for KIND 1.
// Sx: this replica's spins ([x,y,z][site][cell]); r: the replica's global
// index (keys the Philox counters). Adds the energy change to de and the
// accepted proposals to nacc.
template <int KIND, int COLOUR, int LT, bool TRACK, bool MAT = false>
__device__ __forceinline__ void cell_update(float *__restrict__ Sx, int cell, int r, int pass,
                                            float beta_r, float cosc_r, uint32_t sweep,
                                            uint32_t key0, uint32_t key1, float jeff, double &de,
                                            unsigned &nacc,
                                            const float *__restrict__ coup = nullptr,
                                            bool post_rotate = false) {
  constexpr int L = LT, LL = LT * LT, NC = LT * LT * LT;
  float *Sy = Sx + 16 * NC;
  float *Sz = Sy + 16 * NC;
  const int z = cell / LL;
  const int y = (cell / L) % L;
  const int x = cell % L;
  const int X[3] = {x == 0 ? L - 1 : x - 1, x, x == L - 1 ? 0 : x + 1};
  const int Y[3] = {(y == 0 ? L - 1 : y - 1) * L, y * L, (y == L - 1 ? 0 : y + 1) * L};
  const int Z[3] = {(z == 0 ? L - 1 : z - 1) * LL, z * LL, (z == L - 1 ? 0 : z + 1) * LL};
  f3 s[4], sp[4];
  bool wr[4];
  float fx[4], fy[4], fz[4];
  double gx[4], gy[4], gz[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int site = COLOUR + 4 * k;
    field_start<MAT>(coup, COLOUR, fx[k], fy[k], fz[k]);
    gx[k] = gy[k] = gz[k] = 0.0;
#pragma unroll
    for (int n = 0; n < 6; ++n) {
      const Link lk = kLinks[site][n];
      const uint32_t j = 4u * (uint32_t)(lk.site * NC + Z[lk.dz + 1] + Y[lk.dy + 1] + X[lk.dx + 1]);
      const float a = ld_at(Sx, j), b = ld_at(Sy, j), c = ld_at(Sz, j);
      field_add<MAT, KIND == 0 || TRACK>(coup, COLOUR, lk.site & 3, a, b, c, fx[k], fy[k], fz[k],
                                         gx[k], gy[k], gz[k]);
    }
    const uint32_t i = 4u * (uint32_t)(site * NC + cell);
    s[k] = {ld_at(Sx, i), ld_at(Sy, i), ld_at(Sz, i)};
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int site = COLOUR + 4 * k;
    if (KIND == 0) {
      sp[k] = metropolis_site<MAT>(s[k], fx[k], fy[k], fz[k], gx[k], gy[k], gz[k], cell, site,
                                   COLOUR, r, pass, beta_r, cosc_r, sweep, key0, key1, jeff, coup,
                                   post_rotate, wr[k], de, nacc);
    } else {
      const f3 f = {fx[k], fy[k], fz[k]};
      sp[k] = overrelax(f, s[k]);
      wr[k] = true;
      if (TRACK) de += field_de<MAT>(jeff, coup, COLOUR, s[k], sp[k], gx[k], gy[k], gz[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (wr[k]) {
      const uint32_t i = 4u * (uint32_t)((COLOUR + 4 * k) * NC + cell);
      st_at(Sx, i, sp[k].x);
      st_at(Sy, i, sp[k].y);
      st_at(Sz, i, sp[k].z);
    }
  }
}

// The spins one fused over-relaxation pair (colours CA, CA + 1) reads for a
// cell: (site, dx, dy, dz) with offsets in -1..1, built at compile time from
// the link table, and the position in that list of every value the two phases
// use, so the kernel indexes its register copy with constants only. Phase
// CA + 1 needs the new colour-CA values of its neighbours: those of the cell
// come from the first phase, and 3 are of colour-CA sites of neighbouring
// cells ("foreign", listed in fv). With RECOMP they are recomputed from their
// own links, which are then needs too (41 spins for either pair); without it
// the caller supplies them (29 spins).
struct PairNeed { int8_t site, dx, dy, dz; };
template <int CA, bool RECOMP = true>
struct PairNeeds {
  PairNeed v[64] = {};
  int n = 0;
  int a_self[4] = {}, a_nb[4][6] = {};  // colour-CA sites of the cell and their links
  int t_self[4] = {}, t_nb[4][6] = {};  // colour-CA+1 sites and their links (-1: colour CA, new)
  int f_idx[4][6] = {};                 // the link's position in fv, or -1
  PairNeed fv[8] = {};                  // the foreign colour-CA neighbours
  int nf = 0;
  int f_self[8] = {}, f_nb[8][6] = {};  // RECOMP: a foreign neighbour's value and its links
  constexpr int find(int site, int dx, int dy, int dz) const {
    for (int i = 0; i < n; ++i)
      if (v[i].site == site && v[i].dx == dx && v[i].dy == dy && v[i].dz == dz) return i;
    return -1;
  }
  constexpr int add(int site, int dx, int dy, int dz) {
    const int i = find(site, dx, dy, dz);
    if (i >= 0) return i;
    v[n] = {(int8_t)site, (int8_t)dx, (int8_t)dy, (int8_t)dz};
    return n++;
  }
  constexpr PairNeeds() {
    for (int k = 0; k < 4; ++k) {
      const int s = CA + 4 * k;
      a_self[k] = add(s, 0, 0, 0);
      for (int m = 0; m < 6; ++m)
        a_nb[k][m] = add(kLinks[s][m].site, kLinks[s][m].dx, kLinks[s][m].dy, kLinks[s][m].dz);
    }
    for (int k = 0; k < 4; ++k) {
      const int t = CA + 1 + 4 * k;
      t_self[k] = add(t, 0, 0, 0);
      for (int m = 0; m < 6; ++m) {
        const Link lk = kLinks[t][m];
        const bool ca = (lk.site & 3) == CA;
        t_nb[k][m] = (!ca || RECOMP) ? add(lk.site, lk.dx, lk.dy, lk.dz) : -1;
        f_idx[k][m] = -1;
        if (ca && (lk.dx | lk.dy | lk.dz) != 0) {
          f_idx[k][m] = nf;
          fv[nf] = {lk.site, lk.dx, lk.dy, lk.dz};
          if (RECOMP) {
            f_self[nf] = t_nb[k][m];
            for (int q = 0; q < 6; ++q) {
              const Link l2 = kLinks[lk.site][q];
              f_nb[nf][q] = add(l2.site, lk.dx + l2.dx, lk.dy + l2.dy, lk.dz + l2.dz);
            }
          }
          ++nf;
        }
      }
    }
  }
};

// Over-relaxation of the spin at index self from its 6 links (indices nbr)
// among the values vx/vy/vz, with its FP64 neighbour sum when TRACK.
template <bool TRACK, int NN>
__device__ __forceinline__ f3 pair_relax(const float (&vx)[NN], const float (&vy)[NN],
                                         const float (&vz)[NN], int self, const int *nbr,
                                         double &gx, double &gy, double &gz) {
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
  gx = gy = gz = 0.0;
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    const int i = nbr[m];
    fx = fx + vx[i];
    fy = fy + vy[i];
    fz = fz + vz[i];
    if (TRACK) {
      gx += (double)vx[i];
      gy += (double)vy[i];
      gz += (double)vz[i];
    }
  }
  return overrelax({fx, fy, fz}, {vx[self], vy[self], vz[self]});
}

// Phase CA of a cell's pair: na[k] = new spin of site CA + 4k.
template <int CA, bool TRACK, bool RECOMP, int NN>
__device__ __forceinline__ void or_pair_phase_a(const float (&vx)[NN], const float (&vy)[NN],
                                                const float (&vz)[NN], float jeff, double &de,
                                                f3 (&na)[4]) {
  constexpr PairNeeds<CA, RECOMP> nd{};
  static_assert(NN == nd.n, "value arrays hold the pair's needs");
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    double gx, gy, gz;
    na[k] = pair_relax<TRACK>(vx, vy, vz, nd.a_self[k], nd.a_nb[k], gx, gy, gz);
    if (TRACK) {
      const int i = nd.a_self[k];
      de += field_de<false>(jeff, nullptr, CA, {vx[i], vy[i], vz[i]}, na[k], gx, gy, gz);
    }
  }
}

// The new values of the foreign colour-CA neighbours, recomputed with their
// owners' float expressions in their owners' order (over-relaxation is
// deterministic), so they equal what the owners compute bit for bit.
template <int CA, int NN>
__device__ __forceinline__ void or_pair_foreign(const float (&vx)[NN], const float (&vy)[NN],
                                                const float (&vz)[NN], f3 (&fo)[3]) {
  constexpr PairNeeds<CA, true> nd{};
  static_assert(nd.nf == 3, "3 foreign neighbours per pair");
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    double hx, hy, hz;
    fo[f] = pair_relax<false>(vx, vy, vz, nd.f_self[f], nd.f_nb[f], hx, hy, hz);
  }
}

// Phase CA + 1 of a cell's pair from the old values, the cell's new colour-CA
// spins na and the foreign ones fo: nb[k] = new spin of site CA + 1 + 4k.
template <int CA, bool TRACK, bool RECOMP, int NN>
__device__ __forceinline__ void or_pair_phase_b(const float (&vx)[NN], const float (&vy)[NN],
                                                const float (&vz)[NN], float jeff, double &de,
                                                const f3 (&na)[4], const f3 (&fo)[3], f3 (&nb)[4]) {
  constexpr int CB = CA + 1;
  constexpr PairNeeds<CA, RECOMP> nd{};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int t = CB + 4 * k;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      const Link lk = kLinks[t][m];
      f3 w;
      if ((lk.site & 3) == CA) {
        w = nd.f_idx[k][m] < 0 ? na[lk.site >> 2] : fo[nd.f_idx[k][m]];
      } else {
        const int i = nd.t_nb[k][m];
        w = {vx[i], vy[i], vz[i]};
      }
      fx = fx + w.x;
      fy = fy + w.y;
      fz = fz + w.z;
      if (TRACK) {
        gx += (double)w.x;
        gy += (double)w.y;
        gz += (double)w.z;
      }
    }
    const int it = nd.t_self[k];
    const f3 told = {vx[it], vy[it], vz[it]};
    nb[k] = overrelax({fx, fy, fz}, told);
    if (TRACK) de += field_de<false>(jeff, nullptr, CB, told, nb[k], gx, gy, gz);
  }
}

// The arithmetic of two over-relaxation colour phases CA, CA + 1 of one cell,
// scalar model, on the cell's PairNeeds<CA> values vx/vy/vz (old spins; the
// colours below CA already updated this pass), the foreign neighbours
// recomputed: every result equals two separate phase launches bit for bit.
// na, nb: the new colour-CA and colour-CA+1 spins of sites CA + 4k,
// CA + 1 + 4k. TRACK: FP64 energy changes of this cell's updates added to de.
template <int CA, bool TRACK, int NN>
__device__ __forceinline__ void or_pair_math(const float (&vx)[NN], const float (&vy)[NN],
                                             const float (&vz)[NN], float jeff, double &de,
                                             f3 (&na)[4], f3 (&nb)[4]) {
  or_pair_phase_a<CA, TRACK, true>(vx, vy, vz, jeff, de, na);
  f3 fo[3];
  or_pair_foreign<CA>(vx, vy, vz, fo);
  or_pair_phase_b<CA, TRACK, true>(vx, vy, vz, jeff, de, na, fo, nb);
}

// Store of the pair's 8 new spins of one cell into the second buffer.
template <int CA, int NC>
__device__ __forceinline__ void or_pair_store(float *__restrict__ oth, int cell, const f3 (&na)[4],
                                              const f3 (&nb)[4]) {
  constexpr int CB = CA + 1, P = 16 * NC;
  float *oy = oth + P, *oz = oth + 2 * P;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t ia = 4u * (uint32_t)((CA + 4 * k) * NC + cell);
    const uint32_t ib = 4u * (uint32_t)((CB + 4 * k) * NC + cell);
    st_at(oth, ia, na[k].x);
    st_at(oy, ia, na[k].y);
    st_at(oz, ia, na[k].z);
    st_at(oth, ib, nb[k].x);
    st_at(oy, ib, nb[k].y);
    st_at(oz, ib, nb[k].z);
  }
}

// Two over-relaxation colour phases CA, CA + 1 of one cell in one pass over
// the data, scalar model (cur/oth: this replica's spins in the two buffers).
// Colours below CA have already been updated this pass and are read from
// oth; the others are read from cur; the two new colours are written to oth,
// so no thread reads what another thread of the launch writes. All 41 spins
// the cell needs are loaded in one burst before any arithmetic (one memory
// round trip, and the halo lines are used while they are still in L2).
template <int CA, int LT, bool TRACK>
__device__ __forceinline__ void or_pair_cell(const float *__restrict__ cur,
                                             float *__restrict__ oth, int cell, float jeff,
                                             double &de) {
  static_assert(CA == 0 || CA == 2, "colour pairs (0, 1) and (2, 3)");
  constexpr int L = LT, LL = LT * LT, NC = LT * LT * LT, P = 16 * NC;
  constexpr PairNeeds<CA> nd{};
  constexpr int NN = nd.n;
  const int z = cell / LL;
  const int y = (cell / L) % L;
  const int x = cell % L;
  const int X[3] = {x == 0 ? L - 1 : x - 1, x, x == L - 1 ? 0 : x + 1};
  const int Y[3] = {(y == 0 ? L - 1 : y - 1) * L, y * L, (y == L - 1 ? 0 : y + 1) * L};
  const int Z[3] = {(z == 0 ? L - 1 : z - 1) * LL, z * LL, (z == L - 1 ? 0 : z + 1) * LL};
  float vx[NN], vy[NN], vz[NN];
  const float *cy = cur + P, *cz = cur + 2 * P;
  const float *oy = oth + P, *oz = oth + 2 * P;
#pragma unroll
  for (int i = 0; i < NN; ++i) {
    const PairNeed q = nd.v[i];
    const uint32_t j = 4u * (uint32_t)(q.site * NC + Z[q.dz + 1] + Y[q.dy + 1] + X[q.dx + 1]);
    const bool o = (q.site & 3) < CA;
    vx[i] = ld_at(o ? oth : cur, j);
    vy[i] = ld_at(o ? oy : cy, j);
    vz[i] = ld_at(o ? oz : cz, j);
  }
  f3 na[4], nb[4];
  or_pair_math<CA, TRACK>(vx, vy, vz, jeff, de, na, nb);
  or_pair_store<CA, NC>(oth, cell, na, nb);
}

// One site's update (site = colour + 4k, runtime), for kernels that cannot
// afford the registers of cell_update (its 4 concurrent targets): the same
// float expressions in the same order, so the same results bit for bit.
// links: the link table in constant memory of the calling translation unit.
template <int KIND, int LT, bool TRACK, bool MAT = false>
__device__ __forceinline__ void site_update(float *__restrict__ Sx, const Link (*links)[6], int cell,
                                            int site, int r, int pass, float beta_r, float cosc_r,
                                            uint32_t sweep, uint32_t key0, uint32_t key1,
                                            float jeff, double &de, unsigned &nacc,
                                            const float *__restrict__ coup = nullptr,
                                            bool post_rotate = false) {
  constexpr int L = LT, LL = LT * LT, NC = LT * LT * LT;
  float *Sy = Sx + 16 * NC;
  float *Sz = Sy + 16 * NC;
  const int z = cell / LL;
  const int y = (cell / L) % L;
  const int x = cell % L;
  const int bs = site & 3;
  float fx, fy, fz;
  field_start<MAT>(coup, bs, fx, fy, fz);
  double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
  for (int n = 0; n < 6; ++n) {
    const Link lk = links[site][n];
    const int x1 = x + lk.dx, y1 = y + lk.dy, z1 = z + lk.dz;
    const int j = lk.site * NC + (z1 < 0 ? z1 + L : (z1 >= L ? z1 - L : z1)) * LL +
                  (y1 < 0 ? y1 + L : (y1 >= L ? y1 - L : y1)) * L +
                  (x1 < 0 ? x1 + L : (x1 >= L ? x1 - L : x1));
    const float a = Sx[j], b = Sy[j], c = Sz[j];
    field_add<MAT, KIND == 0 || TRACK>(coup, bs, lk.site & 3, a, b, c, fx, fy, fz, gx, gy, gz);
  }
  const int i = site * NC + cell;
  const f3 s = {Sx[i], Sy[i], Sz[i]};
  f3 sp;
  bool wr;
  if (KIND == 0) {
    const u32x4 w = philox4x32_10(
        rng_counter((uint32_t)cell, (uint32_t)site, (uint32_t)pass, kKindMetropolis, (uint32_t)r,
                    sweep),
        key0, key1);
    const f3 h = field_h<MAT>(jeff, fx, fy, fz);
    sp = cone_proposal(s, cosc_r, u01(w.x), u01(w.y));
    wr = metropolis_accept(h, s, sp, beta_r, u01(w.z));
    nacc += wr ? 1u : 0u;
    if (post_rotate) {
      sp = rotate_about(h, wr ? sp : s, u01(w.w));
      wr = true;
    }
  } else {
    const f3 f = {fx, fy, fz};
    sp = overrelax(f, s);
    wr = true;
  }
  if (wr) {
    if (KIND == 0 || TRACK) de += field_de<MAT>(jeff, coup, bs, s, sp, gx, gy, gz);
    Sx[i] = sp.x;
    Sy[i] = sp.y;
    Sz[i] = sp.z;
  }
}

// Block-wide sum of (de, nacc); the result is valid in thread 0. Every thread
// of the block must call it. Uses 2 * (blockDim / 64) words of shared memory.
template <int BLOCK>
__device__ __forceinline__ void block_sum(double &de, unsigned &nacc) {
  constexpr int NW = BLOCK / 64;
  __shared__ double s_de[NW];
  __shared__ unsigned s_acc[NW];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  de = wave_sum_d(de);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) nacc += __shfl_xor(nacc, off, 64);
  if (lane == 0) {
    s_de[wv] = de;
    s_acc[wv] = nacc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    unsigned a = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      t += s_de[w];
      a += s_acc[w];
    }
    de = t;
    nacc = a;
  }
  __syncthreads();
}

}  // namespace pyro
