/* CPU oracle for the pyrochlore Heisenberg Monte Carlo step.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference algorithm,
 * used as the checker by tests/, __graft_entry__.smoke() and the CPU-baseline
 * leg of bench.py. Nothing in the product path includes, links or loads it.
 *
 * Built with gcc -O2 -ffp-contract=off (no fast-math, SSE float arithmetic),
 * so every float expression below rounds exactly once per operation, in the
 * order written. The GPU kernels evaluate the same expressions in the same
 * order; that is the basis of the bitwise trajectory comparison.
 *
 * Each function cites the reference lines it restates. Where the reference is
 * defective (see the deviations list in oracle/README.md) the intended
 * behaviour is restated instead and the deviation is named in the comment.
 */
#include "pyro_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Pyrochlore.cu:152-167 (links16): 16 sites x 6 neighbours, (dx,dy,dz,site). */
static const int8_t LINKS[16][6][4] = {
    {{0, 0, 0, 1}, {0, 0, 0, 2}, {0, 0, 0, 3}, {0, -1, -1, 10}, {-1, -1, 0, 5}, {-1, 0, -1, 15}},
    {{0, 0, 0, 0}, {0, 0, 0, 2}, {0, 0, 0, 3}, {0, 0, 0, 4}, {0, 0, -1, 14}, {0, 0, -1, 11}},
    {{0, 0, 0, 0}, {0, 0, 0, 1}, {0, 0, 0, 3}, {0, 0, 0, 8}, {-1, 0, 0, 13}, {-1, 0, 0, 7}},
    {{0, 0, 0, 0}, {0, 0, 0, 1}, {0, 0, 0, 2}, {0, -1, 0, 9}, {0, -1, 0, 6}, {0, 0, 0, 12}},
    {{0, 0, 0, 5}, {0, 0, 0, 6}, {0, 0, 0, 7}, {0, 0, 0, 1}, {0, 0, -1, 14}, {0, 0, -1, 11}},
    {{0, 0, 0, 4}, {0, 0, 0, 6}, {0, 0, 0, 7}, {1, 1, 0, 0}, {1, 0, -1, 10}, {0, 1, -1, 15}},
    {{0, 0, 0, 4}, {0, 0, 0, 5}, {0, 0, 0, 7}, {0, 1, 0, 3}, {0, 1, 0, 12}, {0, 0, 0, 9}},
    {{0, 0, 0, 4}, {0, 0, 0, 5}, {0, 0, 0, 6}, {1, 0, 0, 2}, {1, 0, 0, 8}, {0, 0, 0, 13}},
    {{0, 0, 0, 9}, {0, 0, 0, 10}, {0, 0, 0, 11}, {-1, 0, 0, 7}, {0, 0, 0, 2}, {-1, 0, 0, 13}},
    {{0, 0, 0, 8}, {0, 0, 0, 10}, {0, 0, 0, 11}, {0, 1, 0, 3}, {0, 0, 0, 6}, {0, 1, 0, 12}},
    {{0, 0, 0, 8}, {0, 0, 0, 9}, {0, 0, 0, 11}, {0, 1, 1, 0}, {-1, 0, 1, 5}, {-1, 1, 0, 15}},
    {{0, 0, 0, 8}, {0, 0, 0, 9}, {0, 0, 0, 10}, {0, 0, 1, 1}, {0, 0, 1, 4}, {0, 0, 0, 14}},
    {{0, 0, 0, 13}, {0, 0, 0, 14}, {0, 0, 0, 15}, {0, 0, 0, 3}, {0, -1, 0, 6}, {0, -1, 0, 9}},
    {{0, 0, 0, 12}, {0, 0, 0, 14}, {0, 0, 0, 15}, {1, 0, 0, 2}, {1, 0, 0, 8}, {0, 0, 0, 7}},
    {{0, 0, 0, 12}, {0, 0, 0, 13}, {0, 0, 0, 15}, {0, 0, 0, 11}, {0, 0, 1, 1}, {0, 0, 1, 4}},
    {{0, 0, 0, 12}, {0, 0, 0, 13}, {0, 0, 0, 14}, {1, 0, 1, 0}, {0, -1, 1, 5}, {1, -1, 0, 10}}};

/* Pyrochlore.cu:36-39 (IndependentSubLattices). */
static const int8_t INDEP[24][4] = {
    {0, 4, 8, 12}, {0, 14, 8, 6}, {0, 14, 7, 9}, {0, 11, 13, 6}, {0, 11, 7, 12}, {0, 4, 13, 9},
    {1, 5, 13, 9}, {1, 15, 8, 6}, {1, 15, 7, 9}, {1, 10, 13, 6}, {1, 10, 7, 12}, {1, 5, 8, 12},
    {2, 10, 14, 6}, {2, 15, 11, 6}, {2, 15, 4, 9}, {2, 10, 4, 12}, {2, 5, 14, 9}, {2, 5, 11, 12},
    {3, 15, 11, 7}, {3, 15, 4, 8}, {3, 10, 14, 7}, {3, 10, 4, 13}, {3, 5, 14, 8}, {3, 5, 11, 13}};

enum { KIND_METROPOLIS = 1, KIND_INIT = 2, KIND_SWAP = 3, KIND_OR_ANGLE = 6, KIND_SETS = 7 };

void orc_links16(int8_t out[16][6][4]) { memcpy(out, LINKS, sizeof(LINKS)); }
void orc_independent_sets(int8_t out[24][4]) { memcpy(out, INDEP, sizeof(INDEP)); }

/* Philox4x32-10, Salmon et al. SC'11, section 4 (the reference used cuRAND and
 * MKL MT19937 here, MonteCarlo.cu:3-13,116; replaced by this counter-based
 * generator on both sides). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Offset added to the replica index of every Metropolis / init counter, and
 * the group index of the swap counters: lets one replica (or one
 * equal-size group) of a heterogeneous batch be replayed on its own. */
static int g_replica_base = 0;
static int g_swap_group = 0;
void orc_set_replica_base(int base) { g_replica_base = base; }
void orc_set_swap_group(int group) { g_swap_group = group; }

static void draw(uint32_t cell, uint32_t site, uint32_t pass, uint32_t kind, uint32_t replica,
                 uint32_t sweep, uint64_t seed, uint32_t out[4]) {
  if (kind == KIND_SWAP)
    replica = (uint32_t)g_swap_group;
  else
    replica += (uint32_t)g_replica_base;
  uint32_t ctr[4] = {cell, site | (pass << 8) | (kind << 24), replica, sweep};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  orc_philox4x32_10(ctr, key, out);
}

static float u01(uint32_t x) { return (float)(x >> 8) * 5.9604644775390625e-8f; }

static double u01_53(uint32_t a, uint32_t b) {
  uint64_t m = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
  return (double)m * 1.1102230246251565e-16;
}

/* exp(x), x <= 0: Cody-Waite reduction with the Cephes two-part ln2, degree-7
 * Taylor polynomial, exponent-field scaling; 0 below -86. Stands in for the
 * reference's EXP in Metropolis (Math.cu:979). */
float orc_exp_neg(float x) {
  if (!(x > -86.0f)) return 0.0f;
  if (x > 0.0f) x = 0.0f;
  float kf = x * 1.44269502162933349609375f + 0.5f;
  float ti = (float)(int)kf;
  if (ti > kf) ti = ti - 1.0f;
  kf = ti;
  float r = x - kf * 0.693359375f;
  r = r - kf * -2.12194440e-4f;
  float p = 1.98412698e-4f;
  p = p * r + 1.38888889e-3f;
  p = p * r + 8.33333333e-3f;
  p = p * r + 4.16666667e-2f;
  p = p * r + 1.66666667e-1f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  union { uint32_t u; float f; } s;
  s.u = (uint32_t)((int)kf + 127) << 23;
  return p * s.f;
}

/* sin(pi a), cos(pi a), a in [0,2]: exact quarter-turn reduction, Taylor
 * polynomials of degree 9 (sin) and 10 (cos). Stands in for SINCOSPI
 * (Math.cu:123,178). */
void orc_sincospi(float a, float *s, float *c) {
  float nf = (float)(int)(a * 2.0f + 0.5f);
  float f = a - nf * 0.5f;
  float t = f * 3.14159274101257324f;
  float t2 = t * t;
  float sp = 2.75573192e-6f;
  sp = sp * t2 - 1.98412698e-4f;
  sp = sp * t2 + 8.33333333e-3f;
  sp = sp * t2 - 1.66666667e-1f;
  sp = sp * t2;
  sp = sp * t + t;
  float cp = -2.75573192e-7f;
  cp = cp * t2 + 2.48015873e-5f;
  cp = cp * t2 - 1.38888889e-3f;
  cp = cp * t2 + 4.16666667e-2f;
  cp = cp * t2 - 0.5f;
  cp = cp * t2 + 1.0f;
  switch (((int)nf) & 3) {
    case 0: *s = sp; *c = cp; break;
    case 1: *s = cp; *c = -sp; break;
    case 2: *s = -sp; *c = -cp; break;
    default: *s = -cp; *c = sp; break;
  }
}

/* Renormalisation of a nearly unit vector (every caller's is a reflection or
 * rotation of a unit spin, off by rounding): one Newton step for 1/sqrt(q)
 * from 1, the factor 1 + c with c = (1 - q)/2 (exact), applied as v + v c
 * (error 3/8 (q - 1)^2, below an FP32 ulp for |q - 1| < 1e-4; the length
 * unbiased to rounding). The reference renormalises with the hardware RSQRT
 * (Math.cu:238); this is a deterministic replacement. */
static void renorm(float v[3]) {
  float q = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  float c = 0.5f * (1.0f - q);
  v[0] = v[0] + v[0] * c;
  v[1] = v[1] + v[1] * c;
  v[2] = v[2] + v[2] * c;
}

void orc_renormalise(const float v[3], float out[3]) {
  out[0] = v[0];
  out[1] = v[1];
  out[2] = v[2];
  renorm(out);
}

/* Cone proposal, Math.cu:244-281: uniform on the cap of half-angle acos(cos_cone)
 * around s via a Householder reflection. Deviation (reference defect, Math.cu:254-260:
 * the s.z < -0.99995 branch builds the cap around +-y, not around s): here the
 * reflection pole is +z for s.z >= 0 and -z otherwise, which is exact for every s. */
void orc_cone_proposal(const float s[3], float cos_cone, float u_cos, float u_phi, float out[3]) {
  float tc = 1.0f - (1.0f - cos_cone) * u_cos;
  float q = 1.0f - tc * tc;
  float ts = sqrtf(q > 0.0f ? q : 0.0f);
  float ps, pc;
  orc_sincospi(2.0f * u_phi, &ps, &pc);
  float r0 = ts * pc, r1 = ts * ps, r2, N2, d;
  if (s[2] >= 0.0f) {
    r2 = tc;
    N2 = s[2] + 1.0f;
    d = 1.0f + s[2];
  } else {
    r2 = -tc;
    N2 = s[2] - 1.0f;
    d = 1.0f - s[2];
  }
  float scal = (s[0] * r0 + s[1] * r1 + N2 * r2) / d;
  out[0] = s[0] * scal - r0;
  out[1] = s[1] * scal - r1;
  out[2] = N2 * scal - r2;
  renorm(out);
}

/* Metropolis.  Math.cu:968-986: DE = (-h.s) - (-h.s'), accept iff u < min(1, exp(DE*beta)). */
int orc_metropolis_accept(const float h[3], const float s[3], const float sp[3], float beta,
                          float u) {
  float eb = -(h[0] * s[0] + h[1] * s[1] + h[2] * s[2]);
  float ea = -(h[0] * sp[0] + h[1] * sp[1] + h[2] * sp[2]);
  float x = (eb - ea) * beta;
  if (x >= 0.0f) return 1;
  return u < orc_exp_neg(x);
}

/* Over-relaxation by pi, Math.cu:220-242: s' = 2 (s.h)/|h|^2 h - s, renormalised
 * (one Newton step from 1 instead of RSQRT). Unchanged when h = 0 (the reference
 * divides by 0). */
void orc_overrelax(const float h[3], const float s[3], float out[3]) {
  float hh = h[0] * h[0] + h[1] * h[1] + h[2] * h[2];
  if (!(hh > 0.0f)) {
    out[0] = s[0]; out[1] = s[1]; out[2] = s[2];
    return;
  }
  float scal = (s[0] * h[0] + s[1] * h[1] + s[2] * h[2]) / hh;
  float t = 2.0f * scal;
  out[0] = t * h[0] - s[0];
  out[1] = t * h[1] - s[1];
  out[2] = t * h[2] - s[2];
  renorm(out);
}

/* Periodic neighbour index, as the CPU TotalEnergy resolves it
 * (Pyrochlore.cu:476-496: x+dx, y+dy, z+dz wrapped independently). The GPU
 * resolver of the reference, CalculateNeighbor3D (Math.cu:988-1009), adds the
 * x offset to y and the y offset to z; that is a defect and is not restated. */
static long nbr_index(int L, int x, int y, int z, const int8_t *lk) {
  int x1 = x + lk[0] + L;
  while (x1 >= L) x1 -= L;
  int y1 = y + lk[1] + L;
  while (y1 >= L) y1 -= L;
  int z1 = z + lk[2] + L;
  while (z1 >= L) z1 -= L;
  long nc = (long)L * L * L;
  return ((long)(z1 * L + y1) * L + x1) + (long)lk[3] * nc;
}

/* Local-field model. NULL: the scalar Heisenberg model, C = -I (J1=J2=1,
 * J3=J4=0), h = -jeff * (neighbour sum). Otherwise the 156 floats of the
 * matrix model (CalculateLocalExchangeField, Pyrochlore.cu:210-275, with the
 * Zeeman term of InitializeMagneticField :414-450): coup[9 (4 b + b') + 3 i + j]
 * = spin_len2 C_{b b'}[i][j], coup[144 + 3 b + i] = 2 spin_len2 B_b[i];
 * h = B_b + sum_n C_{b b_n} s_n, energy -sum_i s_i.(B_b + (1/2) sum_n C s_n). */
static const float *g_coup = NULL;

void orc_set_couplings(const float *coup156) { g_coup = coup156; }

/* Field of site `site` at (x,y,z): FP32 in link order (the field the update
 * uses: the plain neighbour sum for the scalar model, B + sum C s for the
 * matrix model) and FP64 (for the energy bookkeeping: sum s, or sum C s
 * without B). */
static void nbr_sum(const float *S, int L, int x, int y, int z, int site, float f[3],
                    double d[3]) {
  long nc = (long)L * L * L;
  const float *Sx = S, *Sy = S + 16 * nc, *Sz = S + 32 * nc;
  int b = site & 3;
  if (g_coup) {
    f[0] = g_coup[144 + 3 * b];
    f[1] = g_coup[145 + 3 * b];
    f[2] = g_coup[146 + 3 * b];
  } else {
    f[0] = f[1] = f[2] = 0.0f;
  }
  d[0] = d[1] = d[2] = 0.0;
  for (int k = 0; k < 6; k++) {
    long j = nbr_index(L, x, y, z, LINKS[site][k]);
    float a = Sx[j], bb = Sy[j], c = Sz[j];
    if (g_coup) {
      const float *m = g_coup + 9 * (4 * b + (LINKS[site][k][3] & 3));
      f[0] = f[0] + ((m[0] * a + m[1] * bb) + m[2] * c);
      f[1] = f[1] + ((m[3] * a + m[4] * bb) + m[5] * c);
      f[2] = f[2] + ((m[6] * a + m[7] * bb) + m[8] * c);
      d[0] += ((double)m[0] * a + (double)m[1] * bb) + (double)m[2] * c;
      d[1] += ((double)m[3] * a + (double)m[4] * bb) + (double)m[5] * c;
      d[2] += ((double)m[6] * a + (double)m[7] * bb) + (double)m[8] * c;
    } else {
      f[0] = f[0] + a;
      f[1] = f[1] + bb;
      f[2] = f[2] + c;
      d[0] += (double)a;
      d[1] += (double)bb;
      d[2] += (double)c;
    }
  }
}

/* Optional rotation about h after each Metropolis test (Math.cu:168-197,
 * Pyrochlore.cu:872), angle 2 pi u (Rodrigues), renormalised; s unchanged
 * when h = 0. */
static int g_post_rotate = 0;

void orc_set_post_rotate(int on) { g_post_rotate = on != 0; }

void orc_rotate_about(const float h[3], const float s[3], float u, float out[3]) {
  float hh = h[0] * h[0] + h[1] * h[1] + h[2] * h[2];
  if (!(hh > 0.0f)) {
    out[0] = s[0]; out[1] = s[1]; out[2] = s[2];
    return;
  }
  float inv = 1.0f / sqrtf(hh);
  float n[3] = {h[0] * inv, h[1] * inv, h[2] * inv};
  float st, ct;
  orc_sincospi(2.0f * u, &st, &ct);
  float ns = n[0] * s[0] + n[1] * s[1] + n[2] * s[2];
  float k = ns * (1.0f - ct);
  out[0] = (s[0] * ct + (n[1] * s[2] - n[2] * s[1]) * st) + n[0] * k;
  out[1] = (s[1] * ct + (n[2] * s[0] - n[0] * s[2]) * st) + n[1] * k;
  out[2] = (s[2] * ct + (n[0] * s[1] - n[1] * s[0]) * st) + n[2] * k;
  renorm(out);
}

/* Update schedule. g_sets = 0: phase c of a pass updates the colour class
 * {c, c+4, c+8, c+12}. g_sets = 2: phase c updates the colour class perm[c]
 * of a permutation of the four drawn per (sweep, stage, pass) (a random
 * colour order; every site still moves once a pass, so the move draws keep
 * their colour-order counters). g_sets = 1: it
 * updates one of the 24 independent sets,
 * drawn per (sweep, stage, pass, phase) and shared by every replica, as the
 * reference picks IndependentSubLattices[Update[0] % 24] for each launch
 * (Pyrochlore.cu:848, :985, :1021); its Update buffer is never filled
 * (Pyrochlore.cu:84), so the draw here is a Philox word instead. A site can
 * then be updated twice in one pass, so the move draws of a drawn-set phase
 * carry the phase in bits 4-5 of their site field.
 * g_or_angle = 0: over-relaxation by pi (OverRelaxationByPiStep,
 * Pyrochlore.cu:975-1007); 1: rotation about the local field h by a uniform
 * angle 2 pi u (OverRelaxationStep, Pyrochlore.cu:1009-1045), u from its own
 * Philox stream keyed by (cell, site field, over-relaxation pass, replica,
 * sweep). */
static int g_sets = 0;
static int g_or_angle = 0;

void orc_set_update_schedule(int sets, int or_angle) {
  g_sets = (sets == 1 || sets == 2) ? sets : 0;
  g_or_angle = or_angle != 0;
}

/* The colour that phase `phase` of a pass updates under the drawn order: a
 * Fisher-Yates shuffle of (0, 1, 2, 3) whose choices are the mixed-radix
 * digits (base 4, 3, 2) of one Philox word, kind 7, replica field 1. */
int orc_drawn_colour(uint32_t sweep, int stage, int pass, int phase, uint64_t seed) {
  uint32_t ctr[4] = {0u, (uint32_t)stage | ((uint32_t)pass << 8) | ((uint32_t)KIND_SETS << 24), 1u,
                     sweep};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t w[4];
  orc_philox4x32_10(ctr, key, w);
  int perm[4] = {0, 1, 2, 3};
  uint32_t v = w[0];
  for (int i = 3; i > 0; i--) {
    const int j = (int)(v % (uint32_t)(i + 1));
    v /= (uint32_t)(i + 1);
    const int t = perm[i];
    perm[i] = perm[j];
    perm[j] = t;
  }
  return perm[phase];
}

/* The set drawn for phase `phase` of pass `pass` of a stage (0: Metropolis,
 * 1: over-relaxation) of sweep `sweep`. */
int orc_drawn_set(uint32_t sweep, int stage, int pass, int phase, uint64_t seed) {
  uint32_t ctr[4] = {(uint32_t)phase, (uint32_t)stage | ((uint32_t)pass << 8) | ((uint32_t)KIND_SETS << 24),
                     0u, sweep};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t w[4];
  orc_philox4x32_10(ctr, key, w);
  return (int)(w[0] % 24u);
}

/* Target site k of phase `phase` and the site field of its draws. */
static void phase_sites(uint32_t sweep, int stage, int pass, int phase, uint64_t seed, int sites[4],
                        uint32_t tags[4]) {
  if (g_sets == 2) {
    const int c = orc_drawn_colour(sweep, stage, pass, phase, seed);
    for (int k = 0; k < 4; k++) {
      sites[k] = c + 4 * k;
      tags[k] = (uint32_t)sites[k];
    }
  } else if (g_sets == 1) {
    const int set = orc_drawn_set(sweep, stage, pass, phase, seed);
    for (int k = 0; k < 4; k++) {
      sites[k] = INDEP[set][k];
      tags[k] = (uint32_t)sites[k] | ((uint32_t)phase << 4);
    }
  } else {
    for (int k = 0; k < 4; k++) {
      sites[k] = phase + 4 * k;
      tags[k] = (uint32_t)sites[k];
    }
  }
}

/* Order in which a colour phase visits its N/4 updates (cell * 4 + k, target
 * site colour + 4 k): NULL = cell by cell. The updates of a phase are
 * independent, so any order must give the same spins; tests use a shuffled
 * order as the race check of the sublattice-parallel update. */
static const int64_t *g_order = NULL;

void orc_set_update_order(const int64_t *perm) { g_order = perm; }

/* FP64 energy change of s -> sp at site basis b from the FP64 sums d. */
static double move_de(float jeff, int b, const float s[3], const float sp[3], const double d[3]) {
  if (g_coup)
    return -(((double)sp[0] - (double)s[0]) * (d[0] + (double)g_coup[144 + 3 * b]) +
             ((double)sp[1] - (double)s[1]) * (d[1] + (double)g_coup[145 + 3 * b]) +
             ((double)sp[2] - (double)s[2]) * (d[2] + (double)g_coup[146 + 3 * b]));
  return (double)jeff * (((double)sp[0] - (double)s[0]) * d[0] +
                         ((double)sp[1] - (double)s[1]) * d[1] +
                         ((double)sp[2] - (double)s[2]) * d[2]);
}

void orc_init_spins(float *spins, int L, int R, uint64_t seed) {
  long nc = (long)L * L * L;
#pragma omp parallel for collapse(2) schedule(static)
  for (int r = 0; r < R; r++) {
    for (long cell = 0; cell < nc; cell++) {
      float *S = spins + (long)r * 48 * nc;
      for (int site = 0; site < 16; site++) {
        uint32_t w[4];
        draw((uint32_t)cell, (uint32_t)site, 0, KIND_INIT, (uint32_t)r, 0, seed, w);
        /* Math.cu:120-130: azimuth from the first uniform, cos(theta) from the second. */
        float ps, pc;
        orc_sincospi(2.0f * u01(w[0]), &ps, &pc);
        float tc = 2.0f * u01(w[1]) - 1.0f;
        float q = 1.0f - tc * tc;
        float ts = sqrtf(q > 0.0f ? q : 0.0f);
        long i = (long)site * nc + cell;
        S[i] = pc * ts;
        S[16 * nc + i] = ps * ts;
        S[32 * nc + i] = tc;
      }
    }
  }
}

/* One Metropolis colour phase: MCMultipleTemperaturesPerSystem, Pyrochlore.cu:829-885,
 * restated with the cone proposal of Math.cu:244-281 (the reference kernel passes its
 * cone cosine but draws from the full sphere; cos_cone = -1 reproduces that) and one
 * Philox draw per update instead of the pre-filled uniform buffer. The post-accept
 * rotation of Pyrochlore.cu:872 is not applied. */
void orc_metropolis_phase(float *spins, int L, int R, int colour, uint32_t sweep, int pass,
                          const float *beta, const float *cos_cone, float jeff, uint64_t seed,
                          uint64_t *accepted, double *dE) {
  long nc = (long)L * L * L;
  int sites[4];
  uint32_t tags[4];
  phase_sites(sweep, 0, pass, colour, seed, sites, tags);
  for (int r = 0; r < R; r++) {
    float *S = spins + (long)r * 48 * nc;
    float *Sx = S, *Sy = S + 16 * nc, *Sz = S + 32 * nc;
    uint64_t acc = 0;
    double de = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : acc, de)
    for (long t = 0; t < 4 * nc; t++) {
      long upd = g_order ? (long)g_order[t] : t;
      long cell = upd >> 2;
      int x = (int)(cell % L), y = (int)((cell / L) % L), z = (int)(cell / ((long)L * L));
      {
        int k = (int)(upd & 3);
        int site = sites[k];
        long i = (long)site * nc + cell;
        float s[3] = {Sx[i], Sy[i], Sz[i]};
        float f[3];
        double d[3];
        nbr_sum(S, L, x, y, z, site, f, d);
        float h[3] = {-jeff * f[0], -jeff * f[1], -jeff * f[2]};
        if (g_coup) {
          h[0] = f[0];
          h[1] = f[1];
          h[2] = f[2];
        }
        uint32_t w[4];
        draw((uint32_t)cell, tags[k], (uint32_t)pass, KIND_METROPOLIS, (uint32_t)r, sweep, seed, w);
        float sp[3];
        orc_cone_proposal(s, cos_cone[r], u01(w[0]), u01(w[1]), sp);
        int take = orc_metropolis_accept(h, s, sp, beta[r], u01(w[2]));
        if (take) acc++;
        if (g_post_rotate) {
          float rot[3];
          orc_rotate_about(h, take ? sp : s, u01(w[3]), rot);
          sp[0] = rot[0]; sp[1] = rot[1]; sp[2] = rot[2];
          take = 1;
        }
        if (take) {
          de += move_de(jeff, site & 3, s, sp, d);
          Sx[i] = sp[0];
          Sy[i] = sp[1];
          Sz[i] = sp[2];
        }
      }
    }
    if (accepted) accepted[r] += acc;
    if (dE) dE[r] += de;
  }
}

/* One over-relaxation phase: OverRelaxationByPiStep, Pyrochlore.cu:975-1007, or
 * with g_or_angle the random-angle rotation of OverRelaxationStep,
 * Pyrochlore.cu:1009-1045, about h (-jeff f, or f for the matrix model);
 * the set of phase `phase` of over-relaxation pass `pass` by the schedule. */
void orc_overrelax_phase_ex(float *spins, int L, int R, int phase, uint32_t sweep, int pass,
                            float jeff, uint64_t seed, double *dE) {
  long nc = (long)L * L * L;
  int sites[4];
  uint32_t tags[4];
  phase_sites(sweep, 1, pass, phase, seed, sites, tags);
  for (int r = 0; r < R; r++) {
    float *S = spins + (long)r * 48 * nc;
    float *Sx = S, *Sy = S + 16 * nc, *Sz = S + 32 * nc;
    double de = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : de)
    for (long t = 0; t < 4 * nc; t++) {
      long upd = g_order ? (long)g_order[t] : t;
      long cell = upd >> 2;
      int x = (int)(cell % L), y = (int)((cell / L) % L), z = (int)(cell / ((long)L * L));
      {
        int k = (int)(upd & 3);
        int site = sites[k];
        long i = (long)site * nc + cell;
        float s[3] = {Sx[i], Sy[i], Sz[i]};
        float f[3];
        double d[3];
        nbr_sum(S, L, x, y, z, site, f, d);
        float sp[3];
        if (g_or_angle) {
          float h[3] = {-jeff * f[0], -jeff * f[1], -jeff * f[2]};
          if (g_coup) {
            h[0] = f[0];
            h[1] = f[1];
            h[2] = f[2];
          }
          uint32_t w[4];
          draw((uint32_t)cell, tags[k], (uint32_t)pass, KIND_OR_ANGLE, (uint32_t)r, sweep, seed, w);
          orc_rotate_about(h, s, u01(w[0]), sp);
        } else {
          orc_overrelax(f, s, sp);
        }
        de += move_de(jeff, site & 3, s, sp, d);
        Sx[i] = sp[0];
        Sy[i] = sp[1];
        Sz[i] = sp[2];
      }
    }
    if (dE) dE[r] += de;
  }
}

/* The colour phase of the plain schedule (over-relaxation by pi). */
void orc_overrelax_phase(float *spins, int L, int R, int colour, float jeff, double *dE) {
  const int sets = g_sets, angle = g_or_angle;
  g_sets = g_or_angle = 0;
  orc_overrelax_phase_ex(spins, L, R, colour, 0, 0, jeff, 0, dE);
  g_sets = sets;
  g_or_angle = angle;
}

/* Energy: CPU TotalEnergy, Pyrochlore.cu:468-523 (every bond visited from both
 * ends, halved), with the FP64 accumulation of the build; plus the sublattice
 * sums that MagnetizationGPU and Measurement start from (Pyrochlore.cu:277-342,
 * 526-578). */
void orc_observables(const float *spins, int L, int R, float jeff, double *out) {
  long nc = (long)L * L * L;
  for (int r = 0; r < R; r++) {
    const float *S = spins + (long)r * 48 * nc;
    const float *Sx = S, *Sy = S + 16 * nc, *Sz = S + 32 * nc;
    double e = 0.0;
    double sub[12];
    for (int i = 0; i < 12; i++) sub[i] = 0.0;
    for (long cell = 0; cell < nc; cell++) {
      int x = (int)(cell % L), y = (int)((cell / L) % L), z = (int)(cell / ((long)L * L));
      for (int site = 0; site < 16; site++) {
        long i = (long)site * nc + cell;
        float f[3];
        double d[3];
        nbr_sum(S, L, x, y, z, site, f, d);
        int b = site & 3;
        if (g_coup)
          e += (double)Sx[i] * ((double)g_coup[144 + 3 * b] + 0.5 * d[0]) +
               (double)Sy[i] * ((double)g_coup[145 + 3 * b] + 0.5 * d[1]) +
               (double)Sz[i] * ((double)g_coup[146 + 3 * b] + 0.5 * d[2]);
        else
          e += (double)Sx[i] * d[0] + (double)Sy[i] * d[1] + (double)Sz[i] * d[2];
        sub[3 * b + 0] += Sx[i];
        sub[3 * b + 1] += Sy[i];
        sub[3 * b + 2] += Sz[i];
      }
    }
    out[r * 13 + 0] = g_coup ? -e : 0.5 * (double)jeff * e;
    for (int i = 0; i < 12; i++) out[r * 13 + 1 + i] = sub[i];
  }
}

/* MagnetizationGPU, Pyrochlore.cu:277-342: the sum over spins of g_b s_i,
 * from the four sublattice sums s_b (sub12[3 b + k]). The four g-tensors are
 * the reference's g1..g4 (:283-286), whose diagonal is g1 = (2 gxy + gz)/3
 * and whose off-diagonal entries are +-g2, g2 = (gz - gxy)/3, with the sign
 * patterns of :283-286 (the same numbers InitializeMagneticField uses,
 * :416-420). The sum runs as the kernel's: site 0's row products, then
 * sites 1, 2, 3 added in turn (:318-336). */
void orc_magnetization_g(const double sub12[12], double gxy, double gz, double out[3]) {
  const double g1 = (2.0 * gxy + gz) / 3.0, g2 = (gz - gxy) / 3.0;
  /* off-diagonal signs of (0,1), (0,2), (1,2) per basis site, Pyrochlore.cu:283-286 */
  static const int sgn[4][3] = {{1, 1, 1}, {1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}};
  for (int i = 0; i < 3; i++) out[i] = 0.0;
  for (int b = 0; b < 4; b++) {
    double g[3][3];
    g[0][0] = g[1][1] = g[2][2] = g1;
    g[0][1] = g[1][0] = sgn[b][0] * g2;
    g[0][2] = g[2][0] = sgn[b][1] * g2;
    g[1][2] = g[2][1] = sgn[b][2] * g2;
    const double *sb = sub12 + 3 * b;
    for (int i = 0; i < 3; i++) out[i] += g[i][0] * sb[0] + g[i][1] * sb[1] + g[i][2] * sb[2];
  }
}

/* One round of the cone adaptation, after a round of sweeps with acceptance
 * counts accepted[r] of proposed[r]. The reference's CalculateSolidAngle
 * (MonteCarlo.cu:371-426) multiplies the half-angle theta = acos(cos_cone) by
 * 0.9 while the ratio is below the target, never enlarges it, and stops when
 * replica 0 reaches the target; its acceptance counters are never filled by
 * the production kernels (MonteCarlo.cu:346-367 reads DevResults, which
 * MCMultipleTemperaturesPerSystem leaves alone). The library's rule
 * (pyro_adapt_cone), restated here: every replica's theta is multiplied by 0.9 below the target and
 * divided by 0.9 at or above it, clamped to [1e-4, pi], and cos_cone =
 * (float)cos(theta), kept below 1. */
void orc_adapt_cone_step(int R, double target, const uint64_t *accepted,
                         const uint64_t *proposed, float *cos_cone) {
  for (int r = 0; r < R; r++) {
    double ratio = proposed[r] ? (double)accepted[r] / (double)proposed[r] : 0.0;
    double c = (double)cos_cone[r];
    if (c > 1.0) c = 1.0;
    if (c < -1.0) c = -1.0;
    double theta = acos(c);
    if (ratio < target)
      theta *= 0.9;
    else
      theta *= 1.0 / 0.9;
    if (theta > 3.14159265358979323846) theta = 3.14159265358979323846;
    if (theta < 1e-4) theta = 1e-4;
    float f = (float)cos(theta);
    if (f > 0.99999994f) f = 0.99999994f;
    if (f < -1.0f) f = -1.0f;
    cos_cone[r] = f;
  }
}

/* TrySwapingTemperatures, MonteCarlo.cu:105-172. The MKL uniforms become
 * Philox draws (counter = pair index, kind = swap, round), and the gsl sort by
 * descending temperature is kept incrementally in replica_at[]. */
int orc_pt_swap(int R, int parity, uint32_t round, uint64_t seed, const double *energy,
                double *temps, float *beta, float *cos_cone, int *replica_at) {
  int n_acc = 0;
  for (int c = parity ? 1 : 0; c + 1 < R; c += 2) {
    int t1 = replica_at[c], t2 = replica_at[c + 1];
    double T1 = temps[t1], T2 = temps[t2];
    double arg = (1.0 / T1 - 1.0 / T2) * (energy[t1] - energy[t2]);
    uint32_t w[4];
    draw((uint32_t)c, 0, 0, KIND_SWAP, 0, round, seed, w);
    double u = u01_53(w[0], w[1]);
    int acc = (arg >= 0.0) || (u < exp(arg));
    if (acc) {
      temps[t1] = T2;
      temps[t2] = T1;
      float b = beta[t1]; beta[t1] = beta[t2]; beta[t2] = b;
      float cc = cos_cone[t1]; cos_cone[t1] = cos_cone[t2]; cos_cone[t2] = cc;
      replica_at[c] = t2;
      replica_at[c + 1] = t1;
      n_acc++;
    }
  }
  return n_acc;
}

/* Mc, Pyrochlore.cu:74-133, plus the parallel-tempering cadence of
 * ParallelTempering (MonteCarlo.cu:429-470). Deviations: no 1000-sweep cap
 * (Pyrochlore.cu:83,131); by default colours in fixed order 0..3 rather than
 * a set index read from an unfilled buffer (Pyrochlore.cu:84,848), drawn sets
 * with orc_set_update_schedule. */
void orc_run(float *spins, int L, int R, int n_sweeps, uint32_t sweep0, int n_metro, int n_or,
             int pt_every, uint32_t *pt_round, double *temps, float *beta, float *cos_cone,
             int *replica_at, float jeff, uint64_t seed, uint64_t *accepted, double *dE,
             double *anchor, int *n_swaps_accepted) {
  double *obs = NULL;
  double *en = NULL;
  if (pt_every > 0) {
    obs = (double *)malloc(sizeof(double) * 13 * (size_t)R);
    en = (double *)malloc(sizeof(double) * (size_t)R);
  }
  for (int i = 0; i < n_sweeps; i++) {
    uint32_t sweep = sweep0 + (uint32_t)i;
    for (int p = 0; p < n_metro; p++)
      for (int c = 0; c < 4; c++)
        orc_metropolis_phase(spins, L, R, c, sweep, p, beta, cos_cone, jeff, seed, accepted, dE);
    for (int o = 0; o < n_or; o++)
      for (int c = 0; c < 4; c++) orc_overrelax_phase_ex(spins, L, R, c, sweep, o, jeff, seed, dE);
    if (pt_every > 0 && (sweep + 1) % (uint32_t)pt_every == 0) {
      orc_observables(spins, L, R, jeff, obs);
      for (int r = 0; r < R; r++) en[r] = obs[13 * r];
      if (anchor && dE)
        for (int r = 0; r < R; r++) {
          anchor[r] = en[r];
          dE[r] = 0.0;
        }
      int a = orc_pt_swap(R, (int)(*pt_round & 1u), *pt_round, seed, en, temps, beta, cos_cone,
                          replica_at);
      if (n_swaps_accepted) *n_swaps_accepted += a;
      (*pt_round)++;
    }
  }
  free(obs);
  free(en);
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}
