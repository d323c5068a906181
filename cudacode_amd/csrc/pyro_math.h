// Shared host/device arithmetic for the pyrochlore Heisenberg Monte Carlo step.
//
// Everything here is compiled for gfx950 and for the host by hipcc with
// -ffp-contract=off, so the same float operations run in the same order on
// both sides. There is no device math library call on the update path: the
// exponential and sin/cos(pi x) are written out below from published methods,
// and sqrt/division are the IEEE correctly rounded ones. That is what lets a
// CPU restatement reproduce a GPU trajectory bit for bit.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PYRO_HD __host__ __device__ __forceinline__
#else
#define PYRO_HD static inline
#endif

namespace pyro {

// ---------------------------------------------------------------------------
// Lattice tables
// ---------------------------------------------------------------------------
// 16-site conventional cubic cell of the pyrochlore lattice. Site s = 4*t + b:
// b = s % 4 is the basis site of the primitive tetrahedron, t the FCC
// translation. Each site has 6 nearest neighbours, given as the cell offset
// (dx, dy, dz) in {-1,0,1} and the neighbour's site index. The table is the
// neighbour list of Pyrochlore.cu:152-167 (links16); the offsets are applied
// per axis (x+dx, y+dy, z+dz, periodic), as the CPU TotalEnergy does at
// Pyrochlore.cu:476-496.
constexpr int kSitesPerCell = 16;
constexpr int kNeighbours = 6;
constexpr int kColours = 4;          // colour = site % 4, an independent set
constexpr int kSitesPerColour = 4;   // sites c, c+4, c+8, c+12 of a cell

struct Link { int8_t dx, dy, dz, site; };

#define PYRO_LINKS16_INIT                                                    \
  {{{0, 0, 0, 1}, {0, 0, 0, 2}, {0, 0, 0, 3}, {0, -1, -1, 10}, {-1, -1, 0, 5}, {-1, 0, -1, 15}}, \
   {{0, 0, 0, 0}, {0, 0, 0, 2}, {0, 0, 0, 3}, {0, 0, 0, 4}, {0, 0, -1, 14}, {0, 0, -1, 11}},     \
   {{0, 0, 0, 0}, {0, 0, 0, 1}, {0, 0, 0, 3}, {0, 0, 0, 8}, {-1, 0, 0, 13}, {-1, 0, 0, 7}},      \
   {{0, 0, 0, 0}, {0, 0, 0, 1}, {0, 0, 0, 2}, {0, -1, 0, 9}, {0, -1, 0, 6}, {0, 0, 0, 12}},      \
   {{0, 0, 0, 5}, {0, 0, 0, 6}, {0, 0, 0, 7}, {0, 0, 0, 1}, {0, 0, -1, 14}, {0, 0, -1, 11}},     \
   {{0, 0, 0, 4}, {0, 0, 0, 6}, {0, 0, 0, 7}, {1, 1, 0, 0}, {1, 0, -1, 10}, {0, 1, -1, 15}},     \
   {{0, 0, 0, 4}, {0, 0, 0, 5}, {0, 0, 0, 7}, {0, 1, 0, 3}, {0, 1, 0, 12}, {0, 0, 0, 9}},        \
   {{0, 0, 0, 4}, {0, 0, 0, 5}, {0, 0, 0, 6}, {1, 0, 0, 2}, {1, 0, 0, 8}, {0, 0, 0, 13}},        \
   {{0, 0, 0, 9}, {0, 0, 0, 10}, {0, 0, 0, 11}, {-1, 0, 0, 7}, {0, 0, 0, 2}, {-1, 0, 0, 13}},    \
   {{0, 0, 0, 8}, {0, 0, 0, 10}, {0, 0, 0, 11}, {0, 1, 0, 3}, {0, 0, 0, 6}, {0, 1, 0, 12}},      \
   {{0, 0, 0, 8}, {0, 0, 0, 9}, {0, 0, 0, 11}, {0, 1, 1, 0}, {-1, 0, 1, 5}, {-1, 1, 0, 15}},     \
   {{0, 0, 0, 8}, {0, 0, 0, 9}, {0, 0, 0, 10}, {0, 0, 1, 1}, {0, 0, 1, 4}, {0, 0, 0, 14}},       \
   {{0, 0, 0, 13}, {0, 0, 0, 14}, {0, 0, 0, 15}, {0, 0, 0, 3}, {0, -1, 0, 6}, {0, -1, 0, 9}},    \
   {{0, 0, 0, 12}, {0, 0, 0, 14}, {0, 0, 0, 15}, {1, 0, 0, 2}, {1, 0, 0, 8}, {0, 0, 0, 7}},      \
   {{0, 0, 0, 12}, {0, 0, 0, 13}, {0, 0, 0, 15}, {0, 0, 0, 11}, {0, 0, 1, 1}, {0, 0, 1, 4}},     \
   {{0, 0, 0, 12}, {0, 0, 0, 13}, {0, 0, 0, 14}, {1, 0, 1, 0}, {0, -1, 1, 5}, {1, -1, 0, 10}}}

// The 24 sets of 4 mutually non-adjacent sites of a cell (one site per
// tetrahedron), Pyrochlore.cu:36-39. The sweep uses sets {c, c+4, c+8, c+12}.
#define PYRO_INDEPENDENT_SETS_INIT                                                       \
  {{0, 4, 8, 12}, {0, 14, 8, 6}, {0, 14, 7, 9}, {0, 11, 13, 6}, {0, 11, 7, 12},          \
   {0, 4, 13, 9}, {1, 5, 13, 9}, {1, 15, 8, 6}, {1, 15, 7, 9}, {1, 10, 13, 6},           \
   {1, 10, 7, 12}, {1, 5, 8, 12}, {2, 10, 14, 6}, {2, 15, 11, 6}, {2, 15, 4, 9},         \
   {2, 10, 4, 12}, {2, 5, 14, 9}, {2, 5, 11, 12}, {3, 15, 11, 7}, {3, 15, 4, 8},         \
   {3, 10, 14, 7}, {3, 10, 4, 13}, {3, 5, 14, 8}, {3, 5, 11, 13}}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11): 10 rounds of a 32x32->64 multiply bijection with a
// Weyl-sequence key schedule.
// ---------------------------------------------------------------------------
struct u32x4 { uint32_t x, y, z, w; };

PYRO_HD void philox_mulhilo(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

PYRO_HD u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    philox_mulhilo(M0, c.x, &hi0, &lo0);
    philox_mulhilo(M1, c.z, &hi1, &lo1);
    u32x4 n;
    n.x = hi1 ^ c.y ^ k0;
    n.y = lo1;
    n.z = hi0 ^ c.w ^ k1;
    n.w = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// Uniform float in [0, 1) from the top 24 bits: exact in FP32.
PYRO_HD float u01(uint32_t x) { return (float)(x >> 8) * 5.9604644775390625e-8f; }

// Uniform double in [0, 1) from 53 bits of two words.
PYRO_HD double u01_53(uint32_t a, uint32_t b) {
  uint64_t m = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
  return (double)m * 1.1102230246251565e-16;
}

// Counter layout. Word 1 packs the site, a pass index and the stream kind, so
// every (replica, cell, site, sweep, pass, kind) draw is a distinct counter.
enum : uint32_t {
  kKindMetropolis = 1u,
  kKindInit = 2u,
  kKindSwap = 3u,
  kKindLearned = 4u,
  kKindLearnedInit = 5u,
  kKindOrAngle = 6u,  // angle of the random-angle over-relaxation
  kKindSets = 7u,     // independent set of a phase under the drawn-set schedule
};
PYRO_HD u32x4 rng_counter(uint32_t cell, uint32_t site, uint32_t pass, uint32_t kind,
                          uint32_t replica, uint32_t sweep) {
  u32x4 c;
  c.x = cell;
  c.y = site | (pass << 8) | (kind << 24);
  c.z = replica;
  c.w = sweep;
  return c;
}

// ---------------------------------------------------------------------------
// exp(x) for x <= 0, float. Method: Cody-Waite argument reduction
// x = k*ln2 + r with ln2 split into a 9-bit head (so k*head is exact for
// |k| <= 127) and a tail, as in Cephes expf (S. L. Moshier, "Methods and
// Programs for Mathematical Functions", 1989); exp(r), |r| <= ln2/2, by its
// degree-7 Taylor polynomial (truncation error < 6e-9 relative, below half
// an FP32 ulp); scaling by 2^k through the exponent field. Inputs below -86
// return 0 so 2^k stays a normal number on every target.
// ---------------------------------------------------------------------------
PYRO_HD float exp_neg(float x) {
  if (!(x > -86.0f)) return 0.0f;
  if (x > 0.0f) x = 0.0f;
  const float kLog2e = 1.44269502162933349609375f;
  const float kLn2Head = 0.693359375f;
  const float kLn2Tail = -2.12194440e-4f;
  float kf = x * kLog2e + 0.5f;
  kf = (float)(int)kf - (((float)(int)kf > kf) ? 1.0f : 0.0f);  // floor
  float r = x - kf * kLn2Head;
  r = r - kf * kLn2Tail;
  float p = 1.98412698e-4f;      // 1/5040
  p = p * r + 1.38888889e-3f;    // 1/720
  p = p * r + 8.33333333e-3f;    // 1/120
  p = p * r + 4.16666667e-2f;    // 1/24
  p = p * r + 1.66666667e-1f;    // 1/6
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  int k = (int)kf;
  union { uint32_t u; float f; } s;
  s.u = (uint32_t)(k + 127) << 23;
  return p * s.f;
}

// sin(pi*a), cos(pi*a) for a in [0, 2]. Reduction to a = n/2 + f with
// |f| <= 1/4 is exact in FP32; sin and cos of pi*f by Taylor polynomials of
// degree 9 and 10 (truncation < 2e-9 on |pi*f| <= pi/4); quadrant fix-up by
// the angle-addition identities for multiples of pi/2.
PYRO_HD void sincospi(float a, float* s, float* c) {
  float nf = a * 2.0f + 0.5f;
  nf = (float)(int)nf;  // a >= 0: truncation is floor
  float f = a - nf * 0.5f;
  float t = f * 3.14159274101257324f;
  float t2 = t * t;
  float sp = 2.75573192e-6f;       // 1/9!
  sp = sp * t2 - 1.98412698e-4f;   // 1/7!
  sp = sp * t2 + 8.33333333e-3f;   // 1/5!
  sp = sp * t2 - 1.66666667e-1f;   // 1/3!
  sp = sp * t2;
  sp = sp * t + t;
  float cp = -2.75573192e-7f;      // 1/10!
  cp = cp * t2 + 2.48015873e-5f;   // 1/8!
  cp = cp * t2 - 1.38888889e-3f;   // 1/6!
  cp = cp * t2 + 4.16666667e-2f;   // 1/4!
  cp = cp * t2 - 0.5f;
  cp = cp * t2 + 1.0f;
  int n = ((int)nf) & 3;
  float so = sp, co = cp;
  if (n == 1) { so = cp; co = -sp; }
  else if (n == 2) { so = -sp; co = -cp; }
  else if (n == 3) { so = -cp; co = sp; }
  *s = so;
  *c = co;
}

// ---------------------------------------------------------------------------
// Spin helpers
// ---------------------------------------------------------------------------
struct f3 { float x, y, z; };

PYRO_HD float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

PYRO_HD f3 normalise(f3 v) {
  float n = 1.0f / sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
  f3 o;
  o.x = v.x * n;
  o.y = v.y * n;
  o.z = v.z * n;
  return o;
}

// Renormalisation of a vector whose squared length q is within 1e-4 of 1:
// every caller's vector is a reflection or rotation of a unit vector, off by
// FP32 rounding only (|q - 1| of order 1e-7). One Newton step for 1/sqrt(q)
// from 1 gives the factor 1 + c with c = (1 - q)/2, whose error 3/8 (q - 1)^2
// is below 4e-9, far under an FP32 ulp. Both 1 - q (Sterbenz) and the halving
// are exact, and each component is scaled as v + v c, so its only rounding of
// consequence is the last one: the length comes out unbiased to about 3e-9
// (a factor 1 + c rounded to the float grid next to 1 would be coarser above
// 1 than below and bias it), in four IEEE operations and three per component
// instead of a correctly rounded square root and divide. Deterministic, so
// host and device round identically.
PYRO_HD f3 renormalise(f3 v) {
  float q = v.x * v.x + v.y * v.y + v.z * v.z;
  float c = 0.5f * (1.0f - q);
  f3 o;
  o.x = v.x + v.x * c;
  o.y = v.y + v.y * c;
  o.z = v.z + v.z * c;
  return o;
}

// Uniform point on the unit sphere from two uniforms: azimuth 2*pi*u_phi,
// cos(theta) = 2*u_cos - 1 (Math.cu:120-130).
PYRO_HD f3 uniform_sphere(float u_phi, float u_cos) {
  float ps, pc;
  sincospi(2.0f * u_phi, &ps, &pc);
  float tc = 2.0f * u_cos - 1.0f;
  float ts = sqrtf(fmaxf(0.0f, 1.0f - tc * tc));
  f3 o;
  o.x = pc * ts;
  o.y = ps * ts;
  o.z = tc;
  return o;
}

// Uniform point on the spherical cap {v : v.s >= cos_cone} around the unit
// vector s (Math.cu:244-281). A point r is drawn on the cap around +z (or -z
// when s.z < 0) and carried onto s by the Householder reflection that maps
// +z (or -z) to s; choosing the pole on s's side keeps the reflection's
// denominator 1 + |s.z| >= 1. The proposal density depends only on the angle
// between s and s', so it is symmetric and plain Metropolis applies.
PYRO_HD f3 cone_proposal(f3 s, float cos_cone, float u_cos, float u_phi) {
  float tc = 1.0f - (1.0f - cos_cone) * u_cos;
  float ts = sqrtf(fmaxf(0.0f, 1.0f - tc * tc));
  float ps, pc;
  sincospi(2.0f * u_phi, &ps, &pc);
  const bool up = s.z >= 0.0f;
  f3 r;
  r.x = ts * pc;
  r.y = ts * ps;
  r.z = up ? tc : -tc;
  f3 N;
  N.x = s.x;
  N.y = s.y;
  N.z = up ? s.z + 1.0f : s.z - 1.0f;
  float d = up ? 1.0f + s.z : 1.0f - s.z;
  float scal = (N.x * r.x + N.y * r.y + N.z * r.z) / d;
  f3 o;
  o.x = N.x * scal - r.x;
  o.y = N.y * scal - r.y;
  o.z = N.z * scal - r.z;
  return renormalise(o);
}

// Metropolis test (Math.cu:968-986): local energies -h.s before and after,
// accept iff u < min(1, exp(beta * (E_before - E_after))).
PYRO_HD bool metropolis_accept(f3 h, f3 s, f3 sp, float beta, float u) {
  float e_before = -(h.x * s.x + h.y * s.y + h.z * s.z);
  float e_after = -(h.x * sp.x + h.y * sp.y + h.z * sp.z);
  float x = (e_before - e_after) * beta;
  return (x >= 0.0f) | (u < exp_neg(x));
}

// Over-relaxation (Math.cu:220-242): reflect s through the axis of h,
// s' = 2 (s.h) h / |h|^2 - s, then renormalise. Leaves s unchanged when the
// local field vanishes exactly. The result is invariant under h -> a*h, so the
// raw neighbour sum can be passed for h.
PYRO_HD f3 overrelax(f3 h, f3 s) {
  float hh = h.x * h.x + h.y * h.y + h.z * h.z;
  if (!(hh > 0.0f)) return s;
  float scal = (s.x * h.x + s.y * h.y + s.z * h.z) / hh;
  float t = 2.0f * scal;
  f3 v;
  v.x = t * h.x - s.x;
  v.y = t * h.y - s.y;
  v.z = t * h.z - s.z;
  return renormalise(v);
}

// Rotation of s about the local field h by the angle 2 pi u (Rodrigues'
// formula), renormalised. The optional move after each Metropolis test
// (_CudaRotate3DVector, Math.cu:168-197, applied at Pyrochlore.cu:872): it
// keeps h.s, so the local energy, and any rotation angle preserves the
// uniform measure on the circle about h. The reference's angle argument
// works out to 2 pi^2 u (its half angle is passed in units of pi); here it
// is uniform on [0, 2 pi). Unchanged when h = 0.
PYRO_HD f3 rotate_about(f3 h, f3 s, float u) {
  float hh = h.x * h.x + h.y * h.y + h.z * h.z;
  if (!(hh > 0.0f)) return s;
  float inv = 1.0f / sqrtf(hh);
  f3 n = {h.x * inv, h.y * inv, h.z * inv};
  float st, ct;
  sincospi(2.0f * u, &st, &ct);
  float ns = n.x * s.x + n.y * s.y + n.z * s.z;
  float k = ns * (1.0f - ct);
  f3 v;
  v.x = (s.x * ct + (n.y * s.z - n.z * s.y) * st) + n.x * k;
  v.y = (s.y * ct + (n.z * s.x - n.x * s.z) * st) + n.y * k;
  v.z = (s.z * ct + (n.x * s.y - n.y * s.x) * st) + n.z * k;
  return renormalise(v);
}

// The 12 order-parameter components of Measurement (Pyrochlore.cu:580-624)
// from the sublattice sums s[3 b + c] (b = basis site, c = component),
// divided by n_spins: m[0] A2 (all-in-all-out), m[1..2] E, m[3..5] T1,
// m[6..8] ferromagnetic (the magnetisation per spin), m[9..11] T2. With this
// normalisation the ideal all-in-all-out state has |m_A2| = 1.
PYRO_HD void fold_order_params(const double *s, double n_spins, double *m) {
  const double a = 0.5773502691896258, b = 0.816496580927726, c = 0.4082482904638631,
               d = 0.7071067811865475;
  const double *s0 = s, *s1 = s + 3, *s2 = s + 6, *s3 = s + 9;
  m[0] = a * (s0[0] + s0[1] + s0[2]) + a * (-s1[0] - s1[1] + s1[2]) +
         a * (s2[0] - s2[1] - s2[2]) + a * (-s3[0] + s3[1] - s3[2]);
  m[1] = (-b * s0[0] + c * s0[1] + c * s0[2]) + (b * s1[0] - c * s1[1] + c * s1[2]) +
         (-b * s2[0] - c * s2[1] - c * s2[2]) + (b * s3[0] + c * s3[1] - c * s3[2]);
  m[2] = d * (-s0[1] + s0[2] + s1[1] + s1[2] + s2[1] - s2[2] - s3[1] - s3[2]);
  m[3] = d * (-s0[1] + s0[2] - s1[1] - s1[2] + s2[1] - s2[2] + s3[1] + s3[2]);
  m[4] = d * (s0[0] - s0[2] + s1[0] + s1[2] - s2[0] - s2[2] - s3[0] + s3[2]);
  m[5] = d * (-s0[0] + s0[1] + s1[0] - s1[1] + s2[0] + s2[1] - s3[0] - s3[1]);
  m[6] = s0[0] + s1[0] + s2[0] + s3[0];
  m[7] = s0[1] + s1[1] + s2[1] + s3[1];
  m[8] = s0[2] + s1[2] + s2[2] + s3[2];
  m[9] = d * (-s0[1] - s0[2] - s1[1] + s1[2] + s2[1] + s2[2] + s3[1] - s3[2]);
  m[10] = d * (-s0[0] - s0[2] - s1[0] + s1[2] + s2[0] - s2[2] + s3[0] + s3[2]);
  m[11] = d * (-s0[0] - s0[1] + s1[0] + s1[1] + s2[0] - s2[1] - s3[0] + s3[1]);
  for (int q = 0; q < 12; ++q) m[q] = m[q] / n_spins;
}

}  // namespace pyro
