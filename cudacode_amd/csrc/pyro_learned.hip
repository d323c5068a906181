// Learned local proposal: a 6 -> 64 -> 64 -> 4 ReLU network maps the local
// environment (u = beta*h, s) to a von Mises-Fisher proposal vMF(mu, kappa)
// for the new spin; an exact Metropolis-Hastings test keeps the Boltzmann
// distribution stationary. No reference counterpart (the reference's moves are
// Pyrochlore.cu:829-885 and 975-1007); this file is the build's addition.
//
// Inference runs on BF16 MFMA with FP32 accumulation. The network is evaluated
// transposed, D = W * X^T (rows = units, columns = the 64 samples of a wave),
// with v_mfma_f32_32x32x8_bf16: the accumulator layout of one layer (lane l
// holds column l%32, rows 8b + 4(l/32) + r) is exactly the B-operand layout
// of the next layer's 8-wide k-steps, so activations never leave registers.
// Training (forward, backward, Adam) runs in FP32 on the vector units.
#include <cmath>
#include <cstring>
#include <hip/hip_ext.h>

#include "pyro_internal.h"

namespace pyro {

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4v __attribute__((ext_vector_type(4)));

enum : uint32_t { kKindTrain = 6u, kKindEnv = 7u };

__constant__ Link c_links_learned[16][6] = PYRO_LINKS16_INIT;

// Round to nearest even, as the hardware conversion (v_cvt_pk_bf16_f32) does.
__host__ __device__ __forceinline__ uint32_t bf16_bits(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return u >> 16;
}

__device__ __forceinline__ uint32_t cvt2(float lo, float hi) {
  const f32x2 v = {lo, hi};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

__device__ __forceinline__ float relu(float x) { return fmaxf(x, 0.0f); }

// Exact power-of-two headroom of the hidden pre-activations: the packed
// layer-1 operand (weights and bias column) and the layer-2 bias are scaled by
// 2^-kHeadroomLog2 and the packed layer-3 weights by 2^kHeadroomLog2 (layer
// 2's weights are unscaled: its input and output carry the same factor).
// Scaling by a power of two is exact in FP32 and BF16, so every activation is
// the unscaled one times 2^-16, and the network's output is unchanged, as
// long as the hidden pre-activations stay below 2^16 (the trained network's
// are O(10)); beyond that they saturate, which changes the proposal but not
// the exactness of the Metropolis-Hastings test (both directions evaluate the
// same function).
constexpr int kHeadroomLog2 = 16;
constexpr float kHeadroomDown = 1.0f / (float)(1 << kHeadroomLog2);
constexpr float kHeadroomUp = (float)(1 << kHeadroomLog2);

// ReLU and BF16 rounding of two scaled pre-activations in one instruction:
// v_cvt_pk_bf16_f32 with the clamp modifier returns min(max(bf16(x), 0), 1)
// (checked on gfx950 over negative, zero, denormal and large inputs), which
// is ReLU for the scaled values below 1.
//
// The compiler's hazard recognizer does not look inside inline asm, so it
// inserts no wait states between this instruction and the MFMAs around it.
// Three guards supply them (gfx950's rules, as the compiler applies them to
// its own instructions):
//  - a read of an MFMA result needs 12 states: the accumulator is read only
//    after mfma_read_guard;
//  - a write into the destination of an MFMA in flight needs 12 states: the
//    result register is one the compiler sees as free, which excludes every
//    live accumulator register but not the unread rows of an accumulator
//    (layer 3 reads 4 of its 32), so those are kept live by the same guard;
//  - an MFMA reading the result needs 2 states: mfma_operand_guard.
// tests/test_learned_isa.py checks these distances in the built assembly.
__device__ __forceinline__ uint32_t relu_cvt2(float lo, float hi) {
  uint32_t r;
  asm("v_cvt_pk_bf16_f32 %0, %1, %2 clamp" : "=v"(r) : "v"(lo), "v"(hi));
  return r;
}

// Placed after the last MFMA of an accumulator chain: "updates" the whole
// accumulator (ordering every later read after it and keeping all of its
// registers allocated) and waits the 13 states that a VALU access to a 32x32
// MFMA result needs on gfx950 (12, as the compiler itself inserts).
__device__ __forceinline__ void mfma_read_guard(f32x16 &d) { asm volatile("s_nop 12" : "+v"(d)); }

// Placed after the conversions that build an MFMA operand: a VALU write of a
// VGPR that an MFMA reads as SrcA/B/C needs 2 wait states on gfx950 (the
// compiler's own s_nop 1 for its conversions).
__device__ __forceinline__ void mfma_operand_guard(u32x4v &p) { asm volatile("s_nop 1" : "+v"(p)); }

// ReLU of 8 consecutive (guarded) accumulator registers, converted to the 8
// bf16 of one B operand of a 32x32x16 MFMA: four instructions.
__device__ __forceinline__ bf16x8 act8(const f32x16 &d, int base) {
  u32x4v p;
#pragma unroll
  for (int j = 0; j < 4; ++j) p[j] = relu_cvt2(d[base + 2 * j], d[base + 2 * j + 1]);
  mfma_operand_guard(p);
  return __builtin_bit_cast(bf16x8, p);
}

// Hidden unit that 32x32x16 k-step t, lane half h, element j stands for. The
// accumulator of a 32x32 MFMA holds, in lane half h, rows 8b + 4h + r of its
// 32-row tile (register 4b + r); reading 8 consecutive registers 8t'..8t'+7
// as the next layer's B operand therefore feeds the units
//   32 (t / 2) + 8 (2 (t % 2) + j / 4) + 4 h + j % 4,
// and the A operand (weights) is packed with its columns in the same order,
// which leaves the contraction unchanged.
__host__ __device__ __forceinline__ int kstep_unit(int t, int h, int j) {
  return 32 * (t >> 1) + 8 * (2 * (t & 1) + (j >> 2)) + 4 * h + (j & 3);
}

// Weight fragments held in registers by every wave.
struct MlpW {
  s16x4 a1[2];     // layer 1 (32x32x8): [row tile]
  bf16x8 a2[2][4]; // layer 2 (32x32x16): [row tile][k step]
  bf16x8 a3[4];    // layer 3 (32x32x16, rows 4..31 zero): [k step]
  f32x16 b2[2];    // bias of the rows this lane's accumulators hold: the
                   // layer-2 accumulators start from it, so the MFMA adds it
  float b3[4];
};

__device__ __forceinline__ void mlp_load(const uint16_t *__restrict__ pack,
                                         const float *__restrict__ w, MlpW &W) {
  const int lane = threadIdx.x & 63, h = lane >> 5, n = lane & 31;
#pragma unroll
  for (int tj = 0; tj < 2; ++tj)
    W.a1[tj] = __builtin_bit_cast(
        s16x4, *reinterpret_cast<const u32x2 *>(pack + kPackP1 + (32 * tj + n) * 8 + 4 * h));
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int t = 0; t < 4; ++t)
      W.a2[ti][t] = __builtin_bit_cast(
          bf16x8, *reinterpret_cast<const u32x4v *>(pack + kPackP2 + ((ti * 4 + t) * 64 + lane) * 8));
#pragma unroll
  for (int t = 0; t < 4; ++t)
    W.a3[t] = __builtin_bit_cast(
        bf16x8, *reinterpret_cast<const u32x4v *>(pack + kPackP3 + (t * 64 + lane) * 8));
  // b2 of the rows this lane's accumulators hold, 4 consecutive rows a load
  const float4 *b2 = reinterpret_cast<const float4 *>(pack + kPackB2);
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int blk = 0; blk < 4; ++blk) {
      const float4 q = b2[8 * ti + 2 * blk + h];  // rows 32 ti + 8 blk + 4 h + (0..3)
      W.b2[ti][4 * blk + 0] = q.x;
      W.b2[ti][4 * blk + 1] = q.y;
      W.b2[ti][4 * blk + 2] = q.z;
      W.b2[ti][4 * blk + 3] = q.w;
    }
#pragma unroll
  for (int q = 0; q < 4; ++q) W.b3[q] = w[kOffB3 + q];
}

// Forward pass of the 64 samples of a wave (one per lane). Every lane of the
// wave must call it. x: this lane's 6 features; o: this lane's 4 outputs.
// 4 MFMA 32x32x8 (layer 1, bias folded as a constant input) + 16 + 8 MFMA
// 32x32x16 (layers 2 and 3).
__device__ __forceinline__ void mlp_forward(const MlpW &W, const float x[6], float o[4]) {
  const int lane = threadIdx.x & 63, h = lane >> 5, n = lane & 31;
  const uint32_t p0 = cvt2(x[0], x[1]), p1 = cvt2(x[2], x[3]), p2 = cvt2(x[4], x[5]);
  const uint32_t p3 = 0x3F80u;  // bf16(1.0) low, 0 high: the bias column
  s16x4 b1[2];
#pragma unroll
  for (int tn = 0; tn < 2; ++tn) {
    const int src = 32 * tn + n;
    const uint32_t q0 = __shfl(p0, src), q1 = __shfl(p1, src), q2 = __shfl(p2, src);
    const u32x2 v = h ? u32x2{q2, p3} : u32x2{q0, q1};
    b1[tn] = __builtin_bit_cast(s16x4, v);
  }
  const f32x16 zero = {};
  bf16x8 b2[2][4];
#pragma unroll
  for (int tj = 0; tj < 2; ++tj)
#pragma unroll
    for (int tn = 0; tn < 2; ++tn) {
      f32x16 d = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(W.a1[tj], b1[tn], zero, 0, 0, 0);
      mfma_read_guard(d);
      b2[tn][2 * tj] = act8(d, 0);
      b2[tn][2 * tj + 1] = act8(d, 8);
    }
  bf16x8 b3[2][4];
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tn = 0; tn < 2; ++tn) {
      f32x16 acc = W.b2[ti];
#pragma unroll
      for (int t = 0; t < 4; ++t)
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W.a2[ti][t], b2[tn][t], acc, 0, 0, 0);
      mfma_read_guard(acc);
      b3[tn][2 * ti] = act8(acc, 0);
      b3[tn][2 * ti + 1] = act8(acc, 8);
    }
  f32x16 d3[2];
#pragma unroll
  for (int tn = 0; tn < 2; ++tn) {
    f32x16 acc = zero;
#pragma unroll
    for (int t = 0; t < 4; ++t)
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W.a3[t], b3[tn][t], acc, 0, 0, 0);
    mfma_read_guard(acc);
    d3[tn] = acc;
  }
  // Rows 0..3 of column n sit in lane n (upper half zero-padded rows elsewhere).
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float other = __shfl(d3[1][q], n);
    o[q] = (h ? other : d3[0][q]) + W.b3[q];
  }
}

// Hardware reciprocal, reciprocal square root, 2^x and log2 (about 1 ulp).
// The proposal's densities need not round like the host code: the network
// output is whatever it is, and the sampler and its density only have to
// describe the same law to float accuracy.
__device__ __forceinline__ float rcp_f(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float rsq_f(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float exp_f(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504089f); }
__device__ __forceinline__ float log_f(float x) { return __builtin_amdgcn_logf(x) * 0.693147180560f; }

// log(1 + x) for x in [-1, 0] from the hardware log, as log(u) x / (u - 1)
// with u = 1 + x rounded (D. Goldberg, "What every computer scientist should
// know about floating-point arithmetic", 1991, theorem 4): a few ulps where
// the library's log1pf costs about a hundred instructions; x at u = 1;
// -inf at x = -1 like log1p.
__device__ __forceinline__ float log1p_f(float x) {
  const float u = 1.0f + x;
  return u == 1.0f ? x : log_f(u) * (x * rcp_f(u - 1.0f));
}

// expm1(-2 kappa) for kappa >= 0: Taylor to the 7th power below kappa = 1/4
// (truncation < 3e-7 relative), exp(-2 kappa) - 1 above (no cancellation:
// the result is below -0.39).
__device__ __forceinline__ float expm1_neg2k(float kappa) {
  const float x = -2.0f * kappa;
  if (kappa > 0.25f) return exp_f(x) - 1.0f;
  float p = 1.0f / 5040.0f;
  p = p * x + 1.0f / 720.0f;
  p = p * x + 1.0f / 120.0f;
  p = p * x + 1.0f / 24.0f;
  p = p * x + 1.0f / 6.0f;
  p = p * x + 0.5f;
  p = p * x + 1.0f;
  return p * x;
}

__device__ __forceinline__ void features(f3 u, f3 s, float x[6]) {
  const float r2 = u.x * u.x + u.y * u.y + u.z * u.z;
  const float ir = rsq_f(r2);
  const float g = r2 > 0.0f ? log_f(1.0f + r2 * ir) * ir : 1.0f;
  x[0] = u.x * g;
  x[1] = u.y * g;
  x[2] = u.z * g;
  x[3] = s.x;
  x[4] = s.y;
  x[5] = s.z;
}

struct Vmf { f3 mu; float kappa, em2k, log_norm; };  // em2k = expm1(-2 kappa)

__device__ __forceinline__ Vmf decode(const float o[4]) {
  Vmf v;
  const float n2 = o[0] * o[0] + o[1] * o[1] + o[2] * o[2];
  if (n2 > 0.0f) {
    const float in = rsq_f(n2);
    v.mu = {o[0] * in, o[1] * in, o[2] * in};
  } else {
    v.mu = {0.0f, 0.0f, 1.0f};
  }
  const float lk = fminf(fmaxf(o[3], kLogKappaMin), kLogKappaMax);
  v.kappa = exp_f(lk);
  // log of kappa / (2 pi (1 - exp(-2 kappa))); the exp(-kappa) factor of the
  // normaliser is carried by the -kappa |mu - y|^2 / 2 term of the density.
  // expm1 without cancellation: for small kappa it is the normaliser.
  v.em2k = expm1_neg2k(v.kappa);
  v.log_norm = lk - 1.83787706641f - log_f(-v.em2k);
  return v;
}

// log q(y) = log C'(kappa) - kappa |mu - y|^2 / 2: the vMF log density written
// so that y near mu loses no precision.
__device__ __forceinline__ float vmf_log_density(const Vmf &v, f3 y) {
  const float dx = v.mu.x - y.x, dy = v.mu.y - y.y, dz = v.mu.z - y.z;
  return v.log_norm - 0.5f * v.kappa * (dx * dx + dy * dy + dz * dz);
}

// vMF draw: cosine to the axis by the inverse CDF (Wood 1994), uniform
// azimuth, carried onto mu by the Householder reflection through the pole on
// mu's side.
__device__ __forceinline__ f3 vmf_sample(const Vmf &v, float u1, float u2) {
  const float k = v.kappa;
  const float p = 1.0f - u1;  // (0, 1]
  float w = 1.0f + log1p_f((1.0f - p) * v.em2k) * rcp_f(k);
  w = fminf(1.0f, fmaxf(-1.0f, w));
  const float st = sqrtf(fmaxf(0.0f, 1.0f - w * w));
  float ps, pc;
  ps = __builtin_amdgcn_sinf(u2);  // hardware sin/cos take revolutions: sin(2 pi u2)
  pc = __builtin_amdgcn_cosf(u2);
  const f3 m = v.mu;
  const bool up = m.z >= 0.0f;
  const f3 r = {st * pc, st * ps, up ? w : -w};
  const f3 N = {m.x, m.y, up ? m.z + 1.0f : m.z - 1.0f};
  const float d = up ? 1.0f + m.z : 1.0f - m.z;
  const float scal = (N.x * r.x + N.y * r.y + N.z * r.z) * rcp_f(d);
  f3 o = {N.x * scal - r.x, N.y * scal - r.y, N.z * scal - r.z};
  // The stored spin is renormalised as every other update does, so that its
  // length has the same (unbiased) statistics: the over-relaxation passes that
  // renormalise it afterwards would otherwise turn any difference into an
  // untracked energy drift (a hardware rsq here did, by an order of
  // magnitude). This vector also carries the hardware sin, cos and reciprocal
  // errors, larger than rounding, so it takes two steps: the first leaves
  // |q - 1| at the square of its deviation.
  return renormalise(renormalise(o));
}

// log of the Metropolis-Hastings ratio for s -> sp at fixed local field h:
// beta h.(sp - s) + log q(s | u, sp) - log q(sp | u, s).
__device__ __forceinline__ float mh_log_ratio(f3 h, float beta, f3 s, f3 sp, const Vmf &fwd,
                                              const Vmf &rev) {
  const float de = (h.x * sp.x + h.y * sp.y + h.z * sp.z) - (h.x * s.x + h.y * s.y + h.z * s.z);
  return beta * de + vmf_log_density(rev, s) - vmf_log_density(fwd, sp);
}

__device__ __forceinline__ int wrap_l(int v, int L) { return v < 0 ? v + L : (v >= L ? v - L : v); }

// Periodic wrap of v in [-1, L]; a mask when L is a compile-time power of two.
template <int LT>
__device__ __forceinline__ int wrap_t(int v, int L) {
  if constexpr (LT > 0 && (LT & (LT - 1)) == 0) return v & (LT - 1);
  return wrap_l(v, L);
}

// One colour phase of learned proposals on every replica; same geometry as
// the Metropolis phase (block = 4 waves = the 4 sites of the colour, lane =
// cell), with G groups of 64 cells per block so the weight fragments loaded
// into registers serve G forward-pass pairs. grid = (ceil(ncell/(64 G)), R).
//
// Each block writes its energy change and acceptance count to its own slot,
// slots[(4 r + colour) * nslots + blockIdx.x]: a plain store on the first
// learned pass of a sweep, a plain read-modify-write (accumulate) on later
// ones. Within a launch no other block touches the slot, and the launches of
// a replica are ordered on its stream. k_learned_fold moves the slots into dE
// and the acceptance counters once per sweep. Device-scope atomics from every
// block (64 per replica at L = 16, eight replicas to a 64-byte line)
// serialise at the memory side: they took 8 of 47 us per L=16 x 64 launch.
template <int LT, int G>
__global__ __launch_bounds__(kPhaseBlock, 3) void k_learned_phase(
    float *__restrict__ spins, int L_rt, int ncell_rt, int colour, int pass,
    const float *__restrict__ beta, const uint32_t *__restrict__ ctl, uint32_t key0,
    uint32_t key1, float jeff, const uint16_t *__restrict__ pack, const float *__restrict__ w,
    LearnedSlot *__restrict__ slots, int nslots, int accumulate, int r0, uint32_t sweep_delta) {
  const int L = LT > 0 ? LT : L_rt;
  const int ncell = LT > 0 ? LT * LT * LT : ncell_rt;
  const int r = r0 + (int)blockIdx.y;
  const int k = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int site = colour + 4 * k;
  const size_t plane = (size_t)16 * ncell;
  float *Sx = spins + (size_t)r * 3 * plane;
  float *Sy = Sx + plane;
  float *Sz = Sy + plane;
  const float bt = beta[r];
  const uint32_t sweep = ctl[kCtlSweep] + sweep_delta;

  MlpW W;
  mlp_load(pack, w, W);

  // Energy change of this lane's accepted proposals in FP32 (from the FP32
  // neighbour sum the move itself uses: about 1e-7 of |s' - s| |h| per move,
  // unbiased); summed over the block in FP64.
  float de_f = 0.0f;
  unsigned nacc = 0;
  for (int g = 0; g < G; ++g) {
    const int cell = (blockIdx.x * G + g) * kPhaseCells + lane;
    const bool valid = cell < ncell;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    f3 s = {0.0f, 0.0f, 1.0f};
    uint32_t i = 0;
    if (valid) {
      const int LL = L * L;
      const int z = cell / LL;
      const int rem = cell - z * LL;
      const int y = rem / L;
      const int x = rem - y * L;
#pragma unroll
      for (int nb = 0; nb < 6; ++nb) {
        const Link lk = c_links_learned[site][nb];
        const uint32_t j =
            4u * (uint32_t)(lk.site * ncell + (wrap_t<LT>(z + lk.dz, L) * L + wrap_t<LT>(y + lk.dy, L)) * L +
                            wrap_t<LT>(x + lk.dx, L));
        fx = fx + ld_at(Sx, j);
        fy = fy + ld_at(Sy, j);
        fz = fz + ld_at(Sz, j);
      }
      i = 4u * (uint32_t)(site * ncell + cell);
      s = {ld_at(Sx, i), ld_at(Sy, i), ld_at(Sz, i)};
    }
    const f3 hf = {-jeff * fx, -jeff * fy, -jeff * fz};
    const f3 u = {bt * hf.x, bt * hf.y, bt * hf.z};
    float x[6], o[4];
    features(u, s, x);
    mlp_forward(W, x, o);
    const Vmf fwd = decode(o);
    const u32x4 rnd = philox4x32_10(
        rng_counter((uint32_t)cell, (uint32_t)site, (uint32_t)pass, kKindLearned, (uint32_t)r, sweep),
        key0, key1);
    const f3 sp = vmf_sample(fwd, u01(rnd.x), u01(rnd.y));
    features(u, sp, x);
    mlp_forward(W, x, o);
    const Vmf rev = decode(o);
    const float la = mh_log_ratio(hf, bt, s, sp, fwd, rev);
    const bool accept = valid && (u01(rnd.z) < exp_f(fminf(la, 0.0f)));
    if (accept) {
      de_f = de_f + jeff * (((sp.x - s.x) * fx + (sp.y - s.y) * fy) + (sp.z - s.z) * fz);
      st_at(Sx, i, sp.x);
      st_at(Sy, i, sp.y);
      st_at(Sz, i, sp.z);
    }
    nacc += (unsigned)__popcll(__ballot(accept));
  }
  double de = (double)de_f;
  __shared__ double s_de[4];
  __shared__ unsigned s_acc[4];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) de += __shfl_xor(de, off, 64);
  if (lane == 0) {
    s_de[k] = de;
    s_acc[k] = nacc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    LearnedSlot &sl = slots[((size_t)r * kColours + colour) * nslots + blockIdx.x];
    const double t = (s_de[0] + s_de[1]) + (s_de[2] + s_de[3]);
    const unsigned long long a = (unsigned long long)(s_acc[0] + s_acc[1] + s_acc[2] + s_acc[3]);
    if (accumulate) {
      sl.de = sl.de + t;
      sl.acc = sl.acc + a;
    } else {
      sl.de = t;
      sl.acc = a;
    }
  }
}

// Adds each replica's learned-phase slots (nslots per colour) to dE[r] and
// acc[r] and clears them: one block per replica of the launch range, slots
// summed in a fixed order (so the result does not depend on scheduling), one
// atomic per replica (other streams' kernels may update neighbouring
// replicas' words).
__global__ __launch_bounds__(256) void k_learned_fold(LearnedSlot *__restrict__ slots, int nslots,
                                                      double *__restrict__ dE,
                                                      unsigned long long *__restrict__ acc, int r0) {
  const int r = r0 + (int)blockIdx.x;
  LearnedSlot *s = slots + (size_t)r * kColours * nslots;
  nslots *= kColours;
  double de = 0.0;
  unsigned long long a = 0;
  for (int i = threadIdx.x; i < nslots; i += 256) {
    de += s[i].de;
    a += s[i].acc;
    s[i].de = 0.0;
    s[i].acc = 0;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    de += __shfl_xor(de, off, 64);
    a += __shfl_xor(a, off, 64);
  }
  __shared__ double s_de[4];
  __shared__ unsigned long long s_a[4];
  const int k = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_de[k] = de;
    s_a[k] = a;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double t = (s_de[0] + s_de[1]) + (s_de[2] + s_de[3]);
    if (t != 0.0) atomicAdd(&dE[r], t);
    const unsigned long long n = (s_a[0] + s_a[1]) + (s_a[2] + s_a[3]);
    if (n) atomicAdd(&acc[r], n);
  }
}

// Batched evaluation through the same device code as the phase kernel.
// rows: n x 9 floats (u, s, sp). out: n x 9 floats
// (o(s) [4], o(sp) [4], log MH ratio at h = u, beta = 1). grid = ceil(n/64), block = 64.
__global__ __launch_bounds__(64) void k_mlp_eval(const float *__restrict__ rows, int n,
                                                 const uint16_t *__restrict__ pack,
                                                 const float *__restrict__ w,
                                                 float *__restrict__ out) {
  const int idx = blockIdx.x * 64 + threadIdx.x;
  const bool valid = idx < n;
  MlpW W;
  mlp_load(pack, w, W);
  f3 u = {0.0f, 0.0f, 0.0f}, s = {0.0f, 0.0f, 1.0f}, sp = {0.0f, 0.0f, 1.0f};
  if (valid) {
    const float *q = rows + (size_t)idx * 9;
    u = {q[0], q[1], q[2]};
    s = {q[3], q[4], q[5]};
    sp = {q[6], q[7], q[8]};
  }
  float x[6], o1[4], o2[4];
  features(u, s, x);
  mlp_forward(W, x, o1);
  features(u, sp, x);
  mlp_forward(W, x, o2);
  const Vmf fwd = decode(o1), rev = decode(o2);
  const float la = mh_log_ratio(u, 1.0f, s, sp, fwd, rev);
  if (valid) {
    float *p = out + (size_t)idx * 9;
    for (int q = 0; q < 4; ++q) {
      p[q] = o1[q];
      p[4 + q] = o2[q];
    }
    p[8] = la;
  }
}

// vMF draws through the device sampler (for the statistical tests):
// rows n x 4 (mu, kappa) -> out n x 3, uniforms from Philox(key, idx).
__global__ void k_vmf_draw(const float *__restrict__ rows, int n, uint32_t key0, uint32_t key1,
                           float *__restrict__ out) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const float *q = rows + (size_t)idx * 4;
  Vmf v;
  v.mu = {q[0], q[1], q[2]};
  v.kappa = q[3];
  v.em2k = expm1_neg2k(v.kappa);
  v.log_norm = 0.0f;
  const u32x4 rnd = philox4x32_10(rng_counter((uint32_t)idx, 0, 0, kKindLearned, 0, 0), key0, key1);
  const f3 y = vmf_sample(v, u01(rnd.x), u01(rnd.y));
  out[(size_t)idx * 3 + 0] = y.x;
  out[(size_t)idx * 3 + 1] = y.y;
  out[(size_t)idx * 3 + 2] = y.z;
}

// Environments (u = beta h, s) at Philox-chosen (replica, cell, site) of the
// current configuration, appended at row offset `base`. grid = ceil(n/256).
__global__ void k_gather_env(const float *__restrict__ spins, int L, int ncell, int R,
                             int per_replica, uint32_t draw, const float *__restrict__ beta,
                             float jeff, uint32_t key0, uint32_t key1, float *__restrict__ env,
                             int base, int cap) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= per_replica * R || base + idx >= cap) return;
  const int r = idx / per_replica;
  const u32x4 rnd = philox4x32_10(rng_counter((uint32_t)idx, 0, 0, kKindEnv, 0, draw), key0, key1);
  const int cell = (int)(rnd.x % (uint32_t)ncell);
  const int site = (int)(rnd.y & 15u);
  const size_t plane = (size_t)16 * ncell;
  const float *Sx = spins + (size_t)r * 3 * plane;
  const float *Sy = Sx + plane;
  const float *Sz = Sy + plane;
  const int LL = L * L;
  const int z = cell / LL, rem = cell - z * LL, y = rem / L, x = rem - y * L;
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
  for (int nb = 0; nb < 6; ++nb) {
    const Link lk = c_links_learned[site][nb];
    const int j = lk.site * ncell + (wrap_l(z + lk.dz, L) * L + wrap_l(y + lk.dy, L)) * L +
                  wrap_l(x + lk.dx, L);
    fx = fx + Sx[j];
    fy = fy + Sy[j];
    fz = fz + Sz[j];
  }
  const int i = site * ncell + cell;
  const float bt = beta[r];
  float *e = env + (size_t)(base + idx) * 6;
  e[0] = bt * (-jeff * fx);
  e[1] = bt * (-jeff * fy);
  e[2] = bt * (-jeff * fz);
  e[3] = Sx[i];
  e[4] = Sy[i];
  e[5] = Sz[i];
}

// ---------------------------------------------------------------------------
// Training, FP32 on the vector units.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double langevin(double k) {
  if (k < 1e-3) return k / 3.0 - k * k * k / 45.0;
  return 1.0 / tanh(k) - 1.0 / k;
}
__device__ __forceinline__ double langevin_d(double k) {
  if (k < 1e-3) return 1.0 / 3.0 - k * k / 15.0;
  if (k > 350.0) return 1.0 / (k * k);
  const double sh = sinh(k);
  return 1.0 / (k * k) - 1.0 / (sh * sh);
}
__device__ __forceinline__ double log_norm_d(double k) {
  if (k < 1e-6) return -log(4.0 * M_PI);
  return log(k) - log(2.0 * M_PI) - k - log(-expm1(-2.0 * k));
}

// Forward, loss and backward of 64 samples per block (one per thread), then
// the block's partial gradient. Sample n of the batch reads env row
// rowsel[n] (explicit list) or a Philox-chosen row. Activations and deltas
// stay in LDS, [unit][sample] with a padded row stride so that the
// gradient loop (threads = consecutive parameters, all at the same sample)
// reads without bank conflicts; a2 is overwritten by d2 and a1 by d1 once
// the gradients that need them are taken. part[block][p] (p < kMlpParams:
// gradient sums over the block's samples in sample order; p = kMlpParams:
// loss sum). grid = ceil(B/64), block = 64.
constexpr int kGradLen = kMlpParams + 1;
constexpr int kLdsRow = 65;

__device__ __forceinline__ void grad_range(int p0, int p1, const float *__restrict__ sd,
                                           const float *__restrict__ sa, int a_rows,
                                           float *__restrict__ part, int nvalid) {
  // Parameters [p0, p1): weights d_row x a_row (a_rows columns per row), then
  // biases (p >= p0 + rows * a_rows) which sum d alone.
  for (int p = p0 + (int)threadIdx.x; p < p1; p += 64) {
    const int q = p - p0;
    const int rows_w = (p1 - p0) / (a_rows + 1);  // weight rows = bias count
    float t = 0.0f;
    if (q < rows_w * a_rows) {
      const float *d = sd + (q / a_rows) * kLdsRow;
      const float *a = sa + (q % a_rows) * kLdsRow;
      for (int m = 0; m < nvalid; ++m) t += d[m] * a[m];
    } else {
      const float *d = sd + (q - rows_w * a_rows) * kLdsRow;
      for (int m = 0; m < nvalid; ++m) t += d[m];
    }
    part[p] = t;
  }
}

__global__ __launch_bounds__(64) void k_train_samples(const float *__restrict__ env, int n_env,
                                                      const int *__restrict__ rowsel, int B,
                                                      uint32_t step, uint32_t key0, uint32_t key1,
                                                      const float *__restrict__ w,
                                                      float *__restrict__ part_all) {
  __shared__ float sx[6 * kLdsRow], sa1[64 * kLdsRow], sa2[64 * kLdsRow], sd3[4 * kLdsRow],
      sl[64];
  const int li = threadIdx.x;
  const int n = blockIdx.x * 64 + li;
  const int nvalid = min(64, B - (int)blockIdx.x * 64);
  float *part = part_all + (size_t)blockIdx.x * kGradLen;
  float go[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  if (n < B) {
    int row;
    if (rowsel) {
      row = rowsel[n];
    } else {
      const u32x4 rnd = philox4x32_10(rng_counter((uint32_t)n, 0, 0, kKindTrain, 0, step), key0, key1);
      row = (int)(rnd.x % (uint32_t)n_env);
    }
    const float *e = env + (size_t)row * 6;
    const f3 u = {e[0], e[1], e[2]};
    const f3 s = {e[3], e[4], e[5]};
    float x[6];
    features(u, s, x);
    for (int k = 0; k < 6; ++k) sx[k * kLdsRow + li] = x[k];
    for (int j = 0; j < 64; ++j) {
      float t = w[kOffB1 + j];
      for (int k = 0; k < 6; ++k) t += w[kOffW1 + j * 6 + k] * x[k];
      sa1[j * kLdsRow + li] = relu(t);
    }
    for (int j = 0; j < 64; ++j) {
      float t = w[kOffB2 + j];
      for (int k = 0; k < 64; ++k) t += w[kOffW2 + j * 64 + k] * sa1[k * kLdsRow + li];
      sa2[j * kLdsRow + li] = relu(t);
    }
    float o[4];
    for (int q = 0; q < 4; ++q) {
      float t = w[kOffB3 + q];
      for (int k = 0; k < 64; ++k) t += w[kOffW3 + q * 64 + k] * sa2[k * kLdsRow + li];
      o[q] = t;
    }
    // Reverse KL to vMF(u/|u|, |u|) and its gradient wrt o, in FP64.
    const double on = sqrt((double)o[0] * o[0] + (double)o[1] * o[1] + (double)o[2] * o[2]);
    double mu[3] = {0.0, 0.0, 1.0};
    if (on > 0.0) {
      mu[0] = o[0] / on;
      mu[1] = o[1] / on;
      mu[2] = o[2] / on;
    }
    double lk = o[3];
    const bool clamped = lk < (double)kLogKappaMin || lk > (double)kLogKappaMax;
    lk = fmin(fmax(lk, (double)kLogKappaMin), (double)kLogKappaMax);
    const double kap = exp(lk);
    const double ks = sqrt((double)u.x * u.x + (double)u.y * u.y + (double)u.z * u.z);
    double hat[3] = {0.0, 0.0, 0.0};
    if (ks > 0.0) {
      hat[0] = u.x / ks;
      hat[1] = u.y / ks;
      hat[2] = u.z / ks;
    }
    const double c = mu[0] * hat[0] + mu[1] * hat[1] + mu[2] * hat[2];
    const double A = langevin(kap);
    const double kl = log_norm_d(kap) - log_norm_d(ks) + A * (kap - ks * c);
    go[3] = clamped ? 0.0f : (float)(langevin_d(kap) * (kap - ks * c) * kap);
    if (on > 0.0) {
      const double gm[3] = {-A * ks * hat[0], -A * ks * hat[1], -A * ks * hat[2]};
      const double proj = gm[0] * mu[0] + gm[1] * mu[1] + gm[2] * mu[2];
      for (int q = 0; q < 3; ++q) go[q] = (float)((gm[q] - proj * mu[q]) / on);
    }
    sl[li] = (float)kl;
    for (int q = 0; q < 4; ++q) sd3[q * kLdsRow + li] = go[q];
  }
  __syncthreads();
  // Layer 3 (d3 x a2) and the loss.
  grad_range(kOffW3, kMlpParams, sd3, sa2, 64, part, nvalid);
  if (li == 0) {
    float t = 0.0f;
    for (int m = 0; m < nvalid; ++m) t += sl[m];
    part[kMlpParams] = t;
  }
  __syncthreads();
  if (n < B)  // d2 replaces a2 (this thread's column only)
    for (int k = 0; k < 64; ++k) {
      float t = 0.0f;
      for (int q = 0; q < 4; ++q) t += w[kOffW3 + q * 64 + k] * go[q];
      float &a = sa2[k * kLdsRow + li];
      a = a > 0.0f ? t : 0.0f;
    }
  __syncthreads();
  grad_range(kOffW2, kOffW3, sa2, sa1, 64, part, nvalid);  // d2 x a1
  __syncthreads();
  if (n < B) {  // d1 replaces a1
    float d1[64];
#pragma unroll
    for (int k = 0; k < 64; ++k) d1[k] = 0.0f;
    for (int j = 0; j < 64; ++j) {
      const float d2 = sa2[j * kLdsRow + li];
#pragma unroll
      for (int k = 0; k < 64; ++k) d1[k] += w[kOffW2 + j * 64 + k] * d2;
    }
#pragma unroll
    for (int k = 0; k < 64; ++k) {
      float &a = sa1[k * kLdsRow + li];
      a = a > 0.0f ? d1[k] : 0.0f;
    }
  }
  __syncthreads();
  grad_range(kOffW1, kOffW2, sa1, sx, 6, part, nvalid);  // d1 x x
}

// Batch-mean gradient and loss: grad[p] = sum over blocks of part[blk][p] in
// block order (FP64 accumulation) / B. One parameter per thread.
__global__ void k_train_reduce(const float *__restrict__ part, int nblk, int B,
                               float *__restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kGradLen) return;
  double t = 0.0;
  for (int b = 0; b < nblk; ++b) t += (double)part[(size_t)b * kGradLen + p];
  grad[p] = (float)(t / B);
}

// Adam step on every parameter; thread 0 also files the batch loss
// (g[kMlpParams]) into loss_out when given.
__global__ void k_adam(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, float lr, float b1, float b2, float eps,
                       float c1, float c2, float *__restrict__ loss_out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0 && loss_out) *loss_out = g[kMlpParams];
  if (p >= kMlpParams) return;
  const float gp = g[p];
  const float mp = b1 * m[p] + (1.0f - b1) * gp;
  const float vp = b2 * v[p] + (1.0f - b2) * gp * gp;
  m[p] = mp;
  v[p] = vp;
  w[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
}

__global__ void k_mlp_pack(const float *__restrict__ w, uint16_t *__restrict__ pack) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kPackSize) return;
  float val = 0.0f;
  if (p < kPackP2) {  // [64][8]: W1 columns, bias, 0
    const int row = p / 8, col = p % 8;
    val = col < 6 ? w[kOffW1 + row * 6 + col] : (col == 6 ? w[kOffB1 + row] : 0.0f);
    val *= kHeadroomDown;
  } else if (p < kPackP3) {  // [ti][t][lane][j]
    const int q = p - kPackP2, j = q & 7, lane = (q >> 3) & 63, t = (q >> 9) & 3, ti = q >> 11;
    val = w[kOffW2 + (32 * ti + (lane & 31)) * 64 + kstep_unit(t, lane >> 5, j)];
  } else if (p < kPackB2) {  // [t][lane][j], rows 4..31 zero
    const int q = p - kPackP3, j = q & 7, lane = (q >> 3) & 63, t = q >> 9;
    const int row = lane & 31;
    val = row < 4 ? w[kOffW3 + row * 64 + kstep_unit(t, lane >> 5, j)] * kHeadroomUp : 0.0f;
  } else {  // b2 * 2^-16 as FP32, low then high 16 bits of each
    const int q = p - kPackB2;
    const uint32_t bits = __builtin_bit_cast(uint32_t, w[kOffB2 + (q >> 1)] * kHeadroomDown);
    pack[p] = (uint16_t)((q & 1) ? bits >> 16 : bits & 0xFFFFu);
    return;
  }
  pack[p] = (uint16_t)bf16_bits(val);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static inline uint32_t klo(const pyro_ctx *c) { return (uint32_t)c->seed; }
static inline uint32_t khi(const pyro_ctx *c) { return (uint32_t)(c->seed >> 32); }

int mlp_alloc(pyro_ctx *ctx) {
  if (ctx->d_mlp_w) return 0;
  const size_t nb = sizeof(float) * kMlpParams;
  if (hipMalloc((void **)&ctx->d_mlp_w, nb) != hipSuccess ||
      hipMalloc((void **)&ctx->d_mlp_m, nb) != hipSuccess ||
      hipMalloc((void **)&ctx->d_mlp_v, nb) != hipSuccess ||
      hipMalloc((void **)&ctx->d_grad, nb + sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&ctx->d_mlp_pack, sizeof(uint16_t) * kPackSize) != hipSuccess ||
      hipMalloc((void **)&ctx->d_acc_learned, sizeof(unsigned long long) * ctx->R) != hipSuccess) {
    mlp_free(ctx);  // all or nothing: a later call allocates afresh
    return set_error(PYRO_ENOMEM, "learned proposal: device allocation failed");
  }
  ctx->learned_nslots = (ctx->ncell + kPhaseCells - 1) / kPhaseCells;
  const size_t slot_bytes =
      sizeof(LearnedSlot) * (size_t)ctx->R * kColours * (size_t)ctx->learned_nslots;
  if (hipMalloc((void **)&ctx->d_learned_slots, slot_bytes) != hipSuccess) {
    mlp_free(ctx);
    return set_error(PYRO_ENOMEM, "learned proposal: device allocation failed");
  }
  if (hipMemset(ctx->d_learned_slots, 0, slot_bytes) != hipSuccess ||
      hipMemset(ctx->d_mlp_w, 0, nb) != hipSuccess || hipMemset(ctx->d_mlp_m, 0, nb) != hipSuccess ||
      hipMemset(ctx->d_mlp_v, 0, nb) != hipSuccess ||
      hipMemset(ctx->d_acc_learned, 0, sizeof(unsigned long long) * ctx->R) != hipSuccess) {
    mlp_free(ctx);
    return set_error(PYRO_EHIP, "learned proposal: clearing the device state failed");
  }
  ctx->h_proposed_learned.assign(ctx->R, 0);
  return 0;
}

void mlp_free(pyro_ctx *ctx) {
  void *ptrs[] = {ctx->d_mlp_w, ctx->d_mlp_m, ctx->d_mlp_v, ctx->d_grad, ctx->d_mlp_pack,
                  ctx->d_env, ctx->d_train, ctx->d_acc_learned, ctx->d_learned_slots};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  ctx->d_mlp_w = ctx->d_mlp_m = ctx->d_mlp_v = ctx->d_grad = ctx->d_env = ctx->d_train = nullptr;
  ctx->d_mlp_pack = nullptr;
  ctx->d_acc_learned = nullptr;
  ctx->d_learned_slots = nullptr;
  ctx->learned_nslots = 0;
}

int mlp_pack(pyro_ctx *ctx) {
  k_mlp_pack<<<(kPackSize + 255) / 256, 256, 0, ctx->stream>>>(ctx->d_mlp_w, ctx->d_mlp_pack);
  PYRO_HIP(hipGetLastError());
  ctx->mlp_ready = true;
  return 0;
}

// 64-cell groups per block. Several groups per wave share the weight
// fragments a wave loads (22.5 KB). With the per-block slots, 2 and 4 groups
// spill at 168 VGPRs and run 45.9 and 47.9 us per L=16 x 64 launch against
// 39.7 for one. (When every block still ended with two device-scope
// atomics, fewer blocks also meant fewer atomics: 47.0 / 43.0 / 40.6 us for
// 1 / 4 / 8 groups at 216 VGPRs, but slower sweeps, since two 216-VGPR waves
// leave no room for another stream's over-relaxation wave on their SIMD.)
// One group.
constexpr int kLearnedGroups = 1;

template <int LT>
static void launch_learned_t(pyro_ctx *ctx, int colour, int pass, int accumulate) {
  constexpr int G = kLearnedGroups;
  dim3 grid((ctx->ncell + kPhaseCells * G - 1) / (kPhaseCells * G), launch_nrep(ctx));
  if (ctx->time_ev0) {
    hipExtLaunchKernelGGL(k_learned_phase<LT, G>, grid, dim3(kPhaseBlock), 0, launch_stream(ctx),
                          ctx->time_ev0, ctx->time_ev1, 0, ctx->d_spins, ctx->L, ctx->ncell, colour,
                          pass, (const float *)ctx->d_beta, (const uint32_t *)ctx->d_ctl, klo(ctx),
                          khi(ctx), ctx->jeff, (const uint16_t *)ctx->d_mlp_pack,
                          (const float *)ctx->d_mlp_w, ctx->d_learned_slots, ctx->learned_nslots,
                          accumulate, launch_r0(ctx), ctx->launch_delta);
    return;
  }
  k_learned_phase<LT, G><<<grid, kPhaseBlock, 0, launch_stream(ctx)>>>(
      ctx->d_spins, ctx->L, ctx->ncell, colour, pass, ctx->d_beta, ctx->d_ctl, klo(ctx), khi(ctx),
      ctx->jeff, ctx->d_mlp_pack, ctx->d_mlp_w, ctx->d_learned_slots, ctx->learned_nslots,
      accumulate, launch_r0(ctx), ctx->launch_delta);
}

int launch_learned_phase(pyro_ctx *ctx, int colour, int pass, int accumulate) {
  if (!ctx->mlp_ready || !ctx->d_learned_slots)
    return set_error(PYRO_ESTATE, "learned proposal: no weights set");
  if (ctx->hetero) return set_error(PYRO_ESTATE, "learned proposal: needs a uniform lattice size");
  if (ctx->mat) return set_error(PYRO_ESTATE, "learned proposal: it is for the scalar model");
  switch (ctx->L) {
    case 8: launch_learned_t<8>(ctx, colour, pass, accumulate); break;
    case 16: launch_learned_t<16>(ctx, colour, pass, accumulate); break;
    case 64: launch_learned_t<64>(ctx, colour, pass, accumulate); break;
    default: launch_learned_t<0>(ctx, colour, pass, accumulate); break;
  }
  PYRO_HIP(hipGetLastError());
  return 0;
}

int launch_learned_fold(pyro_ctx *ctx) {
  if (!ctx->d_learned_slots) return set_error(PYRO_ESTATE, "learned proposal: not set up");
  k_learned_fold<<<launch_nrep(ctx), 256, 0, launch_stream(ctx)>>>(
      ctx->d_learned_slots, ctx->learned_nslots, ctx->d_dE, ctx->d_acc_learned, launch_r0(ctx));
  PYRO_HIP(hipGetLastError());
  return 0;
}

}  // namespace pyro

using namespace pyro;

extern "C" {

int pyro_mlp_set_weights(pyro_ctx *ctx, const float *w) {
  if (!ctx || !w) return set_error(PYRO_EINVAL, "pyro_mlp_set_weights: null argument");
  for (int p = 0; p < kMlpParams; ++p)
    if (!std::isfinite(w[p])) return set_error(PYRO_EINVAL, "pyro_mlp_set_weights: w[%d] not finite", p);
  PYRO_HIP(hipSetDevice(ctx->device));
  PYRO_TRY(mlp_alloc(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  PYRO_HIP(hipMemcpy(ctx->d_mlp_w, w, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
  PYRO_HIP(hipMemset(ctx->d_mlp_m, 0, sizeof(float) * kMlpParams));
  PYRO_HIP(hipMemset(ctx->d_mlp_v, 0, sizeof(float) * kMlpParams));
  ctx->adam_t = 0;
  PYRO_TRY(mlp_pack(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

int pyro_mlp_adam_step(pyro_ctx *ctx, const float *grad, float lr) {
  if (!ctx || !grad || !(lr > 0.0f)) return set_error(PYRO_EINVAL, "pyro_mlp_adam_step: bad argument");
  if (!ctx->d_mlp_w) return set_error(PYRO_ESTATE, "pyro_mlp_adam_step: no weights set");
  for (int p = 0; p < kMlpParams; ++p)
    if (!std::isfinite(grad[p])) return set_error(PYRO_EINVAL, "pyro_mlp_adam_step: grad[%d] not finite", p);
  PYRO_HIP(hipSetDevice(ctx->device));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  PYRO_HIP(hipMemcpy(ctx->d_grad, grad, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const int t = ++ctx->adam_t;
  k_adam<<<(kMlpParams + 255) / 256, 256, 0, ctx->stream>>>(
      ctx->d_mlp_w, ctx->d_mlp_m, ctx->d_mlp_v, ctx->d_grad, lr, b1, b2, eps,
      (float)(1.0 - std::pow((double)b1, t)), (float)(1.0 - std::pow((double)b2, t)), nullptr);
  PYRO_HIP(hipGetLastError());
  PYRO_TRY(mlp_pack(ctx));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

int pyro_mlp_get_weights(pyro_ctx *ctx, float *w) {
  if (!ctx || !w) return set_error(PYRO_EINVAL, "pyro_mlp_get_weights: null argument");
  if (!ctx->d_mlp_w) return set_error(PYRO_ESTATE, "pyro_mlp_get_weights: no weights");
  PYRO_HIP(hipSetDevice(ctx->device));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  PYRO_HIP(hipMemcpy(w, ctx->d_mlp_w, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
  return 0;
}

int pyro_mlp_eval(pyro_ctx *ctx, const float *rows9, int n, float *out9) {
  if (!ctx || !rows9 || !out9 || n < 1) return set_error(PYRO_EINVAL, "pyro_mlp_eval: bad argument");
  if (!ctx->mlp_ready) return set_error(PYRO_ESTATE, "pyro_mlp_eval: no weights set");
  PYRO_HIP(hipSetDevice(ctx->device));
  float *d_in = nullptr, *d_out = nullptr;
  const size_t nb = sizeof(float) * 9 * (size_t)n;
  if (hipMalloc((void **)&d_in, nb) != hipSuccess || hipMalloc((void **)&d_out, nb) != hipSuccess) {
    if (d_in) (void)hipFree(d_in);
    return set_error(PYRO_ENOMEM, "pyro_mlp_eval: allocation failed");
  }
  int rc = 0;
  if (hipMemcpyAsync(d_in, rows9, nb, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
    rc = set_error(PYRO_EHIP, "pyro_mlp_eval: upload failed");
  if (!rc) {
    k_mlp_eval<<<(n + 63) / 64, 64, 0, ctx->stream>>>(d_in, n, ctx->d_mlp_pack, ctx->d_mlp_w, d_out);
    if (hipGetLastError() != hipSuccess) rc = set_error(PYRO_EHIP, "pyro_mlp_eval: launch failed");
  }
  if (!rc && hipMemcpyAsync(out9, d_out, nb, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
    rc = set_error(PYRO_EHIP, "pyro_mlp_eval: download failed");
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess)
    rc = set_error(PYRO_EHIP, "pyro_mlp_eval: synchronise failed");
  (void)hipFree(d_in);
  (void)hipFree(d_out);
  return rc;
}

int pyro_vmf_draw(pyro_ctx *ctx, const float *rows4, int n, float *out3) {
  if (!ctx || !rows4 || !out3 || n < 1) return set_error(PYRO_EINVAL, "pyro_vmf_draw: bad argument");
  PYRO_HIP(hipSetDevice(ctx->device));
  float *d_in = nullptr, *d_out = nullptr;
  if (hipMalloc((void **)&d_in, sizeof(float) * 4 * (size_t)n) != hipSuccess ||
      hipMalloc((void **)&d_out, sizeof(float) * 3 * (size_t)n) != hipSuccess) {
    if (d_in) (void)hipFree(d_in);
    return set_error(PYRO_ENOMEM, "pyro_vmf_draw: allocation failed");
  }
  int rc = 0;
  if (hipMemcpyAsync(d_in, rows4, sizeof(float) * 4 * (size_t)n, hipMemcpyHostToDevice, ctx->stream) !=
      hipSuccess)
    rc = set_error(PYRO_EHIP, "pyro_vmf_draw: upload failed");
  if (!rc) {
    k_vmf_draw<<<(n + 255) / 256, 256, 0, ctx->stream>>>(d_in, n, klo(ctx), khi(ctx), d_out);
    if (hipGetLastError() != hipSuccess) rc = set_error(PYRO_EHIP, "pyro_vmf_draw: launch failed");
  }
  if (!rc && hipMemcpyAsync(out3, d_out, sizeof(float) * 3 * (size_t)n, hipMemcpyDeviceToHost,
                            ctx->stream) != hipSuccess)
    rc = set_error(PYRO_EHIP, "pyro_vmf_draw: download failed");
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess)
    rc = set_error(PYRO_EHIP, "pyro_vmf_draw: synchronise failed");
  (void)hipFree(d_in);
  (void)hipFree(d_out);
  return rc;
}

// Partial-gradient buffer for batches of up to B samples: one kGradLen row
// per 64-sample block.
static int ensure_train(pyro_ctx *ctx, int B) {
  if (ctx->train_cap >= B) return 0;
  if (ctx->d_train) PYRO_HIP(hipFree(ctx->d_train));
  ctx->d_train = nullptr;
  ctx->train_cap = 0;
  PYRO_HIP(hipMalloc((void **)&ctx->d_train, sizeof(float) * kGradLen * (size_t)((B + 63) / 64)));
  ctx->train_cap = B;
  return 0;
}

int pyro_mlp_loss_grad(pyro_ctx *ctx, const float *env6, int n, double *loss, float *grad) {
  if (!ctx || !env6 || n < 1 || n > (1 << 20))
    return set_error(PYRO_EINVAL, "pyro_mlp_loss_grad: bad argument");
  if (!ctx->d_mlp_w) return set_error(PYRO_ESTATE, "pyro_mlp_loss_grad: no weights set");
  PYRO_HIP(hipSetDevice(ctx->device));
  PYRO_TRY(ensure_train(ctx, n));
  float *d_env = nullptr;
  int *d_sel = nullptr;
  if (hipMalloc((void **)&d_env, sizeof(float) * 6 * (size_t)n) != hipSuccess ||
      hipMalloc((void **)&d_sel, sizeof(int) * (size_t)n) != hipSuccess) {
    if (d_env) (void)hipFree(d_env);
    return set_error(PYRO_ENOMEM, "pyro_mlp_loss_grad: allocation failed");
  }
  std::vector<int> sel(n);
  for (int i = 0; i < n; ++i) sel[i] = i;
  std::vector<float> g(kMlpParams + 1);
  int rc = 0;
  if (hipMemcpy(d_env, env6, sizeof(float) * 6 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_sel, sel.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice) != hipSuccess)
    rc = set_error(PYRO_EHIP, "pyro_mlp_loss_grad: upload failed");
  if (!rc) {
    k_train_samples<<<(n + 63) / 64, 64, 0, ctx->stream>>>(d_env, n, d_sel, n, 0, klo(ctx), khi(ctx),
                                                          ctx->d_mlp_w, ctx->d_train);
    k_train_reduce<<<(kGradLen + 255) / 256, 256, 0, ctx->stream>>>(ctx->d_train, (n + 63) / 64, n,
                                                                     ctx->d_grad);
    if (hipGetLastError() != hipSuccess) rc = set_error(PYRO_EHIP, "pyro_mlp_loss_grad: launch failed");
  }
  if (!rc && (hipStreamSynchronize(ctx->stream) != hipSuccess ||
              hipMemcpy(g.data(), ctx->d_grad, sizeof(float) * (kMlpParams + 1), hipMemcpyDeviceToHost) !=
                  hipSuccess))
    rc = set_error(PYRO_EHIP, "pyro_mlp_loss_grad: download failed");
  (void)hipFree(d_env);
  (void)hipFree(d_sel);
  if (rc) return rc;
  if (grad) std::memcpy(grad, g.data(), sizeof(float) * kMlpParams);
  if (loss) *loss = g[kMlpParams];
  return 0;
}

int pyro_gather_environments(pyro_ctx *ctx, int per_replica, uint32_t draw, int capacity, int *n_total) {
  if (!ctx || per_replica < 1 || capacity < 1) return set_error(PYRO_EINVAL, "pyro_gather_environments: bad argument");
  if (ctx->hetero) return set_error(PYRO_ESTATE, "pyro_gather_environments: needs a uniform lattice size");
  if (ctx->mat) return set_error(PYRO_ESTATE, "pyro_gather_environments: the learned proposal is for the scalar model");
  PYRO_HIP(hipSetDevice(ctx->device));
  if (capacity > ctx->env_cap) {
    PYRO_HIP(hipStreamSynchronize(ctx->stream));
    float *p = nullptr;
    PYRO_HIP(hipMalloc((void **)&p, sizeof(float) * 6 * (size_t)capacity));
    if (ctx->d_env && ctx->env_n > 0)
      PYRO_HIP(hipMemcpy(p, ctx->d_env, sizeof(float) * 6 * (size_t)ctx->env_n, hipMemcpyDeviceToDevice));
    if (ctx->d_env) PYRO_HIP(hipFree(ctx->d_env));
    ctx->d_env = p;
    ctx->env_cap = capacity;
  }
  const int want = per_replica * ctx->R;
  const int take = std::min(want, ctx->env_cap - ctx->env_n);
  if (take > 0) {
    k_gather_env<<<(want + 255) / 256, 256, 0, ctx->stream>>>(
        ctx->d_spins, ctx->L, ctx->ncell, ctx->R, per_replica, draw, ctx->d_beta, ctx->jeff, klo(ctx),
        khi(ctx), ctx->d_env, ctx->env_n, ctx->env_cap);
    PYRO_HIP(hipGetLastError());
    ctx->env_n += take;
  }
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  if (n_total) *n_total = ctx->env_n;
  return 0;
}

int pyro_get_environments(pyro_ctx *ctx, float *env6, int *n) {
  if (!ctx || !n) return set_error(PYRO_EINVAL, "pyro_get_environments: null argument");
  PYRO_HIP(hipSetDevice(ctx->device));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  *n = ctx->env_n;
  if (env6 && ctx->env_n > 0)
    PYRO_HIP(hipMemcpy(env6, ctx->d_env, sizeof(float) * 6 * (size_t)ctx->env_n, hipMemcpyDeviceToHost));
  return 0;
}

int pyro_clear_environments(pyro_ctx *ctx) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  ctx->env_n = 0;
  return 0;
}

int pyro_mlp_train(pyro_ctx *ctx, int steps, int batch, float lr, float lr_final, float *loss_hist) {
  if (!ctx || steps < 0 || batch < 1 || batch > (1 << 20) || !(lr > 0.0f) || !(lr_final > 0.0f))
    return set_error(PYRO_EINVAL, "pyro_mlp_train: bad argument");
  if (!ctx->d_mlp_w) return set_error(PYRO_ESTATE, "pyro_mlp_train: no initial weights");
  if (ctx->env_n < 1) return set_error(PYRO_ESTATE, "pyro_mlp_train: no environments gathered");
  PYRO_HIP(hipSetDevice(ctx->device));
  PYRO_TRY(ensure_train(ctx, batch));
  float *d_loss = nullptr;
  if (loss_hist && steps > 0 && hipMalloc((void **)&d_loss, sizeof(float) * (size_t)steps) != hipSuccess)
    return set_error(PYRO_ENOMEM, "pyro_mlp_train: loss history allocation failed");
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  int rc = 0;
  for (int s = 0; s < steps && rc == 0; ++s) {
    const int t = ++ctx->adam_t;
    const double frac = steps > 1 ? (double)s / (steps - 1) : 0.0;
    const float lr_s = (float)(lr * std::pow((double)lr_final / lr, frac));  // geometric decay
    k_train_samples<<<(batch + 63) / 64, 64, 0, ctx->stream>>>(ctx->d_env, ctx->env_n, nullptr, batch,
                                                              (uint32_t)t, klo(ctx), khi(ctx),
                                                              ctx->d_mlp_w, ctx->d_train);
    k_train_reduce<<<(kGradLen + 255) / 256, 256, 0, ctx->stream>>>(
        ctx->d_train, (batch + 63) / 64, batch, ctx->d_grad);
    k_adam<<<(kMlpParams + 255) / 256, 256, 0, ctx->stream>>>(
        ctx->d_mlp_w, ctx->d_mlp_m, ctx->d_mlp_v, ctx->d_grad, lr_s, b1, b2, eps,
        (float)(1.0 - std::pow((double)b1, t)), (float)(1.0 - std::pow((double)b2, t)),
        d_loss ? d_loss + s : nullptr);
    if (hipGetLastError() != hipSuccess) rc = set_error(PYRO_EHIP, "pyro_mlp_train: launch failed");
  }
  if (rc == 0) rc = mlp_pack(ctx);
  if (rc == 0 && hipStreamSynchronize(ctx->stream) != hipSuccess)
    rc = set_error(PYRO_EHIP, "pyro_mlp_train: synchronise failed");
  if (rc == 0 && d_loss &&
      hipMemcpy(loss_hist, d_loss, sizeof(float) * (size_t)steps, hipMemcpyDeviceToHost) != hipSuccess)
    rc = set_error(PYRO_EHIP, "pyro_mlp_train: loss history download failed");
  if (d_loss) (void)hipFree(d_loss);
  return rc;
}

int pyro_learned_phase(pyro_ctx *ctx, int colour, uint32_t sweep, int pass) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  if (colour < 0 || colour > 3) return set_error(PYRO_EINVAL, "colour %d outside 0..3", colour);
  if (pass < 0 || pass > 255) return set_error(PYRO_EINVAL, "pass %d outside 0..255", pass);
  PYRO_HIP(hipSetDevice(ctx->device));
  PYRO_TRY(launch_set_sweep(ctx, sweep, ctx->h_pt_round));
  PYRO_TRY(launch_learned_phase(ctx, colour, pass, 0));
  PYRO_TRY(launch_learned_fold(ctx));
  PYRO_TRY(launch_set_sweep(ctx, ctx->h_sweep, ctx->h_pt_round));
  for (auto &p : ctx->h_proposed_learned) p += 4ull * (uint64_t)ctx->ncell;
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

int pyro_acceptance_learned(pyro_ctx *ctx, uint64_t *accepted, uint64_t *proposed) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  if (!ctx->d_acc_learned) return set_error(PYRO_ESTATE, "no learned proposal set up");
  PYRO_HIP(hipSetDevice(ctx->device));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  if (accepted)
    PYRO_HIP(hipMemcpy(accepted, ctx->d_acc_learned, sizeof(uint64_t) * ctx->R, hipMemcpyDeviceToHost));
  if (proposed) std::memcpy(proposed, ctx->h_proposed_learned.data(), sizeof(uint64_t) * ctx->R);
  return 0;
}

}  // extern "C"
