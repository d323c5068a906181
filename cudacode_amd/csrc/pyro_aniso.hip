// Matrix local-field model: anisotropic nearest-neighbour exchange between
// the four basis sites and a Zeeman field per basis site (the reference's
// general coupling path: couplings_const and J1..J4, Pyrochlore.cu:139-184;
// InitializeMagneticField :414-450; InitializeCouplings3D :643-711; the field
// of CalculateLocalExchangeField :210-275 and the CPU TotalEnergy :468-523).
//
// Energy: H = -spin_len2 (sum_<ij> s_i^T C_{b_i b_j} s_j + 2 sum_i B_{b_i}.s_i),
// local field h_i = spin_len2 (sum_j C_{b_i b_j} s_j + 2 B_{b_i}); with
// C = -J I and B = 0 this is the scalar Heisenberg model. The colour-phase
// and observables kernels of pyro_phase.h are instantiated here with MAT =
// true, in a translation unit of their own.
#include <cmath>
#include <cstring>
#include <vector>

#include "pyro_phase.h"

namespace pyro {

int launch_phase_mat(pyro_ctx *ctx, int kind, int colour, int pass) {
  if (kind == 0)
    launch_phase<0, true>(ctx, colour, pass);
  else if (kind == 2)
    launch_phase<2, true>(ctx, colour, pass);
  else
    launch_phase<1, true>(ctx, colour, pass);
  PYRO_HIP(hipGetLastError());
  return 0;
}

int launch_observables_partials_mat(pyro_ctx *ctx) {
  launch_observables_partials<true>(ctx);
  PYRO_HIP(hipGetLastError());
  return 0;
}

// Exchange matrices of the 6 basis pairs b < b' in the global frame, entries
// as signed indices: +-k stands for +-J_k. The pair (b', b) is the transpose.
// These are the symmetry-allowed nearest-neighbour couplings of the pyrochlore
// lattice in the J1..J4 parameterisation the reference uses (couplings_const,
// Pyrochlore.cu:169-184); C_{b b} is unused (no bonds within a basis class).
static const int8_t kExchangeUpper[6][2 + 9] = {
    {0, 1, -1, -3, 4, -3, -1, 4, -4, -4, -2},
    {0, 2, -2, -4, -4, 4, -1, -3, 4, -3, -1},
    {0, 3, -1, 4, -3, -4, -2, -4, -3, 4, -1},
    {1, 2, -1, 4, 3, -4, -2, 4, 3, -4, -1},
    {1, 3, -2, -4, 4, 4, -1, 3, -4, 3, -1},
    {2, 3, -1, 3, -4, 3, -1, 4, 4, -4, -2},
};

// Local [111] axis of each basis site (times sqrt 3): the g-tensor of site b
// is g_xy I + (g_z - g_xy) n_b n_b^T (InitializeMagneticField builds the same
// four matrices from g1 = (2 g_xy + g_z)/3 and g2 = (g_z - g_xy)/3).
static const int kLocalAxis[4][3] = {{1, 1, 1}, {1, 1, -1}, {1, -1, -1}, {1, -1, 1}};

// Bohr magneton in meV/T as the reference converts a field in tesla
// (Pyrochlore.cu:447-449).
constexpr double kMuBmeVperT = 0.0578838261305;

// Component (i, j) of the g-tensor of basis site b.
static double g_tensor(int b, int i, int j, double gxy, double gz) {
  return (i == j ? gxy : 0.0) + (gz - gxy) * (double)(kLocalAxis[b][i] * kLocalAxis[b][j]) / 3.0;
}

}  // namespace pyro

using namespace pyro;

extern "C" {

int pyro_exchange_matrices(double J1, double J2, double J3, double J4, double *C144) {
  if (!C144) return set_error(PYRO_EINVAL, "pyro_exchange_matrices: null output");
  const double J[5] = {0.0, J1, J2, J3, J4};
  std::memset(C144, 0, sizeof(double) * 144);
  for (const auto &e : kExchangeUpper) {
    const int b = e[0], bp = e[1];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        const int k = e[2 + 3 * i + j];
        const double v = k < 0 ? -J[-k] : J[k];
        C144[9 * (4 * b + bp) + 3 * i + j] = v;
        C144[9 * (4 * bp + b) + 3 * j + i] = v;  // C_{b' b} = C_{b b'}^T
      }
  }
  return 0;
}

int pyro_zeeman_fields(const double *B3, double gxy, double gz, double *B12) {
  if (!B3 || !B12) return set_error(PYRO_EINVAL, "pyro_zeeman_fields: null argument");
  for (int b = 0; b < 4; ++b)
    for (int i = 0; i < 3; ++i) {
      double t = 0.0;
      for (int j = 0; j < 3; ++j) t += g_tensor(b, i, j, gxy, gz) * B3[j];
      B12[3 * b + i] = t * kMuBmeVperT;
    }
  return 0;
}

int pyro_fold_magnetization_g(const double *sub12, int R, double gxy, double gz, double *M3) {
  if (!sub12 || !M3 || R < 0 || !std::isfinite(gxy) || !std::isfinite(gz))
    return set_error(PYRO_EINVAL, "pyro_fold_magnetization_g: bad argument");
  for (int r = 0; r < R; ++r) {
    const double *sb = sub12 + (size_t)12 * r;
    for (int i = 0; i < 3; ++i) {
      // MagnetizationGPU's order: basis sites 0..3 in turn, each a row of
      // its g-tensor dotted with that site's sum (Pyrochlore.cu:318-336).
      double m = 0.0;
      for (int b = 0; b < 4; ++b) {
        double t = 0.0;
        for (int j = 0; j < 3; ++j) t += g_tensor(b, i, j, gxy, gz) * sb[3 * b + j];
        m += t;
      }
      M3[(size_t)3 * r + i] = m;
    }
  }
  return 0;
}

int pyro_magnetization_g(pyro_ctx *ctx, double gxy, double gz, double *M3) {
  if (!ctx || !M3) return set_error(PYRO_EINVAL, "pyro_magnetization_g: null argument");
  std::vector<double> sub((size_t)ctx->R * 12);
  PYRO_TRY(pyro_sublattice_sums(ctx, sub.data()));
  return pyro_fold_magnetization_g(sub.data(), ctx->R, gxy, gz, M3);
}

int pyro_set_couplings(pyro_ctx *ctx, const double *C144, const double *B12) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  if (C144 && ctx->hetero)
    return set_error(PYRO_ESTATE, "pyro_set_couplings: exchange matrices need a uniform lattice size");
  if (C144 && ctx->exec_mode != 0)
    return set_error(PYRO_ESTATE, "pyro_set_couplings: exchange matrices run in execution mode 0 only");
  if (C144 && ctx->n_learned > 0)
    return set_error(PYRO_ESTATE, "pyro_set_couplings: the learned proposal is for the scalar model");
  float h[kCoupFloats] = {};
  if (C144) {
    double scale = 0.0;
    for (int q = 0; q < 144; ++q) {
      if (!std::isfinite(C144[q])) return set_error(PYRO_EINVAL, "pyro_set_couplings: C[%d] not finite", q);
      scale = std::fmax(scale, std::fabs(C144[q]));
    }
    for (int b = 0; b < 4; ++b)
      for (int bp = 0; bp < 4; ++bp)
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j)
            if (b != bp && std::fabs(C144[9 * (4 * b + bp) + 3 * i + j] -
                                     C144[9 * (4 * bp + b) + 3 * j + i]) > 1e-12 * scale)
              return set_error(PYRO_EINVAL,
                               "pyro_set_couplings: C[%d][%d] is not the transpose of C[%d][%d]", bp,
                               b, b, bp);
    if (B12)
      for (int q = 0; q < 12; ++q)
        if (!std::isfinite(B12[q])) return set_error(PYRO_EINVAL, "pyro_set_couplings: B[%d] not finite", q);
    for (int q = 0; q < 144; ++q) h[q] = (float)(ctx->spin_len2 * C144[q]);
    for (int q = 0; q < 12; ++q) h[144 + q] = B12 ? (float)(2.0 * ctx->spin_len2 * B12[q]) : 0.0f;
  }
  PYRO_HIP(hipSetDevice(ctx->device));
  PYRO_HIP(hipStreamSynchronize(ctx->stream));
  if (C144 && !ctx->d_coup) PYRO_HIP(hipMalloc((void **)&ctx->d_coup, sizeof(float) * kCoupFloats));
  if (C144) PYRO_HIP(hipMemcpy(ctx->d_coup, h, sizeof h, hipMemcpyHostToDevice));
  std::memcpy(ctx->h_coup, h, sizeof h);
  ctx->mat = C144 != nullptr;
  PYRO_TRY(reset_model_state(ctx));
  return 0;
}

int pyro_get_couplings(pyro_ctx *ctx, int *is_matrix, float *coup) {
  if (!ctx) return set_error(PYRO_EINVAL, "null ctx");
  if (is_matrix) *is_matrix = ctx->mat ? 1 : 0;
  if (coup) std::memcpy(coup, ctx->h_coup, sizeof ctx->h_coup);
  return 0;
}

}  // extern "C"
