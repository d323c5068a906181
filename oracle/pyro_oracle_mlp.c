/* CPU oracle for the learned local proposal (test infrastructure only).
 *
 * The reference has no learned component; this restates the build's own
 * definition so the GPU kernels have a checker:
 *   features  x = (u * log1p(|u|)/|u|, s) with u = beta * h  (6 floats)
 *   network   6 -> 64 -> 64 -> 4, ReLU, weights packed as
 *             W1[64][6] b1[64] W2[64][64] b2[64] W3[4][64] b3[4]  (4868 floats)
 *   output    mu = o[0:3]/|o[0:3]|, kappa = exp(clamp(o[3], ln kmin, ln kmax))
 *   proposal  von Mises-Fisher vMF(mu, kappa) on the unit sphere
 *   loss      KL(vMF(mu, kappa) || vMF(u/|u|, |u|)), the exact reverse KL to the
 *             conditional Boltzmann law p(s'|h) ~ exp(beta h.s'), averaged over
 *             environments (u, s) drawn by the plain sampler
 * All arithmetic here is FP64 (forward, backward, loss).
 */
#include <math.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#include "pyro_oracle.h"

#define H1 64
#define H2 64
#define NIN 6
#define NOUT 4
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + H1 * NIN)
#define OFF_W2 (OFF_B1 + H1)
#define OFF_B2 (OFF_W2 + H2 * H1)
#define OFF_W3 (OFF_B2 + H2)
#define OFF_B3 (OFF_W3 + NOUT * H2)
#define NPARAM (OFF_B3 + NOUT)

static const double LOG_KMIN = -4.605170185988091; /* ln 0.01 */
static const double LOG_KMAX = 8.317766166719343;  /* ln 4096 */

int orc_mlp_nparams(void) { return NPARAM; }

void orc_mlp_features(const double u[3], const double s[3], double x[6]) {
  double r = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  double g = r > 0.0 ? log1p(r) / r : 1.0;
  x[0] = u[0] * g;
  x[1] = u[1] * g;
  x[2] = u[2] * g;
  x[3] = s[0];
  x[4] = s[1];
  x[5] = s[2];
}

/* Forward pass; keeps pre-activations for the backward pass. */
static void forward(const double *w, const double x[6], double z1[H1], double a1[H1],
                    double z2[H2], double a2[H2], double o[NOUT]) {
  for (int j = 0; j < H1; j++) {
    double t = w[OFF_B1 + j];
    for (int k = 0; k < NIN; k++) t += w[OFF_W1 + j * NIN + k] * x[k];
    z1[j] = t;
    a1[j] = t > 0.0 ? t : 0.0;
  }
  for (int j = 0; j < H2; j++) {
    double t = w[OFF_B2 + j];
    for (int k = 0; k < H1; k++) t += w[OFF_W2 + j * H1 + k] * a1[k];
    z2[j] = t;
    a2[j] = t > 0.0 ? t : 0.0;
  }
  for (int q = 0; q < NOUT; q++) {
    double t = w[OFF_B3 + q];
    for (int k = 0; k < H2; k++) t += w[OFF_W3 + q * H2 + k] * a2[k];
    o[q] = t;
  }
}

void orc_mlp_forward(const double *w, const double x[6], double o[4]) {
  double z1[H1], a1[H1], z2[H2], a2[H2];
  forward(w, x, z1, a1, z2, a2, o);
}

/* A(k) = coth k - 1/k, the mean cosine of vMF(k) on the sphere, and A'(k). */
static double langevin(double k) {
  if (k < 1e-3) return k / 3.0 - k * k * k / 45.0;
  return 1.0 / tanh(k) - 1.0 / k;
}
static double langevin_d(double k) {
  if (k < 1e-3) return 1.0 / 3.0 - k * k / 15.0;
  double sh = sinh(k);
  if (k > 350.0) return 1.0 / (k * k);
  return 1.0 / (k * k) - 1.0 / (sh * sh);
}

/* log of the vMF normaliser C(k) = k / (4 pi sinh k). */
double orc_vmf_log_norm(double k) {
  if (k < 1e-6) return -log(4.0 * M_PI);
  return log(k) - log(2.0 * M_PI) - k - log(-expm1(-2.0 * k));
}

double orc_vmf_log_density(const double mu[3], double k, const double y[3]) {
  double c = mu[0] * y[0] + mu[1] * y[1] + mu[2] * y[2];
  return orc_vmf_log_norm(k) + k * c;
}

/* Outputs -> (mu, kappa); returns 1 if kappa was clamped. */
static int decode(const double o[4], double mu[3], double *kappa, double *norm) {
  double n = sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
  *norm = n;
  if (n > 0.0) {
    mu[0] = o[0] / n;
    mu[1] = o[1] / n;
    mu[2] = o[2] / n;
  } else {
    mu[0] = 0.0;
    mu[1] = 0.0;
    mu[2] = 1.0;
  }
  double lk = o[3];
  int clamped = 0;
  if (lk < LOG_KMIN) { lk = LOG_KMIN; clamped = 1; }
  if (lk > LOG_KMAX) { lk = LOG_KMAX; clamped = 1; }
  *kappa = exp(lk);
  return clamped;
}

void orc_mlp_decode(const double o[4], double mu[3], double *kappa) {
  double n;
  decode(o, mu, kappa, &n);
}

/* Reverse KL of one environment and its gradient wrt the outputs. */
static double kl_and_grad(const double o[4], const double u[3], double g[4]) {
  double mu[3], k, n;
  int clamped = decode(o, mu, &k, &n);
  double ks = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  double hat[3] = {0.0, 0.0, 0.0};
  if (ks > 0.0) {
    hat[0] = u[0] / ks;
    hat[1] = u[1] / ks;
    hat[2] = u[2] / ks;
  }
  double c = mu[0] * hat[0] + mu[1] * hat[1] + mu[2] * hat[2];
  double A = langevin(k);
  double kl = orc_vmf_log_norm(k) - orc_vmf_log_norm(ks) + A * (k - ks * c);
  /* dKL/dk = A'(k)(k - ks c); dKL/dmu = -A(k) ks hat */
  double dk = langevin_d(k) * (k - ks * c);
  g[3] = clamped ? 0.0 : dk * k;
  if (n > 0.0) {
    double gm[3] = {-A * ks * hat[0], -A * ks * hat[1], -A * ks * hat[2]};
    double proj = gm[0] * mu[0] + gm[1] * mu[1] + gm[2] * mu[2];
    for (int i = 0; i < 3; i++) g[i] = (gm[i] - proj * mu[i]) / n;
  } else {
    g[0] = g[1] = g[2] = 0.0;
  }
  return kl;
}

double orc_mlp_loss_grad(const double *w, const double *env, int n, double *grad) {
  if (grad) memset(grad, 0, sizeof(double) * NPARAM);
  double total = 0.0;
  for (int i = 0; i < n; i++) {
    const double *e = env + 6 * i;
    double x[6], z1[H1], a1[H1], z2[H2], a2[H2], o[NOUT], go[NOUT];
    orc_mlp_features(e, e + 3, x);
    forward(w, x, z1, a1, z2, a2, o);
    total += kl_and_grad(o, e, go);
    if (!grad) continue;
    double d2[H2], d1[H1];
    for (int q = 0; q < NOUT; q++) {
      grad[OFF_B3 + q] += go[q];
      for (int k = 0; k < H2; k++) grad[OFF_W3 + q * H2 + k] += go[q] * a2[k];
    }
    for (int k = 0; k < H2; k++) {
      double t = 0.0;
      for (int q = 0; q < NOUT; q++) t += w[OFF_W3 + q * H2 + k] * go[q];
      d2[k] = z2[k] > 0.0 ? t : 0.0;
    }
    for (int j = 0; j < H2; j++) {
      grad[OFF_B2 + j] += d2[j];
      for (int k = 0; k < H1; k++) grad[OFF_W2 + j * H1 + k] += d2[j] * a1[k];
    }
    for (int k = 0; k < H1; k++) {
      double t = 0.0;
      for (int j = 0; j < H2; j++) t += w[OFF_W2 + j * H1 + k] * d2[j];
      d1[k] = z1[k] > 0.0 ? t : 0.0;
    }
    for (int j = 0; j < H1; j++) {
      grad[OFF_B1 + j] += d1[j];
      for (int k = 0; k < NIN; k++) grad[OFF_W1 + j * NIN + k] += d1[j] * x[k];
    }
  }
  if (grad)
    for (int p = 0; p < NPARAM; p++) grad[p] /= n;
  return total / n;
}

/* vMF sample in 3-D (inverse CDF of the cosine, Wood 1994; uniform azimuth),
 * carried onto mu by the Householder reflection through the pole on mu's side. */
void orc_vmf_sample(const double mu[3], double k, double u1, double u2, double out[3]) {
  double v = 1.0 - u1; /* (0, 1] */
  double w = 1.0 + log1p((1.0 - v) * expm1(-2.0 * k)) / k;
  if (k < 1e-8) w = 2.0 * v - 1.0;
  if (w > 1.0) w = 1.0;
  if (w < -1.0) w = -1.0;
  double st = sqrt(1.0 - w * w);
  double phi = 2.0 * M_PI * u2;
  double r[3] = {st * cos(phi), st * sin(phi), w};
  double N[3] = {mu[0], mu[1], 0.0};
  double d;
  if (mu[2] >= 0.0) {
    N[2] = mu[2] + 1.0;
    d = 1.0 + mu[2];
  } else {
    r[2] = -w;
    N[2] = mu[2] - 1.0;
    d = 1.0 - mu[2];
  }
  double scal = (N[0] * r[0] + N[1] * r[1] + N[2] * r[2]) / d;
  for (int i = 0; i < 3; i++) out[i] = N[i] * scal - r[i];
  double nn = sqrt(out[0] * out[0] + out[1] * out[1] + out[2] * out[2]);
  for (int i = 0; i < 3; i++) out[i] /= nn;
}
