// A C++ host program driving the Monte Carlo step through the C ABI of
// include/pyro.h, in the order of the reference's main loop (Pyrochlore.cu:
// InitializeDevice + InitializeCouplings3D at :1251-1267, the parallel
// tempering loop of MonteCarlo.cu:429-470 around Mc, then TotalEnergyv2 and
// CopyResultsOnCPU): create a ctx on a geometric temperature ladder, adapt the
// proposal cones, run the sweeps (exchanges every pt_every sweeps inside
// pyro_sweep), and print the energy per spin by rung, the sweep and exchange
// counters, and an FNV-1a hash of the spins so that a run can be compared
// with any other caller's run of the same configuration.
//
// usage: pt_driver L R T_min T_max seed n_or pt_every sweeps
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "pyro.h"

static int fail(const char *what) {
  std::fprintf(stderr, "%s: %s\n", what, pyro_last_error());
  return 1;
}

int main(int argc, char **argv) {
  if (argc != 9) {
    std::fprintf(stderr, "usage: %s L R T_min T_max seed n_or pt_every sweeps\n", argv[0]);
    return 2;
  }
  const int L = std::atoi(argv[1]), R = std::atoi(argv[2]);
  const double t_min = std::atof(argv[3]), t_max = std::atof(argv[4]);
  const uint64_t seed = std::strtoull(argv[5], nullptr, 0);
  const int n_or = std::atoi(argv[6]), pt_every = std::atoi(argv[7]), sweeps = std::atoi(argv[8]);
  if (L < 2 || R < 1 || sweeps < 0) return 2;
  std::vector<double> T(R);
  for (int k = 0; k < R; ++k)  // T_k = T_min (T_max / T_min)^(k / (R - 1))
    T[k] = R == 1 ? t_min : t_min * std::pow(t_max / t_min, (double)k / (R - 1));

  for (int k = 0; k < R; ++k) std::printf("ladder %d %.17g\n", k, T[k]);

  pyro_config cfg = {};
  cfg.L = L;
  cfg.n_replicas = R;
  cfg.temperatures = T.data();
  cfg.seed = seed;
  cfg.J = 1.0;
  cfg.spin_len2 = 1.0;
  cfg.n_metro = 1;
  cfg.n_or = n_or;
  cfg.pt_every = pt_every;
  cfg.cos_cone = 0.5f;
  cfg.device = 0;
  pyro_ctx *ctx = nullptr;
  if (pyro_create(&cfg, &ctx)) return fail("pyro_create");
  int rc = 0;
  if (pyro_adapt_cone(ctx, 0.5, 5, 10)) rc = fail("pyro_adapt_cone");
  if (!rc && pyro_sweep(ctx, sweeps, 1)) rc = fail("pyro_sweep");

  std::vector<double> E(R), Tr(R);
  std::vector<int> at(R);
  uint32_t n_sweep = 0, n_round = 0;
  if (!rc && (pyro_energy(ctx, E.data()) || pyro_get_temperatures(ctx, Tr.data(), at.data()) ||
              pyro_get_sweep(ctx, &n_sweep, &n_round)))
    rc = fail("readback");
  const size_t n_floats = (size_t)R * 48 * L * L * L;
  std::vector<float> spins(n_floats);
  if (!rc && pyro_get_spins(ctx, spins.data())) rc = fail("pyro_get_spins");
  if (!rc) {
    uint64_t h = 1469598103934665603ull;  // FNV-1a over the bytes of the spins
    const unsigned char *b = reinterpret_cast<const unsigned char *>(spins.data());
    for (size_t i = 0; i < n_floats * sizeof(float); ++i) h = (h ^ b[i]) * 1099511628211ull;
    const double N = 16.0 * L * L * L;
    std::printf("counters %u %u\n", n_sweep, n_round);
    for (int k = 0; k < R; ++k)  // rungs by descending temperature
      std::printf("rung %d T %.6f E/N %.12f\n", k, Tr[at[k]], E[at[k]] / N);
    std::printf("spins_fnv1a %016llx\n", (unsigned long long)h);
  }
  pyro_destroy(ctx);
  return rc;
}
