/* CPU oracle for the pyrochlore Heisenberg Monte Carlo step.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference's
 * algorithm (taillefu/cudacode: Pyrochlore.cu, MonteCarlo.cu, Math.cu), used by
 * the tests and by the CPU-baseline leg of bench.py as the checker. The product
 * (cudacode_amd/libpyro.so) never links, loads or calls it.
 *
 * Conventions (shared with the GPU path so a trajectory is reproducible bit
 * for bit from the same seed):
 *   spins   FP32, layout [replica][component x,y,z][site 0..15][cell],
 *           cell = (z*L + y)*L + x  (the reference SoA, Math.cu:93,
 *           Pyrochlore.cu:843-845)
 *   energy  H = jeff * sum_<ij> s_i.s_j, jeff = J * spin_len2; local field
 *           h_i = -jeff * sum_j s_j
 *   RNG     Philox4x32-10, key = (seed lo, seed hi), counter =
 *           (cell, site | pass<<8 | kind<<24, replica, sweep)
 *
 * Parity status per function is listed in oracle/README.md.
 */
#ifndef PYRO_ORACLE_H
#define PYRO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

void orc_links16(int8_t out[16][6][4]);
void orc_independent_sets(int8_t out[24][4]);
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float orc_exp_neg(float x);
void orc_sincospi(float a, float *s, float *c);
void orc_cone_proposal(const float s[3], float cos_cone, float u_cos, float u_phi, float out[3]);
void orc_overrelax(const float h[3], const float s[3], float out[3]);
void orc_renormalise(const float v[3], float out[3]);
int orc_metropolis_accept(const float h[3], const float s[3], const float sp[3], float beta,
                          float u);

/* Uniform spins on the sphere from the Philox stream (kind = init). */
void orc_init_spins(float *spins, int L, int R, uint64_t seed);

/* One colour phase of Metropolis (cone proposal) on every replica.
 * beta[r], cos_cone[r] per replica. accepted[r] += accepts; dE[r] += FP64
 * energy change. Either of accepted / dE may be NULL. */
void orc_metropolis_phase(float *spins, int L, int R, int colour, uint32_t sweep, int pass,
                          const float *beta, const float *cos_cone, float jeff, uint64_t seed,
                          uint64_t *accepted, double *dE);

/* One colour phase of over-relaxation by pi on every replica. dE may be NULL. */
void orc_overrelax_phase(float *spins, int L, int R, int colour, float jeff, double *dE);

/* Phase `phase` of over-relaxation pass `pass` of sweep `sweep` under the
 * update schedule below (its set, and the rotation angle when random). */
void orc_overrelax_phase_ex(float *spins, int L, int R, int phase, uint32_t sweep, int pass,
                            float jeff, uint64_t seed, double *dE);

/* Update schedule for every following phase: sets = 1 draws the independent
 * set of each phase from the 24, sets = 2 runs the colour classes of each
 * pass in a drawn order (else phase c updates {c, c+4, c+8, c+12});
 * or_angle = 1 makes over-relaxation a rotation about h by a uniform random
 * angle (else by pi). orc_drawn_set / orc_drawn_colour: the set index or
 * colour of a phase (stage 0 Metropolis, 1 over-relaxation). */
void orc_set_update_schedule(int sets, int or_angle);
int orc_drawn_set(uint32_t sweep, int stage, int pass, int phase, uint64_t seed);
int orc_drawn_colour(uint32_t sweep, int stage, int pass, int phase, uint64_t seed);

/* Per-replica FP64 observables: out[r*13 + 0] = energy, out[r*13 + 1 + 3*b + k]
 * = sum over the N/4 spins of sublattice b (= site % 4) of component k. */
void orc_observables(const float *spins, int L, int R, float jeff, double *out);

/* Parallel-tempering exchange (MonteCarlo.cu:105-172). Rungs are ordered by
 * descending temperature; replica_at[rung] is the replica holding that rung.
 * Pairs (c, c+1) for c even (parity 0) or odd (parity 1) are proposed and
 * accepted with min(1, exp((1/T1 - 1/T2)(E1 - E2))); on acceptance the
 * temperatures (here: the rung assignment, beta and the cone) are exchanged.
 * temps[r] is replica r's temperature. Returns the number accepted. */
int orc_pt_swap(int R, int parity, uint32_t round, uint64_t seed, const double *energy,
                double *temps, float *beta, float *cos_cone, int *replica_at);

/* Local-field model for every following call: NULL = scalar Heisenberg
 * (-jeff * neighbour sum); else 156 floats, the exchange matrices (times
 * spin_len2, 9 per ordered basis pair) and Zeeman fields (times 2 spin_len2,
 * 3 per basis site). The array must outlive its use. */
void orc_set_couplings(const float *coup156);
void orc_magnetization_g(const double sub12[12], double gxy, double gz, double out[3]);
void orc_adapt_cone_step(int R, double target, const uint64_t *accepted,
                         const uint64_t *proposed, float *cos_cone);

/* Rotation of s about h by 2 pi u after each Metropolis test (Math.cu:168-197):
 * the move itself, and the switch for every following Metropolis phase. */
void orc_rotate_about(const float h[3], const float s[3], float u, float out[3]);
void orc_set_post_rotate(int on);

/* Visiting order of a colour phase's updates (cell * 4 + k; NULL = natural),
 * for the order-independence check. The array must outlive its use. */
void orc_set_update_order(const int64_t *perm);

/* Full sweeps: n_metro Metropolis passes + n_or over-relaxation passes, each
 * over colours 0..3; after every pt_every sweeps (if > 0) a swap with parity
 * alternating from the stored round counter. sweep0 is the first sweep index;
 * *pt_round is advanced. anchor optional (NULL): if given, the full energies
 * computed for each swap become the new anchor of the incremental energy
 * (anchor[r] = E_r, dE[r] = 0), as the device library does. */
void orc_run(float *spins, int L, int R, int n_sweeps, uint32_t sweep0, int n_metro, int n_or,
             int pt_every, uint32_t *pt_round, double *temps, float *beta, float *cos_cone,
             int *replica_at, float jeff, uint64_t seed, uint64_t *accepted, double *dE,
             double *anchor, int *n_swaps_accepted);

/* Learned proposal (pyro_oracle_mlp.c), FP64 throughout. env rows are
 * (u = beta*h (3), s (3)); weights are the 4868 packed parameters. */
int orc_mlp_nparams(void);
void orc_mlp_features(const double u[3], const double s[3], double x[6]);
void orc_mlp_forward(const double *w, const double x[6], double o[4]);
void orc_mlp_decode(const double o[4], double mu[3], double *kappa);
double orc_mlp_loss_grad(const double *w, const double *env, int n, double *grad);
double orc_vmf_log_norm(double kappa);
double orc_vmf_log_density(const double mu[3], double kappa, const double y[3]);
void orc_vmf_sample(const double mu[3], double kappa, double u1, double u2, double out[3]);

/* Replay part of a heterogeneous batch: base is added to the replica index of
 * every Metropolis/init counter; group replaces the replica word of the swap
 * counters (the batch swaps within groups of equal lattice size). Both 0 by
 * default. */
void orc_set_replica_base(int base);
void orc_set_swap_group(int group);

/* Threads used by the OpenMP loops (1 when built without OpenMP). */
int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
