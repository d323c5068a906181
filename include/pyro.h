/* pyro.h -- C ABI of the MI355X-native pyrochlore Heisenberg Monte Carlo step.
 *
 * The reference (taillefu/cudacode) has no plugin or FFI layer: its boundary is
 * a set of host C++ functions inside one unity build that take a
 * std::vector<dev>& (Pyrochlore.cu, MonteCarlo.cu, Math.cu). Each entry point
 * below replaces one of them; the comment names the reference function and
 * its file:line.
 *
 * Conventions
 *   - Every function returns int: 0 on success, a negative PYRO_E* code on
 *     failure; pyro_last_error() returns a thread-local message. Nothing calls
 *     exit(). (The reference ignores CUDA status codes, printf's and goes on,
 *     or calls exit(): MonteCarlo.cu:5-7,221-224, Math.cu:676-680.)
 *   - One pyro_ctx per GPU; a ctx is not thread-safe. Calls enqueue on the
 *     ctx's HIP stream; functions that return host data synchronise it.
 *   - Spins are FP32 unit vectors, layout [replica][x,y,z][site 0..15][cell],
 *     cell = (z*L + y)*L + x: the reference SoA (Math.cu:93,
 *     Pyrochlore.cu:843-845). All replicas stay resident in HBM.
 *   - Energy H = J * spin_len2 * sum_<ij> s_i.s_j. spin_len2 = 1 is the unit
 *     scale; spin_len2 = 0.25 is the reference's spin-1/2 scale
 *     (Pyrochlore.cu:254-262,388; T_ref = T / 4).
 */
#ifndef PYRO_H
#define PYRO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PYRO_OK 0
#define PYRO_EINVAL (-1)   /* bad argument */
#define PYRO_EHIP (-2)     /* HIP runtime error */
#define PYRO_ENOMEM (-3)   /* device or host allocation failed */
#define PYRO_ESTATE (-4)   /* call not valid in the ctx's current state */
#define PYRO_EIO (-5)      /* file missing, short, corrupt or not matching the ctx */

typedef struct pyro_ctx pyro_ctx;

typedef struct pyro_config {
  int L;                      /* cubic cells per side; N = 16 L^3 spins per replica */
  int n_replicas;             /* R, all resident on the one GPU */
  const double *temperatures; /* R temperatures, replica r starts at temperatures[r] */
  uint64_t seed;              /* Philox key */
  double J;                   /* exchange, > 0 antiferromagnetic */
  double spin_len2;           /* 1 (unit scale) or 0.25 (reference scale) */
  int n_metro;                /* Metropolis passes per sweep */
  int n_or;                   /* over-relaxation passes per sweep */
  int pt_every;               /* sweeps between parallel-tempering swaps, 0 = off */
  float cos_cone;             /* initial cosine of the proposal cone, -1 = whole sphere */
  int device;                 /* HIP device ordinal */
} pyro_config;

/* Version and error text. */
const char *pyro_version(void);
const char *pyro_last_error(void);

/* Create / destroy a ctx: allocates all device state. Replaces InitializeDevice
 * (missing glue, called at Pyrochlore.cu:1251) and InitializeCouplings3D
 * (Pyrochlore.cu:643-711). */
int pyro_create(const pyro_config *cfg, pyro_ctx **out);
int pyro_destroy(pyro_ctx *ctx);

/* A batch of replicas of different lattice sizes in one ctx (no reference
 * counterpart: the reference runs one size per process, Pyrochlore.cu:1251).
 * L[r] in {4, 8, 12, 16, 24, 32}. Replica r's spins occupy 48*L[r]^3 floats
 * at offset pyro_get_layout()[r] of the spin buffer, in the uniform layout.
 * Parallel tempering exchanges within groups of equal L (rungs grouped by
 * ascending L, each group by descending T). Every function of this header
 * accepts a batch ctx except the single-phase, learned-proposal and
 * phase-timing calls, which need one lattice size (PYRO_ESTATE). */
typedef struct pyro_batch_config {
  int n_replicas;
  const int *L;               /* n_replicas lattice sizes */
  const double *temperatures; /* n_replicas temperatures */
  uint64_t seed;
  double J;
  double spin_len2;
  int n_metro, n_or, pt_every;
  float cos_cone;
  int device;
} pyro_batch_config;
int pyro_create_batch(const pyro_batch_config *cfg, pyro_ctx **out);
int pyro_get_layout(pyro_ctx *ctx, int *L, long long *offset);

/* Sweep execution. mode 0: per-phase kernels (uniform ctx only; the
 * reference's launch structure, Pyrochlore.cu:105-127); mode 1: one
 * persistent kernel per run of sweeps up to the next exchange or measurement,
 * in which a block of `block` threads (256, 512, 1024) sweeps whole replicas
 * through every phase with block barriers only (replicas of L <= 8 staged in
 * LDS), replicas packed into n_bins balanced bins (<= 0: one per block);
 * mode 2: per-phase kernels over a list of tiles of `block` cells (128, 256
 * or 512; 0 means 256; a block of threads per tile) of all replicas.
 * Modes 1 and 2 need L in {4,8,12,16,24,32}. Results are identical in every
 * mode and tile size. Default: 0 for a uniform ctx, 2 (256-cell tiles) for a
 * batch. pyro_get_exec_mode returns the persistent block size in mode 1 and
 * the tile size in mode 2. */
int pyro_set_exec_mode(pyro_ctx *ctx, int mode, int block, int n_bins);
int pyro_get_exec_mode(pyro_ctx *ctx, int *mode, int *block, int *n_bins);

/* Autotuner: time `sweeps` sweeps of each candidate execution (per-phase
 * cell- and site-per-thread kernels on 1, 2 or 3 streams, and cache-blocked
 * in 2, 4, 8, 16 or 32 replica groups, 1, 2 or 3 of them at a time;
 * persistent kernel at 256/512/1024 threads
 * when every replica has L <= 8 and is staged in LDS -- a block owns whole
 * replicas, so larger ones serialise: 5.9-12 ms per sweep for 256 mixed
 * replicas of L up to 24 against 1.58 ms as tile lists -- and tile lists of
 * 128, 256 and 512 cells on 1, 2 or 3 streams, and of 256 cells
 * cache-blocked in 2, 3, 4, 6 or 8 groups, 1, 2 or 3 at a time;
 * pyro_set_group_blocking), captured in
 * one graph as pyro_sweep replays them, restoring the state after each trial,
 * and keep the fastest. results (optional, 6 doubles per candidate, at most
 * PYRO_AUTOTUNE_MAX candidates): mode, block (mode 1) or tile size (mode 2),
 * phase variant, streams, cache-blocking groups (0: none), ms per sweep. */
#define PYRO_AUTOTUNE_MAX 64
int pyro_autotune(pyro_ctx *ctx, int sweeps, double *results, int *n_results);

/* Lattice tables: links16 (Pyrochlore.cu:152-167) and IndependentSubLattices
 * (Pyrochlore.cu:36-43), as (dx, dy, dz, site) and 24 sets of 4 sites. */
int pyro_links16(int8_t out[16][6][4]);
int pyro_independent_sets(int8_t out[24][4]);

/* Philox4x32-10 block function (replaces the cuRAND fill, MonteCarlo.cu:3-13,
 * and the MKL MT19937 stream, MonteCarlo.cu:116). Host-side. */
int pyro_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Spins. init: uniform on the sphere from the Philox stream
 * (GenerateRandomVectorsOnDevice, MonteCarlo.cu:257-293; Generate3DVectorList,
 * Math.cu:132-143). set/get copy R*48*L^3 floats host<->device
 * (CopyResultsOnGPU / CopyResultsOnCPU, missing glue, Pyrochlore.cu:1530,1766).
 * device_ptr exposes the resident buffer (no copy; it holds the state between
 * calls: fused over-relaxation passes use a second buffer only inside a
 * sweep). All reset the energy anchor. */
int pyro_init_random_spins(pyro_ctx *ctx);
int pyro_set_spins(pyro_ctx *ctx, const float *soa);
int pyro_get_spins(pyro_ctx *ctx, float *soa);
int pyro_spins_device_ptr(pyro_ctx *ctx, float **dev_ptr, size_t *n_floats);

/* Temperatures and proposal cones per replica (the host vectors the reference
 * swaps in TrySwapingTemperatures and copies with cudaMemcpyAsync,
 * MonteCarlo.cu:94-98,207-210). replica_at[rung] lists replicas by descending
 * temperature (the gsl_sort_largest_index order, MonteCarlo.cu:120). */
int pyro_set_temperatures(pyro_ctx *ctx, const double *T);
int pyro_get_temperatures(pyro_ctx *ctx, double *T, int *replica_at);
int pyro_set_cos_cone(pyro_ctx *ctx, const float *cos_cone);
int pyro_get_cos_cone(pyro_ctx *ctx, float *cos_cone);

/* Single colour phases (colour = site % 4 in 0..3). sweep and pass key the
 * Philox counters. MCMultipleTemperaturesPerSystem (Pyrochlore.cu:829-885)
 * and OverRelaxationByPiStep (Pyrochlore.cu:975-1007). */
int pyro_metropolis_phase(pyro_ctx *ctx, int colour, uint32_t sweep, int pass);
int pyro_overrelax_phase(pyro_ctx *ctx, int colour);

/* n_sweeps full sweeps (n_metro Metropolis + n_or over-relaxation passes, each
 * over colours 0..3, and a parallel-tempering swap after every pt_every-th
 * sweep). Replaces Mc (Pyrochlore.cu:74-133) driven by MonteCarlo /
 * ParallelTempering (MonteCarlo.cu:176-228,429-470), without Mc's 1000-sweep
 * cap. use_graph != 0 replays a captured hipGraph per block of sweeps. */
int pyro_sweep(pyro_ctx *ctx, int n_sweeps, int use_graph);
int pyro_get_sweep(pyro_ctx *ctx, uint32_t *sweep, uint32_t *pt_round);

/* Sweeps per captured graph of pyro_sweep: a whole number of exchange and
 * measurement intervals, at least 16. A call replays the graph for each whole
 * block that starts at a multiple of it, and runs the sweeps before the first
 * such start and after the last block one launch at a time (no reference
 * counterpart). */
int pyro_get_graph_block(pyro_ctx *ctx, int *sweeps);
int pyro_set_sweep(pyro_ctx *ctx, uint32_t sweep, uint32_t pt_round);

/* One parallel-tempering exchange with the given parity on the current
 * energies (TrySwapingTemperatures, MonteCarlo.cu:105-172). round keys the
 * Philox draw. */
int pyro_pt_swap(pyro_ctx *ctx, int parity, uint32_t round, int *n_accepted);

/* Observables, FP64, per replica (replica order). energy: R doubles
 * (TotalEnergyv2, Math.cu:309-330); magnetization: R*3 (Magnetization,
 * Pyrochlore.cu:453-466, g = 1; pyro_magnetization_g weighs by the
 * g-tensors); order_params: R*12 q=0 components
 * normalised so the ideal ordered state has modulus 1 (Measurement +
 * DoMeasurement, Pyrochlore.cu:526-641,1081-1092); sublattice: R*12 raw
 * sublattice sums (b, xyz). */
int pyro_energy(pyro_ctx *ctx, double *E);
int pyro_magnetization(pyro_ctx *ctx, double *M3);
int pyro_order_params(pyro_ctx *ctx, double *op12);
int pyro_sublattice_sums(pyro_ctx *ctx, double *sub12);

/* Host-only fold of R sublattice-sum rows (R*12) into the 12 order
 * parameters (R*12), divided by n_spins: the arithmetic of pyro_order_params
 * without the device reduction. */
int pyro_fold_order_params(const double *sub12, int R, double n_spins, double *op12);

/* Incremental FP64 energy: the last full energy plus every accepted change
 * since (no reference counterpart; checked against pyro_energy). */
int pyro_energy_incremental(pyro_ctx *ctx, double *E);

/* Metropolis acceptance counters since the last reset (CalculateRatio,
 * MonteCarlo.cu:346-367; the reference kernels never fill them). */
int pyro_acceptance(pyro_ctx *ctx, uint64_t *accepted, uint64_t *proposed);
int pyro_reset_acceptance(pyro_ctx *ctx);

/* Cone adaptation toward a target acceptance (CalculateSolidAngle,
 * MonteCarlo.cu:371-426, which shrinks the angle by 0.9 per round): rounds of
 * sweeps_per_round sweeps, the half-angle of each replica scaled by 0.9 or
 * 1/0.9 toward `target`. Thermalisation only. */
int pyro_adapt_cone(pyro_ctx *ctx, double target, int rounds, int sweeps_per_round);

/* Energy time series, one row per sweep in rung order (incremental energy of
 * the replica holding each rung), for autocorrelation analysis. capacity rows
 * are kept; recording stops when full. */
int pyro_record_start(pyro_ctx *ctx, int capacity);
int pyro_record_fetch(pyro_ctx *ctx, double *rows, int *n_rows);

/* Anisotropic exchange and Zeeman field: the matrix local-field model
 * (InitializeCouplings3D, Pyrochlore.cu:643-711; couplings_const :169-184;
 * the field of CalculateLocalExchangeField :210-275; the energy of the CPU
 * TotalEnergy :468-523). C144[9 (4 b + b') + 3 i + j] is the 3x3 exchange
 * matrix C_{b b'} between basis sites b != b' (C_{b' b} must be C_{b b'}^T;
 * C_{b b} is unused); B12[3 b + i] the Zeeman field on basis site b (NULL:
 * none). Energy H = -spin_len2 (sum_<ij> s_i^T C s_j + 2 sum_i B_b . s_i) and
 * local field h_i = spin_len2 (sum_j C s_j + 2 B_b), so C = -J I, B = 0 is
 * the scalar model of pyro_create. The colour phases and observables use it
 * from this call on (FP32 field, FP64 energies); the incremental energy is
 * re-anchored. C144 = NULL returns to the scalar model. Uniform lattices in
 * execution mode 0 only; not with the learned proposal (PYRO_ESTATE). */
int pyro_set_couplings(pyro_ctx *ctx, const double *C144, const double *B12);
/* is_matrix: 1 when the matrix model is on; coup156: the FP32 values the
 * kernels use (spin_len2 C, then 2 spin_len2 B). Either may be NULL. */
int pyro_get_couplings(pyro_ctx *ctx, int *is_matrix, float *coup156);
/* The exchange matrices of the J1..J4 parameterisation (couplings_const,
 * Pyrochlore.cu:169-184; the reference's defaults J1 = -0.09/0.29,
 * J2 = -0.22/0.29, J3 = -1, J4 = 0 at :139-142). J1 = J2 = 1, J3 = J4 = 0
 * gives C = -I, the Heisenberg antiferromagnet. */
int pyro_exchange_matrices(double J1, double J2, double J3, double J4, double *C144);
/* Zeeman fields per basis site for a field B3 in tesla and a g-tensor with
 * components g_xy, g_z about each site's local [111] axis, in meV
 * (InitializeMagneticField, Pyrochlore.cu:414-450). */
int pyro_zeeman_fields(const double *B3, double gxy, double gz, double *B12);

/* Magnetisation weighted by the g-tensor of each basis site, M = sum_b g_b
 * s_b over the sublattice sums s_b, with g_b = g_xy I + (g_z - g_xy) n_b
 * n_b^T about the site's local [111] axis n_b: MagnetizationGPU
 * (Pyrochlore.cu:277-342, its g1..g4 built from g1 = (2 g_xy + g_z)/3 and
 * g2 = (g_z - g_xy)/3 as InitializeMagneticField does, :416-420) followed by
 * its sum over cells (Magnetization, :453-466). M3: R*3 doubles, replica
 * order; g_xy = g_z = 1 gives pyro_magnetization. The fold takes R rows of
 * pyro_sublattice_sums and needs no device. */
int pyro_magnetization_g(pyro_ctx *ctx, double gxy, double gz, double *M3);
int pyro_fold_magnetization_g(const double *sub12, int R, double gxy, double gz, double *M3);

/* Simulated annealing (SimulatedAnnealing / SimulatedAnnealingThread,
 * MonteCarlo.cu:68-104,295-344): `steps` steps; at step m = 0..steps-1 every
 * replica runs `sweeps_per_step` sweeps at T0 - (T0 - T_r) m / steps, where
 * T_r is its current temperature (the reference starts at T0 = 20 and, like
 * here, stops one step above the targets). No exchanges or measurements
 * during the schedule; afterwards beta = 1/T_r again, the cadences are back
 * and the incremental energy is re-anchored. */
int pyro_anneal(pyro_ctx *ctx, double T_start, int steps, int sweeps_per_step);

/* Measurement series (the per-measurement accumulation of the reference's
 * measurement loop, Pyrochlore.cu:1604-1672, and DoMeasurement :1081-1092):
 * after every `every` sweeps (counted from sweep 0; the first row comes after
 * the first sweep whose index + 1 is a multiple of `every`) the device
 * computes the full FP64 energy and the sublattice sums of every replica and
 * appends one row of PYRO_MEAS_Q doubles per rung, in rung order (descending
 * T), before that sweep's exchange:
 *   E, |m_A2|, |m_E|, |m_T1|, |m_F| (= |M| / N), |m_T2|
 * (order parameters per spin, pyro_fold_order_params). No host round trip:
 * the rows are written inside the captured sweeps. capacity rows are kept;
 * measuring stops when full. rows: n_rows x R x PYRO_MEAS_Q (fetch with
 * rows = NULL to get n_rows). The energies also re-anchor the incremental
 * energy. pyro_measure_stop turns the cadence off and keeps the rows. */
#define PYRO_MEAS_Q 6
int pyro_measure_start(pyro_ctx *ctx, int every, int capacity);
int pyro_measure_stop(pyro_ctx *ctx);
int pyro_measure_fetch(pyro_ctx *ctx, double *rows, int *n_rows);

/* Change the per-sweep schedule (Metropolis and over-relaxation passes; the
 * reference's --McSteps/--OverRelaxation options, Pyrochlore.cu:1161-1216). */
int pyro_set_passes(pyro_ctx *ctx, int n_metro, int n_or);

/* The phases of n replica groups [p R / n, (p + 1) R / n), p < n, can run
 * on n streams (1 <= n <= 4, n <= R) that join at exchanges, measurements
 * and the end of each call, in execution modes 0 and 2; n = 1 runs every
 * replica on the ctx stream. Spins, rungs and counters are identical for
 * every n; energies agree to FP64 rounding (no reference counterpart: a
 * performance setting; each group's kernels fill the GPU while another's
 * drain, hiding part of each short phase kernel's ramp). Default: 3 when
 * group 0 of 3 spans >= 256 workgroups of 256 cells, else 2 when the first
 * half does, else 1; pyro_autotune tries 1, 2 and 3, and 4 where the whole
 * over-relaxation pass runs (L = 16, untracked over-relaxation, fusion level
 * 2). Four streams share the default four hardware queues with the runtime
 * and time erratically: faster than three on some GPUs, slower on others.
 * Learned passes take the same split. */
int pyro_set_streams(pyro_ctx *ctx, int n);
int pyro_get_streams(pyro_ctx *ctx, int *n);

/* Cache blocking of the sweeps: with groups >= 2 each exchange or
 * measurement interval runs group after group on the ctx stream, replica
 * group p ([p R / groups, (p + 1) R / groups)) through all of the interval's
 * sweeps before group p + 1 starts, so that a group whose two spin buffers
 * fit the 256 MB Infinity Cache keeps them there across the interval.
 * groups = 0 (or 1): off, the default. Up to R groups in execution mode 0,
 * up to 8 in the tile-list execution (mode 2); PYRO_ESTATE in the persistent
 * execution. With n > 1 streams (pyro_set_streams), stream s takes groups
 * s, s + n, ... in turn, so n groups are in flight at a time. Spins, rungs, counters and measurement rows are identical to an
 * unblocked run; energies agree to FP64 rounding (no reference counterpart: a
 * performance setting; the reference sweeps every replica in each launch). */
int pyro_set_group_blocking(pyro_ctx *ctx, int groups);
int pyro_get_group_blocking(pyro_ctx *ctx, int *groups);

/* Over-relaxation colours per launch: on = 0 one, on = 1 two, on = 2
 * (default) a whole pass where a kernel for it exists (L = 16 and L = 64
 * without over-relaxation energy tracking), two elsewhere. Fusion applies to
 * execution mode 0 with the cell-per-thread kernels on a uniform lattice, or
 * the tile-list execution mode 2; scalar model. A fused launch reads the
 * spins from one buffer and writes the updated colours to a second one
 * (allocated with the ctx), recomputing the neighbouring values of earlier
 * colours that later ones need, so a pass moves 36 N (two colours a launch)
 * or 24 N (a whole pass) instead of 60 N bytes; results are identical (no
 * reference counterpart: a performance setting). Any on > 0 also fuses the
 * Metropolis phases two colours per launch at L = 16 (scalar model, no
 * post-Metropolis rotation), in place: identical spins and acceptance
 * counts, energies to FP64 rounding. PYRO_EINVAL for on outside 0..2. */
int pyro_set_or_fusion(pyro_ctx *ctx, int on);

/* Spin guard (no reference counterpart; the reference checks nothing): counts
 * the spins that are not finite or whose squared norm differs from 1 by more
 * than tol (FP32 updates keep it within a few 1e-7), and returns the
 * smallest such spin index (-1 if none). Spin i of replica r has index
 * offset_r / 3 + i, offset_r being the replica's float offset in get_spins
 * (pyro_get_layout; r * 48 L^3 for a uniform lattice). */
int pyro_check_spins(pyro_ctx *ctx, double tol, long long *n_bad, long long *first_bad);

/* Optional move after every Metropolis test, accepted or not: rotate the spin
 * about its local field by a Philox-drawn angle, uniform in [0, 2 pi)
 * (_CudaRotate3DVector, Math.cu:168-197, applied at Pyrochlore.cu:872; the
 * reference's angle works out to 2 pi^2 u). It keeps the local energy, so the
 * Boltzmann law is unchanged. Off by default; execution mode 0 on a uniform
 * lattice only. Acceptance counts are those of the Metropolis test. */
int pyro_set_post_rotate(pyro_ctx *ctx, int on);

/* Update schedule of the reference's other kernels. independent_sets: each
 * phase of a pass updates one of the 24 independent sets
 * (pyro_independent_sets), drawn per (sweep, pass, phase) from the Philox
 * stream and shared by every replica, as the reference launches
 * MCMultipleTemperaturesPerSystem / OverRelaxationByPiStep on
 * IndependentSubLattices[Update[0] % 24] (Pyrochlore.cu:848, :985, :1021;
 * its Update buffer is never filled, Pyrochlore.cu:84). A set may repeat
 * within a pass, so a pass need not touch every site; the Boltzmann law is
 * kept. independent_sets = 2: the four colour classes {c, c+4, c+8, c+12}
 * run in an order drawn per (sweep, pass) from the Philox stream, so every
 * site moves once a pass. Default 0: phase c updates {c, c+4, c+8, c+12}.
 * or_random_angle: over-relaxation rotates each spin about its local field
 * by a Philox-drawn angle uniform in [0, 2 pi) (OverRelaxationStep,
 * Pyrochlore.cu:1009-1045) instead of by pi (OverRelaxationByPiStep,
 * Pyrochlore.cu:975-1007, the default); both keep the energy. Either option
 * runs the one-phase-per-launch kernels; execution mode 0 on a uniform
 * lattice only. Learned passes keep the colour classes. */
int pyro_set_update_schedule(pyro_ctx *ctx, int independent_sets, int or_random_angle);

/* Kernel options (no reference counterpart; the reference's block shape is the
 * --BlockSpin option, Pyrochlore.cu:1152-1154). phase_variant 1 (default): one
 * thread per cell and the colour's 4 sites, for L in {4,8,12,16,24,32,64};
 * 0: one thread per site, any L. track_or_energy: also account the FP64
 * energy change of over-relaxation updates (zero up to FP32 rounding) in the
 * incremental energy; off by default. Results are identical either way. */
int pyro_set_kernel_options(pyro_ctx *ctx, int phase_variant, int track_or_energy);

/* ---- Learned local proposal (no reference counterpart) -------------------
 * A 6 -> 64 -> 64 -> 4 ReLU network maps the environment (u = beta*h, s) of a
 * spin to a von Mises-Fisher proposal vMF(mu, kappa); an exact
 * Metropolis-Hastings test keeps the Boltzmann law stationary. 4868 FP32
 * parameters packed W1[64][6] b1[64] W2[64][64] b2[64] W3[4][64] b3[4];
 * inference on BF16 MFMA with FP32 accumulation, training in FP32. */
int pyro_mlp_set_weights(pyro_ctx *ctx, const float *w);
int pyro_mlp_get_weights(pyro_ctx *ctx, float *w);
/* rows9: n x (u[3], s[3], sp[3]); out9: n x (o(u,s)[4], o(u,sp)[4],
 * log MH ratio with h = u and beta = 1), through the phase kernel's code. */
int pyro_mlp_eval(pyro_ctx *ctx, const float *rows9, int n, float *out9);
/* Device vMF sampler: rows4 n x (mu[3], kappa) -> out3 n x 3. */
int pyro_vmf_draw(pyro_ctx *ctx, const float *rows4, int n, float *out3);
/* Mean loss (reverse KL to the Boltzmann conditional) and gradient over n
 * environments env6 = (u[3], s[3]) rows, FP32 training kernels. */
int pyro_mlp_loss_grad(pyro_ctx *ctx, const float *env6, int n, double *loss, float *grad);
/* Environments of the current configuration: per_replica Philox-chosen sites
 * of every replica appended to the device set (capacity rows kept). */
int pyro_gather_environments(pyro_ctx *ctx, int per_replica, uint32_t draw, int capacity,
                             int *n_total);
int pyro_get_environments(pyro_ctx *ctx, float *env6, int *n);
int pyro_clear_environments(pyro_ctx *ctx);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected; moments
 * reset by pyro_mlp_set_weights) with a caller's gradient, e.g. from
 * pyro_mlp_loss_grad; the MFMA weight images are repacked. */
int pyro_mlp_adam_step(pyro_ctx *ctx, const float *grad, float lr);
/* Adam on minibatches drawn from the environment set; learning rate decays
 * geometrically from lr to lr_final. loss_hist (optional) receives the mean
 * batch loss of each step. */
int pyro_mlp_train(pyro_ctx *ctx, int steps, int batch, float lr, float lr_final, float *loss_hist);
/* One colour phase of learned proposals; learned passes per sweep. */
int pyro_learned_phase(pyro_ctx *ctx, int colour, uint32_t sweep, int pass);
int pyro_set_learned_passes(pyro_ctx *ctx, int n_learned);
int pyro_acceptance_learned(pyro_ctx *ctx, uint64_t *accepted, uint64_t *proposed);

/* Checkpoint / resume (own binary format, FNV-1a checksummed): spins,
 * temperatures, cones, rung assignment, sweep and exchange counters, seed,
 * schedule, energy bookkeeping, acceptance counters and the learned
 * proposal's weights. A resumed run continues bit for bit (the random
 * numbers are counter-based). Load checks that the file matches the ctx's
 * replica count and sizes; a short or corrupt file is PYRO_EIO and leaves
 * the ctx unchanged. (The reference's resume path reads InitialConf files,
 * Pyrochlore.cu:1528-1552.) */
int pyro_checkpoint_save(pyro_ctx *ctx, const char *path);
int pyro_checkpoint_load(pyro_ctx *ctx, const char *path);

/* The reference's configuration files (SaveConfigurations /
 * ReadConfigurations, Math.cu:618-701): "<prefix><k>.dat" per temperature
 * rank k (descending T), each record 3N values in AoS order (s0x s0y s0z
 * s1x ...), 8-byte (the reference's TYPE=double) or 4-byte values, one
 * record appended per call when append != 0. Import reads record `rep`; a
 * missing file or a short read is PYRO_EIO (the reference exits or returns 1,
 * Math.cu:676-689). Uniform lattice size only. */
int pyro_export_configurations(pyro_ctx *ctx, const char *prefix, int append, int bytes_per_value);
int pyro_import_configurations(pyro_ctx *ctx, const char *prefix, int rep, int bytes_per_value);

/* Time n_launches back-to-back colour-phase launches of one kind (0 =
 * Metropolis, 1 = over-relaxation, 2 = learned proposal, which needs weights
 * set; colours cycle 0..3; 3 = fused over-relaxation pairs (0,1), (2,3)
 * alternating; 4 = fused Metropolis pairs (0,1), (2,3) alternating, where
 * the sweeps use them; 5 = whole over-relaxation passes, where the sweeps
 * use them) on the ctx stream,
 * each launch carrying its own start/stop HIP events in its dispatch
 * (hipExtLaunchKernelGGL). avg_ms: mean kernel time per launch; span_ms:
 * first start to last stop. Kinds 0-2 and 4 run for real (the state
 * advances) but all at the ctx's current sweep index, so a colour's launches
 * reuse one set of random draws: the state they leave is not an equilibrium
 * sample, and measurements should follow some sweeps later. Kinds 3 and 5
 * write only the second buffer and leave the state as it was. The reference's
 * counterpart is the ENABLE_TIMING cudaEvent block around Mc
 * (MonteCarlo.cu:36-63). */
int pyro_time_phase(pyro_ctx *ctx, int kind, int n_launches, double *avg_ms, double *span_ms);

/* Energies and sublattice sums of R replicas of L^3-cell spins that the caller
 * holds in device memory (uniform layout, scalar model, jeff = J * spin_len2),
 * into the caller's device buffer obs13: per replica E, then the 12
 * sublattice sums (b, xyz) -- the values pyro_energy and pyro_sublattice_sums
 * return for a ctx holding the same spins, bit for bit. Enqueued on `stream`
 * (a hipStream_t; NULL: the null stream) of `device`, nothing synchronised,
 * nothing allocated: the partial sums go to a fixed scratch array in the
 * library, one per GPU, of at most 4 MiB (R * 4 * ceil(L^3/256) * 4 doubles;
 * L = 64 x 32 replicas fits exactly); larger requests are PYRO_EINVAL before
 * any device work. Calls that overlap on one GPU race on that scratch. The
 * scratch address is looked up once per device, at pyro_create or at the
 * first call on that device: a caller that captures graphs makes that first
 * call (or creates a ctx) before capturing.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops). */
int pyro_observables_noalloc(int device, const float *spins, int L, int R, float jeff,
                             double *obs13, void *stream);

/* Synchronise the ctx stream. */
int pyro_sync(pyro_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
