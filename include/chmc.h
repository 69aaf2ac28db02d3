/*
 * chmc.h -- C ABI of the MI355X Lennard-Jones HMC engine (libchmc.so).
 *
 * This is the only door to device code.  All entry points are extern "C", take plain pointers and
 * sizes, return 0 on success and a negative CHMC_E_* code on failure (message via
 * chmc_last_error), never throw, and never hand device pointers to the caller.  Host buffers are
 * caller-owned; device memory is owned by the context; each context owns one HIP stream and is not
 * thread-safe (one context per host thread).
 *
 * A context holds a batch of independent chains ("systems"), each with its own N, box, cutoff,
 * LJ parameters, mass and beta, its own cell grid and its own Metropolis decision; every entry
 * point below acts on all chains at once (one launch per stage for the whole batch).  With one
 * chain it is the reference's single `particles_type`.
 *
 * Units and conventions follow the reference (code/modules/identical_particles.cpp):
 *   forces hold +grad U (Q1, :499-509, :613-633); kicks subtract p -= dt*beta*f (:686-688);
 *   drifts add x += dt*(beta/m)*p and wrap x -= L*floor(x/L) (:665-667); time is beta-scaled.
 * Positions/momenta/forces cross the ABI as AoS double[3] per particle (the layout of
 * code-carsten/config.hh config<coordinate>); inside the context they are SoA in HBM.
 */
#ifndef CHMC_H
#define CHMC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct chmc_ctx chmc_ctx;

enum {
    CHMC_OK = 0,
    CHMC_E_ARG = -1,      /* invalid argument (bad chain index, size, scheme, ...) */
    CHMC_E_GEOMETRY = -2, /* box dimension <= cutoff (identical_particles_binning.cpp:7-10) */
    CHMC_E_HIP = -3,      /* a HIP runtime call failed */
    CHMC_E_NODEV = -4,    /* no usable GPU */
    CHMC_E_STATE = -5     /* call out of order (e.g. bins requested before any binning) */
};

/* One chain; replaces the YAML keys particles.{N,mass,beta,cutoff,eps,sigma} and geometry.{Lx,Ly,Lz}
 * read by particles_type / identical_particles (particles.cpp:17-36, identical_particles.cpp:18-28). */
typedef struct chmc_chain_desc {
    int N;
    double L[3];
    double cutoff, sigma, eps, mass, beta;
} chmc_chain_desc;

/* context flags */
enum {
    CHMC_FORCE_FP32 = 0,       /* MD forces: FP32 pairs, FP32 per-atom sums (default) */
    CHMC_FORCE_FP64 = 1 << 0,  /* MD forces in FP64 (the reference's arithmetic) */
    CHMC_ORDER_ATOM_ID = 1 << 1,/* in-cell order by atom id instead of the reference's stable order */
    CHMC_CELL_MD = 1 << 2      /* FP32 MD of chmc_integrate / chmc_hmc_run through the 27-cell sweep with
                                  a re-binning per force evaluation (the reference's cost model) instead
                                  of the skin neighbour lists */
};

/* integrator schemes: the numbering of code-carsten/integrator.hh:14 (set_integrator).  The lp_*
 * schemes (lp_leapfrog, integrator.hh:51-79; lp_omf4, :156-197) are the leapfrog and OMF4 sequences
 * whose every momentum and position update is followed by rounding of the updated variable to the
 * grid 10^-n_prec (n_prec from chmc_set_lp_exponent); positions are rounded before the wrap. */
enum { CHMC_LEAPFROG = 0, CHMC_LP_LEAPFROG = 1, CHMC_OMF4 = 2, CHMC_LP_OMF4 = 3, CHMC_EULER = 4, CHMC_RUTH = 5,
       CHMC_OMF2 = 6 };

/* ---- lifetime ------------------------------------------------------------------------------ */
/* Replaces the identical_particles constructor + cutoff_binning + parallel_binning_init
 * (identical_particles.cpp:18-87, identical_particles_binning.cpp:4-44): allocates every chain's
 * state and cell grid, permute_vector = identity.  device < 0 picks the current device. */
int chmc_create(chmc_ctx** ctx, int nchains, const chmc_chain_desc* chains, int device, int flags);
int chmc_destroy(chmc_ctx* ctx);
const char* chmc_last_error(const chmc_ctx* ctx);   /* NULL ctx: last error of a failed chmc_create */
int chmc_sync(chmc_ctx* ctx);

/* geometry of one chain: nbin[3] and sizebin[3] as cutoff_binning computes them */
int chmc_get_geometry(const chmc_ctx* ctx, int chain, int nbin[3], double sizebin[3], int* bintot);
int chmc_num_atoms(const chmc_ctx* ctx, int* total, int* nchains);

/* ---- state transfer (AoS double[3]; deep_copy of x / h_x in HMC.cpp:72,102,111,115) --------- */
int chmc_set_positions(chmc_ctx* ctx, int chain, const double* xyz);
int chmc_get_positions(chmc_ctx* ctx, int chain, double* xyz);
int chmc_set_momenta(chmc_ctx* ctx, int chain, const double* pxyz);
int chmc_get_momenta(chmc_ctx* ctx, int chain, double* pxyz);
/* project input generator: n^3 simple-cubic sites (n^3 >= N) at (i+1/2)L/n, first N in
 * lexicographic order, + uniform jitter in [-jitter, jitter) (replaces InitX cold/hot starts,
 * identical_particles.cpp:198-257, whose Kokkos RNG stream cannot be reproduced) */
int chmc_init_lattice(chmc_ctx* ctx, uint64_t seed, double jitter);

/* ---- binning (create_binning, particles.hpp:92; parallel_binning, identical_particles_binning.cpp:97-110) */
int chmc_bin(chmc_ctx* ctx);
/* bincount / binoffsets / permute_vector of one chain, chain-local indices
 * (particles.hpp:45-47); any pointer may be NULL */
int chmc_get_bins(chmc_ctx* ctx, int chain, int* bincount, int* binoffsets, int* permute_vector);

/* ---- LJ sweeps -------------------------------------------------------------------------------- */
/* compute_force (identical_particles.cpp:559-563): binning + force; fp64 = 0 FP32 pairs, 1 FP64 */
int chmc_compute_force(chmc_ctx* ctx, int fp64);
int chmc_get_force(chmc_ctx* ctx, int chain, double* grad_xyz);   /* +grad U of the last chmc_compute_force */
/* compute_potential (identical_particles.cpp:377-385): binning + FP64 V per chain (V[nchains]) */
int chmc_compute_potential(chmc_ctx* ctx, double* V);
/* evaluate_potential (identical_particles.cpp:387-393): FP64 V with the current bins */
int chmc_evaluate_potential(chmc_ctx* ctx, double* V);
/* pair virial W = sum_{i<j} r_ij . F_ij of the current configuration (no reference counterpart) */
int chmc_virial(chmc_ctx* ctx, double* W);
/* candidate pairs of the 27-cell stencil and interacting pairs (r < cutoff), ordered, of one chain in
 * its current configuration: the N_cand and N_int of the LJ sweep's algorithmic operation count */
int chmc_pair_counts(chmc_ctx* ctx, int chain, long long* ncand, long long* nint);
/* compute_kinetic_E (identical_particles.cpp:459-463): K per chain */
int chmc_kinetic(chmc_ctx* ctx, double* K);

/* ---- integrator building blocks --------------------------------------------------------------- */
/* update_momenta (identical_particles.cpp:691-694): force (with binning) + kick p -= dt*beta*f */
int chmc_update_momenta(chmc_ctx* ctx, double dt);
/* update_positions (identical_particles.cpp:671-673): x += dt*(beta/m)*p, wrap */
int chmc_update_positions(chmc_ctx* ctx, double dt);
/* kick with the force of the last chmc_compute_force, without recomputing it */
int chmc_kick(chmc_ctx* ctx, double dt);
/* integrator_type::integrate for LEAP / OMF2 / OMF4 (integrator.cpp:52-131) and code-carsten's euler /
 * ruth (integrator.hh:200-240); lambda is OMF2's (reference: 0.1938, integrator.cpp:72).  Each
 * drift is fused into the binning of the following force evaluation. */
int chmc_integrate(chmc_ctx* ctx, int scheme, double dt, int steps, double lambda);
/* the exponent set_integrator(integs, exponent) hands to lp_leapfrog(n) / lp_omf4(n)
 * (code-carsten/integrator.hh:249, :257): rounding grid 10^-n_prec of the CHMC_LP_* schemes, used by
 * chmc_integrate and chmc_hmc_run; 0 <= n_prec <= 18, default 16 (lp_leapfrog() : n_prec(16), :53) */
int chmc_set_lp_exponent(chmc_ctx* ctx, int n_prec);
int chmc_get_lp_exponent(const chmc_ctx* ctx, int* n_prec);

/* ---- HMC pieces ------------------------------------------------------------------------------- */
/* hb (identical_particles.cpp:283-296): p ~ N(0, m/beta), Philox4x64-10 key {seed, "HEAT"},
 * counter {atom, traj, chain, 0} */
int chmc_heatbath(chmc_ctx* ctx, uint64_t seed, uint64_t traj);
int chmc_save_positions(chmc_ctx* ctx);      /* h_x = x   (HMC.cpp:72,103,107) */
int chmc_restore_positions(chmc_ctx* ctx);   /* x = h_x   (HMC.cpp:111) */
/* params.istart (read_infile.hpp:17, HMC.cpp:66-67): the number of trajectories already run; the next
 * chmc_hmc_run numbers its trajectories istart+1, ... and keys its heat bath on those numbers, so a
 * restart that restores istart (and the seed) continues the same device random stream */
int chmc_set_trajectory(chmc_ctx* ctx, int istart);
int chmc_get_trajectory(const chmc_ctx* ctx, int* istart);

/* HMC_class::run (HMC.cpp:61-136) for every chain of the context, all on the device: heat bath,
 * K_i, MD, V_f (FP64), K_f, dh = beta (K_f + V_f - K_i - V_i), thermalisation, Metropolis with
 * u from a per-chain mt19937_64 seeded seed + chain (gen_random, HMC.cpp:48-50), restore on
 * reject.  use_graph = 1 captures one trajectory in a hipGraph and replays it. */
typedef struct chmc_hmc_params {
    int ntraj, therm, scheme, steps;
    double dt, lambda;
    uint64_t seed;
    int use_graph;
    int randomize;     /* randomize_trajectory (HMC.cpp:82-85): binomial steps, no graph */
    /* surrogate = 1 drives the MD with the learned surrogate's forces (chmc_surrogate_* below); the
     * Metropolis test still uses the exact FP64 LJ energies, so the chain stays exact.  After every
     * window of fallback_window trajectories whose acceptance rate is below fallback_acc the run
     * switches to LJ forces for its remaining trajectories.  No reference counterpart. */
    int surrogate;
    int fallback_window;
    double fallback_acc;
    /* optional caller-drawn randomness (NULL: drawn inside from mt19937_64(seed + chain)):
     * uniforms[chain * ntraj + q] is the Metropolis uniform of the chain's q-th post-thermalisation
     * trajectory of this call (gen_random, HMC.cpp:48-50); traj_steps[t] the MD steps of trajectory t
     * when randomize is set (set_binomial_steps, integrator.cpp:37-49).  A caller that keeps one host
     * mt19937_64 across calls (and saves it, HMC.cpp:52-59) reproduces the reference's single stream. */
    const double* uniforms;
    const int* traj_steps;
} chmc_hmc_params;

/* per trajectory and chain, record index = traj * nchains + chain */
typedef struct chmc_hmc_record {
    double dh, u, k, w;   /* dH; U, K and virial W after the decision (K = K_f if accepted, else the
                             heat bath's K_i: the Ki of HMC.cpp:104 that print_xyz writes) */
    int accepted;
    int surrogate;        /* 1 if the trajectory's MD used the surrogate's forces */
} chmc_hmc_record;

/* records may be NULL; nacc[nchains] (may be NULL) receives the accepted post-thermalisation counts */
int chmc_hmc_run(chmc_ctx* ctx, const chmc_hmc_params* hp, chmc_hmc_record* records, int* nacc);

/* ---- measurements ----------------------------------------------------------------------------- */
/* compute_RDF (identical_particles.cpp:698-776) of one chain: counts[nbins] of ordered pairs and
 * g[nbins] normalised as the reference does; either pointer may be NULL */
int chmc_rdf(chmc_ctx* ctx, int chain, int nbins, double size_bin, unsigned long long* counts, double* g);

/* code-carsten potential_energy (potential_energy.hh:6-15) on the GPU: sum_{i<j}
 * LennardJones(|r_i - r_j|, n, a, b) (lennardjones.hh:5-9) of an AoS double[3] set, open boundaries */
int chmc_potential_energy_open(const double* xyz, int n_particles, int n, double a, double b, double* result);

/* No counterpart in the reference.  chmc_potential_energy_open for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): xyz_dev is a device AoS double[3] array of
 * n_particles atoms, result_dev a device double, stream a hipStream_t (NULL: the null stream).  The work
 * is queued on the stream and the call returns without waiting; the result has the same bits as
 * chmc_potential_energy_open's.  Block partials go to a static device buffer, one per GPU, sized for
 * at most CHMC_OPEN_NOALLOC_MAX atoms (512 KB); larger sets fail with CHMC_E_ARG.  Calls that overlap
 * on one GPU share that buffer and race.  No hipMalloc, hipFree or stream creation is made. */
#define CHMC_OPEN_NOALLOC_MAX 65536
int chmc_potential_energy_open_noalloc(const double* xyz_dev, int n_particles, int n, double a, double b,
                                       double* result_dev, void* stream);

/* ---- learned surrogate (no reference counterpart) ----------------------------------------------
 * A per-atom MLP energy E = sum_i e(G_i) over 32 radial Gaussian descriptors G_i (centres 0.8..2.5,
 * width 0.1, cosine cutoff at the chain's r_c), 32 -> 64 -> 64 -> 1 with SiLU, bfloat16 MFMA layers
 * and FP32 accumulation.  Its forces are -grad E, exactly conservative, so HMC driven by them with the
 * exact LJ Hamiltonian in the accept step samples the LJ distribution.  One parameter vector of
 * CHMC_SURR_NPARAMS floats: W1[64][32], b1[64], W2[64][64], b2[64], w3[64], b3, shift[32], scale[32]
 * (input z = (G - shift) * scale); the last 64 are set by the caller, not trained. */
#define CHMC_SURR_NPARAMS 6401
#define CHMC_SURR_NTRAIN 6337    /* the trained parameters: W1, b1, W2, b2, w3, b3 */
#define CHMC_SURR_NGRAD 6338     /* the trained parameters + one slot for the sum of squared residuals */
#define CHMC_SURR_NFEAT 32
int chmc_surrogate_set_params(chmc_ctx* ctx, const float* params, int n);
int chmc_surrogate_get_params(chmc_ctx* ctx, float* params, int n);
/* Adam moments m[CHMC_SURR_NTRAIN], v[CHMC_SURR_NTRAIN] and step count of the training (either
 * pointer of get may be NULL); with the parameters they checkpoint a training run (set_params resets
 * them to zero) */
int chmc_surrogate_get_adam(chmc_ctx* ctx, float* m, float* v, int* step);
int chmc_surrogate_set_adam(chmc_ctx* ctx, const float* m, const float* v, int step);
/* E of every chain (energy[nchains], may be NULL) at the current positions; store_force keeps +grad E
 * readable through chmc_get_force.  FP32 neighbour-list contexts evaluate with the kernels of
 * surrogate-driven MD (descriptors from the skin lists), cell-MD contexts with the 27-cell pair list. */
int chmc_surrogate_eval(chmc_ctx* ctx, double* energy, int store_force);
/* descriptors G[natoms][32] and dE/dG g[natoms][32] of the last chmc_surrogate_eval, atom order */
int chmc_surrogate_features(chmc_ctx* ctx, int chain, float* G, float* g);
/* training set on the device: room for nconfigs snapshots of every chain; collect appends the
 * descriptors and per-atom LJ energies u_i = 2 eps sum_j sr6 (sr6 - 1) of the current positions */
int chmc_surrogate_reserve(chmc_ctx* ctx, int nconfigs);
int chmc_surrogate_collect(chmc_ctx* ctx, long long* rows);
int chmc_surrogate_training_data(chmc_ctx* ctx, long long row0, int nrows, float* G, float* u);
/* sum over rows of the gradient of (e_i - u_i)^2 / nrows w.r.t. the trained parameters, last slot
 * sum of (e_i - u_i)^2: grad[CHMC_SURR_NGRAD] */
int chmc_surrogate_loss_grad(chmc_ctx* ctx, long long row0, int nrows, float* grad);
/* nsteps Adam steps (lr, beta1 0.9, beta2 0.999, eps 1e-8) on batches of batch_rows consecutive rows
 * at offsets drawn from mt19937_64(seed); loss_history[nsteps] (may be NULL) gets each batch's MSE */
int chmc_surrogate_train(chmc_ctx* ctx, int nsteps, int batch_rows, float lr, uint64_t seed, float* loss_history);
/* Energy and force matching on whole configurations: loss = mean_i (e_i - u_i)^2 + w_force *
 * mean_{i,d} (dE/dx_id - dU/dx_id)^2 over every atom of every chain (U the truncated LJ energy, u_i its
 * per-atom share, dU/dx the LJ gradient).  fm_loss_grad: at the current positions, the gradient w.r.t.
 * the trained parameters grad[CHMC_SURR_NTRAIN] and both mean losses (no update).  reserve_configs /
 * collect_config keep snapshots of the context's positions on the device; train_fm runs nsteps Adam
 * steps (lr, 0.9, 0.999, 1e-8), each on one snapshot drawn by mt19937_64(seed), and writes the mean
 * losses before each step (loss_e, loss_f may be NULL); the positions are left at the last snapshot. */
int chmc_surrogate_fm_loss_grad(chmc_ctx* ctx, double w_force, float* grad, double* loss_e, double* loss_f);
int chmc_surrogate_reserve_configs(chmc_ctx* ctx, int nconfigs);
int chmc_surrogate_collect_config(chmc_ctx* ctx, int* nconfigs);
int chmc_surrogate_train_fm(chmc_ctx* ctx, int nsteps, double w_force, float lr, uint64_t seed, float* loss_e,
                            float* loss_f);

/* ---- tuning ----------------------------------------------------------------------------------- */
/* tiling of the FP32 LJ sweep, encoded i-slots per wavefront * 10 + wavefronts per workgroup, one of
 * 41, 42, 44, 81, 82, 84, 161, 162; 0 runs the autotuner (times every tiling on the current state) */
int chmc_set_force_tiling(chmc_ctx* ctx, int tiling);
int chmc_get_force_tiling(const chmc_ctx* ctx, int* tiling);
/* average device time (ms) of one force evaluation (binning + FP32 sweep + kick) over reps */
int chmc_time_force(chmc_ctx* ctx, int reps, double* ms_bin, double* ms_force);

/* Neighbour-list MD (FP32 contexts without CHMC_CELL_MD): chmc_integrate and chmc_hmc_run keep per-atom
 * lists of the pairs within cutoff + skin and rebuild them (on the device, inside the captured
 * trajectory) only when an atom has moved more than skin / 2 since the last rebuild, so every pair
 * within the cutoff is always listed; each integration / trajectory starts from a fresh rebuild.
 * No reference counterpart (the reference re-bins on every evaluation, identical_particles.cpp:379,560).
 * set_skin reallocates the lists (default 0.6 sigma-units of length, the fastest at N = 32768,
 * rho* = 0.8442).  list_stats: list entries and interacting pairs (within the cutoff) of chain `chain`
 * at the current positions (rebuilds first), and the number of rebuilds run since the context was
 * created. */
int chmc_set_skin(chmc_ctx* ctx, double skin);
int chmc_get_skin(const chmc_ctx* ctx, double* skin);
/* one skin per chain (skins[nchains], each 0 < skin <= 2): every chain's list cells, capacities and
 * rebuild threshold follow its own skin; set_skin sets them all back to one value.  get_skin then
 * reports the largest.  No reference counterpart (a tuning parameter of the list MD above). */
int chmc_set_chain_skins(chmc_ctx* ctx, const double* skins);
int chmc_get_chain_skin(const chmc_ctx* ctx, int chain, double* skin);
int chmc_list_stats(chmc_ctx* ctx, int chain, long long* entries, long long* nint, long long* rebuilds);
/* average device time (ms) of the list sweep with kick + drift (the MD step of the trajectory) and of
 * one rebuild, over reps, at the current positions (the positions do not change: dt = 0) */
int chmc_time_list(chmc_ctx* ctx, int reps, double* ms_sweep, double* ms_rebuild);
/* average device time (ms) of the surrogate's two force-evaluation launches on the lists: descriptors +
 * MLP (forward and input gradient) and the force sweep with kick + drift, over reps each, at the current
 * positions (dt = 0: nothing moves).  Needs surrogate parameters (chmc_surrogate_set_params or training).
 * No reference counterpart (the surrogate has none). */
int chmc_time_surrogate(chmc_ctx* ctx, int reps, double* ms_desc_mlp, double* ms_sweep);
/* shape of the list sweep: 4, 8 or 16 lanes per atom (first unconditional batches of 16, 14 and 7 list
 * entries per lane), or 810 / 811 / 812 / 816 (8 lanes, batches of 10 / 11 / 12 / 16 entries); 0 runs
 * the autotuner (times each at the current positions and keeps the fastest) */
int chmc_set_list_lanes(chmc_ctx* ctx, int lanes);
int chmc_get_list_lanes(const chmc_ctx* ctx, int* lanes);
/* list entry format: 32 (neighbour slot and image code per entry), 16 (segment of the list cell's
 * 27-cell neighbourhood and rank inside that cell: half the bytes, a few more instructions per entry),
 * or 0, the default: 16 when the lists' expected entries at 4 bytes exceed 256 MiB (the last-level
 * cache), else 32.  Both give the same pairs in the same order; the FP64 energy is bitwise the same and
 * the FP32 forces differ only where the sweep shape reads 16-bit entries two at a time.  get returns
 * the format in use once the lists are set up.  No reference counterpart (a storage choice of the list
 * MD above). */
int chmc_set_list_entry_bits(chmc_ctx* ctx, int bits);
int chmc_get_list_entry_bits(const chmc_ctx* ctx, int* bits);
/* rebuild checks of list MD before every `every`-th force evaluation of an integration (1 to 4; default
 * 1), and before its last one.  In between, the rebuild trigger sits lower than skin / 2 by the
 * distance every - 1 drifts can add for momenta |p| <= 10 sqrt(m / beta), so the lists still hold every
 * pair within the cutoff; each drift verifies that bound, and a violation makes the call that reads
 * the flags fail with CHMC_E_STATE.  1: the check before every evaluation.  No reference counterpart. */
int chmc_set_list_check(chmc_ctx* ctx, int every);
int chmc_get_list_check(const chmc_ctx* ctx, int* every);
/* start rebuilds of list MD (default off): a chain whose trajectory was accepted rebuilds its lists at
 * the start of the next trajectory, before the save of the start state (and so do a chmc_hmc_run start
 * and chmc_save_positions), so that every trajectory from those positions -- the accepted state is the
 * start of all the rejected trajectories that follow it -- begins with lists built where it starts.
 * Fewer rebuilds inside trajectories where acceptance is low, one more per accepted trajectory: a gain
 * at low acceptance, a loss at high.  No reference counterpart. */
int chmc_set_list_start_rebuild(chmc_ctx* ctx, int on);
int chmc_get_list_start_rebuild(const chmc_ctx* ctx, int* on);
/* speculative trajectories: a context created with `copies` identical chains (same descriptor, one
 * skin, FP32 neighbour-list MD) runs chmc_hmc_run as ONE chain, chain 0.  Each batch runs the chain's
 * next `copies` trajectories at once, all from its current state (copy k: trajectory t + 1 + k, heat
 * bath and Metropolis uniform of that trajectory); the Metropolis tests then take them in order up to
 * the first accepted one, whose end state every copy takes over, and the later ones are run again in
 * the next batch.  A trajectory of the sequential loop starts from the same state exactly when the
 * ones before it were rejected, so records, positions, momenta and acceptance equal the single-chain
 * run's bit for bit; where acceptance is low, nearly every batch keeps all `copies` trajectories.
 * chmc_hmc_run then writes the chain's records in column 0 (the other columns zero) and its count in
 * nacc[0]; randomize runs are refused (surrogate runs, with their fallback windows, are speculative
 * too).  copies <= 1 switches it off.  No reference
 * counterpart (HMC.cpp:77-117 is the sequential loop this reproduces). */
int chmc_set_speculation(chmc_ctx* ctx, int copies);
int chmc_get_speculation(const chmc_ctx* ctx, int* copies);

#ifdef __cplusplus
}
#endif
#endif
