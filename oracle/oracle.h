/*
 * oracle.h -- FP64 CPU restatement of the urbach/chemHMC Lennard-Jones HMC hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls this code; it is
 * loaded by tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py, and
 * only there as the checker or as the timed CPU baseline.
 *
 * Every function names the reference file:line it restates (paths relative to the reference
 * repository root).  Positions, momenta and forces are AoS double[3] per particle, the layout of
 * code-carsten/config.hh:15-94 (config<coordinate>) and of Kokkos' LayoutRight host views.
 *
 * Pinning (see tests/test_oracle.py): closed forms of the LJ pair, brute-force all-images sums
 * (the reference's own all_neighbour strategy), the 5-point derivative rule of
 * code/test/test_binning.cpp:108-168, the strategy-agreement tolerances of test_binning.cpp,
 * numpy's Philox4x64-10 for the counter RNG, the C++ standard's mt19937_64 known answer, the
 * compiled code-carsten potential_energy() (oracle/_ref), and integrator/statistics invariants.
 */
#ifndef CHMC_ORACLE_H
#define CHMC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- random numbers -------------------------------------------------------------------- */
void orc_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);
typedef struct { uint64_t mt[312]; int mti; } orc_mt64;
void     orc_mt64_seed(orc_mt64* g, uint64_t seed);
uint64_t orc_mt64_next(orc_mt64* g);
double   orc_mt64_uniform(orc_mt64* g);            /* HMC.cpp:48-50 gen_random() */
int      orc_binomial_steps(orc_mt64* g, int average_steps);   /* integrator.cpp:37-49 */

/* ---- project input generator (lattice + jitter, Philox keyed on seed) ------------------ */
void orc_lattice_init(int N, const double L[3], uint64_t seed, int chain, double jitter, double* x);
void orc_heatbath(int N, double mass, double beta, uint64_t seed, int chain, uint64_t traj, double* p);

/* ---- geometry / binning ---------------------------------------------------------------- */
int  orc_cutoff_binning(const double L[3], double cutoff, int nbin[3], double sizebin[3]);
int  orc_which_bin(const double* xi, const int nbin[3], const double sizebin[3]);
void orc_binning(int N, const double* x, const int nbin[3], const double sizebin[3],
                 int* perm, int* bincount, int* binoffsets);
void orc_binning_refalgo(int N, const double* x, const int nbin[3], const double sizebin[3],
                         int* perm, int* perm_tmp, int* bincount, int* binoffsets);

/* ---- LJ sweeps ------------------------------------------------------------------------- */
void   orc_force_binning(int N, const double* x, const double L[3], const int nbin[3],
                         const int* perm, const int* bincount, const int* binoffsets,
                         double cutoff, double sigma, double eps, double* f);
double orc_potential_binning(int N, const double* x, const double L[3], const int nbin[3],
                             const int* perm, const int* bincount, const int* binoffsets,
                             double cutoff, double sigma, double eps);
double orc_virial_binning(int N, const double* x, const double L[3], const int nbin[3],
                          const int* perm, const int* bincount, const int* binoffsets,
                          double cutoff, double sigma, double eps);
void   orc_pair_counts(int N, const double* x, const double L[3], const int nbin[3],
                       const int* perm, const int* bincount, const int* binoffsets,
                       double cutoff, long long* ncand, long long* nint);
double orc_potential_all(int N, const double* x, const double L[3], double cutoff, double sigma, double eps);
void   orc_force_all(int N, const double* x, const double L[3], double cutoff, double sigma, double eps, double* f);
double orc_lennard_jones(double r, int n, double a, double b);
double orc_potential_energy_open(int N, const double* x, int n, double a, double b);

/* ---- kinetic energy, kick, drift ------------------------------------------------------- */
double orc_kinetic(int N, const double* p, double mass);
void   orc_kick(int N, double* p, const double* f, double dt, double coeff_p);
void   orc_drift(int N, double* x, const double* p, double dt, double coeff_x, const double L[3]);
/* the rounded updates of lp_leapfrog / lp_omf4 (scale = 10^n_prec = orc_lp_scale(n_prec)) */
double orc_lp_scale(int n_prec);
void   orc_kick_lp(int N, double* p, const double* f, double dt, double coeff_p, double scale);
void   orc_drift_lp(int N, double* x, const double* p, double dt, double coeff_x, const double L[3], double scale);

/* ---- RDF ------------------------------------------------------------------------------- */
void orc_rdf(int N, const double* x, const double L[3], int nbins, double size_bin,
             unsigned long long* counts, double* rdf);

/* ---- a whole system: integrators and the HMC loop -------------------------------------- */
enum { ORC_LEAPFROG = 0, ORC_LP_LEAPFROG = 1, ORC_OMF4 = 2, ORC_LP_OMF4 = 3, ORC_EULER = 4, ORC_RUTH = 5, ORC_OMF2 = 6 };
enum { ORC_BIN_COUNTING = 0, ORC_BIN_REFALGO = 1 };

typedef struct {
    int N;
    double L[3], cutoff, sigma, eps, mass, beta;
    int nbin[3], bintot;
    double sizebin[3];
    int bin_mode;
    double *x, *p, *f, *x_saved;
    int *perm, *perm_tmp, *bincount, *binoffsets;
    long long force_evals;
    double lp_scale;    /* 10^n_prec of the lp_* integrators (default n_prec = 16) */
    int lp_active;      /* inside an lp_* integration: kicks and drifts round */
} orc_system;

int    orc_system_init(orc_system* s, int N, const double L[3], double cutoff, double sigma,
                       double eps, double mass, double beta, int bin_mode);
void   orc_system_free(orc_system* s);
void   orc_sys_create_binning(orc_system* s);
void   orc_sys_compute_force(orc_system* s);
double orc_sys_compute_potential(orc_system* s);
void   orc_sys_update_momenta(orc_system* s, double dt);
void   orc_sys_update_positions(orc_system* s, double dt);
void   orc_sys_integrate(orc_system* s, int scheme, double dt, int steps, double lambda);

typedef struct {
    int ntraj, therm, scheme, steps, randomize;
    double dt, lambda;
    uint64_t seed;      /* heat-bath Philox key and mt19937_64 seed */
    int chain;          /* chain id fed to the heat-bath counter */
} orc_hmc_params;

/* records (length ntraj each, any may be NULL): dH, accepted, U after the decision,
 * K after the decision (K_f if accepted, else K_i), virial W of the current configuration */
int orc_hmc_run(orc_system* s, const orc_hmc_params* hp, double* rec_dh, int* rec_acc,
                double* rec_u, double* rec_k, double* rec_w);

#ifdef __cplusplus
}
#endif
#endif
