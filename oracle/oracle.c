/*
 * oracle.c -- FP64 CPU restatement of the urbach/chemHMC LJ-HMC hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the CPU baseline of bench.py.
 * Product code never links it.  Built with -ffp-contract=off so that every product kernel that
 * claims bit-exactness (binning, kick, drift, lattice, RDF bins) can be compared bit for bit.
 *
 * Parity status of each function: see the header of tests/test_oracle.py, which holds the check
 * pinning it.  Citations are file:line in the reference repository.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================================= */
/* Philox4x64-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11).  Pinned against numpy.random.Philox, which implements the same published generator. */
static inline uint64_t mulhilo64(uint64_t a, uint64_t b, uint64_t* hi)
{
    unsigned __int128 prod = (unsigned __int128)a * b;
    *hi = (uint64_t)(prod >> 64);
    return (uint64_t)prod;
}

void orc_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4])
{
    uint64_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint64_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; round++) {
        if (round > 0) {
            k0 += 0x9E3779B97F4A7C15ULL;
            k1 += 0xBB67AE8584CAA73BULL;
        }
        uint64_t hi0, hi1;
        uint64_t lo0 = mulhilo64(0xD2E7470EE14C6C93ULL, c0, &hi0);
        uint64_t lo1 = mulhilo64(0xCA5A826395121157ULL, c2, &hi1);
        uint64_t n0 = hi1 ^ c1 ^ k0;
        uint64_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 53-bit uniform in [0,1) and in (0,1) from a 64-bit word */
static inline double u01_closed_open(uint64_t w) { return (double)(w >> 11) * 0x1.0p-53; }
static inline double u01_open(uint64_t w) { return ((double)(w >> 11) + 0.5) * 0x1.0p-53; }

/* ======================================================================================= */
/* mt19937_64 (Matsumoto & Nishimura 2000 64-bit variant), the engine HMC.cpp:38 seeds and
 * HMC.cpp:48-50 / integrator.cpp:45 draw from. */
#define MT_NN 312
#define MT_MM 156
void orc_mt64_seed(orc_mt64* g, uint64_t seed)
{
    g->mt[0] = seed;
    for (int i = 1; i < MT_NN; i++)
        g->mt[i] = 6364136223846793005ULL * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t)i;
    g->mti = MT_NN;
}

uint64_t orc_mt64_next(orc_mt64* g)
{
    const uint64_t UM = 0xFFFFFFFF80000000ULL, LM = 0x7FFFFFFFULL, MA = 0xB5026F5AA96619E9ULL;
    if (g->mti >= MT_NN) {
        int i;
        for (i = 0; i < MT_NN - MT_MM; i++) {
            uint64_t x = (g->mt[i] & UM) | (g->mt[i + 1] & LM);
            g->mt[i] = g->mt[i + MT_MM] ^ (x >> 1) ^ ((x & 1ULL) ? MA : 0ULL);
        }
        for (; i < MT_NN - 1; i++) {
            uint64_t x = (g->mt[i] & UM) | (g->mt[i + 1] & LM);
            g->mt[i] = g->mt[i + (MT_MM - MT_NN)] ^ (x >> 1) ^ ((x & 1ULL) ? MA : 0ULL);
        }
        uint64_t x = (g->mt[MT_NN - 1] & UM) | (g->mt[0] & LM);
        g->mt[MT_NN - 1] = g->mt[MT_MM - 1] ^ (x >> 1) ^ ((x & 1ULL) ? MA : 0ULL);
        g->mti = 0;
    }
    uint64_t x = g->mt[g->mti++];
    x ^= (x >> 29) & 0x5555555555555555ULL;
    x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
    x ^= (x << 37) & 0xFFF7EEE000000000ULL;
    x ^= (x >> 43);
    return x;
}

/* HMC.cpp:48-50: ((double)gen64() - gen64.min()) / (gen64.max() - gen64.min()); min = 0 and
 * max - min = 2^64-1, which converts to the double 2^64. */
double orc_mt64_uniform(orc_mt64* g)
{
    double num = (double)orc_mt64_next(g) - 0.0;
    double den = (double)(0xFFFFFFFFFFFFFFFFULL - 0ULL);
    return num / den;
}

/* integrator.cpp:37-49: binomial(n = 2*average_steps, p = 1/2) from successive uniforms */
int orc_binomial_steps(orc_mt64* g, int average_steps)
{
    int K = 2, n = K * average_steps, steps = 0;
    double pr = 1. / ((double)K);
    for (int i = 0; i < n; i++) {
        double r = orc_mt64_uniform(g);
        if (r < pr) steps++;
    }
    return steps;
}

/* ======================================================================================= */
/* Project-owned input generator (the reference's cold start, identical_particles.cpp:238-248,
 * has a pow(512,1/3.) rounding quirk and no jitter, and its hot start uses the Kokkos XorShift
 * pool, which cannot be reproduced here).  Sites of an n^3 simple-cubic lattice, n the smallest
 * integer with n^3 >= N, at (i + 1/2) * L/n in lexicographic order (x fastest), truncated to the
 * first N sites, plus a uniform jitter in [-jitter, +jitter) per coordinate from Philox4x64-10 with
 * key {seed, 0x4C415454} and counter {i, chain, 0, 0}, wrapped into [0, L). */
static int lattice_side(int N)
{
    int n = (int)llround(cbrt((double)N));
    if (n < 1) n = 1;
    while ((long long)n * n * n < N) n++;
    while (n > 1 && (long long)(n - 1) * (n - 1) * (n - 1) >= N) n--;
    return n;
}

void orc_lattice_init(int N, const double L[3], uint64_t seed, int chain, double jitter, double* x)
{
    int n = lattice_side(N);
    uint64_t key[2] = { seed, 0x4C415454ULL };
    for (int i = 0; i < N; i++) {
        int idx[3] = { i % n, (i / n) % n, i / (n * n) };
        uint64_t ctr[4] = { (uint64_t)i, (uint64_t)chain, 0, 0 }, out[4];
        orc_philox4x64(ctr, key, out);
        for (int d = 0; d < 3; d++) {
            double a = L[d] / (double)n;
            double site = ((double)idx[d] + 0.5) * a;
            double u = u01_closed_open(out[d]);
            double v = site + (2.0 * u - 1.0) * jitter;
            v = v - L[d] * floor(v / L[d]);
            x[3 * i + d] = v;
        }
    }
}

/* Heat bath, restating identical_particles.cpp:287-296 (p = normal() * sqrt(m/beta)) with a
 * counter-based stream: Philox4x64-10, key {seed, 0x48454154}, counter {i, traj, chain, 0};
 * Box-Muller on words (0,1) gives the x,y components and on words (2,3) the z component. */
void orc_heatbath(int N, double mass, double beta, uint64_t seed, int chain, uint64_t traj, double* p)
{
    uint64_t key[2] = { seed, 0x48454154ULL };
    double sd = sqrt(mass / beta);
    const double twopi = 6.283185307179586;
    for (int i = 0; i < N; i++) {
        uint64_t ctr[4] = { (uint64_t)i, traj, (uint64_t)chain, 0 }, out[4];
        orc_philox4x64(ctr, key, out);
        double r0 = sqrt(-2.0 * log(u01_open(out[0])));
        double a0 = twopi * u01_closed_open(out[1]);
        double r1 = sqrt(-2.0 * log(u01_open(out[2])));
        double a1 = twopi * u01_closed_open(out[3]);
        p[3 * i + 0] = (r0 * cos(a0)) * sd;
        p[3 * i + 1] = (r0 * sin(a0)) * sd;
        p[3 * i + 2] = (r1 * cos(a1)) * sd;
    }
}

/* ======================================================================================= */
/* identical_particles_binning.cpp:4-29 */
int orc_cutoff_binning(const double L[3], double cutoff, int nbin[3], double sizebin[3])
{
    for (int i = 0; i < 3; i++) {
        if (L[i] <= cutoff) return -1;
        nbin[i] = (int)floor(L[i] / cutoff);
        sizebin[i] = L[i] / ((double)nbin[i]);
        if (sizebin[i] < cutoff) return -2;
    }
    return 0;
}

/* particles.hpp:100-113 (which_bin / ctolex).  The reference does not clamp; a coordinate that
 * rounds to exactly L after the wrap of identical_particles.cpp:667 would index cell nbin.  Here
 * (and in the HIP kernels) that measure-zero case is clamped to nbin-1. */
static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
int orc_which_bin(const double* xi, const int nbin[3], const double sizebin[3])
{
    int bx = clampi((int)floor(xi[0] / sizebin[0]), 0, nbin[0] - 1);
    int by = clampi((int)floor(xi[1] / sizebin[1]), 0, nbin[1] - 1);
    int bz = clampi((int)floor(xi[2] / sizebin[2]), 0, nbin[2] - 1);
    return bx + nbin[0] * (by + bz * nbin[1]);
}

/* binning.cpp:15-54 + binning.hpp:43-57 (parallel_binning, identical_particles_binning.cpp:97-110):
 * count per cell, exclusive scan, stable scatter of the current permute_vector, copy back.
 * Its effective semantics is a stable counting sort of `perm` by cell; this is the O(N) form. */
void orc_binning(int N, const double* x, const int nbin[3], const double sizebin[3],
                 int* perm, int* bincount, int* binoffsets)
{
    int bintot = nbin[0] * nbin[1] * nbin[2];
    int* cell = (int*)malloc(sizeof(int) * (size_t)(N > 0 ? N : 1));
    int* tmp = (int*)malloc(sizeof(int) * (size_t)(N > 0 ? N : 1));
    int* fill = (int*)calloc((size_t)bintot, sizeof(int));
    memset(bincount, 0, sizeof(int) * (size_t)bintot);
    for (int k = 0; k < N; k++) {
        cell[k] = orc_which_bin(&x[3 * perm[k]], nbin, sizebin);
        bincount[cell[k]]++;
    }
    int run = 0;
    for (int c = 0; c < bintot; c++) { binoffsets[c] = run; run += bincount[c]; }
    for (int k = 0; k < N; k++) {
        int c = cell[k];
        tmp[binoffsets[c] + fill[c]++] = perm[k];
    }
    memcpy(perm, tmp, sizeof(int) * (size_t)N);
    free(cell); free(tmp); free(fill);
}

/* The reference's cost model (binning.cpp:15-54): every cell scans all N particles, twice.
 * Used only as the "ref-algo" CPU baseline; same output as orc_binning. */
void orc_binning_refalgo(int N, const double* x, const int nbin[3], const double sizebin[3],
                         int* perm, int* perm_tmp, int* bincount, int* binoffsets)
{
    int bintot = nbin[0] * nbin[1] * nbin[2];
    #pragma omp parallel for schedule(dynamic, 4)
    for (int ib = 0; ib < bintot; ib++) {           /* Tag_count_bin */
        int cnt = 0;
        for (int i = 0; i < N; i++)
            if (orc_which_bin(&x[3 * perm[i]], nbin, sizebin) == ib) cnt++;
        bincount[ib] = cnt;
    }
    int run = 0;                                     /* Tag_offset */
    for (int c = 0; c < bintot; c++) { binoffsets[c] = run; run += bincount[c]; }
    #pragma omp parallel for schedule(dynamic, 4)
    for (int ib = 0; ib < bintot; ib++) {           /* Tag_set_bin */
        int upd = 0;
        for (int i = 0; i < N; i++)
            if (orc_which_bin(&x[3 * perm[i]], nbin, sizebin) == ib) perm_tmp[binoffsets[ib] + upd++] = perm[i];
    }
    memcpy(perm, perm_tmp, sizeof(int) * (size_t)N);   /* copy_range */
}

/* ======================================================================================= */
/* Neighbour-cell walk of identical_particles.cpp:581-600 (and :407-426): cell index plus the
 * explicit image shift wrap = b*L applied only when the neighbour index wraps. */
static inline void neighbour_cell(const int ib3[3], const int db[3], const int nbin[3], const double L[3],
                                  int* jb, double wrap[3])
{
    int j3[3];
    for (int d = 0; d < 3; d++) {
        j3[d] = ib3[d] + db[d];
        wrap[d] = 0.0;
        if (j3[d] < 0 || j3[d] >= nbin[d]) {
            j3[d] = (ib3[d] + nbin[d] + db[d]) % nbin[d];
            wrap[d] = db[d] * L[d];
        }
    }
    *jb = j3[0] + nbin[0] * (j3[1] + j3[2] * nbin[1]);
}

static inline void lextoc(int ib, const int nbin[3], int b[3])
{
    b[2] = ib / (nbin[0] * nbin[1]);
    b[1] = (ib - b[2] * nbin[0] * nbin[1]) / nbin[0];
    b[0] = ib - nbin[0] * (b[1] + b[2] * nbin[1]);
}

/* identical_particles.cpp:567-639 Tag_force_binning: f_i = 48 eps sum_j sr6 (0.5 - sr6)/r^2 (x_i - x_j - wrap),
 * i.e. f holds +grad_i U; the cutoff test is r^2 < cutoff^2 (:610). */
void orc_force_binning(int N, const double* x, const double L[3], const int nbin[3],
                       const int* perm, const int* bincount, const int* binoffsets,
                       double cutoff, double sigma, double eps, double* f)
{
    (void)N;
    int bintot = nbin[0] * nbin[1] * nbin[2];
    double rc2 = cutoff * cutoff, s2 = sigma * sigma;
    #pragma omp parallel for schedule(dynamic, 1)
    for (int ib = 0; ib < bintot; ib++) {
        int ib3[3];
        lextoc(ib, nbin, ib3);
        for (int ip = 0; ip < bincount[ib]; ip++) {
            int i = perm[ip + binoffsets[ib]];
            double fi[3] = { 0, 0, 0 };
            for (int bx = -1; bx < 2; bx++)
                for (int by = -1; by < 2; by++)
                    for (int bz = -1; bz < 2; bz++) {
                        int db[3] = { bx, by, bz }, jb;
                        double wrap[3];
                        neighbour_cell(ib3, db, nbin, L, &jb, wrap);
                        double fv[3] = { 0, 0, 0 };
                        for (int jp = 0; jp < bincount[jb]; jp++) {
                            int j = perm[jp + binoffsets[jb]];
                            if (i == j && bx == 0 && by == 0 && bz == 0) continue;
                            double d0 = x[3 * i + 0] - x[3 * j + 0] - wrap[0];
                            double d1 = x[3 * i + 1] - x[3 * j + 1] - wrap[1];
                            double d2 = x[3 * i + 2] - x[3 * j + 2] - wrap[2];
                            double r2 = d0 * d0 + d1 * d1 + d2 * d2;
                            if (r2 < rc2) {
                                double sr2 = s2 / r2;
                                double sr6 = sr2 * sr2 * sr2;
                                double g = sr6 * (-sr6 + 0.5) / r2;
                                fv[0] += g * d0; fv[1] += g * d1; fv[2] += g * d2;
                            }
                        }
                        fi[0] += fv[0]; fi[1] += fv[1]; fi[2] += fv[2];
                    }
            f[3 * i + 0] = fi[0] * (48 * eps);
            f[3 * i + 1] = fi[1] * (48 * eps);
            f[3 * i + 2] = fi[2] * (48 * eps);
        }
    }
}

/* identical_particles.cpp:395-454 Tag_potential_binning: V = 2 eps sum_i sum_{j!=i} sr6 (sr6 - 1) with the
 * cutoff test sqrt(r^2) < cutoff (:432-435) and sr = sigma / r (:437-438); summed per cell in a
 * fixed order so the result is deterministic. */
double orc_potential_binning(int N, const double* x, const double L[3], const int nbin[3],
                             const int* perm, const int* bincount, const int* binoffsets,
                             double cutoff, double sigma, double eps)
{
    (void)N;
    int bintot = nbin[0] * nbin[1] * nbin[2];
    double* part = (double*)calloc((size_t)bintot, sizeof(double));
    #pragma omp parallel for schedule(dynamic, 1)
    for (int ib = 0; ib < bintot; ib++) {
        int ib3[3];
        lextoc(ib, nbin, ib3);
        double tempN = 0;
        for (int ip = 0; ip < bincount[ib]; ip++) {
            int i = perm[ip + binoffsets[ib]];
            for (int bx = -1; bx < 2; bx++)
                for (int by = -1; by < 2; by++)
                    for (int bz = -1; bz < 2; bz++) {
                        int db[3] = { bx, by, bz }, jb;
                        double wrap[3];
                        neighbour_cell(ib3, db, nbin, L, &jb, wrap);
                        double tempM = 0;
                        for (int jp = 0; jp < bincount[jb]; jp++) {
                            int j = perm[jp + binoffsets[jb]];
                            if (i == j && bx == 0 && by == 0 && bz == 0) continue;
                            double d0 = x[3 * i + 0] - x[3 * j + 0] - wrap[0];
                            double d1 = x[3 * i + 1] - x[3 * j + 1] - wrap[1];
                            double d2 = x[3 * i + 2] - x[3 * j + 2] - wrap[2];
                            double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
                            if (r < cutoff) {
                                double sr = sigma / r;
                                double sr6 = sr * sr * sr * sr * sr * sr;
                                tempM += sr6 * (sr6 - 1.0);
                            }
                        }
                        tempN += tempM;
                    }
        }
        part[ib] = tempN;
    }
    double V = 0;
    for (int ib = 0; ib < bintot; ib++) V += part[ib];
    free(part);
    return 2 * eps * V;
}

/* Pair virial W = sum_{i<j} r_ij . F_ij = sum_{i<j} 48 eps (sr12 - sr6/2) over the same stencil
 * and cutoff as the force (no reference counterpart; pinned by the two-particle closed form and
 * by -dU/dlnV under affine scaling in tests/test_oracle.py). */
double orc_virial_binning(int N, const double* x, const double L[3], const int nbin[3],
                          const int* perm, const int* bincount, const int* binoffsets,
                          double cutoff, double sigma, double eps)
{
    (void)N;
    int bintot = nbin[0] * nbin[1] * nbin[2];
    double rc2 = cutoff * cutoff, s2 = sigma * sigma;
    double* part = (double*)calloc((size_t)bintot, sizeof(double));
    #pragma omp parallel for schedule(dynamic, 1)
    for (int ib = 0; ib < bintot; ib++) {
        int ib3[3];
        lextoc(ib, nbin, ib3);
        double acc = 0;
        for (int ip = 0; ip < bincount[ib]; ip++) {
            int i = perm[ip + binoffsets[ib]];
            for (int bx = -1; bx < 2; bx++)
                for (int by = -1; by < 2; by++)
                    for (int bz = -1; bz < 2; bz++) {
                        int db[3] = { bx, by, bz }, jb;
                        double wrap[3];
                        neighbour_cell(ib3, db, nbin, L, &jb, wrap);
                        for (int jp = 0; jp < bincount[jb]; jp++) {
                            int j = perm[jp + binoffsets[jb]];
                            if (i == j && bx == 0 && by == 0 && bz == 0) continue;
                            double d0 = x[3 * i + 0] - x[3 * j + 0] - wrap[0];
                            double d1 = x[3 * i + 1] - x[3 * j + 1] - wrap[1];
                            double d2 = x[3 * i + 2] - x[3 * j + 2] - wrap[2];
                            double r2 = d0 * d0 + d1 * d1 + d2 * d2;
                            if (r2 < rc2) {
                                double sr2 = s2 / r2;
                                double sr6 = sr2 * sr2 * sr2;
                                acc += sr6 * (sr6 - 0.5);
                            }
                        }
                    }
        }
        part[ib] = acc;
    }
    double W = 0;
    for (int ib = 0; ib < bintot; ib++) W += part[ib];
    free(part);
    return 24 * eps * W;
}

/* Candidate and interacting pair counts of the 27-cell stencil (the per-configuration N_cand and
 * N_int that the algorithmic flop count of the force sweep is built from). */
void orc_pair_counts(int N, const double* x, const double L[3], const int nbin[3],
                     const int* perm, const int* bincount, const int* binoffsets,
                     double cutoff, long long* ncand, long long* nint)
{
    (void)N;
    int bintot = nbin[0] * nbin[1] * nbin[2];
    double rc2 = cutoff * cutoff;
    long long nc = 0, ni = 0;
    #pragma omp parallel for schedule(dynamic, 1) reduction(+:nc, ni)
    for (int ib = 0; ib < bintot; ib++) {
        int ib3[3];
        lextoc(ib, nbin, ib3);
        for (int ip = 0; ip < bincount[ib]; ip++) {
            int i = perm[ip + binoffsets[ib]];
            for (int bx = -1; bx < 2; bx++)
                for (int by = -1; by < 2; by++)
                    for (int bz = -1; bz < 2; bz++) {
                        int db[3] = { bx, by, bz }, jb;
                        double wrap[3];
                        neighbour_cell(ib3, db, nbin, L, &jb, wrap);
                        for (int jp = 0; jp < bincount[jb]; jp++) {
                            int j = perm[jp + binoffsets[jb]];
                            if (i == j && bx == 0 && by == 0 && bz == 0) continue;
                            nc++;
                            double d0 = x[3 * i + 0] - x[3 * j + 0] - wrap[0];
                            double d1 = x[3 * i + 1] - x[3 * j + 1] - wrap[1];
                            double d2 = x[3 * i + 2] - x[3 * j + 2] - wrap[2];
                            if (d0 * d0 + d1 * d1 + d2 * d2 < rc2) ni++;
                        }
                    }
        }
    }
    *ncand = nc;
    *nint = ni;
}

/* identical_particles.cpp:306-332 Tag_potential_all: 27 image shifts x all j, pow() form. */
double orc_potential_all(int N, const double* x, const double L[3], double cutoff, double sigma, double eps)
{
    double V = 0;
    #pragma omp parallel for schedule(static) reduction(+:V)
    for (int i = 0; i < N; i++) {
        double Vi = 0;
        for (int bx = -1; bx < 2; bx++)
            for (int by = -1; by < 2; by++)
                for (int bz = -1; bz < 2; bz++)
                    for (int j = 0; j < N; j++) {
                        double r = 0, rij;
                        rij = x[3 * i + 0] - (x[3 * j + 0] + bx * L[0]); r = r + rij * rij;
                        rij = x[3 * i + 1] - (x[3 * j + 1] + by * L[1]); r = r + rij * rij;
                        rij = x[3 * i + 2] - (x[3 * j + 2] + bz * L[2]); r = r + rij * rij;
                        r = sqrt(r);
                        if (r < cutoff && !(i == j && bx == 0 && by == 0 && bz == 0))
                            Vi += (pow(sigma / r, 12) - pow(sigma / r, 6));
                    }
        V += Vi;
    }
    return 2 * eps * V;
}

/* identical_particles.cpp:477-510 (force functor, all pairs, 27 images, pow() form). */
void orc_force_all(int N, const double* x, const double L[3], double cutoff, double sigma, double eps, double* f)
{
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        double fi[3] = { 0, 0, 0 };
        for (int bx = -1; bx < 2; bx++)
            for (int by = -1; by < 2; by++)
                for (int bz = -1; bz < 2; bz++)
                    for (int j = 0; j < N; j++) {
                        double d[3] = { x[3 * i + 0] - (x[3 * j + 0] + bx * L[0]),
                                        x[3 * i + 1] - (x[3 * j + 1] + by * L[1]),
                                        x[3 * i + 2] - (x[3 * j + 2] + bz * L[2]) };
                        double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                        if (r < cutoff && !(i == j && bx == 0 && by == 0 && bz == 0)) {
                            double g = (-pow(sigma / r, 12) + 0.5 * pow(sigma / r, 6)) / (r * r);
                            fi[0] += g * d[0]; fi[1] += g * d[1]; fi[2] += g * d[2];
                        }
                    }
        for (int d = 0; d < 3; d++) f[3 * i + d] = fi[d] * (48 * eps);
    }
}

/* code-carsten/lennardjones.hh:5-9 */
double orc_lennard_jones(double r, int n, double a, double b)
{
    return a / pow(r, n) - b / pow(r, 6);
}

/* code-carsten/potential_energy.hh:6-15: sum over i<j of LennardJones(|r_i - r_j|, n, a, b), open
 * boundaries, no cutoff (the reference calls it with the defaults n=12, a=1e-12, b=1e-6). */
double orc_potential_energy_open(int N, const double* x, int n, double a, double b)
{
    double res = 0;
    for (int i = 0; i + 1 < N; i++)
        for (int j = i + 1; j < N; j++) {
            double d0 = x[3 * i] - x[3 * j], d1 = x[3 * i + 1] - x[3 * j + 1], d2 = x[3 * i + 2] - x[3 * j + 2];
            res += orc_lennard_jones(sqrt(d0 * d0 + d1 * d1 + d2 * d2), n, a, b);
        }
    return res;
}

/* ======================================================================================= */
/* identical_particles.cpp:465-468 */
double orc_kinetic(int N, const double* p, double mass)
{
    double K = 0;
    for (int i = 0; i < N; i++)
        K += (p[3 * i] * p[3 * i] + p[3 * i + 1] * p[3 * i + 1] + p[3 * i + 2] * p[3 * i + 2]) / (2 * mass);
    return K;
}

/* identical_particles.cpp:676-690: p -= dt * c * f, evaluated as (dt*c)*f then the subtraction */
void orc_kick(int N, double* p, const double* f, double dt, double coeff_p)
{
    for (int i = 0; i < 3 * N; i++) p[i] -= dt * coeff_p * f[i];
}

/* identical_particles.cpp:651-669: x += dt * c * p; x -= L * floor(x / L) */
void orc_drift(int N, double* x, const double* p, double dt, double coeff_x, const double L[3])
{
    for (int i = 0; i < N; i++)
        for (int d = 0; d < 3; d++) {
            double v = x[3 * i + d];
            v += dt * coeff_x * p[3 * i + d];
            v -= L[d] * floor(v / L[d]);
            x[3 * i + d] = v;
        }
}

/* ======================================================================================= */
/* identical_particles.cpp:739-776 Tag_RDF: per-dimension minimum of (d-L)^2, d^2, (d+L)^2; bin =
 * (int)(r / size_bin); ordered pairs i != j; g = count V / (1.33333333333 pi (rf^3 - ri^3)(N-1)N).
 * Counted over i<j and doubled: the three-candidate minimum is symmetric under d -> -d exactly. */
void orc_rdf(int N, const double* x, const double L[3], int nbins, double size_bin,
             unsigned long long* counts, double* rdf)
{
    memset(counts, 0, sizeof(unsigned long long) * (size_t)nbins);
    #pragma omp parallel
    {
        unsigned long long* loc = (unsigned long long*)calloc((size_t)nbins, sizeof(unsigned long long));
        #pragma omp for schedule(dynamic, 16)
        for (int i = 0; i < N; i++)
            for (int j = i + 1; j < N; j++) {
                double r = 0;
                for (int dir = 0; dir < 3; dir++) {
                    double dd = x[3 * j + dir] - x[3 * i + dir];
                    double ri = (dd - L[dir]) * (dd - L[dir]);
                    double si = dd * dd;
                    if (si < ri) ri = si;
                    si = (dd + L[dir]) * (dd + L[dir]);
                    if (si < ri) ri = si;
                    r += ri;
                }
                r = sqrt(r);
                int b = (int)(r / size_bin);
                if (b >= 0 && b < nbins) loc[b] += 2;
            }
        #pragma omp critical
        for (int b = 0; b < nbins; b++) counts[b] += loc[b];
        free(loc);
    }
    double V = L[0] * L[1] * L[2];
    for (int ib = 0; ib < nbins; ib++) {
        double ri = ib * size_bin;
        double rf = ri + size_bin;
        ri = ri * ri * ri;
        rf = rf * rf * rf;
        rdf[ib] = (double)counts[ib] * V / (1.33333333333 * M_PI * (rf - ri) * (N - 1) * N);
    }
}

/* ======================================================================================= */
int orc_system_init(orc_system* s, int N, const double L[3], double cutoff, double sigma,
                    double eps, double mass, double beta, int bin_mode)
{
    memset(s, 0, sizeof(*s));
    s->N = N;
    for (int d = 0; d < 3; d++) s->L[d] = L[d];
    s->cutoff = cutoff; s->sigma = sigma; s->eps = eps; s->mass = mass; s->beta = beta;
    s->bin_mode = bin_mode;
    int rc = orc_cutoff_binning(L, cutoff, s->nbin, s->sizebin);
    if (rc != 0) return rc;
    s->bintot = s->nbin[0] * s->nbin[1] * s->nbin[2];
    size_t n3 = sizeof(double) * 3 * (size_t)N;
    s->x = (double*)calloc(1, n3); s->p = (double*)calloc(1, n3);
    s->f = (double*)calloc(1, n3); s->x_saved = (double*)calloc(1, n3);
    s->perm = (int*)malloc(sizeof(int) * (size_t)N);
    s->perm_tmp = (int*)malloc(sizeof(int) * (size_t)N);
    s->bincount = (int*)calloc((size_t)s->bintot, sizeof(int));
    s->binoffsets = (int*)calloc((size_t)s->bintot, sizeof(int));
    for (int i = 0; i < N; i++) s->perm[i] = i;       /* identical_particles_binning.cpp:37-44 */
    s->lp_scale = orc_lp_scale(16);                   /* lp_leapfrog() : n_prec(16), integrator.hh:53 */
    return 0;
}

void orc_system_free(orc_system* s)
{
    free(s->x); free(s->p); free(s->f); free(s->x_saved);
    free(s->perm); free(s->perm_tmp); free(s->bincount); free(s->binoffsets);
    memset(s, 0, sizeof(*s));
}

void orc_sys_create_binning(orc_system* s)
{
    if (s->bin_mode == ORC_BIN_REFALGO)
        orc_binning_refalgo(s->N, s->x, s->nbin, s->sizebin, s->perm, s->perm_tmp, s->bincount, s->binoffsets);
    else
        orc_binning(s->N, s->x, s->nbin, s->sizebin, s->perm, s->bincount, s->binoffsets);
}

/* identical_particles.cpp:559-563: binning, then the force sweep */
void orc_sys_compute_force(orc_system* s)
{
    orc_sys_create_binning(s);
    orc_force_binning(s->N, s->x, s->L, s->nbin, s->perm, s->bincount, s->binoffsets,
                      s->cutoff, s->sigma, s->eps, s->f);
    s->force_evals++;
}

/* identical_particles.cpp:377-393 */
double orc_sys_compute_potential(orc_system* s)
{
    orc_sys_create_binning(s);
    return orc_potential_binning(s->N, s->x, s->L, s->nbin, s->perm, s->bincount, s->binoffsets,
                                 s->cutoff, s->sigma, s->eps);
}

/* lp_leapfrog / lp_omf4 (code-carsten/integrator.hh:51-79, 156-197) call round_and_update_momenta and
 * round_and_update_gauge with N = 10^n_prec (:61, :166); their definitions (update_momenta.hh,
 * update_gauge.hh) are not part of the reference tree.  Restated as: the update, then the updated
 * variable rounded to the grid 1/N, v -> round(v N) / N (round half away from zero) -- the
 * bit-reversible integration of Levesque & Verlet, J. Stat. Phys. 72, 519 (1993): with x and p on the
 * grid every update adds a rounded increment, and round(-d) = -round(d) makes the map exactly
 * reversible.  Positions are rounded before the wrap into the box. */
double orc_lp_scale(int n_prec)
{
    double s = 1.0;
    for (int i = 0; i < n_prec; i++) s *= 10.0;      /* exact for n_prec <= 22 */
    return s;
}

static inline double lp_round(double v, double scale) { return round(v * scale) / scale; }

void orc_kick_lp(int N, double* p, const double* f, double dt, double coeff_p, double scale)
{
    for (int i = 0; i < 3 * N; i++) p[i] = lp_round(p[i] - dt * coeff_p * f[i], scale);
}

void orc_drift_lp(int N, double* x, const double* p, double dt, double coeff_x, const double L[3], double scale)
{
    for (int i = 0; i < N; i++)
        for (int d = 0; d < 3; d++) {
            double v = x[3 * i + d];
            v += dt * coeff_x * p[3 * i + d];
            v = lp_round(v, scale);
            v -= L[d] * floor(v / L[d]);
            x[3 * i + d] = v;
        }
}

/* identical_particles.cpp:691-694 (coeff_p = beta, :642-644); rounded inside an lp_* integration */
void orc_sys_update_momenta(orc_system* s, double dt)
{
    orc_sys_compute_force(s);
    if (s->lp_active) orc_kick_lp(s->N, s->p, s->f, dt, s->beta, s->lp_scale);
    else orc_kick(s->N, s->p, s->f, dt, s->beta);
}

/* identical_particles.cpp:671-673 (coeff_x = beta / mass, :646-648); rounded inside an lp_* integration */
void orc_sys_update_positions(orc_system* s, double dt)
{
    if (s->lp_active) orc_drift_lp(s->N, s->x, s->p, dt, s->beta / s->mass, s->L, s->lp_scale);
    else orc_drift(s->N, s->x, s->p, dt, s->beta / s->mass, s->L);
}

/* integrator.cpp:52-131 (LEAP, OMF2, OMF4) and code-carsten/integrator.hh:200-240 (euler, ruth).
 * Loop bounds keep the reference's `size_t i < steps - 1` form for steps >= 1; steps < 1 is
 * rejected by the caller (Q18). OMF4 keeps the reference's i = 1 start (Q11). */
void orc_sys_integrate(orc_system* s, int scheme, double dt, int steps, double lambda)
{
    if (steps < 1) return;
    /* lp_leapfrog / lp_omf4: the leapfrog / OMF4 sequences with rounded updates (integrator.hh:51-79,
     * 156-197; the exponent given to set_integrator, :249, :257) */
    if (scheme == ORC_LP_LEAPFROG || scheme == ORC_LP_OMF4) {
        s->lp_active = 1;
        orc_sys_integrate(s, scheme == ORC_LP_LEAPFROG ? ORC_LEAPFROG : ORC_OMF4, dt, steps, lambda);
        s->lp_active = 0;
        return;
    }
    if (scheme == ORC_LEAPFROG) {
        orc_sys_update_momenta(s, dt / 2.);
        orc_sys_update_positions(s, dt);
        for (int i = 0; i < steps - 1; i++) {
            orc_sys_update_momenta(s, dt);
            orc_sys_update_positions(s, dt);
        }
        orc_sys_update_momenta(s, dt / 2.);
    } else if (scheme == ORC_OMF2) {
        double om2l = 1. - 2. * lambda;
        orc_sys_update_momenta(s, lambda * dt);
        for (int i = 0; i < steps - 1; i++) {
            orc_sys_update_positions(s, dt / 2.);
            orc_sys_update_momenta(s, om2l * dt);
            orc_sys_update_positions(s, dt / 2.);
            orc_sys_update_momenta(s, 2. * lambda * dt);
        }
        orc_sys_update_positions(s, dt / 2.);
        orc_sys_update_momenta(s, om2l * dt);
        orc_sys_update_positions(s, dt / 2.);
        orc_sys_update_momenta(s, lambda * dt);
    } else if (scheme == ORC_OMF4) {
        const double rho = 0.2539785108410595, theta = -0.03230286765269967;
        const double vartheta = 0.08398315262876693, lam = 0.6822365335719091;
        double e[10] = { rho * dt, lam * dt, theta * dt, 0.5 * (1 - 2. * (lam + vartheta)) * dt,
                         (1 - 2. * (theta + rho)) * dt, 0.5 * (1 - 2. * (lam + vartheta)) * dt,
                         theta * dt, lam * dt, rho * dt, 2 * vartheta * dt };
        orc_sys_update_momenta(s, 0.5 * e[9]);
        for (int i = 1; i < steps - 1; i++)
            for (int j = 0; j < 5; j++) {
                orc_sys_update_positions(s, e[2 * j]);
                orc_sys_update_momenta(s, e[2 * j + 1]);
            }
        for (int j = 0; j < 4; j++) {
            orc_sys_update_positions(s, e[2 * j]);
            orc_sys_update_momenta(s, e[2 * j + 1]);
        }
        orc_sys_update_positions(s, e[8]);
        orc_sys_update_momenta(s, 0.5 * e[9]);
    } else if (scheme == ORC_EULER) {
        for (int i = 0; i < steps; i++) {
            orc_sys_update_momenta(s, dt);
            orc_sys_update_positions(s, dt);
        }
    } else if (scheme == ORC_RUTH) {
        for (int i = 0; i < steps; i++) {
            orc_sys_update_momenta(s, dt);
            orc_sys_update_positions(s, -1. / 24. * dt);
            orc_sys_update_momenta(s, -2. / 3. * dt);
            orc_sys_update_positions(s, 3. / 4. * dt);
            orc_sys_update_momenta(s, 2. / 3. * dt);
            orc_sys_update_positions(s, 7. / 24. * dt);
        }
    }
}

/* HMC.cpp:61-136 (HMC_class::run) with istart = 0: heat bath, K_i, MD, V_f, K_f,
 * dh = beta (K_f + V_f - K_i - V_i), accept unconditionally while i < thermalization_steps,
 * otherwise accept iff gen_random() < exp(-dh); a rejection restores x (momenta are redrawn). */
int orc_hmc_run(orc_system* s, const orc_hmc_params* hp, double* rec_dh, int* rec_acc,
                double* rec_u, double* rec_k, double* rec_w)
{
    if (hp->steps < 1) return -1;
    orc_mt64 gen;
    orc_mt64_seed(&gen, hp->seed);
    double Vi = orc_sys_compute_potential(s);
    double Wi = orc_virial_binning(s->N, s->x, s->L, s->nbin, s->perm, s->bincount, s->binoffsets,
                                   s->cutoff, s->sigma, s->eps);
    double beta = s->beta;
    memcpy(s->x_saved, s->x, sizeof(double) * 3 * (size_t)s->N);
    int acc = 0;
    for (int t = 0; t < hp->ntraj; t++) {
        int i = t + 1;                                  /* first_traj = istart + 1 */
        orc_heatbath(s->N, s->mass, s->beta, hp->seed, hp->chain, (uint64_t)i, s->p);
        double Ki = orc_kinetic(s->N, s->p, s->mass);
        int steps = hp->steps;
        if (hp->randomize) steps = orc_binomial_steps(&gen, hp->steps);
        orc_sys_integrate(s, hp->scheme, hp->dt, steps, hp->lambda);
        double Vf = orc_sys_compute_potential(s);
        double Kf = orc_kinetic(s->N, s->p, s->mass);
        double dh = beta * (Kf + Vf - Ki - Vi);
        int accepted;
        if (i < hp->therm) accepted = 1;
        else {
            double r = orc_mt64_uniform(&gen);
            accepted = r < exp(-dh);
            acc += accepted;
        }
        if (accepted) {
            Vi = Vf;
            Wi = orc_virial_binning(s->N, s->x, s->L, s->nbin, s->perm, s->bincount, s->binoffsets,
                                    s->cutoff, s->sigma, s->eps);
            memcpy(s->x_saved, s->x, sizeof(double) * 3 * (size_t)s->N);
        } else {
            memcpy(s->x, s->x_saved, sizeof(double) * 3 * (size_t)s->N);
        }
        if (rec_dh) rec_dh[t] = dh;
        if (rec_acc) rec_acc[t] = accepted;
        if (rec_u) rec_u[t] = Vi;
        if (rec_k) rec_k[t] = accepted ? Kf : Ki;      /* Ki after HMC.cpp:104 (what print_xyz writes) */
        if (rec_w) rec_w[t] = Wi;
    }
    return acc;
}
