// Host side of libchmc: context, state transfer, sequencing of the hot path, the HMC loop (direct
// or hipGraph-captured), the autotuner of the LJ sweep tiling, and the C ABI of include/chmc.h.
#include "chmc.h"
#include "chmc_internal.hpp"
#include "philox.hpp"
#include "surrogate.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#pragma clang fp contract(off)   // the host lattice generator must round like the FP64 oracle

using namespace chmc;

static_assert(sizeof(chmc_hmc_record) == sizeof(HmcRecord), "record layout");

struct chmc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int flags = 0;
    int nchains = 0, Ntot = 0, Ctot = 0, nchunks = 0;
    std::vector<chmc_chain_desc> desc;
    std::vector<ChainDev> hchains;
    std::vector<Chunk> hchunks;
    DevView v{};
    ChainDev* d_chains = nullptr;
    std::vector<void*> allocs;
    int bin_state = 2;           // 0: bins of the current positions, 1: counted (placement pending), 2: stale
    bool have_bins = false;      // a placement exists that the local binning path can start from
    int force_kind = 0;          // 0 none, 1 fp32 stored, 2 fp64 stored
    int islots = 82;            // LJ sweep tiling: i-slots * 10 + waves per workgroup
    double pending_drift = 0.0;  // drift deferred into the next binning (integrator sequencing)
    int lp_exponent = 16;        // n_prec of the lp_* integrators (lp_leapfrog() : n_prec(16))
    int traj_done = 0;           // trajectories completed on the device counter
    int spec = 0;                // > 1: the chains are that many copies of one chain, run speculatively
    int pad_span = 0;            // every row of both list buffers is padded with no-op entries at least this far
    bool skip_next_check = false; // the next checked list evaluation's rebuild kernels already ran (a
                                  // trajectory's start check, before its save)
    std::string err;
    double* d_uniforms = nullptr;
    size_t uniforms_cap = 0;
    HmcRecord* d_rec = nullptr;
    size_t rec_cap = 0;
    unsigned long long* d_rdf = nullptr;
    int rdf_cap = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    // learned surrogate (allocated on first use)
    bool surr_alloc = false, surr_params = false;
    SurrView sv{};
    float *sG = nullptr, *sg = nullptr, *se = nullptr;
    float *Gtrain = nullptr, *utrain = nullptr;
    long long train_cap = 0, train_rows = 0;
    float* spartial = nullptr;
    long long spartial_cap = 0;
    float* sloss = nullptr;
    int adam_step = 0;
    bool md_surr = false;        // MD kicks use the surrogate's forces (set for the trajectories of an HMC run)
    bool surr_lists_eval = false; // the last surrogate evaluation ran on the lists (sG, sg in list slot order)
    // force matching: per-slot residual forces, LJ energies and dL/dg; stored configurations
    float *fmR = nullptr, *fmu = nullptr, *fmv = nullptr, *fmgrad = nullptr;
    double* xtrain = nullptr;
    int xtrain_cap = 0, xtrain_n = 0;
    // neighbour-list MD (allocated on first use)
    double skin = 0.6;           // the skin set for every chain, or the largest per-chain skin
    std::vector<double> chain_skin;  // per-chain skins when set individually (empty: skin for all)
    bool list_ready = false;
    int list_check = 1;          // rebuild checks before every list_check-th force evaluation of list MD
    int entry_bits = 0;          // list entry format: 16, 32, or 0 (chosen by the lists' size)
    int lpar = 0;                // list parity: which FP32 coordinate buffer / rebuild flag is current
    int cpar = 0;                // list count buffer of the next rebuild launch (alternates per launch)
    bool list_valid = false;     // the lists and their rebuild flags account for every position change
    bool saved_lists = false;    // the last save kept a valid list buffer with the positions
    bool lstate = false;         // list MD owns the state: xl, pl, xl_saved (slot order) are current and
                                 // x, p, x_saved (atom order) are not; false at every API return
    bool list_fresh = false;     // a sweep without drift (or the energy) ran since the last position
                                 // change: no rebuild request can be pending at the current parity
    bool rel64_fresh = false;    // rel64 holds the current positions (written by such a sweep)
    std::vector<void*> lallocs;
    std::vector<Chunk> htiles;
};

static thread_local std::string g_create_error;

static int fail(chmc_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}

#define HIPCHK(ctx, call)                                                                            \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess) return fail((ctx), CHMC_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define CHECK_CTX(ctx) do { if (!(ctx)) return CHMC_E_ARG; } while (0)

template <typename T>
static hipError_t dalloc(chmc_ctx* c, T** p, size_t n)
{
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
    if (e == hipSuccess) {
        c->allocs.push_back(q);
        *p = static_cast<T*>(q);
    }
    return e;
}

// identical_particles_binning.cpp:4-29
static int cutoff_binning(const double L[3], double cutoff, int nbin[3], double sizebin[3])
{
    for (int i = 0; i < 3; i++) {
        if (!(L[i] > cutoff)) return -1;
        nbin[i] = (int)floor(L[i] / cutoff);
        sizebin[i] = L[i] / ((double)nbin[i]);
        if (sizebin[i] < cutoff) return -2;
    }
    return 0;
}

static void destroy_ctx(chmc_ctx* c)
{
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (void* p : c->allocs) hipFree(p);
    for (void* p : c->lallocs) hipFree(p);
    if (c->d_uniforms) hipFree(c->d_uniforms);
    if (c->d_rec) hipFree(c->d_rec);
    if (c->d_rdf) hipFree(c->d_rdf);
    if (c->Gtrain) hipFree(c->Gtrain);
    if (c->utrain) hipFree(c->utrain);
    if (c->spartial) hipFree(c->spartial);
    if (c->xtrain) hipFree(c->xtrain);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->ev2) hipEventDestroy(c->ev2);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

extern "C" int chmc_create(chmc_ctx** out, int nchains, const chmc_chain_desc* chains, int device, int flags)
{
    if (!out || nchains < 1 || !chains) return fail(nullptr, CHMC_E_ARG, "chmc_create: bad arguments");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(nullptr, CHMC_E_NODEV, "no HIP device");
    chmc_ctx* c = new chmc_ctx();
    if (device < 0) hipGetDevice(&device);
    if (device >= ndev) { delete c; return fail(nullptr, CHMC_E_ARG, "device index out of range"); }
    c->device = device;
    c->flags = flags;
    c->nchains = nchains;
    c->desc.assign(chains, chains + nchains);
    // ---- per-chain geometry and offsets -------------------------------------------------------
    long long natoms = 0, ncells = 0;
    for (int k = 0; k < nchains; k++) {
        const chmc_chain_desc& d = chains[k];
        if (d.N < 1 || !(d.cutoff > 0) || !(d.mass > 0) || !(d.beta > 0) || !(d.sigma > 0)) {
            delete c;
            return fail(nullptr, CHMC_E_ARG, "chain " + std::to_string(k) + ": N, cutoff, sigma, mass and beta must be positive");
        }
        ChainDev cd{};
        if (cutoff_binning(d.L, d.cutoff, cd.nbin, cd.sizebin) != 0) {
            delete c;
            return fail(nullptr, CHMC_E_GEOMETRY, "chain " + std::to_string(k) + ": every box dimension must exceed the cutoff");
        }
        cd.atom_off = (int)natoms;
        cd.natoms = d.N;
        cd.cell_off = (int)ncells;
        cd.ncells = cd.nbin[0] * cd.nbin[1] * cd.nbin[2];
        for (int i = 0; i < 3; i++) {
            cd.L[i] = d.L[i];
            cd.sizebin_f[i] = (float)cd.sizebin[i];
        }
        cd.rc = d.cutoff; cd.rc2 = d.cutoff * d.cutoff;
        cd.sigma = d.sigma; cd.sigma2 = d.sigma * d.sigma;
        cd.eps = d.eps; cd.mass = d.mass; cd.beta = d.beta;
        cd.coeff_x = cd.beta / cd.mass;
        cd.rc2_f = (float)cd.rc2; cd.sigma2_f = (float)cd.sigma2; cd.eps48_f = (float)(48 * d.eps);
        cd.chunk_first = (int)c->hchunks.size();
        for (int b = 0; b < d.N; b += kRedChunk) c->hchunks.push_back(Chunk{ k, cd.atom_off + b, cd.atom_off + std::min(d.N, b + kRedChunk), 0 });
        cd.nchunks = (int)c->hchunks.size() - cd.chunk_first;
        natoms += d.N;
        ncells += cd.ncells;
        c->hchains.push_back(cd);
    }
    if (natoms > (1LL << 30) || ncells > (1LL << 28)) { delete c; return fail(nullptr, CHMC_E_ARG, "batch too large"); }
    c->Ntot = (int)natoms;
    c->Ctot = (int)ncells;
    c->nchunks = (int)c->hchunks.size();
    const int N = c->Ntot, Cn = c->Ctot;

    auto bail = [&](const char* what, hipError_t e) {
        std::string m = std::string(what) + ": " + hipGetErrorString(e);
        destroy_ctx(c);
        return fail(nullptr, CHMC_E_HIP, m);
    };
    hipError_t e;
    if ((e = hipSetDevice(device)) != hipSuccess) return bail("hipSetDevice", e);
    if ((e = hipStreamCreate(&c->stream)) != hipSuccess) return bail("hipStreamCreate", e);
    hipEventCreate(&c->ev0); hipEventCreate(&c->ev1); hipEventCreate(&c->ev2);
    DevView& v = c->v;
    v.Ntot = N; v.Ctot = Cn; v.nchains = nchains; v.nchunks = c->nchunks;
    int *atom_chain, *cell_chain;
    Chunk* chunks;
    if ((e = dalloc(c, &c->d_chains, nchains)) != hipSuccess) return bail("hipMalloc", e);
    if ((e = dalloc(c, &atom_chain, N)) != hipSuccess) return bail("hipMalloc", e);
    if ((e = dalloc(c, &cell_chain, Cn)) != hipSuccess) return bail("hipMalloc", e);
    if ((e = dalloc(c, &chunks, c->nchunks)) != hipSuccess) return bail("hipMalloc", e);
    if ((e = dalloc(c, &v.x, 3 * (size_t)N)) != hipSuccess) return bail("hipMalloc", e);
    if ((e = dalloc(c, &v.p, 3 * (size_t)N)) != hipSuccess) return bail("hipMalloc", e);
    if ((e = dalloc(c, &v.x_saved, 3 * (size_t)N)) != hipSuccess) return bail("hipMalloc", e);
    if ((e = dalloc(c, &v.f64, 3 * (size_t)N)) != hipSuccess) return bail("hipMalloc", e);
    if ((e = dalloc(c, &v.f32, 3 * (size_t)N)) != hipSuccess) return bail("hipMalloc", e);
    if ((e = dalloc(c, &v.cell_of, N)) != hipSuccess) return bail("hipMalloc", e);
    if ((e = dalloc(c, &v.slot, N)) != hipSuccess) return bail("hipMalloc", e);
    if ((e = dalloc(c, &v.newcell, N)) != hipSuccess) return bail("hipMalloc", e);
    if ((e = dalloc(c, &v.count, Cn)) != hipSuccess) return bail("hipMalloc", e);
    if ((e = dalloc(c, &v.offs_buf, 2 * ((size_t)Cn + 1))) != hipSuccess) return bail("hipMalloc", e);
    if ((e = dalloc(c, &v.bflag, 2)) != hipSuccess) return bail("hipMalloc", e);
    if ((e = dalloc(c, &v.skey, N)) != hipSuccess) return bail("hipMalloc", e);
    if ((e = dalloc(c, &v.satom, N)) != hipSuccess) return bail("hipMalloc", e);
    if ((e = dalloc(c, &v.perm_buf, 2 * (size_t)N)) != hipSuccess) return bail("hipMalloc", e);
    if ((e = dalloc(c, &v.invperm, N)) != hipSuccess) return bail("hipMalloc", e);
    if ((e = dalloc(c, &v.xs, N)) != hipSuccess) return bail("hipMalloc", e);
    if ((e = dalloc(c, &v.red, (size_t)nchains * RED_NSLOT)) != hipSuccess) return bail("hipMalloc", e);
    if ((e = dalloc(c, &v.chunk_part, (size_t)c->nchunks)) != hipSuccess) return bail("hipMalloc", e);
    if ((e = dalloc(c, &v.cell_part, 4 * (size_t)Cn)) != hipSuccess) return bail("hipMalloc", e);
    if ((e = dalloc(c, &v.state, nchains)) != hipSuccess) return bail("hipMalloc", e);
    if ((e = dalloc(c, &v.traj, 1)) != hipSuccess) return bail("hipMalloc", e);
    v.chains = c->d_chains; v.atom_chain = atom_chain; v.cell_chain = cell_chain; v.chunks = chunks;

    std::vector<int> ha(N), hc(Cn), iota(N);
    for (int k = 0; k < nchains; k++) {
        const ChainDev& cd = c->hchains[k];
        for (int i = 0; i < cd.natoms; i++) ha[cd.atom_off + i] = k;
        for (int i = 0; i < cd.ncells; i++) hc[cd.cell_off + i] = k;
    }
    for (int i = 0; i < N; i++) iota[i] = i;
    if ((e = hipMemcpy(c->d_chains, c->hchains.data(), sizeof(ChainDev) * nchains, hipMemcpyHostToDevice)) != hipSuccess) return bail("hipMemcpy", e);
    if ((e = hipMemcpy(atom_chain, ha.data(), sizeof(int) * N, hipMemcpyHostToDevice)) != hipSuccess) return bail("hipMemcpy", e);
    if ((e = hipMemcpy(cell_chain, hc.data(), sizeof(int) * Cn, hipMemcpyHostToDevice)) != hipSuccess) return bail("hipMemcpy", e);
    if ((e = hipMemcpy(chunks, c->hchunks.data(), sizeof(Chunk) * c->nchunks, hipMemcpyHostToDevice)) != hipSuccess) return bail("hipMemcpy", e);
    // permute_vector = identity (parallel_binning_init, identical_particles_binning.cpp:37-44)
    if ((e = hipMemcpy(v.perm_buf, iota.data(), sizeof(int) * N, hipMemcpyHostToDevice)) != hipSuccess) return bail("hipMemcpy", e);
    if ((e = hipMemset(v.bflag, 0, sizeof(int) * 2)) != hipSuccess) return bail("hipMemset", e);
    v.par = 0;
    if ((e = hipMemset(v.cell_of, 0, sizeof(int) * N)) != hipSuccess) return bail("hipMemset", e);
    if ((e = hipMemcpy(v.invperm, iota.data(), sizeof(int) * N, hipMemcpyHostToDevice)) != hipSuccess) return bail("hipMemcpy", e);
    if ((e = hipMemset(v.count, 0, sizeof(int) * Cn)) != hipSuccess) return bail("hipMemset", e);
    if ((e = hipMemset(v.traj, 0, sizeof(int))) != hipSuccess) return bail("hipMemset", e);
    if ((e = hipMemset(v.state, 0, sizeof(ChainState) * nchains)) != hipSuccess) return bail("hipMemset", e);
    if ((e = hipMemset(v.red, 0, sizeof(double) * nchains * RED_NSLOT)) != hipSuccess) return bail("hipMemset", e);
    if ((e = hipMemset(v.x, 0, sizeof(double) * 3 * (size_t)N)) != hipSuccess) return bail("hipMemset", e);
    if ((e = hipMemset(v.p, 0, sizeof(double) * 3 * (size_t)N)) != hipSuccess) return bail("hipMemset", e);
    if ((e = hipDeviceSynchronize()) != hipSuccess) return bail("hipDeviceSynchronize", e);
    *out = c;
    return CHMC_OK;
}

extern "C" int chmc_destroy(chmc_ctx* c)
{
    CHECK_CTX(c);
    destroy_ctx(c);
    return CHMC_OK;
}

extern "C" const char* chmc_last_error(const chmc_ctx* c)
{
    return c ? c->err.c_str() : g_create_error.c_str();
}

extern "C" int chmc_sync(chmc_ctx* c)
{
    CHECK_CTX(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CHMC_OK;
}

extern "C" int chmc_get_geometry(const chmc_ctx* c, int chain, int nbin[3], double sizebin[3], int* bintot)
{
    if (!c || chain < 0 || chain >= c->nchains) return CHMC_E_ARG;
    const ChainDev& cd = c->hchains[chain];
    for (int i = 0; i < 3; i++) {
        if (nbin) nbin[i] = cd.nbin[i];
        if (sizebin) sizebin[i] = cd.sizebin[i];
    }
    if (bintot) *bintot = cd.ncells;
    return CHMC_OK;
}

extern "C" int chmc_num_atoms(const chmc_ctx* c, int* total, int* nchains)
{
    CHECK_CTX(c);
    if (total) *total = c->Ntot;
    if (nchains) *nchains = c->nchains;
    return CHMC_OK;
}

// ---- state transfer ------------------------------------------------------------------------------
static int put_soa(chmc_ctx* c, double* dbase, int chain, const double* aos)
{
    const ChainDev& cd = c->hchains[chain];
    const int n = cd.natoms;
    std::vector<double> h(3 * (size_t)n);
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) h[(size_t)d * n + i] = aos[3 * i + d];
    for (int d = 0; d < 3; d++)
        HIPCHK(c, hipMemcpyAsync(dbase + (size_t)d * c->Ntot + cd.atom_off, h.data() + (size_t)d * n,
                                 sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CHMC_OK;
}

template <typename T>
static int get_soa(chmc_ctx* c, const T* dbase, int chain, double* aos)
{
    const ChainDev& cd = c->hchains[chain];
    const int n = cd.natoms;
    std::vector<T> h(3 * (size_t)n);
    for (int d = 0; d < 3; d++)
        HIPCHK(c, hipMemcpyAsync(h.data() + (size_t)d * n, dbase + (size_t)d * c->Ntot + cd.atom_off,
                                 sizeof(T) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) aos[3 * i + d] = (double)h[(size_t)d * n + i];
    return CHMC_OK;
}

static int flush_drift(chmc_ctx* c);

extern "C" int chmc_set_positions(chmc_ctx* c, int chain, const double* xyz)
{
    if (!c || !xyz || chain < 0 || chain >= c->nchains) return CHMC_E_ARG;
    c->list_valid = false;             // positions change outside the lists
    const ChainDev& cd = c->hchains[chain];
    // check_in_volume (identical_particles.cpp:228-236): every coordinate in [0, L)
    for (int i = 0; i < cd.natoms; i++)
        for (int d = 0; d < 3; d++) {
            const double q = xyz[3 * i + d];
            if (!(q >= 0.0 && q < cd.L[d]))
                return fail(c, CHMC_E_ARG, "particle " + std::to_string(i) + " outside the box in dimension " + std::to_string(d));
        }
    int rc = flush_drift(c);
    if (rc) return rc;
    c->bin_state = 2;
    c->have_bins = false;     // arbitrary new positions: the next binning takes the general path
    c->force_kind = 0;
    return put_soa(c, c->v.x, chain, xyz);
}

extern "C" int chmc_get_positions(chmc_ctx* c, int chain, double* xyz)
{
    if (!c || !xyz || chain < 0 || chain >= c->nchains) return CHMC_E_ARG;
    int rc = flush_drift(c);
    if (rc) return rc;
    return get_soa(c, c->v.x, chain, xyz);
}

extern "C" int chmc_set_momenta(chmc_ctx* c, int chain, const double* pxyz)
{
    if (!c || !pxyz || chain < 0 || chain >= c->nchains) return CHMC_E_ARG;
    int rc = flush_drift(c);
    if (rc) return rc;
    return put_soa(c, c->v.p, chain, pxyz);
}

extern "C" int chmc_get_momenta(chmc_ctx* c, int chain, double* pxyz)
{
    if (!c || !pxyz || chain < 0 || chain >= c->nchains) return CHMC_E_ARG;
    return get_soa(c, c->v.p, chain, pxyz);
}

static int lattice_side(int N)
{
    int n = (int)llround(cbrt((double)N));
    if (n < 1) n = 1;
    while ((long long)n * n * n < N) n++;
    while (n > 1 && (long long)(n - 1) * (n - 1) * (n - 1) >= N) n--;
    return n;
}

extern "C" int chmc_init_lattice(chmc_ctx* c, uint64_t seed, double jitter)
{
    CHECK_CTX(c);
    c->list_valid = false;             // positions change outside the lists
    if (!(jitter >= 0.0)) return CHMC_E_ARG;
    for (int k = 0; k < c->nchains; k++) {
        const ChainDev& cd = c->hchains[k];
        const int n = lattice_side(cd.natoms);
        std::vector<double> x(3 * (size_t)cd.natoms);
        for (int i = 0; i < cd.natoms; i++) {
            const int idx[3] = { i % n, (i / n) % n, i / (n * n) };
            const u64x4 r = philox4x64_10((uint64_t)i, (uint64_t)k, 0, 0, seed, kKeyLattice);
            for (int d = 0; d < 3; d++) {
                const double a = cd.L[d] / (double)n;
                const double site = ((double)idx[d] + 0.5) * a;
                const double u = u01_co(r.v[d]);
                double q = site + (2.0 * u - 1.0) * jitter;
                q = q - cd.L[d] * floor(q / cd.L[d]);
                x[3 * (size_t)i + d] = q;
            }
        }
        int rc = chmc_set_positions(c, k, x.data());
        if (rc) return rc;
    }
    return CHMC_OK;
}

// ---- hot-path sequencing ------------------------------------------------------------------------
static bool fp64_md(const chmc_ctx* c) { return (c->flags & CHMC_FORCE_FP64) != 0; }
static bool id_order(const chmc_ctx* c) { return (c->flags & CHMC_ORDER_ATOM_ID) != 0; }

static int ensure_bins(chmc_ctx* c);

// a count whose placement is still pending is placed before positions change again, so that every
// count compares against the cells of a completed placement
static int settle_bins(chmc_ctx* c)
{
    return c->bin_state == 1 ? ensure_bins(c) : CHMC_OK;
}

static int flush_drift(chmc_ctx* c)
{
    if (c->pending_drift != 0.0) {
        HIPCHK(c, launch_drift(c->v, c->pending_drift, c->stream));
        c->pending_drift = 0.0;
        c->bin_state = 2;
    }
    return CHMC_OK;
}

// bring the bins up to date with the current positions: a deferred drift is applied inside the count
// kernel; a counted state only needs the placement
static int ensure_bins(chmc_ctx* c)
{
    if (c->pending_drift != 0.0 || c->bin_state == 2) {
        HIPCHK(c, launch_count(c->v, c->pending_drift, c->stream));
        c->pending_drift = 0.0;
        c->bin_state = 1;
    }
    if (c->bin_state == 1) {
        HIPCHK(c, launch_scan(c->v, c->stream));
        HIPCHK(c, launch_place(c->v, c->have_bins && !id_order(c), id_order(c), c->stream));
        c->v.par ^= 1;                 // the placement wrote the other perm/offs buffer
        c->bin_state = 0;
        c->have_bins = true;
    }
    return CHMC_OK;
}

// create_binning: always a fresh binning of the current positions
static int do_bin(chmc_ctx* c)
{
    if (c->bin_state == 0) c->bin_state = 2;
    return ensure_bins(c);
}

static int do_drift(chmc_ctx* c, double dt)
{
    c->list_valid = false;             // positions change outside the lists
    int rc = flush_drift(c);
    if (rc) return rc;
    if ((rc = settle_bins(c))) return rc;
    c->pending_drift = dt;
    c->bin_state = 2;
    c->force_kind = 0;
    return CHMC_OK;
}

// update_momenta (force with fresh bins, then the kick) in one sweep; with fuse_drift the sweep also
// applies the following update_positions(drift_dt) to each atom and counts it into its new cell
static int do_force_kick(chmc_ctx* c, double dt, bool storef, bool fuse_drift, double drift_dt)
{
    int rc = ensure_bins(c);
    if (rc) return rc;
    const bool fuse = fuse_drift && !fp64_md(c);
    const int mode = LJ_FORCE | LJ_KICK | (storef ? LJ_STOREF : 0) | (fuse ? LJ_DRIFT : 0);
    HIPCHK(c, launch_lj(c->v, mode, fp64_md(c), c->islots, dt, fuse ? drift_dt : 0.0, c->stream));
    c->force_kind = (storef && !fuse) ? (fp64_md(c) ? 2 : 1) : 0;
    if (fuse) c->bin_state = 1;
    else if (fuse_drift) return do_drift(c, drift_dt);
    return CHMC_OK;
}

// FP64 energy / virial / kinetic sweep; `sums` adds the per-chain reduction into red[] (the HMC
// trajectory leaves it to the Metropolis kernel)
static int do_energy(chmc_ctx* c, bool rebin, bool sums = true)
{
    int rc = flush_drift(c);
    if (rc) return rc;
    rc = rebin ? do_bin(c) : ensure_bins(c);
    if (rc) return rc;
    HIPCHK(c, launch_lj(c->v, LJ_ENERGY, true, c->islots, 0.0, 0.0, c->stream));
    if (sums) HIPCHK(c, launch_reduce(c->v, SUM_CELLS, c->stream));
    return CHMC_OK;
}

static int read_red(chmc_ctx* c, int slot, double* out)
{
    std::vector<double> h((size_t)c->nchains * RED_NSLOT);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->v.red, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < c->nchains; k++) out[k] = h[(size_t)k * RED_NSLOT + slot];
    return CHMC_OK;
}

extern "C" int chmc_bin(chmc_ctx* c)
{
    CHECK_CTX(c);
    int rc = do_bin(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CHMC_OK;
}

extern "C" int chmc_get_bins(chmc_ctx* c, int chain, int* bincount, int* binoffsets, int* permute_vector)
{
    if (!c || chain < 0 || chain >= c->nchains) return CHMC_E_ARG;
    if (c->bin_state != 0 || c->pending_drift != 0.0) return fail(c, CHMC_E_STATE, "no binning of the current positions");
    const ChainDev& cd = c->hchains[chain];
    const int par = c->v.par;
    const int* offs_d = c->v.offs_buf + (size_t)par * (c->Ctot + 1);
    const int* perm_d = c->v.perm_buf + (size_t)par * c->Ntot;
    std::vector<int> offs(cd.ncells + 1), perm(cd.natoms);
    HIPCHK(c, hipMemcpyAsync(offs.data(), offs_d + cd.cell_off, sizeof(int) * (cd.ncells + 1), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(perm.data(), perm_d + cd.atom_off, sizeof(int) * cd.natoms, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < cd.ncells; k++) {
        if (bincount) bincount[k] = offs[k + 1] - offs[k];
        if (binoffsets) binoffsets[k] = offs[k] - cd.atom_off;
    }
    if (permute_vector)
        for (int k = 0; k < cd.natoms; k++) permute_vector[k] = perm[k] - cd.atom_off;
    return CHMC_OK;
}

extern "C" int chmc_compute_force(chmc_ctx* c, int fp64)
{
    CHECK_CTX(c);
    int rc = do_bin(c);
    if (rc) return rc;
    HIPCHK(c, launch_lj(c->v, LJ_FORCE | LJ_STOREF, fp64 != 0, c->islots, 0.0, 0.0, c->stream));
    c->force_kind = fp64 ? 2 : 1;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CHMC_OK;
}

extern "C" int chmc_get_force(chmc_ctx* c, int chain, double* grad)
{
    if (!c || !grad || chain < 0 || chain >= c->nchains) return CHMC_E_ARG;
    if (c->force_kind == 0) return fail(c, CHMC_E_STATE, "no force computed for the current positions");
    if (c->force_kind == 2) return get_soa(c, c->v.f64, chain, grad);
    return get_soa(c, c->v.f32, chain, grad);
}

extern "C" int chmc_compute_potential(chmc_ctx* c, double* V)
{
    if (!c || !V) return CHMC_E_ARG;
    int rc = do_energy(c, true);
    if (rc) return rc;
    return read_red(c, RED_V, V);
}

extern "C" int chmc_evaluate_potential(chmc_ctx* c, double* V)
{
    if (!c || !V) return CHMC_E_ARG;
    int rc = flush_drift(c);
    if (rc) return rc;
    if (!c->have_bins || c->bin_state == 1) {
        // evaluate_potential uses the bins as they are (identical_particles.cpp:387-393): positions moved
        // since the last binning keep their old cells; a binning must exist at all
        rc = ensure_bins(c);
        if (rc) return rc;
    }
    HIPCHK(c, launch_lj(c->v, LJ_ENERGY, true, c->islots, 0.0, 0.0, c->stream));
    HIPCHK(c, launch_reduce(c->v, SUM_CELLS, c->stream));
    return read_red(c, RED_V, V);
}

extern "C" int chmc_virial(chmc_ctx* c, double* W)
{
    if (!c || !W) return CHMC_E_ARG;
    int rc = do_energy(c, false);
    if (rc) return rc;
    return read_red(c, RED_W, W);
}

extern "C" int chmc_pair_counts(chmc_ctx* c, int chain, long long* ncand, long long* nint)
{
    if (!c || chain < 0 || chain >= c->nchains) return CHMC_E_ARG;
    int rc = do_energy(c, false);
    if (rc) return rc;
    std::vector<double> ni(c->nchains);
    rc = read_red(c, RED_NINT, ni.data());
    if (rc) return rc;
    // candidates of the 27-cell stencil: sum over cells of count(c) * sum_{27 neighbours} count(n), minus self
    const ChainDev& cd = c->hchains[chain];
    std::vector<int> offs(cd.ncells + 1);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int* offs_d = c->v.offs_buf + (size_t)c->v.par * (c->Ctot + 1);
    HIPCHK(c, hipMemcpy(offs.data(), offs_d + cd.cell_off, sizeof(int) * (cd.ncells + 1), hipMemcpyDeviceToHost));
    const int nx = cd.nbin[0], ny = cd.nbin[1], nz = cd.nbin[2];
    long long cand = 0;
    for (int bz = 0; bz < nz; bz++)
        for (int by = 0; by < ny; by++)
            for (int bx = 0; bx < nx; bx++) {
                const int ci = bx + nx * (by + ny * bz);
                const long long ni_c = offs[ci + 1] - offs[ci];
                long long nb = 0;
                for (int dz = -1; dz <= 1; dz++)
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++) {
                            const int j = ((bx + dx + nx) % nx) + nx * (((by + dy + ny) % ny) + ny * ((bz + dz + nz) % nz));
                            nb += offs[j + 1] - offs[j];
                        }
                cand += ni_c * nb - ni_c;
            }
    if (ncand) *ncand = cand;
    if (nint) *nint = (long long)llround(ni[chain]);
    return CHMC_OK;
}

extern "C" int chmc_kinetic(chmc_ctx* c, double* K)
{
    if (!c || !K) return CHMC_E_ARG;
    HIPCHK(c, launch_kinetic(c->v, c->stream));
    HIPCHK(c, launch_reduce(c->v, SUM_CHUNKS_K, c->stream));
    return read_red(c, RED_K, K);
}

extern "C" int chmc_update_momenta(chmc_ctx* c, double dt)
{
    CHECK_CTX(c);
    int rc = flush_drift(c);
    if (rc) return rc;
    rc = do_force_kick(c, dt, true, false, 0.0);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CHMC_OK;
}

extern "C" int chmc_update_positions(chmc_ctx* c, double dt)
{
    CHECK_CTX(c);
    int rc = flush_drift(c);
    if (rc) return rc;
    if ((rc = settle_bins(c))) return rc;
    HIPCHK(c, launch_drift(c->v, dt, c->stream));
    c->bin_state = 2;
    c->force_kind = 0;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CHMC_OK;
}

extern "C" int chmc_kick(chmc_ctx* c, double dt)
{
    CHECK_CTX(c);
    if (c->force_kind == 0) return fail(c, CHMC_E_STATE, "no force computed for the current positions");
    HIPCHK(c, launch_kick(c->v, dt, c->force_kind == 2, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CHMC_OK;
}

// ---- neighbour-list MD -------------------------------------------------------------------------
static bool list_md(const chmc_ctx* c) { return !fp64_md(c) && (c->flags & CHMC_CELL_MD) == 0; }

template <typename T>
static hipError_t lalloc(chmc_ctx* c, T** p, size_t n)
{
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, sizeof(T) * (n > 0 ? n : 1));
    if (e == hipSuccess) c->lallocs.push_back(q);
    *p = static_cast<T*>(q);
    return e;
}

// per-chain list grids (cells of side >= rc + skin of the chain; the skin shrinks where a box side is
// shorter than rc + skin), 32-slot tiles, and the list buffers.  skins: one per chain, or empty for
// `skin` in every chain.
static int list_setup(chmc_ctx* c, double skin, const std::vector<double>& skins)
{
    if (c->lstate) return fail(c, CHMC_E_STATE, "internal: list buffers replaced while list MD owns the state");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (void* p : c->lallocs) hipFree(p);
    c->lallocs.clear();
    c->list_ready = false;
    c->htiles.clear();
    long long lcells = 0, dense = 0;
    double nbmax = 0.0;
    std::vector<int> hlc;
    for (int k = 0; k < c->nchains; k++) {
        ChainDev& cd = c->hchains[k];
        double rl = cd.rc + (skins.empty() ? skin : skins[k]);
        for (int d = 0; d < 3; d++) rl = std::min(rl, cd.L[d]);
        const double sk = rl - cd.rc;
        double vol = 1.0;
        for (int d = 0; d < 3; d++) {
            cd.lnbin[d] = std::max(1, (int)std::floor(cd.L[d] / rl));
            cd.lsize[d] = cd.L[d] / cd.lnbin[d];
            cd.lsize_f[d] = (float)cd.lsize[d];
            vol *= cd.lsize[d] / cd.sigma;
        }
        // displacements are measured from the FP32 rebuild reference (within lsize * 2^-24 per
        // coordinate of the FP64 one): the trigger sits that much below skin / 2, so it is never late
        double lmax = 0.0;
        for (int d = 0; d < 3; d++) {
            lmax = std::max(lmax, cd.lsize[d]);
            if (cd.lnbin[d] > 1023) return fail(c, CHMC_E_GEOMETRY, "chain " + std::to_string(k) + ": more than 1023 list cells along a side");
        }
        const double half = std::max(0.0, 0.5 * sk - 4.0e-7 * lmax);
        cd.half_skin2 = half * half;
        cd.p2max = 100.0 * cd.mass / cd.beta;      // |p| <= 10 sqrt(m / beta): 10 thermal standard deviations
        cd.half_skin = std::sqrt(cd.half_skin2);
        cd.vmax = cd.coeff_x * std::sqrt(cd.p2max);
        cd.rl2_f = (float)(rl * rl * (1.0 + 1e-5));   // a margin over the FP32 rounding of r^2 at the rebuild
        // capacities from a local density bound of max(1.5 / sigma^3, 1.5 x the chain's mean density)
        const double rho = cd.natoms / (cd.L[0] * cd.L[1] * cd.L[2]) * cd.sigma * cd.sigma * cd.sigma;
        const double dens = std::max(1.5, 1.5 * rho);
        cd.lcap = (int)std::ceil(dens * vol) + 16;
        cd.lcap = std::min(cd.lcap, cd.natoms);     // a cell never holds more than the chain (small boxes)
        if (cd.lcap > 256)                          // one build thread per member of a cell
            return fail(c, CHMC_E_GEOMETRY, "chain " + std::to_string(k) + ": list cells too large for the list sweep (use CHMC_CELL_MD)");
        const double rls = rl / cd.sigma;
        nbmax = std::max(nbmax, 4.0 / 3.0 * M_PI * rls * rls * rls * dens + 16.0);
        cd.lcell_off = (int)lcells;
        cd.lncells = cd.lnbin[0] * cd.lnbin[1] * cd.lnbin[2];
        cd.ldense_off = (int)dense;
        for (int i = 0; i < cd.lncells; i++) hlc.push_back(k);
        lcells += cd.lncells;
        dense += (long long)cd.lncells * cd.lcap;
        cd.tile_first = (int)c->htiles.size();
        for (int b = 0; b < cd.natoms; b += kTile)
            c->htiles.push_back(Chunk{ k, cd.atom_off + b, cd.atom_off + std::min(cd.natoms, b + kTile), 0 });
        cd.ntiles = (int)c->htiles.size() - cd.tile_first;
    }
    if (lcells > (1LL << 28) || dense > (1LL << 30)) return fail(c, CHMC_E_ARG, "list grids too large");
    // entry format: 16-bit segment-relative entries halve the rows' bytes and pay where the rows no
    // longer fit the 256 MiB last-level cache (the sweep reads every row once per force evaluation);
    // below that the 32-bit slot entries decode in fewer instructions
    double row_bytes = 0.0;                   // 4 B per expected entry: pairs within rc + skin at the chain's density
    for (int k = 0; k < c->nchains; k++) {
        const ChainDev& cd = c->hchains[k];
        double rl = cd.rc + (skins.empty() ? skin : skins[k]);
        for (int d = 0; d < 3; d++) rl = std::min(rl, cd.L[d]);
        const double rho = cd.natoms / (cd.L[0] * cd.L[1] * cd.L[2]);
        row_bytes += 4.0 * cd.natoms * (4.0 / 3.0 * M_PI * rl * rl * rl * rho);
    }
    // (32-bit entries hold slot << 6: fewer than 2^25 slots)
    const bool slots32 = (size_t)c->Ntot < ((size_t)1 << (31 - kImageBits));
    if (c->entry_bits == 32 && !slots32) return fail(c, CHMC_E_ARG, "32-bit list entries: fewer than 2^25 atoms");
    const bool e16 = c->entry_bits == 16 || (c->entry_bits == 0 && (row_bytes > 256.0 * (1 << 20) || !slots32));
    const int nbcap = std::max(((int)std::ceil(nbmax) + 31) / 32 * 32, (kListSpan + 31) / 32 * 32);
    const size_t N = (size_t)c->Ntot;
    if ((double)N * nbcap >= 2147483647.0) return fail(c, CHMC_E_ARG, "neighbour lists too large (reduce the skin)");
    if (N >= (size_t)1 << 30) return fail(c, CHMC_E_ARG, "too many atoms for the list slots");
    DevView& v = c->v;
    v.entry16 = e16 ? 1 : 0;
    v.Cl = (int)lcells;
    v.ntiles = (int)c->htiles.size();
    v.nbcap = nbcap;
    v.lscan_in_build = 1;
    v.lcap_max = 1;
    int max_cells = 0;
    for (const ChainDev& cd : c->hchains) {
        if (cd.lncells > kScanCells) v.lscan_in_build = 0;
        v.lcap_max = std::max(v.lcap_max, cd.lcap);
        max_cells = std::max(max_cells, cd.lncells);
    }
    // a batch of large chains (speculative copies): each build workgroup scanning its chain's counts
    // costs more than a scan launch of its own (one chain keeps the in-build scan: its idle rebuild
    // checks would pay the extra launch)
    v.lscan_chains = v.lscan_in_build && c->nchains > 1 && max_cells >= kScanChainCells ? 1 : 0;
    int* lcell_chain = nullptr;
    Chunk* tiles = nullptr;
    HIPCHK(c, lalloc(c, &lcell_chain, (size_t)lcells));
    HIPCHK(c, lalloc(c, &tiles, (size_t)v.ntiles));
    HIPCHK(c, lalloc(c, &v.lcount, 2 * (size_t)lcells));
    // the per-slot list state twice (current / trajectory-start buffers)
    HIPCHK(c, lalloc(c, &v.loffs, 2 * ((size_t)lcells + 1)));
    HIPCHK(c, lalloc(c, &v.lmem, (size_t)dense));
    HIPCHK(c, lalloc(c, &v.lrank, (size_t)dense));
    HIPCHK(c, lalloc(c, &v.lmemx, 3 * (size_t)dense));
    HIPCHK(c, lalloc(c, &v.lmemp, 3 * (size_t)dense));
    HIPCHK(c, lalloc(c, &v.xl, 3 * N));
    HIPCHK(c, lalloc(c, &v.pl, 3 * N));
    HIPCHK(c, lalloc(c, &v.xl_saved, 3 * N));
    v.ldense = (int)dense;
    HIPCHK(c, lalloc(c, &v.lperm, 2 * N));
    HIPCHK(c, lalloc(c, &v.linv, 2 * N));
    v.nb_stride = N * nbcap * (e16 ? 2 : 4);
    HIPCHK(c, lalloc(c, &v.nb, 2 * v.nb_stride));
    HIPCHK(c, lalloc(c, &v.nbn, 2 * N));
    HIPCHK(c, lalloc(c, &v.lcs, 2 * N));
    HIPCHK(c, lalloc(c, &v.ctab, 2 * (size_t)lcells * kSegStride));
    HIPCHK(c, lalloc(c, &v.lbuf, 3 * (size_t)c->nchains));
    HIPCHK(c, lalloc(c, &v.tcur, (size_t)v.ntiles));
    HIPCHK(c, lalloc(c, &v.lflag, 4));
    HIPCHK(c, lalloc(c, &v.lreq, 2 * ((size_t)c->nchains + 1)));
    HIPCHK(c, lalloc(c, &v.lscan_tmp, (size_t)lcells));
    HIPCHK(c, lalloc(c, &v.lref, 2 * N));
    HIPCHK(c, lalloc(c, &v.rel64, 4 * N));
    HIPCHK(c, lalloc(c, &v.xsl, 2 * 2 * N));
    HIPCHK(c, lalloc(c, &v.tile_part, 4 * (size_t)v.ntiles));
    v.lcell_chain = lcell_chain;
    v.tiles = tiles;
    HIPCHK(c, hipMemcpy(lcell_chain, hlc.data(), sizeof(int) * hlc.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(tiles, c->htiles.data(), sizeof(Chunk) * c->htiles.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_chains, c->hchains.data(), sizeof(ChainDev) * c->nchains, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(v.lcount, 0, sizeof(int) * 2 * lcells));
    // before the first rebuild every slot's cell is its chain's first and every segment starts at slot 0
    HIPCHK(c, hipMemset(v.lcs, 0, sizeof(int) * 2 * N));
    HIPCHK(c, hipMemset(v.ctab, 0, sizeof(int) * 2 * (size_t)lcells * kSegStride));
    HIPCHK(c, hipMemset(v.lflag, 0, sizeof(int) * 4));
    HIPCHK(c, hipMemset(v.lreq, 0, sizeof(int) * 2 * (c->nchains + 1)));
    HIPCHK(c, hipMemset(v.lbuf, 0, sizeof(int) * 2 * c->nchains));
    HIPCHK(c, launch_fill_int(v.lbuf + 2 * c->nchains, -1, c->nchains, c->stream));
    HIPCHK(c, hipMemset(v.tcur, 0, sizeof(int) * v.ntiles));
    // loffs of a chain's first cell is its atom offset; the end entry is Ntot
    std::vector<int> lo0(lcells + 1, 0);
    for (const ChainDev& cd : c->hchains)
        for (int i = 0; i < cd.lncells; i++) lo0[cd.lcell_off + i] = cd.atom_off;
    lo0[lcells] = c->Ntot;
    for (int b = 0; b < 2; b++)
        HIPCHK(c, hipMemcpy(v.loffs + (size_t)b * (lcells + 1), lo0.data(), sizeof(int) * lo0.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(v.nbn, 0, sizeof(int) * 2 * N));
    HIPCHK(c, launch_list_pad_init(v, c->stream));      // every entry a no-op entry
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // valid atom ids in every slot before the first rebuild (a slot is rewritten by each rebuild)
    std::vector<int> iota(N);
    std::vector<float4> xs0(N);
    for (size_t i = 0; i < N; i++) {
        iota[i] = (int)i;
        int ii = (int)i;
        float w;
        std::memcpy(&w, &ii, sizeof(float));
        xs0[i] = make_float4(0.f, 0.f, 0.f, w);
    }
    for (int b = 0; b < 2; b++) {
        HIPCHK(c, hipMemcpy(v.lperm + b * N, iota.data(), sizeof(int) * N, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(v.linv + b * N, iota.data(), sizeof(int) * N, hipMemcpyHostToDevice));
        for (int q = 0; q < 2; q++)
            HIPCHK(c, hipMemcpy(v.xsl + (2 * b + q) * N, xs0.data(), sizeof(float4) * N, hipMemcpyHostToDevice));
    }
    HIPCHK(c, hipMemset(v.lref, 0, sizeof(float4) * 2 * N));
    HIPCHK(c, hipDeviceSynchronize());
    c->skin = skins.empty() ? skin : *std::max_element(skins.begin(), skins.end());
    c->chain_skin = skins;
    c->lpar = 0;
    c->cpar = 0;
    c->list_valid = false;
    if (v.lanes == 0) v.lanes = 8;
    v.lspan = std::min(sweep_span(v.lanes), nbcap);
    c->pad_span = v.lspan;      // whole rows are padded now, and every build pads to lspan
    c->list_ready = true;
    return CHMC_OK;
}

// the list-sweep shape; rows are padded to its unconditional span (longer rows built for another shape
// are fine; shorter ones are padded further here, in both buffers, before any sweep reads them)
static int set_list_shape(chmc_ctx* c, int lanes)
{
    const int span = std::min(sweep_span(lanes), c->v.nbcap);
    if (span > c->pad_span) HIPCHK(c, launch_list_repad(c->v, span, c->stream));
    c->pad_span = span;                                  // builds from here on pad to span
    c->v.lanes = lanes;
    c->v.lspan = span;
    return CHMC_OK;
}

static int list_ensure(chmc_ctx* c)
{
    return c->list_ready ? CHMC_OK : list_setup(c, c->skin, c->chain_skin);
}

// list MD takes the state over: x, p in the current buffer's slot order (xl, pl); h_x is taken over by
// the next save
static int list_import(chmc_ctx* c)
{
    if (c->lstate) return CHMC_OK;
    int rc = list_ensure(c);
    if (rc) return rc;
    HIPCHK(c, launch_list_import(c->v, c->stream));
    c->lstate = true;
    c->list_fresh = c->rel64_fresh = false;
    return CHMC_OK;
}

// hand the state back in atom order; saved: also h_x (a device HMC loop's last save)
static int list_export(chmc_ctx* c, bool saved)
{
    if (!c->lstate) return CHMC_OK;
    HIPCHK(c, launch_list_export(c->v, saved, c->stream));
    c->lstate = false;
    return CHMC_OK;
}

// the rebuild kernels of the current list parity (a no-op for chains without a request); every launch
// alternates the count buffer
static hipError_t list_rebuild(chmc_ctx* c)
{
    const hipError_t e = launch_list_rebuild(c->v, c->lpar, c->cpar, c->stream);
    c->cpar ^= 1;
    return e;
}

// request a rebuild before the next list sweep (positions changed outside the list path)
static int list_invalidate(chmc_ctx* c)
{
    c->list_fresh = c->rel64_fresh = false;
    HIPCHK(c, launch_fill_int(c->v.lreq + (size_t)c->lpar * (c->nchains + 1), 1, c->nchains + 1, c->stream));
    return CHMC_OK;
}

// rel64 (without fuse_drift and storef): the sweep also writes the FP64 positions the energy reads.
// checked = false skips the rebuild check (the caller's check interval guarantees valid lists)
// the rebuild kernels of a checked evaluation, unless a trajectory's start check just ran them
static int list_check_rebuild(chmc_ctx* c)
{
    if (c->skip_next_check) {
        c->skip_next_check = false;
        return CHMC_OK;
    }
    HIPCHK(c, list_rebuild(c));
    return CHMC_OK;
}

static int list_force_kick(chmc_ctx* c, double dt, bool storef, bool fuse_drift, double drift_dt, bool rel64 = false,
                           bool checked = true)
{
    int rc;
    if (checked && (rc = list_check_rebuild(c))) return rc;
    const bool r64 = rel64 && !fuse_drift && !storef;
    const int mode = LJ_FORCE | LJ_KICK | (storef ? LJ_STOREF : 0) | (fuse_drift ? LJ_DRIFT : 0) | (r64 ? LJ_REL64 : 0);
    HIPCHK(c, launch_list_sweep(c->v, mode, c->lpar, dt, fuse_drift ? drift_dt : 0.0, checked, c->stream));
    if (fuse_drift) c->lpar ^= 1;
    c->list_fresh = !fuse_drift;
    c->rel64_fresh = r64;
    c->force_kind = (storef && !fuse_drift) ? 1 : 0;
    c->bin_state = 2;
    return CHMC_OK;
}

// the same update with the learned surrogate's forces: descriptors and MLP of the current list state,
// then the surrogate sweep with the list sweep's kick / drift / rebuild-request tail
static int list_surr_force_kick(chmc_ctx* c, double dt, bool storef, bool fuse_drift, double drift_dt, bool rel64 = false,
                                bool checked = true)
{
    int rc;
    if (checked && (rc = list_check_rebuild(c))) return rc;
    const bool r64 = rel64 && !fuse_drift && !storef;
    const int mode = LJ_FORCE | LJ_KICK | (storef ? LJ_STOREF : 0) | (fuse_drift ? LJ_DRIFT : 0) | (r64 ? LJ_REL64 : 0);
    HIPCHK(c, launch_lsurr_desc_mlp(c->v, c->sv, c->lpar, c->sg, nullptr, nullptr, c->stream));
    HIPCHK(c, launch_lsurr_sweep(c->v, c->sg, mode, c->lpar, dt, fuse_drift ? drift_dt : 0.0, checked, c->stream));
    if (fuse_drift) c->lpar ^= 1;
    c->list_fresh = !fuse_drift;
    c->rel64_fresh = r64;
    c->force_kind = (storef && !fuse_drift) ? 1 : 0;
    c->bin_state = 2;
    return CHMC_OK;
}

static int list_drift(chmc_ctx* c, double dt)
{
    c->list_fresh = c->rel64_fresh = false;
    HIPCHK(c, launch_list_drift(c->v, c->lpar, dt, c->stream));
    c->lpar ^= 1;
    c->force_kind = 0;
    c->bin_state = 2;
    return CHMC_OK;
}

// FP64 energy / virial / kinetic per tile at the current positions (lists brought up to date first)
static int list_energy(chmc_ctx* c, bool sums)
{
    // after a sweep without drift the lists already hold the current positions (its rebuild check ran)
    if (!c->list_fresh) HIPCHK(c, list_rebuild(c));
    HIPCHK(c, launch_list_energy(c->v, c->lpar, !c->rel64_fresh, c->stream));
    c->list_fresh = c->rel64_fresh = true;
    if (sums) HIPCHK(c, launch_reduce(c->v, SUM_TILES, c->stream));
    return CHMC_OK;
}

static int list_check_overflow(chmc_ctx* c)
{
    int f[4] = { 0, 0, 0, 0 };
    HIPCHK(c, hipMemcpyAsync(f, c->v.lflag, sizeof(f), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (f[2] & 3) return fail(c, CHMC_E_STATE, "neighbour list capacity exceeded (cell members or list entries); results invalid");
    if (f[2] & 4)
        return fail(c, CHMC_E_STATE, "an atom's momentum exceeded 10 sqrt(m / beta), the bound of the list check interval: "
                                     "results invalid (chmc_set_list_check(ctx, 1) checks before every force evaluation)");
    return CHMC_OK;
}

// integrator_type::integrate (integrator.cpp:52-131; code-carsten integrator.hh:200-240) as a list of
// update_momenta (M) / update_positions (X) calls.  Every scheme alternates M and X; an M followed by an
// X runs as one sweep (kick, then the drift and the count of the next binning); a trailing X is
// deferred into the next binning.  storef keeps the force of the last evaluation readable.
struct MdOp { bool kick; double dt; };

static int build_ops(int scheme, double dt, int steps, double lambda, std::vector<MdOp>& ops)
{
    auto M = [&](double h) { ops.push_back(MdOp{ true, h }); };
    auto X = [&](double h) { ops.push_back(MdOp{ false, h }); };
    if (scheme == CHMC_LEAPFROG) {
        M(dt / 2.);
        X(dt);
        for (int i = 0; i < steps - 1; i++) { M(dt); X(dt); }
        M(dt / 2.);
    } else if (scheme == CHMC_OMF2) {
        const double om2l = 1. - 2. * lambda;
        M(lambda * dt);
        for (int i = 0; i < steps - 1; i++) { X(dt / 2.); M(om2l * dt); X(dt / 2.); M(2. * lambda * dt); }
        X(dt / 2.); M(om2l * dt); X(dt / 2.); M(lambda * dt);
    } else if (scheme == CHMC_OMF4) {
        const double rho = 0.2539785108410595, theta = -0.03230286765269967;
        const double vartheta = 0.08398315262876693, lam = 0.6822365335719091;
        const double e[10] = { rho * dt, lam * dt, theta * dt, 0.5 * (1 - 2. * (lam + vartheta)) * dt,
                               (1 - 2. * (theta + rho)) * dt, 0.5 * (1 - 2. * (lam + vartheta)) * dt,
                               theta * dt, lam * dt, rho * dt, 2 * vartheta * dt };
        M(0.5 * e[9]);
        for (int i = 1; i < steps - 1; i++)      // the reference starts at i = 1 (integrator.cpp:118)
            for (int j = 0; j < 5; j++) { X(e[2 * j]); M(e[2 * j + 1]); }
        for (int j = 0; j < 4; j++) { X(e[2 * j]); M(e[2 * j + 1]); }
        X(e[8]);
        M(0.5 * e[9]);
    } else if (scheme == CHMC_EULER) {
        for (int i = 0; i < steps; i++) { M(dt); X(dt); }
    } else if (scheme == CHMC_RUTH) {
        for (int i = 0; i < steps; i++) {
            M(dt); X(-1. / 24. * dt); M(-2. / 3. * dt); X(3. / 4. * dt); M(2. / 3. * dt); X(7. / 24. * dt);
        }
    } else {
        return -1;
    }
    return 0;
}

static int do_surr_force_kick(chmc_ctx* c, double dt, bool storef);
static int surr_geometry(chmc_ctx* c);
static int surr_check_flag(chmc_ctx* c);

static bool lp_scheme(int s) { return s == CHMC_LP_LEAPFROG || s == CHMC_LP_OMF4; }

static double lp_scale_of(int n_prec)
{
    double s = 1.0;
    for (int i = 0; i < n_prec; i++) s *= 10.0;      // exact for n_prec <= 22
    return s;
}

static int integrate_seq(chmc_ctx* c, int scheme, double dt, int steps, double lambda, bool storef, bool rel64_last = false);

// lp_leapfrog / lp_omf4 (code-carsten/integrator.hh:51-79, 156-197): the leapfrog / OMF4 sequence with
// every kernel that kicks or drifts rounding the updated variable (DevView::lp_scale); a drift left
// pending for the next binning is applied before the rounding is switched off again
static int integrate_lp(chmc_ctx* c, int scheme, double dt, int steps, double lambda, bool storef, bool rel64_last)
{
    c->v.lp_scale = lp_scale_of(c->lp_exponent);
    int rc = integrate_seq(c, scheme == CHMC_LP_LEAPFROG ? CHMC_LEAPFROG : CHMC_OMF4, dt, steps, lambda, storef,
                           rel64_last);
    if (!rc) rc = flush_drift(c);
    c->v.lp_scale = 0.0;
    return rc;
}

// rel64_last: a closing kick also writes the FP64 positions of the energy that follows
static int integrate_seq(chmc_ctx* c, int scheme, double dt, int steps, double lambda, bool storef, bool rel64_last)
{
    if (lp_scheme(scheme)) return integrate_lp(c, scheme, dt, steps, lambda, storef, rel64_last);
    std::vector<MdOp> ops;
    if (build_ops(scheme, dt, steps, lambda, ops) != 0) return fail(c, CHMC_E_ARG, "unknown integrator scheme");
    if (list_md(c)) {
        // neighbour-list MD on the slot-ordered state (LJ or, in a surrogate-driven trajectory, surrogate forces)
        int rc = flush_drift(c);
        if (rc) return rc;
        if (!c->lstate) return fail(c, CHMC_E_STATE, "internal: list MD without the slot-ordered state");
        if (!c->list_valid && (rc = list_invalidate(c))) return rc;
        c->list_valid = true;
        // rebuild checks before every list_check-th force evaluation, and before the last one and every
        // evaluation without a following drift; the trigger in between is lowered by what the skipped
        // drifts can add (lcheck_dt)
        const int m = c->list_check;
        double max_drift = 0.0;
        int nevals = 0;
        for (const MdOp& op : ops) {
            if (!op.kick) max_drift = std::max(max_drift, std::fabs(op.dt));
            else nevals++;
        }
        c->v.lcheck_dt = m > 1 ? (m - 1) * max_drift : 0.0;
        int ev = 0;
        for (size_t k = 0; k < ops.size() && !rc; k++) {
            if (ops[k].kick) {
                const bool next_x = k + 1 < ops.size() && !ops[k + 1].kick;
                const bool checked = m == 1 || ev % m == 0 || !next_x || ev == nevals - 1;
                rc = (c->md_surr ? list_surr_force_kick : list_force_kick)(c, ops[k].dt, storef, next_x,
                                                                          next_x ? ops[k + 1].dt : 0.0,
                                                                          rel64_last && k + 1 == ops.size(), checked);
                ev++;
                if (next_x) k++;
            } else {
                rc = list_drift(c, ops[k].dt);
            }
        }
        c->v.lcheck_dt = 0.0;
        return rc;
    }
    c->list_valid = false;
    for (size_t k = 0; k < ops.size(); k++) {
        int rc;
        if (ops[k].kick && c->md_surr) {
            rc = do_surr_force_kick(c, ops[k].dt, storef);
        } else if (ops[k].kick) {
            const bool next_x = k + 1 < ops.size() && !ops[k + 1].kick;
            rc = do_force_kick(c, ops[k].dt, storef, next_x, next_x ? ops[k + 1].dt : 0.0);
            if (next_x) k++;
        } else {
            rc = do_drift(c, ops[k].dt);
        }
        if (rc) return rc;
    }
    return CHMC_OK;
}

static bool scheme_ok(int s)
{
    return s == CHMC_LEAPFROG || s == CHMC_OMF2 || s == CHMC_OMF4 || s == CHMC_EULER || s == CHMC_RUTH || lp_scheme(s);
}

extern "C" int chmc_set_lp_exponent(chmc_ctx* c, int n_prec)
{
    CHECK_CTX(c);
    if (n_prec < 0 || n_prec > 18) return fail(c, CHMC_E_ARG, "set_lp_exponent: n_prec must be in [0, 18]");
    c->lp_exponent = n_prec;
    return CHMC_OK;
}

extern "C" int chmc_get_lp_exponent(const chmc_ctx* c, int* n_prec)
{
    if (!c || !n_prec) return CHMC_E_ARG;
    *n_prec = c->lp_exponent;
    return CHMC_OK;
}

extern "C" int chmc_integrate(chmc_ctx* c, int scheme, double dt, int steps, double lambda)
{
    CHECK_CTX(c);
    if (steps < 1 || !scheme_ok(scheme)) return fail(c, CHMC_E_ARG, "integrate: steps must be >= 1 and the scheme known");
    int rc = flush_drift(c);
    if (rc) return rc;
    if (list_md(c) && (rc = list_import(c))) return rc;
    rc = integrate_seq(c, scheme, dt, steps, lambda, true);
    if (rc) return rc;
    if ((rc = list_export(c, false))) return rc;
    rc = flush_drift(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (list_md(c)) return list_check_overflow(c);
    return CHMC_OK;
}

extern "C" int chmc_heatbath(chmc_ctx* c, uint64_t seed, uint64_t traj)
{
    CHECK_CTX(c);
    HIPCHK(c, launch_heatbath(c->v, seed, (long long)traj, false, false, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CHMC_OK;
}

extern "C" int chmc_save_positions(chmc_ctx* c)
{
    CHECK_CTX(c);
    int rc = flush_drift(c);
    if (rc) return rc;
    // with valid lists the save also keeps their buffer, so a restore needs no rebuild; with start
    // rebuilds the lists are first rebuilt at the saved positions (as a trajectory's start check does
    // after an accepted trajectory), so that every trajectory from them keeps the whole skin
    c->saved_lists = list_md(c) && c->list_ready && c->list_valid;
    if (c->saved_lists && c->v.start_rebuild) {
        if ((rc = list_import(c)) || (rc = list_invalidate(c))) return rc;
        HIPCHK(c, list_rebuild(c));
        HIPCHK(c, launch_fill_int(c->v.lreq + (size_t)c->lpar * (c->nchains + 1), 0, c->nchains + 1, c->stream));
        if ((rc = list_export(c, false))) return rc;
    }
    HIPCHK(c, launch_save_restore(c->v, 0, c->saved_lists ? c->lpar : -1, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CHMC_OK;
}

extern "C" int chmc_restore_positions(chmc_ctx* c)
{
    CHECK_CTX(c);
    int rc = flush_drift(c);
    if (rc) return rc;
    if ((rc = settle_bins(c))) return rc;
    const bool lists = c->saved_lists && list_md(c) && c->list_ready;
    c->list_valid = lists;             // the saved list buffer is valid for the restored positions
    HIPCHK(c, launch_save_restore(c->v, 1, lists ? c->lpar : -1, c->stream));
    c->bin_state = 2;
    c->force_kind = 0;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CHMC_OK;
}

extern "C" int chmc_set_trajectory(chmc_ctx* c, int istart)
{
    CHECK_CTX(c);
    if (istart < 0) return fail(c, CHMC_E_ARG, "set_trajectory: istart must be >= 0");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->v.traj, &istart, sizeof(int), hipMemcpyHostToDevice));
    c->traj_done = istart;
    return CHMC_OK;
}

extern "C" int chmc_get_trajectory(const chmc_ctx* c, int* istart)
{
    if (!c || !istart) return CHMC_E_ARG;
    *istart = c->traj_done;
    return CHMC_OK;
}

// ---- HMC ----------------------------------------------------------------------------------------
// mt19937_64 of HMC.cpp:38 (gen64), host side, used to draw the Metropolis uniforms up front
struct Mt64 {
    uint64_t mt[312];
    int mti;
    explicit Mt64(uint64_t seed)
    {
        mt[0] = seed;
        for (int i = 1; i < 312; i++) mt[i] = 6364136223846793005ULL * (mt[i - 1] ^ (mt[i - 1] >> 62)) + (uint64_t)i;
        mti = 312;
    }
    uint64_t next()
    {
        const uint64_t UM = 0xFFFFFFFF80000000ULL, LM = 0x7FFFFFFFULL, MA = 0xB5026F5AA96619E9ULL;
        if (mti >= 312) {
            int i;
            for (i = 0; i < 156; i++) {
                uint64_t x = (mt[i] & UM) | (mt[i + 1] & LM);
                mt[i] = mt[i + 156] ^ (x >> 1) ^ ((x & 1ULL) ? MA : 0ULL);
            }
            for (; i < 311; i++) {
                uint64_t x = (mt[i] & UM) | (mt[i + 1] & LM);
                mt[i] = mt[i - 156] ^ (x >> 1) ^ ((x & 1ULL) ? MA : 0ULL);
            }
            uint64_t x = (mt[311] & UM) | (mt[0] & LM);
            mt[311] = mt[155] ^ (x >> 1) ^ ((x & 1ULL) ? MA : 0ULL);
            mti = 0;
        }
        uint64_t x = mt[mti++];
        x ^= (x >> 29) & 0x5555555555555555ULL;
        x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
        x ^= (x << 37) & 0xFFF7EEE000000000ULL;
        x ^= (x >> 43);
        return x;
    }
    // gen_random, HMC.cpp:48-50
    double uniform() { return ((double)next() - 0.0) / (double)(0xFFFFFFFFFFFFFFFFULL - 0ULL); }
    // set_binomial_steps, integrator.cpp:37-49
    int binomial(int average_steps)
    {
        int s = 0;
        for (int i = 0; i < 2 * average_steps; i++)
            if (uniform() < 0.5) s++;
        return s;
    }
};

// The start check of a list-MD trajectory (with start rebuilds, chmc_set_list_start_rebuild): the
// rebuild kernels (a chain rebuilds when a request is pending: after an accepted trajectory, whose end
// positions its lists were not built at, or after positions changed outside the lists), then the save
// of the start state, so that the trajectory and every rejected one after it start from lists built at
// the start positions and keep the whole skin for their motion.  The first force evaluation then
// consumes the requests without running the rebuild kernels again.  It pays where most trajectories
// are rejected (each start serves many); at high acceptance the extra rebuild per accepted trajectory
// costs more than it saves.
static int list_start_check(chmc_ctx* c)
{
    HIPCHK(c, list_rebuild(c));
    HIPCHK(c, launch_save_restore_slots(c->v, 0, c->lpar, c->stream));
    c->skip_next_check = true;
    return CHMC_OK;
}

// one trajectory after the heat bath's counter is known: everything between hb() and the copy of
// HMC.cpp:103/111 (HMC.cpp:77-117), enqueued on the context stream
static int enqueue_trajectory(chmc_ctx* c, const chmc_hmc_params& hp, int steps, int ufirst, int rec_base)
{
    int rc;
    if (list_md(c)) {
        // the start check; heat bath + K_i (momenta and K_i partials in the start lists' slot order);
        // neighbour-list MD (LJ or surrogate forces); FP64 LJ V_f, W_f, K_f over the lists
        if (!c->list_valid && (rc = list_invalidate(c))) return rc;      // positions moved outside the lists
        c->list_valid = true;
        if (c->v.start_rebuild && (rc = list_start_check(c))) return rc;
        HIPCHK(c, launch_heatbath(c->v, hp.seed, -1, false, true, c->stream));
        if (steps > 0 && (rc = integrate_seq(c, hp.scheme, hp.dt, steps, hp.lambda, false, true))) return rc;
        c->skip_next_check = false;
        if (steps == 0 && !c->list_valid && (rc = list_invalidate(c))) return rc;
        c->list_valid = true;
        if ((rc = list_energy(c, false))) return rc;
        HIPCHK(c, launch_metropolis(c->v, c->d_uniforms, hp.ntraj, hp.therm, ufirst, c->d_rec, rec_base, c->md_surr ? 1 : 0,
                                    SUM_TILES, c->stream));
        HIPCHK(c, launch_save_restore_slots(c->v, 2, c->lpar, c->stream));
        c->list_fresh = c->rel64_fresh = false;       // a rejected chain returned to its start positions
        c->force_kind = 0;
        c->bin_state = 2;
        return CHMC_OK;
    }
    // heat bath + K_i + the count phase of the binning of the (accepted or restored) positions
    HIPCHK(c, launch_heatbath(c->v, hp.seed, -1, true, false, c->stream));
    c->bin_state = 1;
    // a zero-length binomial draw leaves x and p unchanged (the reference's size_t loop would not end, Q18)
    if (steps > 0 && (rc = integrate_seq(c, hp.scheme, hp.dt, steps, hp.lambda, false))) return rc;
    // V_f: after LEAP/OMF2/OMF4 the last binning is of the final positions; after a closing drift
    // (euler/ruth) it is redone -- the reference rebins in either case (identical_particles.cpp:379)
    if ((rc = do_energy(c, false, false))) return rc;               // V_f, W_f and K_f partials
    HIPCHK(c, launch_metropolis(c->v, c->d_uniforms, hp.ntraj, hp.therm, ufirst, c->d_rec, rec_base, c->md_surr ? 1 : 0,
                                SUM_CELLS, c->stream));
    HIPCHK(c, launch_save_restore(c->v, 2, -1, c->stream));
    // the positions moved outside the list path: lists rebuild before the next list sweep
    if (list_md(c) && c->list_ready && (rc = list_invalidate(c))) return rc;
    c->force_kind = 0;
    c->bin_state = 2;
    return CHMC_OK;
}

// One speculative batch: the chains are copies of one chain at its state after trajectory t0 (the
// device counter); copy k runs trajectory t0 + 1 + k (heat bath of that trajectory, MD, FP64 energy),
// the decisions consume them in order up to the first accepted one (or the run's last trajectory tend),
// and the accepted copy's state goes to every copy.  The chain is the one the sequential loop of
// enqueue_trajectory produces, trajectory for trajectory: a rejected trajectory leaves the state
// unchanged, so every trajectory up to the first accepted one starts from the same state.
static int enqueue_spec_batch(chmc_ctx* c, const chmc_hmc_params& hp, int ufirst, int rec_base)
{
    int rc;
    if (c->v.start_rebuild && (rc = list_start_check(c))) return rc;
    HIPCHK(c, launch_heatbath_spec(c->v, hp.seed, c->stream));
    if ((rc = integrate_seq(c, hp.scheme, hp.dt, hp.steps, hp.lambda, false, true))) return rc;
    c->skip_next_check = false;
    c->list_valid = true;
    if ((rc = list_energy(c, false))) return rc;
    HIPCHK(c, launch_reduce(c->v, SUM_TILES | SUM_CHUNKS_KI, c->stream));
    HIPCHK(c, launch_spec_restore(c->v, c->d_uniforms, hp.therm, ufirst, c->d_rec, rec_base, c->md_surr ? 1 : 0, c->lpar,
                                  c->stream));
    HIPCHK(c, launch_spec_broadcast(c->v, -1, true, true, c->stream));
    c->list_fresh = c->rel64_fresh = false;
    c->force_kind = 0;
    c->bin_state = 2;
    return CHMC_OK;
}

// a captured trajectory sequence: valid for replays that start at binning parity `par`, list parity
// `lpar` and list count buffer `cpar` in MD mode `surr` (spec: speculative batches)
struct TrajGraph {
    hipGraphExec_t exec;
    int par, lpar, cpar, per_graph;
    bool surr, spec;
};

// count units (trajectories or speculative batches) enqueued by enq(), by graph replays
template <typename Enq>
static int run_units(chmc_ctx* c, int count, bool spec, std::vector<TrajGraph>& graphs, Enq enq)
{
    int rc;
    int done = 0;
    const TrajGraph* g = nullptr;
    for (const TrajGraph& k : graphs)
        if (k.par == c->v.par && k.lpar == c->lpar && k.cpar == c->cpar && k.surr == c->md_surr && k.spec == spec) g = &k;
    if (!g) {
        // one unit per graph when the parities baked into the captured launches (binning, list,
        // list count buffer) end where they started, two otherwise.  The host state after the capture
        // is that of per_graph units, which every replay reproduces.
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        HIPCHK(c, hipStreamSynchronize(c->stream));
        const int par0 = c->v.par, lpar0 = c->lpar, cpar0 = c->cpar;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        rc = enq();
        int per_graph = 1;
        if (!rc && (c->v.par != par0 || c->lpar != lpar0 || c->cpar != cpar0)) {
            rc = enq();
            per_graph = 2;
        }
        hipError_t ec = hipStreamEndCapture(c->stream, &graph);
        if (rc) { if (graph) hipGraphDestroy(graph); return rc; }
        HIPCHK(c, ec);
        hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        HIPCHK(c, ei);
        graphs.push_back(TrajGraph{ exec, par0, lpar0, cpar0, per_graph, c->md_surr, spec });
        g = &graphs.back();
        if (count < per_graph) {
            // nothing of the capture ran; the host state is restored by running the leftover directly
            // from the capture's start parities
            c->v.par = par0;
            c->lpar = lpar0;
            c->cpar = cpar0;
        } else {
            HIPCHK(c, hipGraphLaunch(g->exec, c->stream));
            done = per_graph;
        }
    }
    for (; done + g->per_graph <= count; done += g->per_graph) HIPCHK(c, hipGraphLaunch(g->exec, c->stream));
    for (; done < count; done++)
        if ((rc = enq())) return rc;
    return CHMC_OK;
}

// trajectories t0 .. t0+count-1 of a run in the context's current MD mode, by graph replays where
// allowed and directly otherwise
static int run_segment(chmc_ctx* c, const chmc_hmc_params& hp, int t0, int count, int ufirst, int rec_base,
                       const std::vector<int>& rsteps, std::vector<TrajGraph>& graphs)
{
    int rc;
    if (!hp.use_graph || hp.randomize) {
        for (int t = t0; t < t0 + count; t++) {
            const int steps = hp.randomize ? rsteps[t] : hp.steps;
            if ((rc = enqueue_trajectory(c, hp, steps, ufirst, rec_base))) return rc;
        }
        return CHMC_OK;
    }
    return run_units(c, count, false, graphs, [&]() { return enqueue_trajectory(c, hp, hp.steps, ufirst, rec_base); });
}

// a speculative run: batches of nchains trajectories until the device counter reaches tend.  A batch
// consumes between one trajectory and nchains, so the host enqueues as many batches as the remaining
// trajectories need when none is accepted, reads the counter and repeats.
static int run_spec(chmc_ctx* c, const chmc_hmc_params& hp, int ufirst, int rec_base, int done0, int tend,
                    std::vector<TrajGraph>& graphs)
{
    const int K = c->nchains;
    int rc;
    // the window's last trajectory, read by the decisions of every batch (graph replays included), and
    // its first: the device counter holds the trajectories committed so far
    HIPCHK(c, launch_fill_int(c->v.spec_ctl + 3, tend, 1, c->stream));
    int done = done0;
    while (done < tend) {
        const int nb = (tend - done + K - 1) / K;
        auto enq = [&]() { return enqueue_spec_batch(c, hp, ufirst, rec_base); };
        if (hp.use_graph) {
            if ((rc = run_units(c, nb, true, graphs, enq))) return rc;
        } else {
            for (int b = 0; b < nb; b++)
                if ((rc = enq())) return rc;
        }
        HIPCHK(c, hipStreamSynchronize(c->stream));
        const int before = done;
        HIPCHK(c, hipMemcpy(&done, c->v.traj, sizeof(int), hipMemcpyDeviceToHost));
        if (done <= before || done > tend)     // every batch consumes at least one trajectory
            return fail(c, CHMC_E_STATE, "hmc_run: speculative batches did not advance the trajectory counter");
    }
    return CHMC_OK;
}

extern "C" int chmc_hmc_run(chmc_ctx* c, const chmc_hmc_params* hp, chmc_hmc_record* records, int* nacc)
{
    if (!c || !hp) return CHMC_E_ARG;
    int rc;
    if (hp->ntraj < 1 || hp->steps < 1 || !scheme_ok(hp->scheme) || hp->therm < 0)
        return fail(c, CHMC_E_ARG, "hmc_run: ntraj >= 1, steps >= 1, therm >= 0 and a known scheme are required");
    if (hp->surrogate && !c->surr_params) return fail(c, CHMC_E_STATE, "hmc_run: surrogate parameters not set");
    const bool spec = c->spec > 1;
    if (spec) {
        if (hp->randomize)
            return fail(c, CHMC_E_ARG, "hmc_run: speculative trajectories need fixed-length trajectories");
        for (double sk : c->chain_skin)
            if (sk != c->chain_skin[0]) return fail(c, CHMC_E_ARG, "hmc_run: speculative copies need one skin");
    }
    if (list_md(c) && (rc = list_ensure(c))) return rc;          // allocations stay outside graph capture
    if (hp->surrogate && (rc = surr_geometry(c))) return rc;
    const int nch = c->nchains, ntraj = hp->ntraj;
    rc = flush_drift(c);
    if (rc) return rc;
    // ---- Metropolis uniforms: chain k draws from mt19937_64(seed + k) in the reference's order ----
    const int rec_base = c->traj_done;
    const int ufirst = std::max(hp->therm, rec_base + 1);
    std::vector<Mt64> gens;
    for (int k = 0; k < nch; k++) gens.emplace_back(hp->seed + (uint64_t)k);
    std::vector<int> rsteps;
    std::vector<double> u((size_t)nch * ntraj, 2.0);
    if (hp->randomize && nch != 1) return fail(c, CHMC_E_ARG, "randomize_trajectory needs a single-chain context");
    if (hp->uniforms) {
        std::copy(hp->uniforms, hp->uniforms + (size_t)nch * ntraj, u.begin());
        if (hp->randomize) {
            if (!hp->traj_steps) return fail(c, CHMC_E_ARG, "hmc_run: caller uniforms with randomize need traj_steps");
            rsteps.assign(hp->traj_steps, hp->traj_steps + ntraj);
        }
    } else if (!hp->randomize) {
        for (int k = 0; k < nch; k++)
            for (int t = 0; t < ntraj; t++) u[(size_t)k * ntraj + t] = gens[k].uniform();
    } else {
        int q = 0;
        for (int t = 0; t < ntraj; t++) {
            rsteps.push_back(hp->traj_steps ? hp->traj_steps[t] : gens[0].binomial(hp->steps));
            if (rec_base + t + 1 >= hp->therm) u[q++] = gens[0].uniform();
        }
    }
    for (int st : rsteps)
        if (st < 0) return fail(c, CHMC_E_ARG, "hmc_run: negative trajectory length");
    if (c->uniforms_cap < u.size()) {
        if (c->d_uniforms) hipFree(c->d_uniforms);
        c->d_uniforms = nullptr;
        HIPCHK(c, hipMalloc(&c->d_uniforms, sizeof(double) * u.size()));
        c->uniforms_cap = u.size();
    }
    const size_t nrec = (size_t)nch * ntraj;
    if (c->rec_cap < nrec) {
        if (c->d_rec) hipFree(c->d_rec);
        c->d_rec = nullptr;
        HIPCHK(c, hipMalloc(&c->d_rec, sizeof(HmcRecord) * nrec));
        c->rec_cap = nrec;
    }
    HIPCHK(c, hipMemcpyAsync(c->d_uniforms, u.data(), sizeof(double) * u.size(), hipMemcpyHostToDevice, c->stream));

    // ---- V_i, W_i of the start configuration (HMC.cpp:64), h_x = x (HMC.cpp:72) ----
    const bool slot_run = list_md(c);                          // the run starts on the slot-ordered state
    if (slot_run) {
        // with start rebuilds the run starts from lists built at its start positions (as after an
        // accepted trajectory)
        const bool rebuild_all = c->v.start_rebuild || !c->list_valid;
        if ((rc = list_import(c)) || (rebuild_all && (rc = list_invalidate(c)))) return rc;
        // speculative copies start from chain 0's state (positions, lists, rebuild requests); its list
        // rows only when the copies do not all rebuild next
        if (spec) HIPCHK(c, launch_spec_broadcast(c->v, 0, false, !rebuild_all, c->stream));
        if ((rc = list_energy(c, true))) return rc;
        c->list_valid = true;
    } else if ((rc = do_energy(c, true))) {
        return rc;
    }
    std::vector<ChainState> st(nch);
    HIPCHK(c, launch_state_init(c->v, c->stream));
    if (slot_run) {
        HIPCHK(c, launch_save_restore_slots(c->v, 0, c->lpar, c->stream));
    } else {
        const bool lists = list_md(c) && c->list_ready && c->list_valid;
        HIPCHK(c, launch_save_restore(c->v, 0, lists ? c->lpar : -1, c->stream));
    }

    // ---- trajectories, in windows when the surrogate drives the MD (acceptance-based fallback) ----
    std::vector<TrajGraph> graphs;
    c->md_surr = hp->surrogate != 0;
    const int window = (c->md_surr && hp->fallback_window > 0) ? hp->fallback_window : ntraj;
    std::vector<HmcRecord> wrec;
    // speculative runs: the chain's records in column 0, the other columns zero
    if (spec) HIPCHK(c, hipMemsetAsync(c->d_rec, 0, sizeof(HmcRecord) * nrec, c->stream));
    for (int t0 = 0; t0 < ntraj && !rc; t0 += window) {
        const int count = std::min(window, ntraj - t0);
        if (list_md(c) && !c->lstate) {          // LJ and surrogate MD both run on the slot-ordered state
            c->list_valid = false;
            if ((rc = list_import(c))) break;
        }
        rc = spec ? run_spec(c, *hp, ufirst, rec_base, rec_base + t0, rec_base + t0 + count, graphs)
                  : run_segment(c, *hp, t0, count, ufirst, rec_base, rsteps, graphs);
        if (rc || !c->md_surr) continue;
        if ((rc = surr_check_flag(c))) break;
        if (t0 + count >= ntraj) break;
        wrec.resize((size_t)count * nch);
        hipError_t e = hipMemcpy(wrec.data(), c->d_rec + (size_t)t0 * nch, sizeof(HmcRecord) * wrec.size(),
                                 hipMemcpyDeviceToHost);
        if (e != hipSuccess) { rc = fail(c, CHMC_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e)); break; }
        int acc = 0;
        for (const HmcRecord& r : wrec) acc += r.acc;
        const double nwin = spec ? (double)count : (double)wrec.size();   // speculative: one chain, column 0
        if ((double)acc < hp->fallback_acc * nwin) c->md_surr = false;     // LJ forces from here on
    }
    for (TrajGraph& g : graphs) hipGraphExecDestroy(g.exec);
    c->md_surr = false;
    if (rc) return rc;
    if ((rc = list_export(c, true))) return rc;
    c->saved_lists = false;
    // the chain's momenta after the run are the post-MD momenta of its last trajectory, which copy
    // spec_ctl[2] ran (the positions are the same in every copy)
    if (spec) HIPCHK(c, launch_spec_final_momenta(c->v, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (list_md(c) && (rc = list_check_overflow(c))) return rc;
    c->traj_done += ntraj;
    c->bin_state = 2;
    if (records) HIPCHK(c, hipMemcpy(records, c->d_rec, sizeof(HmcRecord) * nrec, hipMemcpyDeviceToHost));
    if (nacc) {
        HIPCHK(c, hipMemcpy(st.data(), c->v.state, sizeof(ChainState) * nch, hipMemcpyDeviceToHost));
        for (int k = 0; k < nch; k++) nacc[k] = st[k].nacc;
    }
    return CHMC_OK;
}

// ---- measurements --------------------------------------------------------------------------------
extern "C" int chmc_rdf(chmc_ctx* c, int chain, int nbins, double size_bin, unsigned long long* counts, double* g)
{
    if (!c || chain < 0 || chain >= c->nchains || nbins < 1 || nbins > 4000 || !(size_bin > 0)) return CHMC_E_ARG;
    int rc = flush_drift(c);
    if (rc) return rc;
    if (c->rdf_cap < nbins) {
        if (c->d_rdf) hipFree(c->d_rdf);
        c->d_rdf = nullptr;
        HIPCHK(c, hipMalloc(&c->d_rdf, sizeof(unsigned long long) * nbins));
        c->rdf_cap = nbins;
    }
    const ChainDev& cd = c->hchains[chain];
    HIPCHK(c, hipMemsetAsync(c->d_rdf, 0, sizeof(unsigned long long) * nbins, c->stream));
    HIPCHK(c, launch_rdf(c->v, chain, cd.natoms, nbins, size_bin, c->d_rdf, c->stream));
    std::vector<unsigned long long> h(nbins);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->d_rdf, sizeof(unsigned long long) * nbins, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double V = cd.L[0] * cd.L[1] * cd.L[2];
    const int N = cd.natoms;
    for (int ib = 0; ib < nbins; ib++) {
        if (counts) counts[ib] = h[ib];
        if (g) {
            double ri = ib * size_bin;
            double rf = ri + size_bin;
            ri = ri * ri * ri;
            rf = rf * rf * rf;
            // identical_particles.cpp:768-774, including its 1.33333333333 constant
            g[ib] = (double)h[ib] * V / (1.33333333333 * M_PI * (rf - ri) * (N - 1) * N);
        }
    }
    return CHMC_OK;
}

extern "C" int chmc_potential_energy_open(const double* xyz, int n_particles, int n, double a, double b, double* result)
{
    if (!xyz || !result || n_particles < 0) return CHMC_E_ARG;
    *result = 0.0;
    if (n_particles < 2) return CHMC_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(nullptr, CHMC_E_NODEV, "no HIP device");
    const int np = n_particles;
    const int T = (np + 255) / 256;
    std::vector<double> soa(3 * (size_t)np);
    for (int i = 0; i < np; i++)
        for (int d = 0; d < 3; d++) soa[(size_t)d * np + i] = xyz[3 * i + d];
    double *dx = nullptr, *dp = nullptr;
    hipStream_t s = nullptr;
    auto cleanup = [&]() {
        if (dx) hipFree(dx);
        if (dp) hipFree(dp);
        if (s) hipStreamDestroy(s);
    };
    hipError_t e = hipStreamCreate(&s);
    if (e == hipSuccess) e = hipMalloc(&dx, sizeof(double) * soa.size());
    if (e == hipSuccess) e = hipMalloc(&dp, sizeof(double) * (size_t)T * T);
    if (e == hipSuccess) e = hipMemcpyAsync(dx, soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice, s);
    int nparts = 0;
    if (e == hipSuccess) e = launch_open_potential(dx, np, n, a, b, dp, &nparts, s);
    std::vector<double> parts((size_t)T * T);
    if (e == hipSuccess) e = hipMemcpyAsync(parts.data(), dp, sizeof(double) * parts.size(), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    cleanup();
    if (e != hipSuccess) return fail(nullptr, CHMC_E_HIP, std::string("potential_energy_open: ") + hipGetErrorString(e));
    double r = 0.0;
    for (double q : parts) r += q;
    *result = r;
    return CHMC_OK;
}


// ---- the same sum without allocation ---------------------------------------------------------------
// Block partials of chmc_potential_energy_open_noalloc live in this file-scope array, one copy per GPU,
// held until the process ends.  Host code never passes the array's name to a kernel (on the host it is
// only a stand-in); it uses the device address that hipGetSymbolAddress returns for the current device.
static constexpr int kOpenScratchAtoms = CHMC_OPEN_NOALLOC_MAX;
static constexpr int kOpenScratchTiles = (kOpenScratchAtoms + 255) / 256;
static __device__ __align__(256) double scratch[kOpenScratchTiles * kOpenScratchTiles];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on the current device, looked up once per device
static hipError_t open_scratch(double** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<double*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

extern "C" int chmc_potential_energy_open_noalloc(const double* xyz_dev, int n_particles, int n, double a, double b,
                                                  double* result_dev, void* stream)
{
    if (!result_dev || n_particles < 0 || (n_particles > 0 && !xyz_dev)) return fail(nullptr, CHMC_E_ARG, "potential_energy_open_noalloc: null pointer or negative size");
    if (n_particles > kOpenScratchAtoms)
        return fail(nullptr, CHMC_E_ARG, "potential_energy_open_noalloc: n_particles " + std::to_string(n_particles) +
                                             " exceeds the scratch capacity of " + std::to_string(kOpenScratchAtoms) + " atoms");
    double* part = nullptr;
    hipError_t e = open_scratch(&part);
    if (e == hipSuccess)
        e = launch_open_potential_aos(n_particles >= 2 ? xyz_dev : nullptr, n_particles >= 2 ? n_particles : 0, n, a, b,
                                      part, result_dev, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail(nullptr, CHMC_E_HIP, std::string("potential_energy_open_noalloc: ") + hipGetErrorString(e));
    return CHMC_OK;
}

// ---- learned surrogate ---------------------------------------------------------------------------
static int surr_ensure(chmc_ctx* c)
{
    if (c->surr_alloc) return CHMC_OK;
    const size_t N = (size_t)c->Ntot;
    HIPCHK(c, dalloc(c, &c->sv.params, kSurrNParams));
    HIPCHK(c, dalloc(c, &c->sv.w1b, kSurrHid * kSurrFeat));
    HIPCHK(c, dalloc(c, &c->sv.w2b, kSurrHid * kSurrHid));
    HIPCHK(c, dalloc(c, &c->sv.w1tb, kSurrHid * kSurrFeat));
    HIPCHK(c, dalloc(c, &c->sv.w2tb, kSurrHid * kSurrHid));
    HIPCHK(c, dalloc(c, &c->sv.adam_m, kSurrNTrain));
    HIPCHK(c, dalloc(c, &c->sv.adam_v, kSurrNTrain));
    HIPCHK(c, dalloc(c, &c->sv.nb_n, N));
    HIPCHK(c, dalloc(c, &c->sv.nb_j, N * kSurrMaxNb));
    HIPCHK(c, dalloc(c, &c->sv.nb_d, N * kSurrMaxNb));
    HIPCHK(c, dalloc(c, &c->sv.nb_f, N * kSurrMaxNb));
    HIPCHK(c, dalloc(c, &c->sv.flag, 1));
    HIPCHK(c, dalloc(c, &c->sG, N * kSurrFeat));
    HIPCHK(c, dalloc(c, &c->sg, N * kSurrFeat));
    HIPCHK(c, dalloc(c, &c->se, N));
    HIPCHK(c, dalloc(c, &c->sloss, 1 << 16));
    HIPCHK(c, hipMemset(c->sv.flag, 0, sizeof(int)));
    HIPCHK(c, hipMemset(c->sv.adam_m, 0, sizeof(float) * kSurrNTrain));
    HIPCHK(c, hipMemset(c->sv.adam_v, 0, sizeof(float) * kSurrNTrain));
    c->surr_alloc = true;
    return CHMC_OK;
}

// the descriptors count each pair once only if no chain's box holds two images within the cutoff
static int surr_geometry(chmc_ctx* c)
{
    for (const ChainDev& cd : c->hchains)
        for (int d = 0; d < 3; d++)
            if (cd.L[d] < 2.0 * cd.rc) return fail(c, CHMC_E_GEOMETRY, "surrogate: every box side must be >= 2 cutoff");
    return CHMC_OK;
}

static int surr_check_flag(chmc_ctx* c)
{
    int f = 0;
    HIPCHK(c, hipMemcpyAsync(&f, c->sv.flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (f) return fail(c, CHMC_E_STATE, "surrogate neighbour list overflow (more than 128 pairs per atom or 768 stencil entries)");
    return CHMC_OK;
}

// pair list + descriptors of the current positions (bins brought up to date first)
static int surr_descriptors(chmc_ctx* c, float* G)
{
    int rc = flush_drift(c);
    if (rc) return rc;
    if ((rc = ensure_bins(c))) return rc;
    HIPCHK(c, launch_surr_pairs(c->v, c->sv, c->stream));
    HIPCHK(c, launch_surr_desc(c->v, c->sv, G, c->stream));
    return CHMC_OK;
}

// update_momenta with the surrogate's gradient
static int do_surr_force_kick(chmc_ctx* c, double dt, bool storef)
{
    int rc = surr_descriptors(c, c->sG);
    if (rc) return rc;
    HIPCHK(c, launch_surr_mlp(c->v, c->sv, c->sG, c->sg, nullptr, c->stream));
    HIPCHK(c, launch_surr_force(c->v, c->sv, c->sg, dt, storef, c->stream));
    c->force_kind = storef ? 1 : 0;
    return CHMC_OK;
}

extern "C" int chmc_surrogate_set_params(chmc_ctx* c, const float* params, int n)
{
    if (!c || !params || n != kSurrNParams) return CHMC_E_ARG;
    int rc = surr_ensure(c);
    if (rc) return rc;
    HIPCHK(c, hipMemcpy(c->sv.params, params, sizeof(float) * kSurrNParams, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(c->sv.adam_m, 0, sizeof(float) * kSurrNTrain));
    HIPCHK(c, hipMemset(c->sv.adam_v, 0, sizeof(float) * kSurrNTrain));
    HIPCHK(c, launch_surr_refresh(c->sv, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->adam_step = 0;
    c->surr_params = true;
    return CHMC_OK;
}

extern "C" int chmc_surrogate_get_params(chmc_ctx* c, float* params, int n)
{
    if (!c || !params || n != kSurrNParams) return CHMC_E_ARG;
    if (!c->surr_params) return fail(c, CHMC_E_STATE, "surrogate parameters not set");
    HIPCHK(c, hipMemcpy(params, c->sv.params, sizeof(float) * kSurrNParams, hipMemcpyDeviceToHost));
    return CHMC_OK;
}

extern "C" int chmc_surrogate_eval(chmc_ctx* c, double* energy, int store_force)
{
    if (!c) return CHMC_E_ARG;
    if (!c->surr_params) return fail(c, CHMC_E_STATE, "surrogate parameters not set");
    int rc = surr_geometry(c);
    if (rc) return rc;
    if (list_md(c)) {
        // the kernels of surrogate-driven MD: descriptors and MLP over the neighbour lists, then the
        // surrogate sweep storing the force (no kick); the state returns to atom order afterwards
        if ((rc = list_import(c))) return rc;
        if (!c->list_valid && (rc = list_invalidate(c))) return rc;
        c->list_valid = true;
        HIPCHK(c, list_rebuild(c));
        HIPCHK(c, launch_lsurr_desc_mlp(c->v, c->sv, c->lpar, c->sg, c->se, c->sG, c->stream));
        if (store_force) HIPCHK(c, launch_lsurr_sweep(c->v, c->sg, LJ_FORCE | LJ_STOREF, c->lpar, 0.0, 0.0, true, c->stream));
        c->list_fresh = store_force != 0;
        if ((rc = list_export(c, false))) return rc;
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if ((rc = list_check_overflow(c))) return rc;
        c->surr_lists_eval = true;
    } else {
        rc = surr_descriptors(c, c->sG);
        if (rc) return rc;
        HIPCHK(c, launch_surr_mlp(c->v, c->sv, c->sG, c->sg, c->se, c->stream));
        HIPCHK(c, launch_surr_force(c->v, c->sv, c->sg, 0.0, store_force != 0, c->stream));
        if ((rc = surr_check_flag(c))) return rc;
        c->surr_lists_eval = false;
    }
    if (store_force) c->force_kind = 1;
    if (energy) {
        std::vector<float> e(c->Ntot);
        HIPCHK(c, hipMemcpy(e.data(), c->se, sizeof(float) * c->Ntot, hipMemcpyDeviceToHost));
        for (int k = 0; k < c->nchains; k++) {
            const ChainDev& cd = c->hchains[k];
            double sum = 0.0;
            for (int i = 0; i < cd.natoms; i++) sum += e[cd.atom_off + i];     // sorted slots of chain k
            energy[k] = sum;
        }
    }
    return CHMC_OK;
}

// Adam state of the trained parameters (first and second moments, step count): a checkpoint of a
// training run is these with the parameters
extern "C" int chmc_surrogate_get_adam(chmc_ctx* c, float* m, float* v, int* step)
{
    CHECK_CTX(c);
    if (!c->surr_params) return fail(c, CHMC_E_STATE, "surrogate parameters not set");
    if (m) HIPCHK(c, hipMemcpy(m, c->sv.adam_m, sizeof(float) * kSurrNTrain, hipMemcpyDeviceToHost));
    if (v) HIPCHK(c, hipMemcpy(v, c->sv.adam_v, sizeof(float) * kSurrNTrain, hipMemcpyDeviceToHost));
    if (step) *step = c->adam_step;
    return CHMC_OK;
}

extern "C" int chmc_surrogate_set_adam(chmc_ctx* c, const float* m, const float* v, int step)
{
    CHECK_CTX(c);
    if (!m || !v || step < 0) return fail(c, CHMC_E_ARG, "surrogate_set_adam: m, v required, step >= 0");
    if (!c->surr_params) return fail(c, CHMC_E_STATE, "surrogate parameters not set");
    HIPCHK(c, hipMemcpy(c->sv.adam_m, m, sizeof(float) * kSurrNTrain, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->sv.adam_v, v, sizeof(float) * kSurrNTrain, hipMemcpyHostToDevice));
    c->adam_step = step;
    return CHMC_OK;
}

// descriptors G and input gradients g = dE/dG of the last chmc_surrogate_eval, in atom order
extern "C" int chmc_surrogate_features(chmc_ctx* c, int chain, float* G, float* g)
{
    if (!c || chain < 0 || chain >= c->nchains) return CHMC_E_ARG;
    if (!c->surr_alloc) return fail(c, CHMC_E_STATE, "surrogate not evaluated");
    const ChainDev& cd = c->hchains[chain];
    std::vector<int> perm(cd.natoms);
    if (c->surr_lists_eval) {       // slot order of the chain's current list buffer
        int cur = 0;
        HIPCHK(c, hipMemcpy(&cur, c->v.lbuf + chain, sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(perm.data(), c->v.lperm + (size_t)cur * c->Ntot + cd.atom_off, sizeof(int) * cd.natoms,
                            hipMemcpyDeviceToHost));
    } else {                        // order of the current binning
        HIPCHK(c, hipMemcpy(perm.data(), c->v.perm_buf + (size_t)c->v.par * c->Ntot + cd.atom_off, sizeof(int) * cd.natoms,
                            hipMemcpyDeviceToHost));
    }
    std::vector<float> hG((size_t)cd.natoms * kSurrFeat), hg((size_t)cd.natoms * kSurrFeat);
    HIPCHK(c, hipMemcpy(hG.data(), c->sG + (size_t)cd.atom_off * kSurrFeat, sizeof(float) * hG.size(), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(hg.data(), c->sg + (size_t)cd.atom_off * kSurrFeat, sizeof(float) * hg.size(), hipMemcpyDeviceToHost));
    for (int k = 0; k < cd.natoms; k++) {
        const int a = perm[k] - cd.atom_off;
        for (int f = 0; f < kSurrFeat; f++) {
            if (G) G[(size_t)a * kSurrFeat + f] = hG[(size_t)k * kSurrFeat + f];
            if (g) g[(size_t)a * kSurrFeat + f] = hg[(size_t)k * kSurrFeat + f];
        }
    }
    return CHMC_OK;
}

extern "C" int chmc_surrogate_reserve(chmc_ctx* c, int nconfigs)
{
    if (!c || nconfigs < 1) return CHMC_E_ARG;
    int rc = surr_ensure(c);
    if (rc) return rc;
    const long long cap = (long long)nconfigs * c->Ntot;
    if (c->Gtrain) hipFree(c->Gtrain);
    if (c->utrain) hipFree(c->utrain);
    c->Gtrain = nullptr;
    c->utrain = nullptr;
    HIPCHK(c, hipMalloc(&c->Gtrain, sizeof(float) * (size_t)cap * kSurrFeat));
    HIPCHK(c, hipMalloc(&c->utrain, sizeof(float) * (size_t)cap));
    c->train_cap = cap;
    c->train_rows = 0;
    return CHMC_OK;
}

// appends the current configuration of every chain (descriptors and per-atom LJ energies) to the
// device-resident training set
extern "C" int chmc_surrogate_collect(chmc_ctx* c, long long* rows)
{
    if (!c) return CHMC_E_ARG;
    if (!c->Gtrain) return fail(c, CHMC_E_STATE, "call chmc_surrogate_reserve first");
    if (c->train_rows + c->Ntot > c->train_cap) return fail(c, CHMC_E_ARG, "training set full");
    int rc = surr_geometry(c);
    if (rc) return rc;
    rc = surr_descriptors(c, c->Gtrain + (size_t)c->train_rows * kSurrFeat);
    if (rc) return rc;
    HIPCHK(c, launch_surr_targets(c->v, c->sv, c->utrain + c->train_rows, c->stream));
    if ((rc = surr_check_flag(c))) return rc;
    c->train_rows += c->Ntot;
    if (rows) *rows = c->train_rows;
    return CHMC_OK;
}

extern "C" int chmc_surrogate_training_data(chmc_ctx* c, long long row0, int nrows, float* G, float* u)
{
    if (!c || row0 < 0 || nrows < 0 || row0 + nrows > c->train_rows) return CHMC_E_ARG;
    if (G) HIPCHK(c, hipMemcpy(G, c->Gtrain + (size_t)row0 * kSurrFeat, sizeof(float) * (size_t)nrows * kSurrFeat, hipMemcpyDeviceToHost));
    if (u) HIPCHK(c, hipMemcpy(u, c->utrain + row0, sizeof(float) * (size_t)nrows, hipMemcpyDeviceToHost));
    return CHMC_OK;
}

static int surr_partial_room(chmc_ctx* c, int nrows)
{
    const long long need = (long long)((nrows + 31) / 32) * kSurrNGrad;
    if (need > c->spartial_cap) {
        if (c->spartial) hipFree(c->spartial);
        c->spartial = nullptr;
        HIPCHK(c, hipMalloc(&c->spartial, sizeof(float) * (size_t)need));
        c->spartial_cap = need;
    }
    return CHMC_OK;
}

// summed parameter gradient (and, in the last slot, the sum of squared residuals) of one batch of
// training rows, without an update: the gradient check against the oracle
extern "C" int chmc_surrogate_loss_grad(chmc_ctx* c, long long row0, int nrows, float* grad)
{
    if (!c || !grad || nrows < 1 || row0 < 0 || row0 + nrows > c->train_rows) return CHMC_E_ARG;
    if (!c->surr_params) return fail(c, CHMC_E_STATE, "surrogate parameters not set");
    int rc = surr_partial_room(c, nrows);
    if (rc) return rc;
    HIPCHK(c, launch_surr_grad_only(c->sv, c->Gtrain, c->utrain, row0, nrows, c->spartial, c->sloss, c->stream));
    HIPCHK(c, hipMemcpy(grad, c->sloss, sizeof(float) * kSurrNGrad, hipMemcpyDeviceToHost));
    return CHMC_OK;
}

// nsteps Adam steps on batches of batch_rows consecutive training rows starting at random offsets;
// loss_history[t] (may be NULL) receives the batch's mean squared residual before step t's update
extern "C" int chmc_surrogate_train(chmc_ctx* c, int nsteps, int batch_rows, float lr, uint64_t seed, float* loss_history)
{
    if (!c || nsteps < 1 || batch_rows < 1 || nsteps > (1 << 16)) return CHMC_E_ARG;
    if (!c->surr_params) return fail(c, CHMC_E_STATE, "surrogate parameters not set");
    if (c->train_rows < batch_rows) return fail(c, CHMC_E_STATE, "not enough training rows");
    int rc = surr_partial_room(c, batch_rows);
    if (rc) return rc;
    Mt64 rng(seed);
    const long long span = c->train_rows - batch_rows + 1;
    for (int t = 0; t < nsteps; t++) {
        const long long row0 = (long long)(rng.next() % (uint64_t)span);
        c->adam_step++;
        HIPCHK(c, launch_surr_train_step(c->sv, c->Gtrain, c->utrain, row0, batch_rows, c->spartial, lr, c->adam_step,
                                         c->sloss + t, c->stream));
    }
    if (loss_history) {
        std::vector<float> h(nsteps);
        HIPCHK(c, hipMemcpy(h.data(), c->sloss, sizeof(float) * nsteps, hipMemcpyDeviceToHost));
        for (int t = 0; t < nsteps; t++) loss_history[t] = h[t] / (float)batch_rows;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CHMC_OK;
}

// ---- force matching on whole configurations -----------------------------------------------------
static int surr_fm_ensure(chmc_ctx* c)
{
    const size_t N = (size_t)c->Ntot;
    if (!c->fmR) {
        HIPCHK(c, dalloc(c, &c->fmR, 4 * N));
        HIPCHK(c, dalloc(c, &c->fmu, N));
        HIPCHK(c, dalloc(c, &c->fmv, N * kSurrFeat));
        HIPCHK(c, dalloc(c, &c->fmgrad, kSurrNTrain + 2));
    }
    const long long need = (long long)surr_fm_nparts(c->Ntot) * (kSurrNTrain + 2);
    if (need > c->spartial_cap) {
        if (c->spartial) hipFree(c->spartial);
        c->spartial = nullptr;
        HIPCHK(c, hipMalloc(&c->spartial, sizeof(float) * (size_t)need));
        c->spartial_cap = need;
    }
    return CHMC_OK;
}

// descriptors, g = dE/dG and the force-matching pass at the current positions; grad_out != NULL: the
// summed gradient and loss sums without an update, else one Adam step with the loss sums in loss_out
static int surr_fm_pass(chmc_ctx* c, double w_force, float* grad_out, float lr, float* loss_out)
{
    int rc = surr_descriptors(c, c->sG);
    if (rc) return rc;
    HIPCHK(c, launch_surr_mlp(c->v, c->sv, c->sG, c->sg, nullptr, c->stream));
    if (!grad_out) c->adam_step++;
    HIPCHK(c, launch_surr_fm(c->v, c->sv, c->sG, c->sg, reinterpret_cast<float4*>(c->fmR), c->fmu, c->fmv,
                             (float)w_force, c->spartial, grad_out, lr, c->adam_step, loss_out, c->stream));
    return CHMC_OK;
}

extern "C" int chmc_surrogate_fm_loss_grad(chmc_ctx* c, double w_force, float* grad, double* loss_e, double* loss_f)
{
    CHECK_CTX(c);
    if (!grad || w_force < 0.0) return fail(c, CHMC_E_ARG, "surrogate_fm_loss_grad: grad required, w_force >= 0");
    if (!c->surr_params) return fail(c, CHMC_E_STATE, "surrogate parameters not set");
    int rc = surr_geometry(c);
    if (!rc) rc = surr_fm_ensure(c);
    if (!rc) rc = surr_fm_pass(c, w_force, c->fmgrad, 0.f, nullptr);
    if (rc) return rc;
    if ((rc = surr_check_flag(c))) return rc;
    std::vector<float> h(kSurrNTrain + 2);
    HIPCHK(c, hipMemcpy(h.data(), c->fmgrad, sizeof(float) * h.size(), hipMemcpyDeviceToHost));
    std::copy(h.begin(), h.begin() + kSurrNTrain, grad);
    if (loss_e) *loss_e = (double)h[kSurrNTrain] / c->Ntot;
    if (loss_f) *loss_f = (double)h[kSurrNTrain + 1] / (3.0 * c->Ntot);
    return CHMC_OK;
}

extern "C" int chmc_surrogate_reserve_configs(chmc_ctx* c, int nconfigs)
{
    CHECK_CTX(c);
    if (nconfigs < 1) return fail(c, CHMC_E_ARG, "surrogate_reserve_configs: nconfigs >= 1");
    int rc = surr_ensure(c);
    if (rc) return rc;
    if (c->xtrain) hipFree(c->xtrain);
    c->xtrain = nullptr;
    c->xtrain_cap = c->xtrain_n = 0;
    HIPCHK(c, hipMalloc(&c->xtrain, sizeof(double) * 3 * (size_t)c->Ntot * (size_t)nconfigs));
    c->xtrain_cap = nconfigs;
    return CHMC_OK;
}

// appends the current positions of every chain (one snapshot of the context) to the configurations
extern "C" int chmc_surrogate_collect_config(chmc_ctx* c, int* nconfigs)
{
    CHECK_CTX(c);
    if (!c->xtrain) return fail(c, CHMC_E_STATE, "call chmc_surrogate_reserve_configs first");
    if (c->xtrain_n >= c->xtrain_cap) return fail(c, CHMC_E_ARG, "configuration store full");
    int rc = flush_drift(c);
    if (rc) return rc;
    const size_t n3 = 3 * (size_t)c->Ntot;
    HIPCHK(c, hipMemcpyAsync(c->xtrain + (size_t)c->xtrain_n * n3, c->v.x, sizeof(double) * n3, hipMemcpyDeviceToDevice,
                             c->stream));
    c->xtrain_n++;
    if (nconfigs) *nconfigs = c->xtrain_n;
    return CHMC_OK;
}

// nsteps Adam steps, each on one stored snapshot drawn from mt19937_64(seed): loss = mean (e - u)^2 +
// w_force mean (dE/dx - dU/dx)^2 over its atoms.  The context's positions are left at the last snapshot.
extern "C" int chmc_surrogate_train_fm(chmc_ctx* c, int nsteps, double w_force, float lr, uint64_t seed,
                                       float* loss_e, float* loss_f)
{
    CHECK_CTX(c);
    if (nsteps < 1 || nsteps > (1 << 15) || w_force < 0.0) return fail(c, CHMC_E_ARG, "surrogate_train_fm: 1 <= nsteps <= 32768, w_force >= 0");
    if (!c->surr_params) return fail(c, CHMC_E_STATE, "surrogate parameters not set");
    if (c->xtrain_n < 1) return fail(c, CHMC_E_STATE, "no stored configurations");
    int rc = surr_geometry(c);
    if (!rc) rc = surr_fm_ensure(c);
    if (!rc) rc = flush_drift(c);
    if (rc) return rc;
    Mt64 rng(seed);
    const size_t n3 = 3 * (size_t)c->Ntot;
    for (int t = 0; t < nsteps; t++) {
        const int k = (int)(rng.next() % (uint64_t)c->xtrain_n);
        HIPCHK(c, hipMemcpyAsync(c->v.x, c->xtrain + (size_t)k * n3, sizeof(double) * n3, hipMemcpyDeviceToDevice, c->stream));
        c->list_valid = false;
        c->bin_state = 2;
        c->have_bins = false;
        c->force_kind = 0;
        if ((rc = surr_fm_pass(c, w_force, nullptr, lr, c->sloss + 2 * t))) return rc;
    }
    if ((rc = surr_check_flag(c))) return rc;
    std::vector<float> h(2 * (size_t)nsteps);
    HIPCHK(c, hipMemcpy(h.data(), c->sloss, sizeof(float) * h.size(), hipMemcpyDeviceToHost));
    for (int t = 0; t < nsteps; t++) {
        if (loss_e) loss_e[t] = h[2 * t] / (float)c->Ntot;
        if (loss_f) loss_f[t] = h[2 * t + 1] / (3.0f * (float)c->Ntot);
    }
    return CHMC_OK;
}

// ---- tuning ---------------------------------------------------------------------------------------
static int time_force_is(chmc_ctx* c, int islots, int reps, double* ms_bin, double* ms_force)
{
    int rc = do_bin(c);
    if (rc) return rc;
    const bool local = !id_order(c);
    auto rebin = [&]() -> hipError_t {
        hipError_t e = launch_count(c->v, 0.0, c->stream);
        if (e == hipSuccess) e = launch_scan(c->v, c->stream);
        if (e == hipSuccess) e = launch_place(c->v, local, id_order(c), c->stream);
        c->v.par ^= 1;
        return e;
    };
    // warm-up
    HIPCHK(c, rebin());
    HIPCHK(c, launch_lj(c->v, LJ_FORCE | LJ_STOREF, false, islots, 0.0, 0.0, c->stream));
    float tb = 0.f, tf = 0.f;
    for (int r = 0; r < reps; r++) {
        HIPCHK(c, hipEventRecord(c->ev0, c->stream));
        HIPCHK(c, rebin());
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
        HIPCHK(c, launch_lj(c->v, LJ_FORCE | LJ_STOREF, false, islots, 0.0, 0.0, c->stream));
        HIPCHK(c, hipEventRecord(c->ev2, c->stream));
        HIPCHK(c, hipEventSynchronize(c->ev2));
        float a = 0.f, b = 0.f;
        HIPCHK(c, hipEventElapsedTime(&a, c->ev0, c->ev1));
        HIPCHK(c, hipEventElapsedTime(&b, c->ev1, c->ev2));
        tb += a;
        tf += b;
    }
    c->bin_state = 0;
    c->force_kind = 1;
    if (ms_bin) *ms_bin = tb / reps;
    if (ms_force) *ms_force = tf / reps;
    return CHMC_OK;
}

static const int kTilings[8] = { 41, 42, 44, 81, 82, 84, 161, 162 };

extern "C" int chmc_set_force_tiling(chmc_ctx* c, int islots)
{
    CHECK_CTX(c);
    for (int t : kTilings)
        if (islots == t) {
            c->islots = islots;
            return CHMC_OK;
        }
    if (islots != 0) return fail(c, CHMC_E_ARG, "tiling must be 0 (autotune) or i-slots*10+waves from {41,42,44,81,82,84,161,162}");
    // autotuner: time each tiling of the FP32 sweep on the current configuration, keep the fastest
    const int (&cand)[8] = kTilings;
    double best = 1e30;
    int best_is = c->islots;
    for (int is : cand) {
        double tf = 0.0;
        int rc = time_force_is(c, is, 5, nullptr, &tf);
        if (rc) return rc;
        if (tf < best) { best = tf; best_is = is; }
    }
    c->islots = best_is;
    return CHMC_OK;
}

extern "C" int chmc_get_force_tiling(const chmc_ctx* c, int* islots)
{
    CHECK_CTX(c);
    if (islots) *islots = c->islots;
    return CHMC_OK;
}

extern "C" int chmc_time_force(chmc_ctx* c, int reps, double* ms_bin, double* ms_force)
{
    CHECK_CTX(c);
    if (reps < 1) return CHMC_E_ARG;
    return time_force_is(c, c->islots, reps, ms_bin, ms_force);
}

extern "C" int chmc_set_skin(chmc_ctx* c, double skin)
{
    CHECK_CTX(c);
    if (!(skin > 0.0) || skin > 2.0) return fail(c, CHMC_E_ARG, "set_skin: 0 < skin <= 2 required");
    int rc = flush_drift(c);
    if (rc) return rc;
    return list_setup(c, skin, {});
}

extern "C" int chmc_get_skin(const chmc_ctx* c, double* skin)
{
    if (!c || !skin) return CHMC_E_ARG;
    *skin = c->skin;
    return CHMC_OK;
}

extern "C" int chmc_set_chain_skins(chmc_ctx* c, const double* skins)
{
    CHECK_CTX(c);
    if (!skins) return CHMC_E_ARG;
    std::vector<double> sk(skins, skins + c->nchains);
    for (int k = 0; k < c->nchains; k++)
        if (!(sk[k] > 0.0) || sk[k] > 2.0)
            return fail(c, CHMC_E_ARG, "set_chain_skins: 0 < skin <= 2 required (chain " + std::to_string(k) + ")");
    int rc = flush_drift(c);
    if (rc) return rc;
    return list_setup(c, c->skin, sk);
}

extern "C" int chmc_get_chain_skin(const chmc_ctx* c, int chain, double* skin)
{
    if (!c || !skin || chain < 0 || chain >= c->nchains) return CHMC_E_ARG;
    *skin = c->chain_skin.empty() ? c->skin : c->chain_skin[chain];
    return CHMC_OK;
}

extern "C" int chmc_list_stats(chmc_ctx* c, int chain, long long* entries, long long* nint, long long* rebuilds)
{
    if (!c || chain < 0 || chain >= c->nchains) return CHMC_E_ARG;
    int rc = flush_drift(c);
    if (rc) return rc;
    if ((rc = list_import(c)) || (rc = list_invalidate(c)) || (rc = list_energy(c, true)) || (rc = list_export(c, false))) return rc;
    std::vector<double> nintv(c->nchains);
    if ((rc = read_red(c, RED_NINT, nintv.data()))) return rc;
    const ChainDev& cd = c->hchains[chain];
    std::vector<int> n(cd.natoms);
    int cur = 0;
    HIPCHK(c, hipMemcpy(&cur, c->v.lbuf + chain, sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(n.data(), c->v.nbn + (size_t)cur * c->Ntot + cd.atom_off, sizeof(int) * cd.natoms,
                        hipMemcpyDeviceToHost));
    int f[4];
    HIPCHK(c, hipMemcpy(f, c->v.lflag, sizeof(f), hipMemcpyDeviceToHost));
    long long tot = 0;
    for (int k : n) tot += k;
    if (entries) *entries = tot;
    if (nint) *nint = (long long)std::llround(nintv[chain]);
    if (rebuilds) *rebuilds = f[3];
    return f[2] ? list_check_overflow(c) : CHMC_OK;
}

extern "C" int chmc_time_list(chmc_ctx* c, int reps, double* ms_sweep, double* ms_rebuild);

extern "C" int chmc_set_list_lanes(chmc_ctx* c, int lanes)
{
    CHECK_CTX(c);
    if (lanes != 0 && lanes != 4 && lanes != 8 && lanes != 16 && lanes != 810 && lanes != 811 && lanes != 812 &&
        lanes != 816)
        return fail(c, CHMC_E_ARG, "set_list_lanes: 0 (autotune), 4, 8, 16, 810, 811, 812 or 816");
    int rc = list_ensure(c);
    if (rc) return rc;
    if (lanes) return set_list_shape(c, lanes);
    double best = 1e30;
    int pick = 8;
    for (int lp : { 4, 8, 16, 810, 811, 812, 816 }) {
        if ((rc = set_list_shape(c, lp))) return rc;
        double ms = 0.0;
        if ((rc = chmc_time_list(c, 20, &ms, nullptr))) return rc;
        if (ms < best) { best = ms; pick = lp; }
    }
    return set_list_shape(c, pick);
}

extern "C" int chmc_set_list_start_rebuild(chmc_ctx* c, int on)
{
    CHECK_CTX(c);
    c->v.start_rebuild = on ? 1 : 0;
    return CHMC_OK;
}

extern "C" int chmc_get_list_start_rebuild(const chmc_ctx* c, int* on)
{
    if (!c || !on) return CHMC_E_ARG;
    *on = c->v.start_rebuild;
    return CHMC_OK;
}

extern "C" int chmc_set_speculation(chmc_ctx* c, int copies)
{
    CHECK_CTX(c);
    if (copies <= 1) {
        c->spec = 0;
        return CHMC_OK;
    }
    if (copies != c->nchains) return fail(c, CHMC_E_ARG, "set_speculation: the context must hold exactly `copies` chains");
    if (!list_md(c))
        return fail(c, CHMC_E_STATE, "set_speculation: speculative trajectories run on the FP32 neighbour-list MD only");
    const chmc_chain_desc& d0 = c->desc[0];
    for (int k = 1; k < c->nchains; k++) {
        const chmc_chain_desc& d = c->desc[k];
        if (d.N != d0.N || d.L[0] != d0.L[0] || d.L[1] != d0.L[1] || d.L[2] != d0.L[2] || d.cutoff != d0.cutoff ||
            d.sigma != d0.sigma || d.eps != d0.eps || d.mass != d0.mass || d.beta != d0.beta)
            return fail(c, CHMC_E_ARG, "set_speculation: the chains must be identical copies of one chain");
    }
    if (!c->v.spec_ctl) {
        HIPCHK(c, dalloc(c, &c->v.spec_ctl, 4));
        HIPCHK(c, dalloc(c, &c->v.spec_st, 1));
        HIPCHK(c, hipMemset(c->v.spec_ctl, 0, 4 * sizeof(int)));
        HIPCHK(c, hipMemset(c->v.spec_st, 0, sizeof(ChainState)));
    }
    c->spec = copies;
    return CHMC_OK;
}

extern "C" int chmc_get_speculation(const chmc_ctx* c, int* copies)
{
    if (!c || !copies) return CHMC_E_ARG;
    *copies = c->spec > 1 ? c->spec : 1;
    return CHMC_OK;
}

extern "C" int chmc_set_list_check(chmc_ctx* c, int every)
{
    CHECK_CTX(c);
    if (every < 1 || every > 4) return fail(c, CHMC_E_ARG, "set_list_check: 1 to 4 force evaluations");
    c->list_check = every;
    return CHMC_OK;
}

extern "C" int chmc_get_list_check(const chmc_ctx* c, int* every)
{
    if (!c || !every) return CHMC_E_ARG;
    *every = c->list_check;
    return CHMC_OK;
}

extern "C" int chmc_set_list_entry_bits(chmc_ctx* c, int bits)
{
    CHECK_CTX(c);
    if (bits != 0 && bits != 16 && bits != 32) return fail(c, CHMC_E_ARG, "set_list_entry_bits: 0, 16 or 32");
    c->entry_bits = bits;
    if (!c->list_ready) return CHMC_OK;
    int rc = flush_drift(c);
    if (rc) return rc;
    return list_setup(c, c->skin, c->chain_skin);
}

extern "C" int chmc_get_list_entry_bits(const chmc_ctx* c, int* bits)
{
    if (!c || !bits) return CHMC_E_ARG;
    *bits = c->list_ready ? (c->v.entry16 ? 16 : 32) : c->entry_bits;
    return CHMC_OK;
}

extern "C" int chmc_get_list_lanes(const chmc_ctx* c, int* lanes)
{
    if (!c || !lanes) return CHMC_E_ARG;
    *lanes = c->v.lanes ? c->v.lanes : 8;
    return CHMC_OK;
}

// the surrogate's force evaluation on the lists: `reps` descriptor + MLP launches, then `reps` force
// sweeps with kick and drift at dt = 0 (positions and momenta unchanged, so the lists stay valid)
extern "C" int chmc_time_surrogate(chmc_ctx* c, int reps, double* ms_desc_mlp, double* ms_sweep)
{
    CHECK_CTX(c);
    if (reps < 1) return CHMC_E_ARG;
    if (fp64_md(c)) return fail(c, CHMC_E_STATE, "time_surrogate: surrogate MD runs on the FP32 list path");
    if (!c->surr_params) return fail(c, CHMC_E_STATE, "time_surrogate: surrogate parameters not set");
    int rc = surr_geometry(c);
    if (rc || (rc = list_ensure(c)) || (rc = flush_drift(c)) || (rc = list_import(c)) || (rc = list_invalidate(c))) return rc;
    HIPCHK(c, list_rebuild(c));
    HIPCHK(c, launch_lsurr_desc_mlp(c->v, c->sv, c->lpar, c->sg, nullptr, nullptr, c->stream));     // warm-up
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    for (int r = 0; r < reps; r++) HIPCHK(c, launch_lsurr_desc_mlp(c->v, c->sv, c->lpar, c->sg, nullptr, nullptr, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    for (int r = 0; r < reps; r++) {
        HIPCHK(c, launch_lsurr_sweep(c->v, c->sg, LJ_FORCE | LJ_KICK | LJ_DRIFT, c->lpar, 0.0, 0.0, true, c->stream));
        c->lpar ^= 1;
    }
    HIPCHK(c, hipEventRecord(c->ev2, c->stream));
    if ((rc = list_export(c, false))) return rc;
    HIPCHK(c, hipEventSynchronize(c->ev2));
    float a = 0.f, b = 0.f;
    HIPCHK(c, hipEventElapsedTime(&a, c->ev0, c->ev1));
    HIPCHK(c, hipEventElapsedTime(&b, c->ev1, c->ev2));
    c->force_kind = 0;
    if (ms_desc_mlp) *ms_desc_mlp = a / reps;
    if (ms_sweep) *ms_sweep = b / reps;
    if ((rc = surr_check_flag(c))) return rc;
    return list_check_overflow(c);
}

extern "C" int chmc_time_list(chmc_ctx* c, int reps, double* ms_sweep, double* ms_rebuild)
{
    CHECK_CTX(c);
    if (reps < 1) return CHMC_E_ARG;
    if (fp64_md(c)) return fail(c, CHMC_E_STATE, "time_list: the list sweep is the FP32 path");
    int rc = flush_drift(c);
    if (rc) return rc;
    if ((rc = list_import(c))) return rc;
    // rebuilds: `reps` forced rebuilds back to back (request flags set by a fill kernel inside the span)
    if ((rc = list_invalidate(c))) return rc;
    HIPCHK(c, list_rebuild(c));            // warm-up
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    for (int r = 0; r < reps; r++) {
        if ((rc = list_invalidate(c))) return rc;
        HIPCHK(c, list_rebuild(c));
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    // the trajectory's MD step alone: force, kick and drift with dt = 0 (positions and momenta
    // unchanged, so the lists stay valid), `reps` sweeps back to back
    for (int r = 0; r < reps; r++) {
        HIPCHK(c, launch_list_sweep(c->v, LJ_FORCE | LJ_KICK | LJ_DRIFT, c->lpar, 0.0, 0.0, true, c->stream));
        c->lpar ^= 1;
    }
    HIPCHK(c, hipEventRecord(c->ev2, c->stream));
    if ((rc = list_export(c, false))) return rc;
    HIPCHK(c, hipEventSynchronize(c->ev2));
    float a = 0.f, b = 0.f;
    HIPCHK(c, hipEventElapsedTime(&a, c->ev0, c->ev1));
    HIPCHK(c, hipEventElapsedTime(&b, c->ev1, c->ev2));
    c->force_kind = 0;
    if (ms_sweep) *ms_sweep = b / reps;
    if (ms_rebuild) *ms_rebuild = a / reps;
    return list_check_overflow(c);
}
