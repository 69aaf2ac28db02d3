// Neighbour-list MD for the FP32 force path: the same truncated LJ forces as the cell sweep of
// k_lj.hip (every pair with r^2 < rc^2 under the 27-cell image rules of identical_particles.cpp:
// 584-600), but the candidate set is a per-atom list of the pairs within rc + skin, rebuilt only when
// an atom has moved more than skin / 2 since the last rebuild (so no pair can have entered rc
// unlisted).  The reference re-bins on every force evaluation (identical_particles.cpp:379,560); here
// a force evaluation between rebuilds is one sweep over ~(4/3) pi (rc + skin)^3 rho candidates per atom
// instead of 27 cells' worth.
//
// Rebuild (kernels that skip every chain whose request flag of the current list parity is clear, so
// that they can sit in a captured trajectory and cost a launch when idle; chains rebuild independently,
// so a chain of a batch evolves exactly as it does alone):
//   k_lcount  atoms -> list cells (floor(x / lsize), clamped), arrival order into a dense member array
//             that carries each member's FP32 cell-relative position
//   k_lscan   only for chains of more than kScanCells cells: one workgroup scans the cell counts and
//             the others rank each list cell's members by atom id once, for every build workgroup
//             that stages them (k_lscan_chains: the same for batches of chains of kScanChainCells cells
//             or more, one scan workgroup per chain)
//   k_lbuild  one workgroup per list cell: (small chains: the counts scanned in LDS) the cell's members
//             ranked by atom id give the compact slots and their slot data (x0, r0, FP32 xsl); the 27
//             neighbour cells' members, ranked the same way, are staged in LDS in slot order with the
//             image shift dd * lsize folded in (the way k_lj stages its j-list), and each own atom's
//             list of the entries within rc + skin is appended in slot order, as 32-bit entries (slot,
//             image code) or, in contexts of large lists, 16-bit ones (neighbour segment, rank in that
//             cell; the cell's segment starts go to ctab).  Rows are then padded past their count to
//             the sweep span with no-op entries (image code kFarCode, or segment kPadSeg)
// Sweep: 32 slots x LP lanes per workgroup (LP = 4, 8 or 16, tuned); each lane reads its first kBatch
//   entries without a count test (the padding), gathers the neighbours' FP32 coordinates (16-bit
//   entries: slot = segment start from the slot's LDS segment row + rank), adds the image shift from an
//   LDS table by code or segment (padding: 1e18 away), walks the rest of the list,
//   and the partial sums meet in a butterfly.  Three lanes of an atom then kick p in FP64, optionally
//   drift x in FP64 (identical_particles.cpp:651-673), and lane 0 writes the FP32 coordinates of the
//   other parity and raises the rebuild request when the atom is beyond skin / 2 from x0.
// Energy: FP64 coordinates relative to the rebuild cell origins (rel64 = r0 + wrap(x - x0)), the
//   potential's r < rc test (identical_particles.cpp:432-435), entries in batches of 4 per lane,
//   per-tile fixed-order partials.
#include "chmc_internal.hpp"
#include "surrogate.hpp"

#pragma clang fp contract(off)

namespace chmc {

__device__ __forceinline__ int clamp_l(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// list cell of a position (global list-cell index), and its grid coordinates
__device__ __forceinline__ int list_cell(const ChainDev& C, const double x3[3], int b[3])
{
#pragma unroll
    for (int d = 0; d < 3; d++) b[d] = clamp_l((int)floor(x3[d] / C.lsize[d]), 0, C.lnbin[d] - 1);
    return C.lcell_off + b[0] + C.lnbin[0] * (b[1] + b[2] * C.lnbin[1]);
}

// ---- rebuild -------------------------------------------------------------------------------------
// request flags of parity q: [0] any chain, [1 + k] chain k
__device__ __forceinline__ int* req(const DevView& v, int q, int chain) { return v.lreq + q * (v.nchains + 1) + 1 + chain; }
__device__ __forceinline__ int* req_any(const DevView& v, int q) { return v.lreq + q * (v.nchains + 1); }
__device__ __forceinline__ void request(const DevView& v, int q, int chain)
{
    atomicOr(req(v, q, chain), 1);
    atomicOr(req_any(v, q), 1);
}

// Counts: two buffers lcount[cb][Cl]; the host alternates cb with every rebuild launch.  A launch
// counts into buffer cb and clears buffer cb ^ 1 (the previous launch's counts), idle or not, so buffer
// cb is zero whenever a launch starts and no later kernel has to clear the counts the build reads.
// Each member is stored with its FP32 position relative to its cell origin (x - b * lsize rounded once,
// the slot coordinates of the rebuild) and its atom id, and its FP64 position and momentum.  The count
// runs over the slots of the current buffer (the slot-ordered state xl, pl).
__global__ __launch_bounds__(256) void k_lcount(DevView v, int q, int cb)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    int* __restrict__ prev = v.lcount + (size_t)(cb ^ 1) * v.Cl;
    for (int c = s; c < v.Cl; c += gridDim.x * blockDim.x) prev[c] = 0;
    if (*req_any(v, q) == 0) return;
    // lanes of slots in requesting chains take part; whole waves stay together for the ballots below
    const int ch = s < v.Ntot ? v.atom_chain[s] : 0;     // a chain's slots span the same range as its atoms
    const bool part = s < v.Ntot && *req(v, q, ch) != 0;
    if (__ballot(part) == 0) return;
    const ChainDev& C = v.chains[ch];
    const int N = v.Ntot;
    int a = 0, c = -1;
    double x3[3] = { 0.0, 0.0, 0.0 };
    float f3[3] = { 0.f, 0.f, 0.f };
    if (part) {
        a = list_buf(v, *list_cur(v, ch)).lperm[s];
#pragma unroll
        for (int d = 0; d < 3; d++) x3[d] = v.xl[d * N + s];
        int b[3];
        c = list_cell(C, x3, b);
#pragma unroll
        for (int d = 0; d < 3; d++) f3[d] = (float)(x3[d] - (double)b[d] * C.lsize[d]);
    }
    // one atomic per distinct cell of the wave: consecutive slots (the previous rebuild's cell order)
    // mostly fall into a few new cells.  Lanes of a group take consecutive member positions.
    unsigned long long left = __ballot(part);
    int lead = 0, rank = 0, gsize = 0;
    while (left) {
        const int l = __ffsll((long long)left) - 1;
        const int cl = __shfl(c, l, 64);
        const unsigned long long m = __ballot(part && c == cl) & left;
        if (part && c == cl) {
            lead = l;
            rank = __popcll(m & ((1ull << lane) - 1ull));
        }
        if (lane == l) gsize = __popcll(m);
        left &= ~m;
    }
    int base = 0;
    if (part && lane == lead) base = atomicAdd(&v.lcount[(size_t)cb * v.Cl + c], gsize);
    base = __shfl(base, lead, 64);
    if (!part) return;
    const int pos = base + rank;
    if (pos < C.lcap) {
        const size_t m = C.ldense_off + (size_t)(c - C.lcell_off) * C.lcap + pos;
        v.lmem[m] = make_float4(f3[0], f3[1], f3[2], __int_as_float(a));
#pragma unroll
        for (int d = 0; d < 3; d++) {
            v.lmemx[(size_t)d * v.ldense + m] = x3[d];
            v.lmemp[(size_t)d * v.ldense + m] = v.pl[d * N + s];
        }
    } else {
        atomicOr(&v.lflag[2], 1);
    }
}

// the members of one list cell of a rebuilding chain ranked by atom id (one wave; ids staged in LDS)
__device__ __forceinline__ void rank_cell_members(const DevView& v, int q, int cb, int cell, int lane, int* ids)
{
    const int k = v.lcell_chain[cell];
    if (*req(v, q, k) == 0) return;
    const ChainDev& C = v.chains[k];
    const int nk = min(v.lcount[(size_t)cb * v.Cl + cell], C.lcap);
    const size_t base = C.ldense_off + (size_t)(cell - C.lcell_off) * C.lcap;
    for (int i = lane; i < nk; i += 64) ids[i] = __float_as_int(v.lmem[base + i].w);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int i = lane; i < nk; i += 64) {
        const int a = ids[i];
        int r = 0;
#pragma unroll 8
        for (int j = 0; j < nk; j++) r += ids[j] < a ? 1 : 0;
        v.lrank[base + i] = r;
    }
}

constexpr int kRankWaves = 16;     // cells ranked per 1024-thread workgroup of k_lscan and k_lscan_chains

// chains with more list cells than kScanCells: workgroup 0 scans all counts (zero in chains that did
// not count) and each rebuilding chain's cells get atom_off + the scan relative to the chain's first
// cell, in the list offsets of the buffer being rebuilt.  It runs on one CU, so every global access is
// coalesced: rounds of 8192 counts go through LDS, each thread scans 8 contiguous ones, a block scan
// over the threads and a running carry join the rounds.  The workgroups after it rank the members of
// every list cell of a rebuilding chain for the build.
__global__ __launch_bounds__(1024) void k_lscan(DevView v, int q, int cb)
{
    if (*req_any(v, q) == 0) return;
    if (blockIdx.x > 0) {                  // workgroups after the first rank the members of kRankWaves cells
        __shared__ int rank_ids[kRankWaves][256];
        const int wid = threadIdx.x >> 6;
        const int cell = ((int)blockIdx.x - 1) * kRankWaves + wid;
        if (cell < v.Cl) rank_cell_members(v, q, cb, cell, threadIdx.x & 63, rank_ids[wid]);
        return;
    }
    constexpr int T = 1024, PT = 8;
    __shared__ int buf[T * PT];
    __shared__ int wsum[16];
    __shared__ int carry;
    const int* __restrict__ cnt = v.lcount + (size_t)cb * v.Cl;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int n = v.Cl;
    if (tid == 0) carry = 0;
    for (int r0 = 0; r0 < n; r0 += T * PT) {
#pragma unroll
        for (int k = 0; k < PT; k++) {
            const int c = r0 + k * T + tid;
            buf[k * T + tid] = c < n ? cnt[c] : 0;
        }
        __syncthreads();
        int vals[PT];
        int s = 0;
#pragma unroll
        for (int k = 0; k < PT; k++) {
            vals[k] = buf[tid * PT + k];
            s += vals[k];
        }
        int incl = s;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        if (lane == 63) wsum[wid] = incl;
        __syncthreads();
        if (wid == 0) {
            int w = lane < 16 ? wsum[lane] : 0;
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                const int t = __shfl_up(w, o, 64);
                if (lane >= o) w += t;
            }
            if (lane < 16) wsum[lane] = w;
        }
        __syncthreads();
        int run = carry + incl - s + (wid > 0 ? wsum[wid - 1] : 0);
#pragma unroll
        for (int k = 0; k < PT; k++) {
            buf[tid * PT + k] = run;
            run += vals[k];
        }
        __syncthreads();                      // every thread has read carry; buf holds the round's scan
        if (tid == T - 1) carry = run;
#pragma unroll
        for (int k = 0; k < PT; k++) {
            const int c = r0 + k * T + tid;
            if (c < n) v.lscan_tmp[c] = buf[k * T + tid];
        }
        __syncthreads();                      // buf free, carry final
    }
    // list offsets of the rebuilding chains (chains are contiguous in cell space: a thread's cells change
    // chain rarely, so the chain's flag, buffer and base are looked up once per change)
    int kprev = -1, base = 0;
    bool rq = false;
    int* loffs = nullptr;
#pragma unroll 4
    for (int c = tid; c < n; c += T) {
        const int k = v.lcell_chain[c];
        const int sc = v.lscan_tmp[c];
        if (k != kprev) {
            kprev = k;
            rq = *req(v, q, k) != 0;
            const ChainDev& C = v.chains[k];
            base = C.atom_off - v.lscan_tmp[C.lcell_off];
            loffs = list_buf(v, *list_start(v, k) ^ 1).loffs;
        }
        if (rq) loffs[c] = base + sc;
    }
}

// batches of large chains (speculative copies): workgroups 0 .. nchains - 1, one per rebuilding chain,
// scan its counts (each thread a contiguous run) into the list offsets of the buffer being rebuilt, so
// that the build workgroups of a chain read 27 offsets instead of each scanning the chain's counts in
// LDS; the workgroups after them rank the members of every list cell of a rebuilding chain by atom id
// (one wave per cell), so that the build workgroups do not rank each neighbour cell again
__global__ __launch_bounds__(1024) void k_lscan_chains(DevView v, int q, int cb)
{
    __shared__ int rank_ids[kRankWaves][256];
    if (*req_any(v, q) == 0) return;
    if ((int)blockIdx.x >= v.nchains) {
        const int wid = threadIdx.x >> 6;
        const int cell = ((int)blockIdx.x - v.nchains) * kRankWaves + wid;
        if (cell < v.Cl) rank_cell_members(v, q, cb, cell, threadIdx.x & 63, rank_ids[wid]);
        return;
    }
    const int k = blockIdx.x;
    if (*req(v, q, k) == 0) return;
    const ChainDev& C = v.chains[k];
    constexpr int T = 1024;
    __shared__ int wsum[16];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int n = C.lncells;
    const int* __restrict__ cnt = v.lcount + (size_t)cb * v.Cl + C.lcell_off;
    int* __restrict__ loffs = list_buf(v, *list_start(v, k) ^ 1).loffs + C.lcell_off;
    const int per = (n + T - 1) / T;
    const int beg = min(tid * per, n), end = min(beg + per, n);
    int s = 0;
    for (int i = beg; i < end; i++) s += cnt[i];
    int incl = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    int run = C.atom_off + incl - s;
    for (int w = 0; w < wid; w++) run += wsum[w];
    for (int i = beg; i < end; i++) {
        loffs[i] = run;
        run += cnt[i];
    }
}

constexpr int kLMaxJ = 1024;      // staged neighbour members per pass (a cell never spans two passes)
constexpr int kJPad = 64;         // far entries past a pass's members: one iteration of the pair test
constexpr int kBuildWaves = 4;    // waves of a build workgroup

// exclusive scan of n <= kScanCells values in LDS by one 256-thread workgroup
__device__ __forceinline__ void block_scan_256(int* a, int n, int* wsum)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int per = (n + 255) / 256;
    const int beg = min(tid * per, n), end = min(beg + per, n);
    int s = 0;
    for (int i = beg; i < end; i++) s += a[i];
    int incl = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int run = incl - s;
    for (int w = 0; w < wave; w++) run += wsum[w];
    for (int i = beg; i < end; i++) {
        const int t = a[i];
        a[i] = run;
        run += t;
    }
    __syncthreads();
}

// U groups of 64 staged members from sj[j0] against the wave's own atom: the wave's hit mask straight
// from the compares (4: ordered less, 2: ordered greater), the positions from mbcnt, the count in a
// scalar.  Row offsets in 32 bits: Ntot * nbcap < 2^31 (checked at setup).  Entries of 16 or 32 bits (W16).
template <int U, bool W16>
__device__ __forceinline__ void build_step(const float4* sj, int j0, int lane, const float4& xi, float rl2,
                                           int rowoff, int nbcap, char* __restrict__ nbb, int& count)
{
    float4 p[U];
    float r2[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        p[u] = sj[j0 + u * 64 + lane];
        const float dx = xi.x - p[u].x, dy = xi.y - p[u].y, dz = xi.z - p[u].z;
        r2[u] = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        const unsigned long long hits = __builtin_amdgcn_fcmpf(r2[u], rl2, 4) & __builtin_amdgcn_fcmpf(r2[u], 0.f, 2);
        const int pos = count + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(hits >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)hits, 0u));
        if (r2[u] < rl2 && r2[u] > 0.f && pos < nbcap) {
            if constexpr (W16) *reinterpret_cast<uint16_t*>(nbb + ((unsigned)(rowoff + pos) << 1)) = (uint16_t)__float_as_int(p[u].w);
            else *reinterpret_cast<int*>(nbb + ((unsigned)(rowoff + pos) << 2)) = __float_as_int(p[u].w);
        }
        count += __popcll(hits);
    }
}

// Pair tests of one staged pass: the own atoms si[0, ni) (slot in w, entry counts in icount) against the
// staged members sj[0, jn), padded with kJPad far entries.  Each wave takes one own atom at a time and
// its 64 lanes test 64 members each step, branch-free (the far padding fails every test); the hits are
// appended in staging order.  With one own atom per wave the hit mask is wave-uniform (~12 VALU per
// group of 64 tested pairs; two own atoms x 32 members took ~18 and 6-9 % more build time).  The loop
// reads two groups of 64 while the second starts inside the members, then at most one more, so no read
// passes jn + 63 and the padding stays one group.
template <bool W16>
__device__ __forceinline__ void build_pass(const float4* sj, const float4* si, int* icount, int ni, int jn,
                                           float rl2, int nbcap, char* __restrict__ nbb)
{
    static_assert(kJPad >= 64, "every step reads fewer than 64 entries past the members");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int i = wave; i < ni; i += kBuildWaves) {
        const float4 xi = si[i];
        int count = icount[i];
        const int rowoff = __float_as_int(xi.w) * nbcap;
        int j0 = 0;
        for (; j0 + 64 < jn; j0 += 128) build_step<2, W16>(sj, j0, lane, xi, rl2, rowoff, nbcap, nbb, count);
        if (j0 < jn) build_step<1, W16>(sj, j0, lane, xi, rl2, rowoff, nbcap, nbb, count);
        if (lane == 0) icount[i] = count;
    }
}

// One workgroup per list cell of a rebuilding chain; it writes the buffer that is not the trajectory's
// start buffer and makes it current.
//   offsets: SCAN = the chain's counts scanned in LDS (chains of at most kScanCells cells), else the
//            offsets k_lscan wrote;
//   staging: the 27 neighbour cells' members (FP32 cell-relative positions, loaded once into
//            registers), ranked by atom id inside each cell, are staged in LDS in slot order with the
//            image shift dd * lsize folded in (the way k_lj stages its j-list);
//   own cell: the ranks of its members give their compact slots and slot data (x0 and r0 = x0 - cell
//            origin from the FP64 positions the count stored with the members, FP32 coordinates of
//            parity q, lperm, linv, tile marks);
//   lists:   each own atom appends the entries within rc + skin in staging order, so the list order
//            is fixed by the slot order and not by the arrival order of the count.
// A neighbourhood of more than kLMaxJ members is staged in passes of whole cells; the own cell is
// loaded with the first pass.
template <bool SCAN, bool W16, bool RANKED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_lbuild(DevView v, int q, int cb)
{
    const int c = xcd_block(blockIdx.x, v.Cl);
    const int ch = v.lcell_chain[c];
    if (*req_any(v, q) == 0 || *req(v, q, ch) == 0) return;
    constexpr int NT = 256, W = kBuildWaves, PT = (kLMaxJ + NT - 1) / NT;
    // staged j in slot order: coordinates relative to the i-cell origin, list entry in w; padded past the
    // pass's members to a whole iteration of the test loop with far-away entries
    __shared__ float4 sj[kLMaxJ + kJPad];
    __shared__ int sid[kLMaxJ];            // staged members' atom ids in arrival order (ranking)
    // own atoms by slot (FP32 coordinates, slot in w), their entry counts and their ids in arrival
    // order: lcap_max of each, in the dynamic LDS, so that the workgroup's LDS stays small
    extern __shared__ float4 dyn_lds[];
    float4* si = dyn_lds;
    int* icount = reinterpret_cast<int*>(dyn_lds + v.lcap_max);
    int* sido = icount + v.lcap_max;
    __shared__ int seg_cell[27], seg_beg[27], seg_cnt[27], seg_pref[28], seg_code[27], wsum[4];
    __shared__ float seg_sh[27][3];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr bool e16 = W16;              // 16-bit segment-relative entries
    const ChainDev& C = v.chains[ch];
    const int tb = *list_start(v, ch) ^ 1;
    const ListBuf L = list_buf(v, tb);
    const int nx = C.lnbin[0], ny = C.lnbin[1], nz = C.lnbin[2];
    const int lc = c - C.lcell_off;
    const int bz = lc / (nx * ny), by = (lc - bz * nx * ny) / nx, bx = lc - nx * (by + bz * ny);
    const int* __restrict__ cnt = v.lcount + (size_t)cb * v.Cl + C.lcell_off;
    const float4* __restrict__ mem = v.lmem + C.ldense_off;
    const double* __restrict__ memx = v.lmemx + C.ldense_off;
    const double* __restrict__ memp = v.lmemp + C.ldense_off;
    const size_t xstride = (size_t)v.ldense;
    const int N = v.Ntot;
    if (tid == 0) {
        *list_cur(v, ch) = tb;
        if (lc == 0) atomicAdd(&v.lflag[3], 1);        // chain rebuilds run (statistics)
    }
    // the own cell's members and their FP64 positions (one per thread, in bounds for every thread
    // below lcap): issued first, their latency overlaps the offsets
    float4 meo = make_float4(0.f, 0.f, 0.f, 0.f);
    double xo[3] = { 0.0, 0.0, 0.0 }, po[3] = { 0.0, 0.0, 0.0 };
    int orank = 0;                                   // RANKED: the member's rank by atom id (k_lscan_chains)
    if (tid < C.lcap) {
        const size_t mi = (size_t)lc * C.lcap + tid;
        meo = mem[mi];
        if constexpr (RANKED) orank = v.lrank[C.ldense_off + mi];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            xo[d] = memx[d * xstride + mi];
            po[d] = memp[d * xstride + mi];
        }
    }
    if (wave == 0) {
        int nk = 0;
        if (lane < 27) {
            const int dd[3] = { lane % 3 - 1, (lane / 3) % 3 - 1, lane / 9 - 1 };
            const int b3[3] = { bx, by, bz }, n3[3] = { nx, ny, nz };
            int j3[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                j3[d] = b3[d] + dd[d];
                if (j3[d] < 0 || j3[d] >= n3[d]) j3[d] = (b3[d] + n3[d] + dd[d]) % n3[d];
                seg_sh[lane][d] = (float)dd[d] * C.lsize_f[d];
            }
            const int jl = j3[0] + nx * (j3[1] + j3[2] * ny);
            seg_cell[lane] = jl;
            seg_code[lane] = image_code(dd[0], dd[1], dd[2]);
            if constexpr (!SCAN) {
                seg_beg[lane] = L.loffs[C.lcell_off + jl];
                nk = min(L.loffs[C.lcell_off + jl + 1] - L.loffs[C.lcell_off + jl], C.lcap);
            } else {
                nk = min(cnt[jl], C.lcap);
            }
            seg_cnt[lane] = nk;
        }
        // staging offsets of the 27 segments: a scan over lanes 0..26
        int incl = nk;
#pragma unroll
        for (int o = 1; o < 32; o <<= 1) {
            const int t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        if (lane < 27) seg_pref[lane + 1] = incl;
        if (lane == 0) seg_pref[0] = 0;
    }
    int* scan = reinterpret_cast<int*>(sj);          // SCAN: sj is free until the first pass is staged
    if constexpr (SCAN) {
        for (int i = tid; i < C.lncells; i += NT) scan[i] = cnt[i];
    }
    __syncthreads();
    const int ni = seg_cnt[13];                      // segment 13: the cell itself (ni <= lcap <= 256)
    if (ni == 0) return;

    const float rl2 = C.rl2_f;
    const int nbcap = v.nbcap;
    int ibeg = 0;
    for (int k0 = 0; k0 < 27;) {
        const bool first = k0 == 0;
        if (!first) __syncthreads();           // sid, sj free again
        int k1 = k0 + 1;
        while (k1 < 27 && seg_pref[k1 + 1] - seg_pref[k0] <= kLMaxJ) k1++;
        const int jlo = seg_pref[k0], jn = seg_pref[k1] - jlo;
        // members of this pass, loaded once (PT per thread); with the first pass the loads are in
        // flight during the scan of the counts, the ids go to LDS for the ranking after it
        float4 me[PT];
        int kk[PT], rk[PT];
#pragma unroll
        for (int u = 0; u < PT; u++) {
            const int t = tid + NT * u;
            kk[u] = k0;
            rk[u] = 0;
            me[u] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (t < jn) {
                int k = k0;
#pragma unroll
                for (int step = 16; step > 0; step >>= 1)
                    if (k + step < k1 && seg_pref[k + step] <= jlo + t) k += step;
                kk[u] = k;
                const size_t mi = (size_t)seg_cell[k] * C.lcap + (jlo + t - seg_pref[k]);
                me[u] = mem[mi];
                if constexpr (RANKED) rk[u] = v.lrank[C.ldense_off + mi];
            }
        }
        if constexpr (SCAN) {
            if (first) {
                block_scan_256(scan, C.lncells, wsum);     // ends with a barrier: scan[] final
                if (tid < 27) seg_beg[tid] = C.atom_off + scan[seg_cell[tid]];
            }
        }
        if constexpr (!RANKED) {
            if (first && tid < ni) sido[tid] = __float_as_int(meo.w);
#pragma unroll
            for (int u = 0; u < PT; u++)
                if (tid + NT * u < jn) sid[tid + NT * u] = __float_as_int(me[u].w);
        }
        __syncthreads();                       // seg_beg, sid, sido; the scan area is free again
        if (first) ibeg = seg_beg[13];
        // 16-bit entries: the cell's segment starts (the padding segments: its own first slot)
        if (e16 && first && tid < kSegStride) L.ctab[(size_t)c * kSegStride + tid] = seg_beg[tid < 27 ? tid : 13];
        if (first && tid < ni) {
            // own slot data: slots by atom id; the rebuild reference (FP32 cell-relative position, cell)
            const int a = __float_as_int(meo.w);
            int r = orank;
            if constexpr (!RANKED) {
#pragma unroll 8
                for (int k = 0; k < ni; k++) r += sido[k] < a ? 1 : 0;
            }
            const int slot = ibeg + r;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                v.xl[d * N + slot] = xo[d];
                v.pl[d * N + slot] = po[d];
            }
            L.lref[slot] = make_float4(meo.x, meo.y, meo.z, __int_as_float(bx | (by << 10) | (bz << 20)));
            L.xsl[(size_t)q * N + slot] = meo;
            L.lperm[slot] = a;
            L.linv[a] = slot;
            if (e16) L.lcs[slot] = lc;
            tile_cur_mark(v, C, slot, tb);
            si[r] = make_float4(meo.x, meo.y, meo.z, __int_as_float(slot));
            icount[r] = 0;
        }
#pragma unroll
        for (int u = 0; u < PT; u++) {
            const int t = tid + NT * u;
            if (t < jn) {
                const int k = kk[u];
                const int s0 = seg_pref[k] - jlo, nk = seg_cnt[k];
                const int a = __float_as_int(me[u].w);
                int r = rk[u];
                if constexpr (!RANKED) {
#pragma unroll 8
                    for (int w = 0; w < nk; w++) r += sid[s0 + w] < a ? 1 : 0;
                }
                // w holds the finished list entry: segment k << kRankBits | rank r, or j slot <<
                // kImageBits | image code
                const int e = e16 ? ((k << kRankBits) | r) : (((seg_beg[k] + r) << kImageBits) | seg_code[k]);
                sj[s0 + r] = make_float4(me[u].x + seg_sh[k][0], me[u].y + seg_sh[k][1], me[u].z + seg_sh[k][2],
                                         __int_as_float(e));
            }
        }
        if (tid < kJPad) sj[jn + tid] = make_float4(kFarShift, kFarShift, kFarShift, 0.f);
        __syncthreads();
        build_pass<W16>(sj, si, icount, ni, jn, rl2, nbcap, L.nb);
        k0 = k1;
    }
    __syncthreads();
    for (int r = tid; r < ni; r += NT) {
        const int n = icount[r];
        L.nbn[ibeg + r] = min(n, nbcap);
        if (n > nbcap) atomicOr(&v.lflag[2], 2);
    }
    // rows padded to the sweep shape's unconditional span with no-op entries
    const int span = v.lspan;
    for (int r = wave; r < ni; r += W) {
        const int slot = __float_as_int(si[r].w);
        if (e16) {
            uint16_t* __restrict__ row = list_row<uint16_t>(v, L, slot);
            for (int k = icount[r] + lane; k < span; k += 64) row[k] = kPadEntry;
        } else {
            int* __restrict__ row = list_row<int>(v, L, slot);
            const int pad = (slot << kImageBits) | kFarCode;
            for (int k = icount[r] + lane; k < span; k += 64) row[k] = pad;
        }
    }
}

// every list row of both buffers filled with no-op entries (before the first rebuild: 32-bit entries
// name the slot itself; with 16-bit ones the segment starts and cells are zero then, so the padding
// reads slot 0)
__global__ __launch_bounds__(256) void k_lpad_init(DevView v)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t per = (size_t)v.Ntot * v.nbcap;
    if (i >= 2 * per) return;
    if (v.entry16) {
        reinterpret_cast<uint16_t*>(v.nb)[i] = kPadEntry;
    } else {
        const int slot = (int)((i % per) / v.nbcap);
        reinterpret_cast<int*>(v.nb)[i] = (slot << kImageBits) | kFarCode;
    }
}

// rows of both buffers padded from their counts up to span (a sweep shape with a longer first batch
// than the span the rows were built with)
__global__ __launch_bounds__(256) void k_lrepad(DevView v, int span)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * (size_t)v.Ntot) return;
    const int b = (int)(i / v.Ntot), slot = (int)(i % v.Ntot);
    const ListBuf L = list_buf(v, b);
    if (v.entry16) {
        uint16_t* __restrict__ row = list_row<uint16_t>(v, L, slot);
        for (int k = max(L.nbn[slot], 0); k < span; k++) row[k] = kPadEntry;
    } else {
        int* __restrict__ row = list_row<int>(v, L, slot);
        const int pad = (slot << kImageBits) | kFarCode;
        for (int k = max(L.nbn[slot], 0); k < span; k++) row[k] = pad;
    }
}

hipError_t launch_list_repad(const DevView& v, int span, hipStream_t s)
{
    const size_t n = 2 * (size_t)v.Ntot;
    hipLaunchKernelGGL(k_lrepad, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, v, min(span, v.nbcap));
    return hipGetLastError();
}

hipError_t launch_list_pad_init(const DevView& v, hipStream_t s)
{
    const size_t n = 2 * (size_t)v.Ntot * v.nbcap;
    hipLaunchKernelGGL(k_lpad_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, v);
    return hipGetLastError();
}

hipError_t launch_list_rebuild(const DevView& v, int q, int cb, hipStream_t s)
{
    hipLaunchKernelGGL(k_lcount, dim3((v.Ntot + 255) / 256), dim3(256), 0, s, v, q, cb);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const size_t dyn = (size_t)v.lcap_max * (sizeof(float4) + 2 * sizeof(int));
    if (v.lscan_chains) {
        const int grid = v.nchains + (v.Cl + kRankWaves - 1) / kRankWaves;
        hipLaunchKernelGGL(k_lscan_chains, dim3(grid), dim3(1024), 0, s, v, q, cb);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if (v.entry16) hipLaunchKernelGGL((k_lbuild<false, true, true>), dim3(v.Cl), dim3(256), dyn, s, v, q, cb);
        else hipLaunchKernelGGL((k_lbuild<false, false, true>), dim3(v.Cl), dim3(256), dyn, s, v, q, cb);
        return hipGetLastError();
    }
    if (v.lscan_in_build) {
        if (v.entry16) hipLaunchKernelGGL((k_lbuild<true, true, false>), dim3(v.Cl), dim3(256), dyn, s, v, q, cb);
        else hipLaunchKernelGGL((k_lbuild<true, false, false>), dim3(v.Cl), dim3(256), dyn, s, v, q, cb);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_lscan, dim3(1 + (v.Cl + kRankWaves - 1) / kRankWaves), dim3(1024), 0, s, v, q, cb);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (v.entry16) hipLaunchKernelGGL((k_lbuild<false, true, true>), dim3(v.Cl), dim3(256), dyn, s, v, q, cb);
    else hipLaunchKernelGGL((k_lbuild<false, false, true>), dim3(v.Cl), dim3(256), dyn, s, v, q, cb);
    return hipGetLastError();
}

// ---- sweep ---------------------------------------------------------------------------------------
// drift of one atom by dt (FP64, wrap into [0, L)), its FP32 list coordinates of the next parity and
// the rebuild request when it is beyond skin / 2 from its rebuild position
// Rebuild trigger: displacement^2 from the rebuild position beyond which a drift requests a rebuild.
// With a check before every force evaluation it is (skin / 2)^2.  With checks every m-th evaluation the
// trigger sits lower by the distance m - 1 drifts can add, lcheck_dt * (beta / m) * |p|max with
// |p|max = sqrt(p2max), so the unchecked evaluations in between still find every pair within rc in
// the lists; every drift verifies |p|^2 <= p2max and flags a violation (the run then reports an
// error).  A trigger at or below zero requests a rebuild at every check.
__device__ __forceinline__ double trigger2(const DevView& v, const ChainDev& C)
{
    if (v.lcheck_dt == 0.0) return C.half_skin2;
    const double h = C.half_skin - v.lcheck_dt * C.vmax;
    return h > 0.0 ? h * h : -1.0;
}

__device__ __forceinline__ void list_drift_slot(const DevView& v, const ListBuf& L, const ChainDev& C, int ch,
                                                int slot, int q, double dt, const double pn[3])
{
    const int N = v.Ntot;
    const double coeff_x = C.coeff_x;              // compute_coeff_position, identical_particles.cpp:646-648
    float f3[3];
    double disp2 = 0.0;
    const float4 ref = L.lref[slot];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double xv = v.xl[d * N + slot];
        xv += dt * coeff_x * pn[d];
        xv = lp_round(xv, v.lp_scale);
        xv = wrap_box(xv, C.L[d]);
        v.xl[d * N + slot] = xv;
        const double r = list_rel(C, ref, d, xv);
        const double e = r - (double)ref_coord(ref, d);
        disp2 += e * e;
        f3[d] = (float)r;
    }
    L.xsl[(size_t)(q ^ 1) * N + slot] = make_float4(f3[0], f3[1], f3[2], 0.f);
    if (disp2 > trigger2(v, C)) request(v, q ^ 1, ch);
    if (v.lcheck_dt != 0.0 && pn[0] * pn[0] + pn[1] * pn[1] + pn[2] * pn[2] > C.p2max) atomicOr(&v.lflag[2], 4);
}

// The per-atom tail of a list sweep, on lanes sub = 0, 1, 2 of the slot (one dimension each; fd is the
// force component of the lane's dimension): store the force in atom order, kick p in FP64, optionally
// drift x in FP64 (identical_particles.cpp:651-673), then lane 0 writes the FP32 coordinates of the other
// parity and raises the chain's rebuild request when the atom is beyond the trigger distance of its
// rebuild position.  Thread 0 clears the requests of this evaluation's parity when its rebuild check ran
// (checked), and otherwise carries a pending one over to the next parity.
template <int MODE, int LP>
__device__ __forceinline__ void list_tail(const DevView& v, const ListBuf& L, const ChainDev& C, const Chunk& tl, int s,
                                          bool val, int sub, int d, int atom, float fd, double p_d, double x_d,
                                          float4 ref, double kick_dt, double drift_dt, int q, int checked)
{
    const int N = v.Ntot;
    double e = 0.0, pn = 0.0;
    float rel = 0.f;
    if (val && sub < 3) {
        if constexpr ((MODE & LJ_STOREF) != 0) v.f32[d * N + atom] = fd;
        pn = p_d;
        if constexpr ((MODE & LJ_KICK) != 0) {
            const double kdt = kick_dt * C.beta;       // dt * coeff_p, coeff_p = beta (:642-644)
            pn = lp_round(pn - kdt * (double)fd, v.lp_scale);
            v.pl[d * N + s] = pn;
        }
        if constexpr ((MODE & LJ_DRIFT) != 0) {
            const double coeff_x = C.coeff_x;          // compute_coeff_position, identical_particles.cpp:646-648
            double xv = x_d;
            xv += drift_dt * coeff_x * pn;
            xv = lp_round(xv, v.lp_scale);
            xv = wrap_box(xv, C.L[d]);
            v.xl[d * N + s] = xv;
            const double r = list_rel(C, ref, d, xv);
            e = r - (double)ref_coord(ref, d);
            rel = (float)r;
        }
        if constexpr ((MODE & LJ_REL64) != 0) v.rel64[4 * (size_t)s + d] = list_rel(C, ref, d, x_d);
    }
    if constexpr ((MODE & LJ_DRIFT) != 0) {
        // lane 0 of the slot gathers the y, z coordinates and the squared displacement and momentum over
        // lanes 0-2 (lane 3 adds zeros: (ex^2 + ey^2) + (ez^2 + 0) is ex^2 + ey^2 + ez^2 to the bit)
        const float ry = dpp_f<kDppFrom1>(rel), rz = dpp_f<kDppFrom2>(rel);
        const double e2 = group_sum<4>(e * e), p2 = group_sum<4>(pn * pn);
        if (val && sub == 0) {
            L.xsl[(size_t)(q ^ 1) * N + s] = make_float4(rel, ry, rz, 0.f);
            if (e2 > trigger2(v, C)) request(v, q ^ 1, tl.chain);
            if (v.lcheck_dt != 0.0 && p2 > C.p2max) atomicOr(&v.lflag[2], 4);
        }
    }
    if (threadIdx.x == 0) {
        if (tl.begin == C.atom_off) {
            if (checked) {
                *req(v, q, tl.chain) = 0;                 // this evaluation's rebuild is done
            } else if (*req(v, q, tl.chain) != 0) {       // no check ran: the request waits for the next one
                *req(v, q, tl.chain) = 0;
                request(v, q ^ 1, tl.chain);
            }
        }
        if (blockIdx.x == 0) *req_any(v, q) = 0;
    }
}

// LP lanes per slot, kTile slots per workgroup.  Each lane first loads up to kBatch of its list entries
// and gathers their coordinates, so that one memory latency covers most lists; the per-atom tail (kick,
// drift, next-parity coordinates, displacement test) runs on three lanes, one per dimension.  The
// workgroup's list buffer comes from tcur[blockIdx], so nothing delays the first list load but the
// tile's own two scalar loads.  32-bit entries (W16 false) carry the slot and the image code, so the
// gathers are issued before the barrier that publishes the shift table.  16-bit entries (W16) are read
// two per 32-bit load where kBatch is even, while the slot's lanes copy the segment starts of its list
// cell to an LDS row; the gathers (slot = segment start + rank) follow the barrier.
// KB entries per lane in the first batch: KB * LP should cover most rows (at 8 lanes, 14 covers the
// typical 105 entries at skin 0.6 and N = 32768; 12 the 94 of skin 0.5, 16 the 116 of skin 0.7); the
// sweep shape (LP, KB) is chosen by the autotuner
template <int MODE, int LP, int KB, bool W16>
__global__ __launch_bounds__(kTile * LP) void k_lsweep(DevView v, double kick_dt, double drift_dt, int q, int checked)
{
    constexpr int kBatch = KB;
    static_assert(kBatch * LP <= kListSpan, "the first batch reads padded entries only");
    constexpr int kShifts = W16 ? kSegStride : 64;
    __shared__ float4 shift_tab[kShifts];                // image shift by segment (W16) or 6-bit image code
    __shared__ int4 seg_tab[W16 ? kTile : 1][kSegStride / 4];   // W16: segment starts of each slot's cell
    const int tile = xcd_block(blockIdx.x, v.ntiles);
    const Chunk tl = v.tiles[tile];
    const ListBuf L = list_buf(v, v.tcur[tile]);
    const ChainDev& C = v.chains[tl.chain];
    const int N = v.Ntot;
    const int t = threadIdx.x / LP, sub = threadIdx.x % LP;
    const int s = tl.begin + t;
    const bool val = s < tl.end;
    const int d = sub < 3 ? sub : 0;                      // the per-dimension tail runs on lanes 0, 1, 2
    const float4* __restrict__ xq = L.xsl + (size_t)q * N;
    const int* segs = reinterpret_cast<const int*>(seg_tab[W16 ? t : 0]);
    float ax = 0.f, ay = 0.f, az = 0.f;
    float4 xi = make_float4(0.f, 0.f, 0.f, 0.f);
    int atom = 0;
    double p_d = 0.0, x_d = 0.0;
    float4 ref = make_float4(0.f, 0.f, 0.f, 0.f);
    // neighbour slot and shift-table index of an entry
    auto slot_of = [&](unsigned e) -> unsigned {
        if constexpr (W16) return (unsigned)entry_slot(segs, e);
        else return e >> kImageBits;
    };
    auto shift_of = [&](unsigned e) -> unsigned {
        if constexpr (W16) return e >> kRankBits;
        else return e & 63u;
    };
    auto pair = [&](const float4 pj, unsigned e) {
        const float4 sh = shift_tab[shift_of(e)];
        const float dx = xi.x - (pj.x + sh.x);
        const float dy = xi.y - (pj.y + sh.y);
        const float dz = xi.z - (pj.z + sh.z);
        const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
        const float inv = __builtin_amdgcn_rcpf(r2);
        const float s2 = C.sigma2_f * inv;
        const float s6 = s2 * s2 * s2;
        float gq = s6 * (0.5f - s6) * inv;            // sr6 (0.5 - sr6) / r^2, identical_particles.cpp:611-613
        gq = r2 < C.rc2_f ? gq : 0.f;                 // list entries never hold the atom itself
        ax = __builtin_fmaf(gq, dx, ax);
        ay = __builtin_fmaf(gq, dy, ay);
        az = __builtin_fmaf(gq, dz, az);
    };
    if (threadIdx.x < kShifts) {
        const int k = threadIdx.x;
        if constexpr (W16) {
            shift_tab[k] = make_float4(seg_shift(k, 0, C.lsize_f[0]), seg_shift(k, 1, C.lsize_f[1]),
                                       seg_shift(k, 2, C.lsize_f[2]), 0.f);
        } else {
            shift_tab[k] = k == kFarCode ? make_float4(kFarShift, kFarShift, kFarShift, 0.f)
                                         : make_float4(image_shift(k, 0, C.lsize_f[0]), image_shift(k, 1, C.lsize_f[1]),
                                                       image_shift(k, 2, C.lsize_f[2]), 0.f);
        }
    }
    unsigned ent[kBatch];
    float4 pj[W16 ? 1 : kBatch];                          // 32-bit entries: gathered before the barrier
    int n = 0;
    if (val) {
        // the first kBatch * LP entries are read unconditionally, before the count is known: past the
        // count the build pads them with no-op entries (nbcap >= kListSpan)
        if constexpr (W16) {
            const uint16_t* __restrict__ nbs = list_row<uint16_t>(v, L, s);
            if constexpr (kBatch % 2 == 0) {
                const unsigned* __restrict__ nb2 = reinterpret_cast<const unsigned*>(nbs);   // nbcap is even
#pragma unroll
                for (int u = 0; u < kBatch / 2; u++) {
                    const unsigned w = nb2[sub + u * LP];
                    ent[2 * u] = w & 0xffffu;
                    ent[2 * u + 1] = w >> 16;
                }
            } else {
#pragma unroll
                for (int u = 0; u < kBatch; u++) ent[u] = nbs[sub + u * LP];
            }
            load_seg_row<LP>(L, C, s, sub, seg_tab[t]);
        } else {
            const int* __restrict__ nbs = list_row<int>(v, L, s);
#pragma unroll
            for (int u = 0; u < kBatch; u++) ent[u] = (unsigned)nbs[sub + u * LP];
        }
        if constexpr ((MODE & LJ_STOREF) != 0) atom = L.lperm[s];     // the stored force is in atom order
        xi = xq[s];
        n = L.nbn[s];
        if constexpr (!W16) {
#pragma unroll
            for (int u = 0; u < kBatch; u++) pj[u] = xq[slot_of(ent[u])];
        }
    }
    __syncthreads();                                      // shift_tab, seg_tab
    if (val) {
        if constexpr (W16) {
            float4 pw[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; u++) pw[u] = xq[slot_of(ent[u])];
#pragma unroll
            for (int u = 0; u < kBatch; u++) pair(pw[u], ent[u]);
        } else {
#pragma unroll
            for (int u = 0; u < kBatch; u++) pair(pj[u], ent[u]);
        }
        for (int k = sub + kBatch * LP; k < n; k += LP) {
            const unsigned ee = W16 ? (unsigned)list_row<uint16_t>(v, L, s)[k] : (unsigned)list_row<int>(v, L, s)[k];
            pair(xq[slot_of(ee)], ee);
        }
    }
    // the tail's operands, fetched after the pairs (fewer registers live through the pair loop)
    if (val && sub < 3) {
        p_d = v.pl[d * N + s];
        if constexpr ((MODE & (LJ_DRIFT | LJ_REL64)) != 0) {
            x_d = v.xl[d * N + s];
            ref = L.lref[s];
        }
    }
    ax = group_sum<LP>(ax);
    ay = group_sum<LP>(ay);
    az = group_sum<LP>(az);
    list_tail<MODE, LP>(v, L, C, tl, s, val, sub, d, atom, C.eps48_f * (d == 0 ? ax : (d == 1 ? ay : az)), p_d, x_d, ref,
                        kick_dt, drift_dt, q, checked);
}

// Surrogate forces on the lists (same tiles, lanes, entries and tail as k_lsweep): for every listed pair
// within rc, dE/dx_i += [sum_k (g_k(i) + g_k(j)) phi'_k(r)] (x_i - x_j) / r, with g = dE/dG of both atoms
// (slot order, from k_lsurr_desc_mlp), phi_k(r) = exp(-(r - mu_k)^2 / 2w^2) fc(r), fc the cosine cutoff.
// Two passes, so that no lane idles through the 32-feature work of a pair beyond rc: the LP lanes of a
// slot first test its list entries (stage_pairs: the first 16 per lane loaded together) and append the
// pairs within rc to an LDS list, then walk that list densely, each lane holding g(i) and gathering g(j)
// per pair.  The
// Gaussians of a pair come from two exponentials per block of 8 centres: with t = r - mu_k,
// exp(-(t - D)^2 c) = exp(-t^2 c) rho, rho = exp(2 c D t - c D^2), and rho advances by exp(-2 c D^2).
constexpr int kSurrPairCap = 104;   // pairs within rc per atom staged in LDS (3 workgroups per CU)

template <int MODE, int LP, bool W16>
__global__ __launch_bounds__(kTile * LP) void k_lsurr_sweep(DevView v, const float* __restrict__ g, double kick_dt,
                                                            double drift_dt, int q, int checked)
{
    __shared__ float4 spd[kTile][kSurrPairCap];          // (dx, dy, dz, neighbour slot bits) of the pairs within rc
    __shared__ int4 seg_tab[W16 ? kTile : 1][kSegStride / 4];   // W16: segment starts of each slot's cell
    const int tile = xcd_block(blockIdx.x, v.ntiles);
    const Chunk tl = v.tiles[tile];
    const ListBuf L = list_buf(v, v.tcur[tile]);
    const ChainDev& C = v.chains[tl.chain];
    const int N = v.Ntot;
    const int t = threadIdx.x / LP, sub = threadIdx.x % LP;
    const int s = tl.begin + t;
    const bool val = s < tl.end;
    const int d = sub < 3 ? sub : 0;
    const float4* __restrict__ xq = L.xsl + (size_t)q * N;
    float ax = 0.f, ay = 0.f, az = 0.f;
    int atom = 0;
    double p_d = 0.0, x_d = 0.0;
    float4 ref = make_float4(0.f, 0.f, 0.f, 0.f);
    int np = 0;
    if (val) {
        if constexpr ((MODE & LJ_STOREF) != 0) atom = L.lperm[s];
        if (sub < 3) {
            p_d = v.pl[d * N + s];
            if constexpr ((MODE & (LJ_DRIFT | LJ_REL64)) != 0) {
                x_d = v.xl[d * N + s];
                ref = L.lref[s];
            }
        }
        // pass 1: the pairs within rc, in entry order (16-bit entries: the slot's lanes share a wave, and
        // a wave barrier orders the segment row's writes before its reads)
        if constexpr (W16) {
            load_seg_row<LP>(L, C, s, sub, seg_tab[t]);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        np = stage_pairs<LP, 16, true, W16>(C, xq, L.nb, v.nbcap, s, reinterpret_cast<const int*>(seg_tab[W16 ? t : 0]),
                                            L.nbn[s], xq[s], sub, &spd[t][0], nullptr, kSurrPairCap);
        if (np > kSurrPairCap) {
            if (sub == 0) atomicOr(&v.lflag[2], 8);
            np = kSurrPairCap;
        }
    }
    __syncthreads();
    if (val) {
        float gi[kSurrFeat];
        const float4* gis = reinterpret_cast<const float4*>(g + (size_t)s * kSurrFeat);
#pragma unroll
        for (int b = 0; b < kSurrFeat / 4; b++) {
            const float4 q4 = gis[b];
            gi[4 * b] = q4.x; gi[4 * b + 1] = q4.y; gi[4 * b + 2] = q4.z; gi[4 * b + 3] = q4.w;
        }
        const float pirc = 3.14159265358979f / sqrtf(C.rc2_f);
        const float c2 = 1.0f / (2.0f * kSurrWidth * kSurrWidth), iw2 = 1.0f / (kSurrWidth * kSurrWidth);
        constexpr float kD = (kSurrMu1 - kSurrMu0) / (float)(kSurrFeat - 1);
        const float cq = __expf(-2.0f * c2 * kD * kD);
        auto load_g = [&](const float4 pd, float gj[kSurrFeat]) {
            const float4* gjs = reinterpret_cast<const float4*>(g + (size_t)__float_as_int(pd.w) * kSurrFeat);
#pragma unroll
            for (int b = 0; b < kSurrFeat / 4; b++) {
                const float4 q4 = gjs[b];
                gj[4 * b] = q4.x; gj[4 * b + 1] = q4.y; gj[4 * b + 2] = q4.z; gj[4 * b + 3] = q4.w;
            }
        };
        auto pair_force = [&](const float4 pd, const float gj[kSurrFeat]) {
            const float r = sqrtf(__builtin_fmaf(pd.z, pd.z, __builtin_fmaf(pd.y, pd.y, pd.x * pd.x)));
            const float fc = 0.5f * (__cosf(pirc * r) + 1.0f), dfc = -0.5f * pirc * __sinf(pirc * r);
            // phi_k'(r) = g_k(r) (dfc - (r - mu_k) fc / w^2) = g_k(r) (P + Q mu_k): one FMA per centre
            const float Q = iw2 * fc, P = __builtin_fmaf(-Q, r, dfc);
            float acc = 0.f;
#pragma unroll
            for (int b = 0; b < kSurrFeat / 8; b++) {
                const float tb = r - mu_k(8 * b);
                float ex = __expf(-tb * tb * c2);
                float rho = __expf(c2 * kD * (2.0f * tb - kD));
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int f = 8 * b + u;
                    const float dphi = ex * __builtin_fmaf(Q, mu_k(f), P);
                    acc = __builtin_fmaf(gi[f] + gj[f], dphi, acc);
                    ex *= rho;
                    rho *= cq;
                }
            }
            const float coef = acc / r;
            ax = __builtin_fmaf(coef, pd.x, ax);
            ay = __builtin_fmaf(coef, pd.y, ay);
            az = __builtin_fmaf(coef, pd.z, az);
        };
        // pass 2: the staged pairs, densely, two per step with both neighbours' g loaded together
        for (int k = sub; k < np; k += 2 * LP) {
            const bool two = k + LP < np;
            const float4 pd0 = spd[t][k];
            const float4 pd1 = two ? spd[t][k + LP] : pd0;
            float g0[kSurrFeat], g1[kSurrFeat];
            load_g(pd0, g0);
            load_g(pd1, g1);
            pair_force(pd0, g0);
            if (two) pair_force(pd1, g1);
        }
    }
    ax = group_sum<LP>(ax);
    ay = group_sum<LP>(ay);
    az = group_sum<LP>(az);
    list_tail<MODE, LP>(v, L, C, tl, s, val, sub, d, atom, d == 0 ? ax : (d == 1 ? ay : az), p_d, x_d, ref, kick_dt,
                        drift_dt, q, checked);
}

__global__ __launch_bounds__(256) void k_ldrift(DevView v, double dt, int q)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= v.Ntot) return;
    const int ch = v.atom_chain[s];          // a chain's slots span the same range as its atoms
    const ChainDev& C = v.chains[ch];
    const ListBuf L = list_buf(v, *list_cur(v, ch));
    if (s == 0) *req_any(v, q) = 0;
    if (s == C.atom_off) {
        // a pending request of parity q carries over to q ^ 1 (nothing consumed it)
        const int pend = *req(v, q, ch);
        *req(v, q, ch) = 0;
        if (pend) request(v, q ^ 1, ch);
    }
    const int N = v.Ntot;
    const double pn[3] = { v.pl[s], v.pl[N + s], v.pl[2 * N + s] };
    list_drift_slot(v, L, C, ch, s, q, dt, pn);
}

// ---- FP64 energy ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_lrel64(DevView v)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= v.Ntot) return;
    const int ch = v.atom_chain[s];
    const ChainDev& C = v.chains[ch];
    const ListBuf L = list_buf(v, *list_cur(v, ch));
    const int N = v.Ntot;
    const float4 ref = L.lref[s];
#pragma unroll
    for (int d = 0; d < 3; d++) v.rel64[4 * (size_t)s + d] = list_rel(C, ref, d, v.xl[d * N + s]);
}

__device__ __forceinline__ double tile_sum_256(double s, double* part)
{
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) s += __shfl_xor(s, m, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    return ((part[0] + part[1]) + part[2]) + part[3];
}

template <bool W16>
__global__ __launch_bounds__(256) void k_lenergy(DevView v, int q)
{
    constexpr int kShifts = W16 ? kSegStride : 64;
    __shared__ double wpart[4];
    __shared__ double shift64[kShifts][3];              // image shift by segment or image code, exact in FP64
    __shared__ int4 seg_tab[W16 ? kTile : 1][kSegStride / 4];   // W16: segment starts of each slot's cell
    const int tile = xcd_block(blockIdx.x, v.ntiles);
    const Chunk tl = v.tiles[tile];
    const ChainDev& C = v.chains[tl.chain];
    const ListBuf L = list_buf(v, *list_cur(v, tl.chain));
    const int t = threadIdx.x >> 3, sub = threadIdx.x & 7;
    const int s = tl.begin + t;
    const bool val = s < tl.end;
    if (threadIdx.x < kShifts) {
        const int k = threadIdx.x;
#pragma unroll
        for (int d = 0; d < 3; d++)
            shift64[k][d] = W16 ? seg_shift64(k, d, C.lsize[d]) : (double)(((k >> (2 * d)) & 3) - 1) * C.lsize[d];
    }
    if constexpr (W16) {
        if (val) load_seg_row<8>(L, C, s, sub, seg_tab[t]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (tl.begin == C.atom_off) *req(v, q, tl.chain) = 0;
        if (blockIdx.x == 0) *req_any(v, q) = 0;
    }
    const int N = v.Ntot;
    const int* segs = reinterpret_cast<const int*>(seg_tab[W16 ? t : 0]);
    double au = 0.0, aw = 0.0, an = 0.0, ak = 0.0;
    if (val) {
        const double2* __restrict__ r64 = reinterpret_cast<const double2*>(v.rel64);   // [Ntot][4]: x, y | z, -
        const double2 ia = r64[2 * (size_t)s], ib = r64[2 * (size_t)s + 1];
        const double xi = ia.x, yi = ia.y, zi = ib.x;
        const int n = L.nbn[s];
        using E = std::conditional_t<W16, uint16_t, int>;
        const E* __restrict__ nbs = list_row<E>(v, L, s);
        // sqrt(r2) < rc (the reference's test) decided without the square root away from the cut-off:
        // the FP64 square root is within one ulp, far inside a 1e-12 relative band
        const double lo2 = C.rc2 * (1.0 - 1e-12), hi2 = C.rc2 * (1.0 + 1e-12);
        // batches of kEB entries per lane (4: 66 VGPRs, 7 waves per SIMD; 8 took 99 VGPRs and 4 waves and
        // ran the trajectory 1.7 % slower with eight speculative copies): the batch's entries, then its
        // neighbours' coordinates, are
        // loaded together (two memory round trips per batch); each lane still adds its entries in
        // increasing order.  Entries past the count read the row's last one (with 16-bit entries only
        // those below the count, and the padding, decode to slots of this buffer).
        constexpr int kEB = 4;
        for (int k0 = sub; k0 < n; k0 += 8 * kEB) {
            unsigned ent[kEB];
#pragma unroll
            for (int u = 0; u < kEB; u++) ent[u] = (unsigned)nbs[min(k0 + 8 * u, n - 1)];
            double2 ja[kEB];
            double jz[kEB];
#pragma unroll
            for (int u = 0; u < kEB; u++) {
                const size_t j = W16 ? (size_t)entry_slot(segs, ent[u]) : (size_t)(ent[u] >> kImageBits);
                ja[u] = r64[2 * j];
                jz[u] = v.rel64[4 * j + 2];
            }
#pragma unroll
            for (int u = 0; u < kEB; u++) {
                const bool live = k0 + 8 * u < n;
                const double* sh = shift64[W16 ? ent[u] >> kRankBits : ent[u] & 63u];
                const double dx = xi - (ja[u].x + sh[0]);
                const double dy = yi - (ja[u].y + sh[1]);
                const double dz = zi - (jz[u] + sh[2]);
                const double r2 = dx * dx + dy * dy + dz * dz;
                bool inside = r2 < lo2;
                // the square root only where some lane of the wave is within the band (a uniform branch)
                const bool band = live && !inside && r2 < hi2;
                if (__builtin_amdgcn_ballot_w64(band) != 0ull)
                    if (band) inside = sqrt(r2) < C.rc;
                if (live && inside && r2 > 0.0) {              // identical_particles.cpp:432-435
                    const double s2 = C.sigma2 / r2;
                    const double s6 = s2 * s2 * s2;
                    au += s6 * (s6 - 1.0);
                    aw += s6 * (s6 - 0.5);
                    an += 1.0;
                }
            }
        }
        if (sub == 0) {
            const double px = v.pl[s], py = v.pl[N + s], pz = v.pl[2 * N + s];
            ak = (px * px + py * py + pz * pz) / (2 * C.mass);
        }
    }
    au = tile_sum_256(au, wpart);
    aw = tile_sum_256(aw, wpart);
    an = tile_sum_256(an, wpart);
    ak = tile_sum_256(ak, wpart);
    if (threadIdx.x == 0) {
        double* tp = &v.tile_part[4 * tile];
        tp[0] = au; tp[1] = aw; tp[2] = an; tp[3] = ak;
    }
}

// ---- slot-ordered state <-> atom order ------------------------------------------------------------
__global__ __launch_bounds__(256) void k_limport(DevView v)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= v.Ntot) return;
    const int ch = v.atom_chain[s];
    const int a = list_buf(v, *list_cur(v, ch)).lperm[s];
    const int N = v.Ntot;
    if (s == v.chains[ch].atom_off) *list_pord(v, ch) = -1;       // pl in the current buffer's order
#pragma unroll
    for (int d = 0; d < 3; d++) {
        v.xl[d * N + s] = v.x[d * N + a];
        v.pl[d * N + s] = v.p[d * N + a];
    }
}

// saved: also h_x (xl_saved, in the start buffer's slot order) into x_saved
__global__ __launch_bounds__(256) void k_lexport(DevView v, int saved)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= v.Ntot) return;
    const int ch = v.atom_chain[s];
    const int cur = *list_cur(v, ch), pord = *list_pord(v, ch);
    const int a = list_buf(v, cur).lperm[s];
    const int ap = pord >= 0 ? list_buf(v, pord).lperm[s] : a;     // pl after a rejection: its own order
    const int N = v.Ntot;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        v.x[d * N + a] = v.xl[d * N + s];
        v.p[d * N + ap] = v.pl[d * N + s];
    }
    if (saved) {
        const int as = list_buf(v, *list_start(v, ch)).lperm[s];
#pragma unroll
        for (int d = 0; d < 3; d++) v.x_saved[d * N + as] = v.xl_saved[d * N + s];
    }
}

hipError_t launch_list_import(const DevView& v, hipStream_t s)
{
    hipLaunchKernelGGL(k_limport, dim3((v.Ntot + 255) / 256), dim3(256), 0, s, v);
    return hipGetLastError();
}

hipError_t launch_list_export(const DevView& v, bool saved, hipStream_t s)
{
    hipLaunchKernelGGL(k_lexport, dim3((v.Ntot + 255) / 256), dim3(256), 0, s, v, saved ? 1 : 0);
    return hipGetLastError();
}

// ---- launchers -----------------------------------------------------------------------------------
// sweep shape: 4, 8, 16 lanes per slot with first batches of 16, 14, 7 entries, or 810 / 811 / 812 / 816
// (8 lanes with batches of 10 / 11 / 12 / 16)
template <int MODE, bool W16>
static hipError_t sweep_lp_fmt(const DevView& v, int shape, int q, double kick_dt, double drift_dt, int ck, hipStream_t s)
{
    const dim3 g(v.ntiles);
    switch (shape) {
    case 4: hipLaunchKernelGGL((k_lsweep<MODE, 4, 16, W16>), g, dim3(kTile * 4), 0, s, v, kick_dt, drift_dt, q, ck); break;
    case 8: hipLaunchKernelGGL((k_lsweep<MODE, 8, 14, W16>), g, dim3(kTile * 8), 0, s, v, kick_dt, drift_dt, q, ck); break;
    case 16: hipLaunchKernelGGL((k_lsweep<MODE, 16, 7, W16>), g, dim3(kTile * 16), 0, s, v, kick_dt, drift_dt, q, ck); break;
    case 810: hipLaunchKernelGGL((k_lsweep<MODE, 8, 10, W16>), g, dim3(kTile * 8), 0, s, v, kick_dt, drift_dt, q, ck); break;
    case 811: hipLaunchKernelGGL((k_lsweep<MODE, 8, 11, W16>), g, dim3(kTile * 8), 0, s, v, kick_dt, drift_dt, q, ck); break;
    case 812: hipLaunchKernelGGL((k_lsweep<MODE, 8, 12, W16>), g, dim3(kTile * 8), 0, s, v, kick_dt, drift_dt, q, ck); break;
    case 816: hipLaunchKernelGGL((k_lsweep<MODE, 8, 16, W16>), g, dim3(kTile * 8), 0, s, v, kick_dt, drift_dt, q, ck); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int MODE>
static hipError_t sweep_lp(const DevView& v, int shape, int q, double kick_dt, double drift_dt, int ck, hipStream_t s)
{
    static_assert(sweep_span(4) == 4 * 16 && sweep_span(8) == 8 * 14 && sweep_span(16) == 16 * 7 &&
                  sweep_span(810) == 8 * 10 && sweep_span(811) == 8 * 11 && sweep_span(812) == 8 * 12 &&
                  sweep_span(816) == 8 * 16, "shape spans");
    if (sweep_span(shape) > v.lspan) return hipErrorInvalidValue;    // rows not padded that far
    return v.entry16 ? sweep_lp_fmt<MODE, true>(v, shape, q, kick_dt, drift_dt, ck, s)
                     : sweep_lp_fmt<MODE, false>(v, shape, q, kick_dt, drift_dt, ck, s);
}

hipError_t launch_list_sweep(const DevView& v, int mode, int q, double kick_dt, double drift_dt, bool checked,
                             hipStream_t s)
{
    const int lp = v.lanes;
    const int ck = checked ? 1 : 0;
    switch (mode) {
    case LJ_FORCE | LJ_KICK: return sweep_lp<LJ_FORCE | LJ_KICK>(v, lp, q, kick_dt, drift_dt, ck, s);
    case LJ_FORCE | LJ_KICK | LJ_STOREF: return sweep_lp<LJ_FORCE | LJ_KICK | LJ_STOREF>(v, lp, q, kick_dt, drift_dt, ck, s);
    case LJ_FORCE | LJ_KICK | LJ_DRIFT: return sweep_lp<LJ_FORCE | LJ_KICK | LJ_DRIFT>(v, lp, q, kick_dt, drift_dt, ck, s);
    case LJ_FORCE | LJ_KICK | LJ_STOREF | LJ_DRIFT:
        return sweep_lp<LJ_FORCE | LJ_KICK | LJ_STOREF | LJ_DRIFT>(v, lp, q, kick_dt, drift_dt, ck, s);
    case LJ_FORCE | LJ_STOREF: return sweep_lp<LJ_FORCE | LJ_STOREF>(v, lp, q, kick_dt, drift_dt, ck, s);
    case LJ_FORCE | LJ_KICK | LJ_REL64: return sweep_lp<LJ_FORCE | LJ_KICK | LJ_REL64>(v, lp, q, kick_dt, drift_dt, ck, s);
    default: return hipErrorInvalidValue;
    }
}

// the surrogate sweep runs 8 lanes per slot whatever the LJ sweep's tuning
template <bool W16>
static hipError_t lsurr_sweep_fmt(const DevView& v, const float* g, int mode, int q, double kick_dt, double drift_dt,
                                  int ck, hipStream_t s)
{
    const dim3 grid(v.ntiles), block(kTile * 8);
    switch (mode) {
    case LJ_FORCE | LJ_KICK:
        hipLaunchKernelGGL((k_lsurr_sweep<LJ_FORCE | LJ_KICK, 8, W16>), grid, block, 0, s, v, g, kick_dt, drift_dt, q, ck);
        break;
    case LJ_FORCE | LJ_KICK | LJ_STOREF:
        hipLaunchKernelGGL((k_lsurr_sweep<LJ_FORCE | LJ_KICK | LJ_STOREF, 8, W16>), grid, block, 0, s, v, g, kick_dt, drift_dt,
                           q, ck);
        break;
    case LJ_FORCE | LJ_KICK | LJ_DRIFT:
        hipLaunchKernelGGL((k_lsurr_sweep<LJ_FORCE | LJ_KICK | LJ_DRIFT, 8, W16>), grid, block, 0, s, v, g, kick_dt, drift_dt,
                           q, ck);
        break;
    case LJ_FORCE | LJ_KICK | LJ_STOREF | LJ_DRIFT:
        hipLaunchKernelGGL((k_lsurr_sweep<LJ_FORCE | LJ_KICK | LJ_STOREF | LJ_DRIFT, 8, W16>), grid, block, 0, s, v, g, kick_dt,
                           drift_dt, q, ck);
        break;
    case LJ_FORCE | LJ_STOREF:
        hipLaunchKernelGGL((k_lsurr_sweep<LJ_FORCE | LJ_STOREF, 8, W16>), grid, block, 0, s, v, g, kick_dt, drift_dt, q, ck);
        break;
    case LJ_FORCE | LJ_KICK | LJ_REL64:
        hipLaunchKernelGGL((k_lsurr_sweep<LJ_FORCE | LJ_KICK | LJ_REL64, 8, W16>), grid, block, 0, s, v, g, kick_dt, drift_dt,
                           q, ck);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_lsurr_sweep(const DevView& v, const float* g, int mode, int q, double kick_dt, double drift_dt,
                              bool checked, hipStream_t s)
{
    const int ck = checked ? 1 : 0;
    return v.entry16 ? lsurr_sweep_fmt<true>(v, g, mode, q, kick_dt, drift_dt, ck, s)
                     : lsurr_sweep_fmt<false>(v, g, mode, q, kick_dt, drift_dt, ck, s);
}

hipError_t launch_list_drift(const DevView& v, int q, double dt, hipStream_t s)
{
    hipLaunchKernelGGL(k_ldrift, dim3((v.Ntot + 255) / 256), dim3(256), 0, s, v, dt, q);
    return hipGetLastError();
}

hipError_t launch_list_energy(const DevView& v, int q, bool rel64, hipStream_t s)
{
    if (rel64) {
        hipLaunchKernelGGL(k_lrel64, dim3((v.Ntot + 255) / 256), dim3(256), 0, s, v);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (v.entry16) hipLaunchKernelGGL(k_lenergy<true>, dim3(v.ntiles), dim3(256), 0, s, v, q);
    else hipLaunchKernelGGL(k_lenergy<false>, dim3(v.ntiles), dim3(256), 0, s, v, q);
    return hipGetLastError();
}

}  // namespace chmc
