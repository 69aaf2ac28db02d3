// Device helpers shared by the kernels that bin atoms (k_bin.hip, and the fused drift + count of the
// FP32 force sweep and of the heat bath): cell index with the reference's floor(x / sizebin)
// (particles.hpp:108-113) clamped to the grid, the 27-cell neighbourhood test, the count phase of one
// atom and the one-workgroup exclusive scan of the per-cell counts.
#pragma once
#include "chmc_internal.hpp"

namespace chmc {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Exclusive scan of the per-cell counts into the next offsets (and the counts cleared), by one
// workgroup of 1024 threads: rounds of 8192 counts are loaded into LDS with coalesced reads, each thread
// scans 8 contiguous ones, a block scan over the threads and a running carry join the rounds, and the
// offsets go back with coalesced writes.
__device__ inline void scan_counts_block(const DevView& v)
{
    constexpr int T = 1024, PT = 8;
    __shared__ int buf[T * PT];
    __shared__ int wsum[16];
    __shared__ int carry;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int n = v.Ctot;
    int* count = v.count;
    int* offs = offs_next(v);
    if (tid == 0) carry = 0;
    for (int r0 = 0; r0 < n; r0 += T * PT) {
#pragma unroll
        for (int k = 0; k < PT; k++) {
            const int c = r0 + k * T + tid;
            buf[k * T + tid] = c < n ? count[c] : 0;
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
            if (c < n) {
                offs[c] = buf[k * T + tid];
                count[c] = 0;
            }
        }
        __syncthreads();                      // buf free, carry final
    }
    if (tid == 0) offs[n] = carry;
}

// cell of a position, and whether it lies outside the 27-cell neighbourhood of cell `old` (periodic)
__device__ __forceinline__ int cell_index(const ChainDev& C, const double x3[3], int b[3])
{
#pragma unroll
    for (int d = 0; d < 3; d++) b[d] = clampi((int)floor(x3[d] / C.sizebin[d]), 0, C.nbin[d] - 1);
    return C.cell_off + b[0] + C.nbin[0] * (b[1] + b[2] * C.nbin[1]);
}

__device__ __forceinline__ bool far_from(const ChainDev& C, const int b[3], int old_global)
{
    const int lc = old_global - C.cell_off;
    const int oz = lc / (C.nbin[0] * C.nbin[1]);
    const int oy = (lc - oz * C.nbin[0] * C.nbin[1]) / C.nbin[0];
    const int ox = lc - C.nbin[0] * (oy + oz * C.nbin[1]);
    const int o3[3] = { ox, oy, oz };
    bool far = false;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int n = C.nbin[d];
        const int dd = ((b[d] - o3[d]) % n + n) % n;
        far |= !(dd == 0 || dd == 1 || dd == n - 1);
    }
    return far;
}

// count phase of one atom at its current position (any drift already applied): new cell at the atom's
// previous sorted slot (local placement), the cell and arrival slot per atom (general placement), and the
// flag raised when the atom left the 27-cell neighbourhood of its previous cell
__device__ __forceinline__ void count_atom(const DevView& v, const ChainDev& C, int a, int prev_slot,
                                           int prev_cell, const double x3[3])
{
    int b[3];
    const int cell = cell_index(C, x3, b);
    v.newcell[prev_slot] = cell;
    if (far_from(C, b, prev_cell)) atomicOr(&v.bflag[v.par ^ 1], 1);
    v.cell_of[a] = cell;
    // one atomic per distinct cell among the wave's active lanes (lanes that run in slot order mostly
    // share a few cells); after kGroups groups the remaining lanes take an atomic each.  The arrival slot
    // only has to be unique within the cell: the general placement ranks the cell's entries by key.
    constexpr int kGroups = 8;
    const int lane = threadIdx.x & 63;
    unsigned long long left = __ballot(1);
    int slot = -1;
    for (int g = 0; g < kGroups && left; g++) {
        const int l = __ffsll((long long)left) - 1;
        const int cl = __shfl(cell, l, 64);
        const unsigned long long m = __ballot(cell == cl) & left;
        int base = 0;
        if (lane == l) base = atomicAdd(&v.count[cell], __popcll(m));
        base = __shfl(base, l, 64);
        if (cell == cl) slot = base + __popcll(m & ((1ull << lane) - 1ull));
        left &= ~m;
    }
    if (slot < 0) slot = atomicAdd(&v.count[cell], 1);
    v.slot[a] = slot;
}

}  // namespace chmc
