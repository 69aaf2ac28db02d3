// Device helpers shared by the kernels that bin atoms (k_bin.hip, and the fused drift + count of the
// FP32 force sweep and of the heat bath): cell index with the reference's floor(x / sizebin)
// (particles.hpp:108-113) clamped to the grid, the 27-cell neighbourhood test, the count phase of one
// atom and the one-workgroup exclusive scan of the per-cell counts.
#pragma once
#include "chmc_internal.hpp"

namespace chmc {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ inline void scan_counts_block(const DevView& v)
{
    __shared__ int wsum[16];
    const int nt = blockDim.x, tid = threadIdx.x;
    const int Ctot = v.Ctot;
    int* count = v.count;
    int* offs = offs_next(v);
    const int per = (Ctot + nt - 1) / nt;
    const int beg = min(tid * per, Ctot), end = min(beg + per, Ctot);
    int s = 0;
    for (int c = beg; c < end; c++) s += count[c];
    const int lane = tid & 63, wid = tid >> 6, nw = nt >> 6;
    int incl = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    if (wid == 0) {
        int w = lane < nw ? wsum[lane] : 0;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            int t = __shfl_up(w, o, 64);
            if (lane >= o) w += t;
        }
        if (lane < nw) wsum[lane] = w;
    }
    __syncthreads();
    int run = incl - s + (wid > 0 ? wsum[wid - 1] : 0);
    for (int c = beg; c < end; c++) {
        const int n = count[c];
        offs[c] = run;
        run += n;
        count[c] = 0;
    }
    if (tid == nt - 1) offs[Ctot] = wsum[nw - 1];
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
    v.slot[a] = atomicAdd(&v.count[cell], 1);
}

}  // namespace chmc
