// Cell binning: the replacement of parallel_binning (code/modules/identical_particles_binning.cpp:97-110,
// code/modules/binning.cpp:15-54) and of the discarded quicksort (code/modules/quick_sort.cpp:509-547).
//
// The reference's binning is, in effect, a stable counting sort of the previous permute_vector by
// cell id, computed with two O(N * cells) sweeps.  Here it is O(N) in four launches:
//   k_drift_count : optional fused drift + wrap (identical_particles.cpp:651-669), cell id
//                   (particles.hpp:108-113), arrival slot from an atomic per-cell counter
//   k_scan        : one-workgroup exclusive scan of the counts; clears the counts for the next call
//   k_scatter     : atom -> (cell offset + arrival slot), with its sort key
//   k_cell_rank   : 16 lanes per cell rank the cell's entries by key, which restores the stable
//                   order the reference produces, and write perm / invperm / the FP32 sorted coordinates
// Sort key = position in the previous permutation (reference semantics, bit-exact permute_vector) or
// the atom id (order independent of history, so forces are a function of positions alone).
#include "chmc_internal.hpp"

#pragma clang fp contract(off)   // drift and cell index must match the FP64 oracle bit for bit

namespace chmc {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__global__ __launch_bounds__(256) void k_drift_count(DevView v, double drift_dt, int do_drift)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= v.Ntot) return;
    const ChainDev& C = v.chains[v.atom_chain[a]];
    const int N = v.Ntot;
    int b[3];
    double coeff_x = C.beta / C.mass;                  // compute_coeff_position, identical_particles.cpp:646-648
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double xv = v.x[d * N + a];
        if (do_drift) {
            xv += drift_dt * coeff_x * v.p[d * N + a];
            xv -= C.L[d] * floor(xv / C.L[d]);
            v.x[d * N + a] = xv;
        }
        b[d] = clampi((int)floor(xv / C.sizebin[d]), 0, C.nbin[d] - 1);
    }
    const int cell = C.cell_off + b[0] + C.nbin[0] * (b[1] + b[2] * C.nbin[1]);
    v.cell_of[a] = cell;
    v.slot[a] = atomicAdd(&v.count[cell], 1);
}

__global__ __launch_bounds__(256) void k_drift(DevView v, double dt)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= v.Ntot) return;
    const ChainDev& C = v.chains[v.atom_chain[a]];
    const int N = v.Ntot;
    double coeff_x = C.beta / C.mass;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double xv = v.x[d * N + a];
        xv += dt * coeff_x * v.p[d * N + a];
        xv -= C.L[d] * floor(xv / C.L[d]);
        v.x[d * N + a] = xv;
    }
}

// Exclusive scan over all cells of all chains by one workgroup (cells <= a few 1e5).
__global__ __launch_bounds__(1024) void k_scan(int* __restrict__ count, int* __restrict__ offs, int Ctot)
{
    __shared__ int wsum[16];
    const int tid = threadIdx.x;
    const int per = (Ctot + 1023) / 1024;
    const int beg = min(tid * per, Ctot), end = min(beg + per, Ctot);
    int s = 0;
    for (int c = beg; c < end; c++) s += count[c];
    // inclusive wave scan
    const int lane = tid & 63, wid = tid >> 6;
    int incl = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    if (wid == 0) {
        int w = lane < 16 ? wsum[lane] : 0;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            int t = __shfl_up(w, o, 64);
            if (lane >= o) w += t;
        }
        if (lane < 16) wsum[lane] = w;
    }
    __syncthreads();
    int run = incl - s + (wid > 0 ? wsum[wid - 1] : 0);
    for (int c = beg; c < end; c++) {
        int n = count[c];
        offs[c] = run;
        run += n;
        count[c] = 0;
    }
    if (tid == 1023) offs[Ctot] = wsum[15];
}

__global__ __launch_bounds__(256) void k_scatter(DevView v, int atom_id_order)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= v.Ntot) return;
    const int dest = v.offs[v.cell_of[a]] + v.slot[a];
    v.skey[dest] = atom_id_order ? a : v.invperm[a];
    v.satom[dest] = a;
}

// 16 lanes per cell: each entry's final position is its rank among the cell's keys (keys are unique).
__global__ __launch_bounds__(256) void k_cell_rank(DevView v)
{
    const int c = blockIdx.x * 16 + (threadIdx.x >> 4);
    const int sub = threadIdx.x & 15;
    if (c >= v.Ctot) return;
    const int beg = v.offs[c];
    const int n = v.offs[c + 1] - beg;
    const ChainDev& C = v.chains[v.cell_chain[c]];
    const int lc = c - C.cell_off;
    const int bz = lc / (C.nbin[0] * C.nbin[1]);
    const int by = (lc - bz * C.nbin[0] * C.nbin[1]) / C.nbin[0];
    const int bx = lc - C.nbin[0] * (by + bz * C.nbin[1]);
    const double o0 = bx * C.sizebin[0], o1 = by * C.sizebin[1], o2 = bz * C.sizebin[2];
    const int N = v.Ntot;
    for (int e = sub; e < n; e += 16) {
        const int key = v.skey[beg + e];
        const int atom = v.satom[beg + e];
        int rank = 0;
        for (int q = 0; q < n; q++) rank += (v.skey[beg + q] < key) ? 1 : 0;
        const int pos = beg + rank;
        v.perm[pos] = atom;
        v.invperm[atom] = pos;
        float4 r;
        r.x = (float)(v.x[atom] - o0);
        r.y = (float)(v.x[N + atom] - o1);
        r.z = (float)(v.x[2 * N + atom] - o2);
        r.w = __int_as_float(atom);
        v.xs[pos] = r;
    }
}

__global__ void k_fill_int(int* p, int value, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

hipError_t launch_fill_int(int* p, int value, int n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_int, dim3((n + 255) / 256), dim3(256), 0, s, p, value, n);
    return hipGetLastError();
}

hipError_t launch_bin(const DevView& v, double drift_dt, bool atom_id_order, hipStream_t s)
{
    const int nb = (v.Ntot + 255) / 256;
    hipLaunchKernelGGL(k_drift_count, dim3(nb), dim3(256), 0, s, v, drift_dt, drift_dt != 0.0 ? 1 : 0);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, v.count, v.offs, v.Ctot);
    hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(256), 0, s, v, atom_id_order ? 1 : 0);
    hipLaunchKernelGGL(k_cell_rank, dim3((v.Ctot + 15) / 16), dim3(256), 0, s, v);
    return hipGetLastError();
}

hipError_t launch_drift(const DevView& v, double dt, hipStream_t s)
{
    hipLaunchKernelGGL(k_drift, dim3((v.Ntot + 255) / 256), dim3(256), 0, s, v, dt);
    return hipGetLastError();
}

}  // namespace chmc
