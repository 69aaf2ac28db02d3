// Cell binning: the replacement of parallel_binning (code/modules/identical_particles_binning.cpp:97-110,
// code/modules/binning.cpp:15-54) and of the discarded quicksort (code/modules/quick_sort.cpp:509-547).
//
// The reference's binning is, in effect, a stable counting sort of the previous permute_vector by
// cell id, computed with two O(N * cells) sweeps.  Here it is O(N), and one of two paths:
//
// general (first binning, or in-cell order by atom id):
//   k_count          : optional fused drift + wrap (identical_particles.cpp:651-669), cell id
//                      (particles.hpp:108-113), arrival slot from an atomic per-cell counter, and the new
//                      cell at the atom's previous sorted slot (used by the local path)
//   k_scan           : one workgroup scans the counts into the next offsets (and clears them)
//   k_scatter        : atom -> (cell offset + arrival slot), with its sort key
//   k_cell_rank      : 16 lanes per cell rank the cell's entries by key -> the stable order
// local (every later binning):
//   count + scan     : as above; the count is done by k_count, by the heat bath, or inside the FP32 force
//                      sweep right after the kick and the following drift of each atom
//   k_collect        : one wave per new cell gathers its atoms from the (deduplicated, index-sorted)
//                      previous neighbour cells in previous-slot order, which IS the stable order, so no
//                      sort is needed.  Atoms move far less than a cell per step; an atom that left the
//                      previous 27-cell neighbourhood raises a flag and the wave then scans all of its
//                      chain's atoms instead (same result, O(N) per cell).
// Both write the next perm/offs buffers, invperm, cell_of and the FP32 cell-relative coordinates; the
// host then flips the parity argument that selects the current buffers.
#include "chmc_internal.hpp"
#include "bin_device.hpp"

#pragma clang fp contract(off)   // drift and cell index must match the FP64 oracle bit for bit

namespace chmc {

__global__ __launch_bounds__(256) void k_count(DevView v, double drift_dt, int do_drift)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < v.Ntot) {
        const ChainDev& C = v.chains[v.atom_chain[a]];
        const int N = v.Ntot;
        const double coeff_x = C.coeff_x;              // compute_coeff_position, identical_particles.cpp:646-648
        double x3[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double xv = v.x[d * N + a];
            if (do_drift) {
                xv += drift_dt * coeff_x * v.p[d * N + a];
                xv = lp_round(xv, v.lp_scale);
                xv = wrap_box(xv, C.L[d]);
                v.x[d * N + a] = xv;
            }
            x3[d] = xv;
        }
        count_atom(v, C, a, v.invperm[a], v.cell_of[a], x3);
    }
}

__global__ __launch_bounds__(1024) void k_scan(DevView v)
{
    scan_counts_block(v);
}

__global__ __launch_bounds__(256) void k_drift(DevView v, double dt)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= v.Ntot) return;
    const ChainDev& C = v.chains[v.atom_chain[a]];
    const int N = v.Ntot;
    double coeff_x = C.coeff_x;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double xv = v.x[d * N + a];
        xv += dt * coeff_x * v.p[d * N + a];
        xv = lp_round(xv, v.lp_scale);
        xv = wrap_box(xv, C.L[d]);
        v.x[d * N + a] = xv;
    }
}

__global__ __launch_bounds__(256) void k_scatter(DevView v, int atom_id_order)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= v.Ntot) return;
    const int dest = offs_next(v)[v.cell_of[a]] + v.slot[a];
    v.skey[dest] = atom_id_order ? a : v.invperm[a];
    v.satom[dest] = a;
}

__device__ __forceinline__ void place_atom(const DevView& v, int* perm_n, int pos, int atom, int c,
                                           double o0, double o1, double o2)
{
    const int N = v.Ntot;
    perm_n[pos] = atom;
    v.invperm[atom] = pos;
    v.cell_of[atom] = c;
    float4 r;
    r.x = (float)(v.x[atom] - o0);
    r.y = (float)(v.x[N + atom] - o1);
    r.z = (float)(v.x[2 * N + atom] - o2);
    r.w = __int_as_float(atom);
    v.xs[pos] = r;
}

__device__ __forceinline__ void cell_origin(const ChainDev& C, int c, int b[3], double o[3])
{
    const int lc = c - C.cell_off;
    b[2] = lc / (C.nbin[0] * C.nbin[1]);
    b[1] = (lc - b[2] * C.nbin[0] * C.nbin[1]) / C.nbin[0];
    b[0] = lc - C.nbin[0] * (b[1] + b[2] * C.nbin[1]);
    for (int d = 0; d < 3; d++) o[d] = b[d] * C.sizebin[d];
}

// the flag of the placement after this one (target = the buffer current now) starts clear; no workgroup
// of this kernel reads it, and the next count phase runs after this kernel
__device__ __forceinline__ void clear_next_flag(DevView& v)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) v.bflag[v.par] = 0;
}

// general path: 16 lanes per cell, each entry's position is its rank among the cell's (unique) keys
__global__ __launch_bounds__(256) void k_cell_rank(DevView v)
{
    const int c = blockIdx.x * 16 + (threadIdx.x >> 4);
    const int sub = threadIdx.x & 15;
    if (c < v.Ctot) {
        const int* offs = offs_next(v);
        int* perm_n = perm_next(v);
        const int beg = offs[c];
        const int n = offs[c + 1] - beg;
        const ChainDev& C = v.chains[v.cell_chain[c]];
        int b[3];
        double o[3];
        cell_origin(C, c, b, o);
        for (int e = sub; e < n; e += 16) {
            const int key = v.skey[beg + e];
            const int atom = v.satom[beg + e];
            int rank = 0;
            for (int q = 0; q < n; q++) rank += (v.skey[beg + q] < key) ? 1 : 0;
            place_atom(v, perm_n, beg + rank, atom, c, o[0], o[1], o[2]);
        }
    }
    clear_next_flag(v);
}

// local path: one wave per new cell, 4 cells per workgroup
__global__ __launch_bounds__(256) void k_collect(DevView v)
{
    __shared__ int seg_beg[4][27], seg_pref[4][28];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int c = blockIdx.x * 4 + wave;
    if (c < v.Ctot) {
        const int* offs_o = offs_cur(v);
        const int* offs_n = offs_next(v);
        const int* perm_o = perm_cur(v);
        int* perm_n = perm_next(v);
        const ChainDev& C = v.chains[v.cell_chain[c]];
        int b[3];
        double o[3];
        cell_origin(C, c, b, o);
        const int dst = offs_n[c];
        const int want = offs_n[c + 1] - dst;
        if (want > 0) {
            if (v.bflag[v.par ^ 1] == 0) {
                // deduplicated neighbour cells (nbin < 3 folds the stencil), sorted by index so that the
                // concatenated previous ranges are in increasing previous-slot order
                int jc = -1;
                if (lane < 27) {
                    const int dd[3] = { lane % 3 - 1, (lane / 3) % 3 - 1, lane / 9 - 1 };
                    bool dup = false;
                    int j3[3];
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        const int n = C.nbin[d];
                        dup |= (n == 1 && dd[d] != 0) || (n == 2 && dd[d] == -1);
                        j3[d] = (b[d] + dd[d] + n) % n;
                    }
                    if (!dup) jc = C.cell_off + j3[0] + C.nbin[0] * (j3[1] + j3[2] * C.nbin[1]);
                }
                // rank of jc among the valid entries (cell indices are distinct after deduplication)
                int rank = 0, nvalid = 0;
                for (int q = 0; q < 27; q++) {
                    const int other = __shfl(jc, q, 64);
                    nvalid += other >= 0 ? 1 : 0;
                    rank += (other >= 0 && other < jc) ? 1 : 0;
                }
                if (jc >= 0) {
                    seg_beg[wave][rank] = offs_o[jc];
                    seg_pref[wave][rank + 1] = offs_o[jc + 1] - offs_o[jc];   // count, scanned below
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                int incl = lane < nvalid ? seg_pref[wave][lane + 1] : 0;
#pragma unroll
                for (int o = 1; o < 32; o <<= 1) {
                    const int t = __shfl_up(incl, o, 64);
                    if (lane >= o) incl += t;
                }
                __builtin_amdgcn_wave_barrier();
                if (lane < nvalid) seg_pref[wave][lane + 1] = incl;
                if (lane == 0) seg_pref[wave][0] = 0;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                const int total = seg_pref[wave][nvalid];
                int placed = 0;
                for (int t0 = 0; t0 < total; t0 += 64) {
                    const int t = t0 + lane;
                    bool hit = false;
                    int k = 0;
                    if (t < total) {
                        int s = 0;
                        for (int step = 16; step > 0; step >>= 1)
                            if (s + step < nvalid && seg_pref[wave][s + step] <= t) s += step;
                        k = seg_beg[wave][s] + (t - seg_pref[wave][s]);
                        hit = v.newcell[k] == c;
                    }
                    const unsigned long long m = __ballot(hit);
                    if (hit) {
                        const int r = __popcll(m & ((1ull << lane) - 1ull));
                        place_atom(v, perm_n, dst + placed + r, perm_o[k], c, o[0], o[1], o[2]);
                    }
                    placed += __popcll(m);
                }
            } else {
                // an atom moved beyond the previous neighbourhood: scan the whole chain in previous order
                int placed = 0;
                for (int t0 = 0; t0 < C.natoms; t0 += 64) {
                    const int k = C.atom_off + t0 + lane;
                    const bool hit = (t0 + lane < C.natoms) && v.newcell[k] == c;
                    const unsigned long long m = __ballot(hit);
                    if (hit) {
                        const int r = __popcll(m & ((1ull << lane) - 1ull));
                        place_atom(v, perm_n, dst + placed + r, perm_o[k], c, o[0], o[1], o[2]);
                    }
                    placed += __popcll(m);
                }
            }
        }
    }
    clear_next_flag(v);
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

hipError_t launch_count(const DevView& v, double drift_dt, hipStream_t s)
{
    hipLaunchKernelGGL(k_count, dim3((v.Ntot + 255) / 256), dim3(256), 0, s, v, drift_dt, drift_dt != 0.0 ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_scan(const DevView& v, hipStream_t s)
{
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, v);
    return hipGetLastError();
}

hipError_t launch_place(const DevView& v, bool local, bool atom_id_order, hipStream_t s)
{
    if (local) {
        hipLaunchKernelGGL(k_collect, dim3((v.Ctot + 3) / 4), dim3(256), 0, s, v);
    } else {
        hipLaunchKernelGGL(k_scatter, dim3((v.Ntot + 255) / 256), dim3(256), 0, s, v, atom_id_order ? 1 : 0);
        hipLaunchKernelGGL(k_cell_rank, dim3((v.Ctot + 15) / 16), dim3(256), 0, s, v);
    }
    return hipGetLastError();
}

hipError_t launch_drift(const DevView& v, double dt, hipStream_t s)
{
    hipLaunchKernelGGL(k_drift, dim3((v.Ntot + 255) / 256), dim3(256), 0, s, v, dt);
    return hipGetLastError();
}

}  // namespace chmc
