// All-pairs kernels.
//  k_rdf: radial distribution histogram, the replacement of Tag_RDF (code/modules/identical_particles.cpp:739-776).
//    The reference runs one team per bin over all N^2 ordered pairs (O(bins * N^2)); here each pair is
//    visited once: 256 x 256 tiles of the upper triangle, FP64 per-dimension minimum image exactly as
//    the reference (min of (d-L)^2, d^2, (d+L)^2), bin = (int)(r / size_bin), and an LDS histogram
//    private to each wavefront, merged into per-block sums and then into global 64-bit counters.
//    Unordered pairs are counted twice, which equals the reference's ordered count exactly.
//  k_open_potential: code-carsten/potential_energy.hh:6-15, sum_{i<j} a/r^n - b/r^6 with open
//    boundaries, FP64, per-block partial sums (summed in block order on the host).
#include "chmc_internal.hpp"

#pragma clang fp contract(off)

namespace chmc {

__global__ __launch_bounds__(256) void k_rdf(DevView v, int chain, int nbins, double size_bin,
                                             unsigned long long* __restrict__ counts)
{
    extern __shared__ unsigned int hist[];     // [4 waves][nbins]
    __shared__ double jx[256], jy[256], jz[256];
    const int I = blockIdx.y, J = blockIdx.x;
    if (J < I) return;
    const ChainDev& C = v.chains[chain];
    const int n = C.natoms, off = C.atom_off, N = v.Ntot;
    const int tid = threadIdx.x, wave = tid >> 6;
    for (int b = tid; b < 4 * nbins; b += 256) hist[b] = 0u;
    const int i = I * 256 + tid;
    const int j0 = J * 256 + tid;
    double xi = 0, yi = 0, zi = 0;
    if (i < n) { xi = v.x[off + i]; yi = v.x[N + off + i]; zi = v.x[2 * N + off + i]; }
    if (j0 < n) { jx[tid] = v.x[off + j0]; jy[tid] = v.x[N + off + j0]; jz[tid] = v.x[2 * N + off + j0]; }
    __syncthreads();
    const double L0 = C.L[0], L1 = C.L[1], L2 = C.L[2];
    if (i < n) {
        const int jend = min(256, n - J * 256);
        const int jstart = (I == J) ? tid + 1 : 0;
        unsigned int* h = &hist[wave * nbins];
        for (int jj = jstart; jj < jend; jj++) {
            double r = 0.0, dd, ri, si;
            dd = jx[jj] - xi;
            ri = (dd - L0) * (dd - L0); si = dd * dd; if (si < ri) ri = si;
            si = (dd + L0) * (dd + L0); if (si < ri) ri = si; r += ri;
            dd = jy[jj] - yi;
            ri = (dd - L1) * (dd - L1); si = dd * dd; if (si < ri) ri = si;
            si = (dd + L1) * (dd + L1); if (si < ri) ri = si; r += ri;
            dd = jz[jj] - zi;
            ri = (dd - L2) * (dd - L2); si = dd * dd; if (si < ri) ri = si;
            si = (dd + L2) * (dd + L2); if (si < ri) ri = si; r += ri;
            r = sqrt(r);
            const int b = (int)(r / size_bin);
            if (b < nbins) atomicAdd(&h[b], 1u);
        }
    }
    __syncthreads();
    for (int b = tid; b < nbins; b += 256) {
        const unsigned long long s = (unsigned long long)hist[b] + hist[nbins + b] + hist[2 * nbins + b] + hist[3 * nbins + b];
        if (s) atomicAdd(&counts[b], 2ull * s);
    }
}

// coordinate d of atom i is xyz[i * si + d * sd]: SoA (si = 1, sd = n) or AoS double[3] (si = 3, sd = 1)
__global__ __launch_bounds__(256) void k_open_potential(const double* __restrict__ xyz, int n, int si, int sd,
                                                        int nexp, double a, double b, double* __restrict__ partials)
{
    __shared__ double jx[256], jy[256], jz[256];
    __shared__ double wpart[4];
    const int I = blockIdx.y, J = blockIdx.x;
    const int tid = threadIdx.x;
    const int pid = I * gridDim.x + J;
    if (J < I) {
        if (tid == 0) partials[pid] = 0.0;
        return;
    }
    const int i = I * 256 + tid, j0 = J * 256 + tid;
    double xi = 0, yi = 0, zi = 0;
    if (i < n) { xi = xyz[i * si]; yi = xyz[i * si + sd]; zi = xyz[i * si + 2 * sd]; }
    if (j0 < n) { jx[tid] = xyz[j0 * si]; jy[tid] = xyz[j0 * si + sd]; jz[tid] = xyz[j0 * si + 2 * sd]; }
    __syncthreads();
    double s = 0.0;
    if (i < n) {
        const int jend = min(256, n - J * 256);
        const int jstart = (I == J) ? tid + 1 : 0;
        for (int jj = jstart; jj < jend; jj++) {
            const double d0 = xi - jx[jj], d1 = yi - jy[jj], d2 = zi - jz[jj];
            const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            s += a / pow(r, (double)nexp) - b / pow(r, 6.0);     // lennardjones.hh:5-9
        }
    }
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) s += __shfl_xor(s, m, 64);
    if ((tid & 63) == 0) wpart[tid >> 6] = s;
    __syncthreads();
    if (tid == 0) partials[pid] = ((wpart[0] + wpart[1]) + wpart[2]) + wpart[3];
}

hipError_t launch_rdf(const DevView& v, int chain, int natoms, int nbins, double size_bin,
                      unsigned long long* counts, hipStream_t s)
{
    const int T = (natoms + 255) / 256;
    if (T == 0) return hipSuccess;
    hipLaunchKernelGGL(k_rdf, dim3(T, T), dim3(256), sizeof(unsigned int) * 4 * nbins, s, v, chain, nbins,
                       size_bin, counts);
    return hipGetLastError();
}

hipError_t launch_open_potential(const double* xyz_soa, int n, int nexp, double a, double b, double* partials,
                                 int* nparts, hipStream_t s)
{
    const int T = (n + 255) / 256;
    *nparts = T * T;
    if (T == 0) return hipSuccess;
    hipLaunchKernelGGL(k_open_potential, dim3(T, T), dim3(256), 0, s, xyz_soa, n, 1, n, nexp, a, b, partials);
    return hipGetLastError();
}

// the partials of k_open_potential summed by one thread in block order (the order of the host sum of
// chmc_potential_energy_open, so both entry points return the same bits); the lower-triangle blocks
// hold +0.0 and are skipped, which leaves every intermediate sum unchanged
__global__ void k_open_sum(const double* __restrict__ partials, int T, double* __restrict__ result)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double r = 0.0;
    for (int I = 0; I < T; I++)
        for (int J = I; J < T; J++) r += partials[I * T + J];
    *result = r;
}

hipError_t launch_open_potential_aos(const double* xyz_aos, int n, int nexp, double a, double b, double* partials,
                                     double* result, hipStream_t s)
{
    const int T = (n + 255) / 256;
    if (T > 0) {
        hipLaunchKernelGGL(k_open_potential, dim3(T, T), dim3(256), 0, s, xyz_aos, n, 3, 1, nexp, a, b, partials);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_open_sum, dim3(1), dim3(64), 0, s, partials, T, result);
    return hipGetLastError();
}

}  // namespace chmc
