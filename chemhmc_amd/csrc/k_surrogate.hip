// Learned surrogate of the LJ force field (no reference counterpart): per-atom MLP on Gaussian radial
// descriptors, used to drive HMC trajectories while the exact FP64 LJ energy judges the Metropolis step.
//
//   descriptor  G_k(i) = sum_j exp(-(r_ij - mu_k)^2 / (2 w^2)) fc(r_ij), fc = (cos(pi r / rc) + 1) / 2,
//               32 centres mu_k uniform on [0.8, 2.5], w = 0.1, neighbours j within rc (27-cell stencil)
//   MLP         z = (G - shift) * scale;  h1 = silu(W1 z + b1);  h2 = silu(W2 h1 + b2);  e = w3 . h2 + b3
//               (32-64-64-1), weights and layer inputs in bfloat16, FP32 accumulation on the MFMA units
//   gradient    dE/dx_i = sum_j [sum_k (g_k(i) + g_k(j)) phi'_k(r_ij)] (x_i - x_j) / r_ij, g = dE/dG
//
// Kernels (all per sorted atom slot, so neighbour references stay valid across them):
//   k_surr_pairs  one workgroup per cell: 27-cell j-list in LDS (as in the LJ sweep); the pairs within rc
//                 of each i-atom are appended in j-list order (wave ballot, deterministic) with
//                 (dx, dy, dz, r, fc, fc') to a fixed-capacity per-atom list
//   k_surr_desc   32 lanes per atom (one per feature): G_k over the atom's list
//   k_surr_mlp    one wave per 16 atoms: forward and input gradient with v_mfma_f32_16x16x32_bf16
//                 (4 + 8 MFMAs forward, 8 + 4 backward); weights staged in LDS in both orientations
//   k_surr_force  32 lanes per atom: per-lane partial of the pair sum, one butterfly per atom; kicks p
//   k_surr_grad   training: MFMA forward on 32 atoms per workgroup, per-atom energy residuals, FP32
//                 parameter gradients accumulated in LDS -> per-workgroup partials
//   k_surr_adam   sums the partials in workgroup order and applies Adam (Kingma & Ba) to the FP32
//                 master weights, refreshing the bfloat16 copies
#include "chmc_internal.hpp"
#include "surrogate.hpp"

#include <algorithm>

namespace chmc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float silu_f(float a) { return a / (1.0f + __expf(-a)); }
__device__ __forceinline__ float dsilu_f(float a)
{
    const float s = 1.0f / (1.0f + __expf(-a));
    return s * (1.0f + a * (1.0f - s));
}

// --------------------------------------------------------------------------------------------------
// pair list
// --------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_surr_pairs(DevView v, SurrView s)
{
    constexpr int IS = 4, JG = 16, NT = 256, W = 4;
    __shared__ float4 sj[kMaxJ];
    __shared__ int sidx[kMaxJ];
    __shared__ int seg_beg[27], seg_pref[28];
    __shared__ float seg_sh[27][3];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int c = blockIdx.x;
    if (c >= v.Ctot) return;
    const ChainDev& C = v.chains[v.cell_chain[c]];
    const int nx = C.nbin[0], ny = C.nbin[1], nz = C.nbin[2];
    const int lc = c - C.cell_off;
    const int bz = lc / (nx * ny), by = (lc - bz * nx * ny) / nx, bx = lc - nx * (by + bz * ny);
    const int* offs = offs_cur(v);
    const int ibeg = offs[c], ni = offs[c + 1] - ibeg;
    if (ni == 0) return;
    if (wave == 0) {
        int cnt = 0;
        if (lane < 27) {
            const int dd[3] = { lane % 3 - 1, (lane / 3) % 3 - 1, lane / 9 - 1 };
            const int b3[3] = { bx, by, bz }, n3[3] = { nx, ny, nz };
            int j3[3];
            for (int d = 0; d < 3; d++) {
                j3[d] = b3[d] + dd[d];
                if (j3[d] < 0 || j3[d] >= n3[d]) j3[d] = (b3[d] + n3[d] + dd[d]) % n3[d];
                seg_sh[lane][d] = (float)dd[d] * C.sizebin_f[d];
            }
            const int jc = C.cell_off + j3[0] + nx * (j3[1] + j3[2] * ny);
            seg_beg[lane] = offs[jc];
            cnt = offs[jc + 1] - offs[jc];
        }
        int incl = cnt;
        for (int o = 1; o < 32; o <<= 1) {
            const int t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        if (lane < 27) seg_pref[lane + 1] = incl;
        if (lane == 0) seg_pref[0] = 0;
    }
    __syncthreads();
    const int nj = min(seg_pref[27], kMaxJ);          // lists beyond kMaxJ raise the overflow flag
    if (seg_pref[27] > kMaxJ && tid == 0) atomicOr(s.flag, 2);
    for (int t = tid; t < nj; t += NT) {
        int k = 0;
        for (int step = 16; step > 0; step >>= 1)
            if (k + step < 27 && seg_pref[k + step] <= t) k += step;
        const int src = seg_beg[k] + (t - seg_pref[k]);
        const float4 q = v.xs[src];
        sj[t] = make_float4(q.x + seg_sh[k][0], q.y + seg_sh[k][1], q.z + seg_sh[k][2], 0.f);
        sidx[t] = src;
    }
    __syncthreads();
    const float rc2 = C.rc2_f, rc = sqrtf(C.rc2_f);
    const int ii = lane % IS, g = lane / IS;
    unsigned long long pattern = 0;
    for (int q = 0; q < JG; q++) pattern |= 1ull << (q * IS + ii);
    const int nchunk = (ni + IS - 1) / IS;
    for (int chunk = wave; chunk < nchunk; chunk += W) {
        const int i = chunk * IS + ii;
        const bool ival = i < ni;
        const int slot = ibeg + (ival ? i : ni - 1);
        const float4 xi = v.xs[slot];
        int count = 0;
        for (int j0 = 0; j0 < nj; j0 += JG) {
            const int j = j0 + g;
            bool hit = false;
            float dx = 0, dy = 0, dz = 0, r2 = 0;
            if (ival && j < nj) {
                const float4 q = sj[j];
                dx = xi.x - q.x; dy = xi.y - q.y; dz = xi.z - q.z;
                r2 = dx * dx + dy * dy + dz * dz;
                hit = r2 < rc2 && r2 > 0.f;
            }
            const unsigned long long m = __ballot(hit) & pattern;
            if (hit) {
                const int pos = count + __popcll(m & ((1ull << lane) - 1ull));
                if (pos < kSurrMaxNb) {
                    const float r = sqrtf(r2);
                    const float a = 3.14159265358979f * r / rc;
                    const size_t e = (size_t)slot * kSurrMaxNb + pos;
                    s.nb_d[e] = make_float4(dx, dy, dz, r);
                    s.nb_f[e] = make_float2(0.5f * (__cosf(a) + 1.0f), -0.5f * 3.14159265358979f / rc * __sinf(a));
                    s.nb_j[e] = sidx[j];
                }
            }
            count += __popcll(m);
        }
        if (ival && g == 0) {
            s.nb_n[slot] = min(count, kSurrMaxNb);
            if (count > kSurrMaxNb) atomicOr(s.flag, 1);
        }
    }
}

// --------------------------------------------------------------------------------------------------
// descriptors: 32 lanes per atom
// --------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_surr_desc(DevView v, SurrView s, float* __restrict__ G_out)
{
    const int a = blockIdx.x * 8 + (threadIdx.x >> 5);
    const int f = threadIdx.x & 31;
    if (a >= v.Ntot) return;
    const float mu = mu_k(f), c2 = 1.0f / (2.0f * kSurrWidth * kSurrWidth);
    const int n = s.nb_n[a];
    float acc = 0.f;
    for (int q = 0; q < n; q++) {
        const size_t e = (size_t)a * kSurrMaxNb + q;
        const float r = s.nb_d[e].w, fc = s.nb_f[e].x;
        const float t = r - mu;
        acc += __expf(-t * t * c2) * fc;
    }
    G_out[(size_t)a * kSurrFeat + f] = acc;
}

// --------------------------------------------------------------------------------------------------
// MLP forward + input gradient on the MFMA units
// --------------------------------------------------------------------------------------------------
struct alignas(16) MlpSmem {
    __bf16 W1[kSurrHid][kSurrFeat];      // [out][in]
    __bf16 W1T[kSurrFeat][kSurrHid];
    __bf16 W2[kSurrHid][kSurrHid];
    __bf16 W2T[kSurrHid][kSurrHid];
    float b1[kSurrHid], b2[kSurrHid], w3[kSurrHid], shift[kSurrFeat], scale[kSurrFeat];
    float b3;
    __bf16 act[4][16][kSurrHid];         // per-wave transpose buffer (D layout -> A layout)
};

__device__ void load_weights(MlpSmem& sm, const SurrView& s)
{
    // the four bf16 matrices (both orientations are kept in global memory) in 16-byte pieces
    auto copy16 = [&](void* dst, const void* src, int bytes) {
        const uint4* q = static_cast<const uint4*>(src);
        uint4* d = static_cast<uint4*>(dst);
        for (int t = threadIdx.x; t < bytes / 16; t += blockDim.x) d[t] = q[t];
    };
    copy16(&sm.W1[0][0], s.w1b, kSurrHid * kSurrFeat * 2);
    copy16(&sm.W1T[0][0], s.w1tb, kSurrHid * kSurrFeat * 2);
    copy16(&sm.W2[0][0], s.w2b, kSurrHid * kSurrHid * 2);
    copy16(&sm.W2T[0][0], s.w2tb, kSurrHid * kSurrHid * 2);
    const float* P = s.params;
    for (int t = threadIdx.x; t < kSurrHid; t += blockDim.x) {
        sm.b1[t] = P[kOffB1 + t];
        sm.b2[t] = P[kOffB2 + t];
        sm.w3[t] = (float)(__bf16)P[kOffW3 + t];
    }
    for (int t = threadIdx.x; t < kSurrFeat; t += blockDim.x) {
        sm.shift[t] = P[kOffShift + t];
        sm.scale[t] = P[kOffScale + t];
    }
    if (threadIdx.x == 0) sm.b3 = P[kOffB3];
}

// A fragment (16 rows x 32 k) from a row-major bf16 tile in LDS: lane l holds row l%16, k = 8*(l/16)..+7
__device__ __forceinline__ bf16x8 frag_rows(const __bf16* base, int ld, int k0)
{
    const int l = threadIdx.x & 63;
    const __bf16* p = base + (l & 15) * ld + k0 + 8 * (l >> 4);
    bf16x8 r;
#pragma unroll
    for (int i = 0; i < 8; i++) r[i] = p[i];
    return r;
}

// One wave: 16 atoms, forward (per-atom energy) and optionally the backward pass.  mode 0: inference
// (de = 1 per atom, produce g = dE/dG); mode 1: training (de = 2 (e - u) / n, produce d2 and d1 in LDS
// for the parameter gradients).  Outputs are in D layout: lane l holds rows 4*(l/16)+i, col l%16 (+16 t).
struct WaveOut {
    f32x4 a1[4], a2[4];     // pre-activations (col tiles of 16)
    float e[4];             // per-atom energies of the lane's 4 rows (valid in lanes with l%16 == 0 after reduce)
};

__device__ void mlp_forward_wave(MlpSmem& sm, const float* __restrict__ G, int a0, int natoms, WaveOut& o)
{
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    // z = bf16((G - shift) * scale), A layout: row m = l%16 (atom a0+m), k = 8*(l/16)..+7
    bf16x8 za;
    {
        const int m = l & 15, k0 = 8 * (l >> 4);
        const int a = a0 + m;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const float gv = a < natoms ? G[(size_t)a * kSurrFeat + k0 + i] : 0.f;
            za[i] = (__bf16)((gv - sm.shift[k0 + i]) * sm.scale[k0 + i]);
        }
    }
    // layer 1: 4 column tiles of 16 hidden units, K = 32
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const bf16x8 wb = frag_rows(&sm.W1[16 * t][0], kSurrFeat, 0);
        f32x4 acc = { 0.f, 0.f, 0.f, 0.f };
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(za, wb, acc, 0, 0, 0);
        const int n = 16 * t + (l & 15);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            acc[i] += sm.b1[n];
            sm.act[w][4 * (l >> 4) + i][n] = (__bf16)silu_f(acc[i]);
        }
        o.a1[t] = acc;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // layer 2: K = 64 (two k tiles), 4 column tiles
    const bf16x8 h1a0 = frag_rows(&sm.act[w][0][0], kSurrHid, 0);
    const bf16x8 h1a1 = frag_rows(&sm.act[w][0][0], kSurrHid, 32);
    float part[4] = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
    for (int t = 0; t < 4; t++) {
        f32x4 acc = { 0.f, 0.f, 0.f, 0.f };
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h1a0, frag_rows(&sm.W2[16 * t][0], kSurrHid, 0), acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h1a1, frag_rows(&sm.W2[16 * t][0], kSurrHid, 32), acc, 0, 0, 0);
        const int n = 16 * t + (l & 15);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            acc[i] += sm.b2[n];
            const float h2 = (float)(__bf16)silu_f(acc[i]);
            part[i] += h2 * sm.w3[n];
        }
        o.a2[t] = acc;
    }
    // e = sum over the 64 columns: reduce across the 16 lanes holding one row
#pragma unroll
    for (int i = 0; i < 4; i++) {
#pragma unroll
        for (int m = 1; m < 16; m <<= 1) part[i] += __shfl_xor(part[i], m, 64);
        o.e[i] = part[i] + sm.b3;
    }
}

// backward from d2 = de * w3 * silu'(a2) to d1 = (d2 W2) * silu'(a1), d2 and d1 in D layout
__device__ void mlp_backward_wave(MlpSmem& sm, const WaveOut& o, const float de[4], f32x4 d2[4], f32x4 d1[4])
{
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int n = 16 * t + (l & 15);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            d2[t][i] = de[i] * sm.w3[n] * dsilu_f(o.a2[t][i]);
            sm.act[w][4 * (l >> 4) + i][n] = (__bf16)d2[t][i];
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const bf16x8 d2a0 = frag_rows(&sm.act[w][0][0], kSurrHid, 0);
    const bf16x8 d2a1 = frag_rows(&sm.act[w][0][0], kSurrHid, 32);
#pragma unroll
    for (int t = 0; t < 4; t++) {
        f32x4 acc = { 0.f, 0.f, 0.f, 0.f };
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(d2a0, frag_rows(&sm.W2T[16 * t][0], kSurrHid, 0), acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(d2a1, frag_rows(&sm.W2T[16 * t][0], kSurrHid, 32), acc, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; i++) d1[t][i] = acc[i] * dsilu_f(o.a1[t][i]);
    }
}

__global__ __launch_bounds__(256) void k_surr_mlp(DevView v, SurrView s, const float* __restrict__ G,
                                                   float* __restrict__ g_out, float* __restrict__ e_out)
{
    __shared__ MlpSmem sm;
    load_weights(sm, s);
    __syncthreads();
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int a0 = (blockIdx.x * 4 + w) * 16;
    if (a0 >= v.Ntot) return;      // wave-uniform; no block barrier follows
    WaveOut o;
    mlp_forward_wave(sm, G, a0, v.Ntot, o);
    if ((l & 15) == 0) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int a = a0 + 4 * (l >> 4) + i;
            if (a < v.Ntot && e_out) e_out[a] = o.e[i];
        }
    }
    const float one[4] = { 1.f, 1.f, 1.f, 1.f };
    f32x4 d2[4], d1[4];
    mlp_backward_wave(sm, o, one, d2, d1);
    // g = (d1 W1) * scale: d1 -> bf16 A layout, B = W1T rows (feature) x 64 hidden
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int n = 16 * t + (l & 15);
#pragma unroll
        for (int i = 0; i < 4; i++) sm.act[w][4 * (l >> 4) + i][n] = (__bf16)d1[t][i];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const bf16x8 d1a0 = frag_rows(&sm.act[w][0][0], kSurrHid, 0);
    const bf16x8 d1a1 = frag_rows(&sm.act[w][0][0], kSurrHid, 32);
#pragma unroll
    for (int t = 0; t < 2; t++) {
        f32x4 acc = { 0.f, 0.f, 0.f, 0.f };
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(d1a0, frag_rows(&sm.W1T[16 * t][0], kSurrHid, 0), acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(d1a1, frag_rows(&sm.W1T[16 * t][0], kSurrHid, 32), acc, 0, 0, 0);
        const int n = 16 * t + (l & 15);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int a = a0 + 4 * (l >> 4) + i;
            if (a < v.Ntot) g_out[(size_t)a * kSurrFeat + n] = acc[i] * sm.scale[n];
        }
    }
}

// --------------------------------------------------------------------------------------------------
// surrogate gradient + kick: 32 lanes per atom
// --------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_surr_force(DevView v, SurrView s, const float* __restrict__ g,
                                                    double kick_dt, int store)
{
    const int a = blockIdx.x * 8 + (threadIdx.x >> 5);
    const int f = threadIdx.x & 31;
    if (a >= v.Ntot) return;
    const float mu = mu_k(f), c2 = 1.0f / (2.0f * kSurrWidth * kSurrWidth), iw2 = 1.0f / (kSurrWidth * kSurrWidth);
    const float gi = g[(size_t)a * kSurrFeat + f];
    const int n = s.nb_n[a];
    float fx = 0.f, fy = 0.f, fz = 0.f;
    for (int q = 0; q < n; q++) {
        const size_t e = (size_t)a * kSurrMaxNb + q;
        const float4 d = s.nb_d[e];
        const float2 fc = s.nb_f[e];
        const int j = s.nb_j[e];
        const float t = d.w - mu;
        const float ex = __expf(-t * t * c2);
        const float dphi = ex * (fc.y - t * iw2 * fc.x);
        const float term = (gi + g[(size_t)j * kSurrFeat + f]) * dphi / d.w;
        fx += term * d.x;
        fy += term * d.y;
        fz += term * d.z;
    }
#pragma unroll
    for (int m = 1; m < 32; m <<= 1) {
        fx += __shfl_xor(fx, m, 64);
        fy += __shfl_xor(fy, m, 64);
        fz += __shfl_xor(fz, m, 64);
    }
    if (f == 0) {
        const int atom = __float_as_int(v.xs[a].w);
        const ChainDev& C = v.chains[v.atom_chain[atom]];
        const int N = v.Ntot;
        const float f3[3] = { fx, fy, fz };
        const double kdt = kick_dt * C.beta;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            if (store) v.f32[d * N + atom] = f3[d];
            if (kick_dt != 0.0) v.p[d * N + atom] = lp_round(v.p[d * N + atom] - kdt * (double)f3[d], v.lp_scale);
        }
    }
}

// --------------------------------------------------------------------------------------------------
// training
// --------------------------------------------------------------------------------------------------
constexpr int kGradRows = 32;          // atoms per training workgroup (2 waves x 16)

struct GradSmem {
    MlpSmem m;
    float d2[kGradRows][kSurrHid];
    float d1[kGradRows][kSurrHid];
    __bf16 h1[kGradRows][kSurrHid];
    __bf16 h2[kGradRows][kSurrHid];
    __bf16 z[kGradRows][kSurrFeat];
    float res[kGradRows];
};

// G, u: training rows; rows [row0, row0 + nrows) of this batch; partial gradients per workgroup
__global__ __launch_bounds__(128) void k_surr_grad(SurrView s, const float* __restrict__ G, const float* __restrict__ u,
                                                   long long row0, int nrows, float* __restrict__ partial)
{
    __shared__ GradSmem sm;
    load_weights(sm.m, s);
    __syncthreads();
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r0 = blockIdx.x * kGradRows + w * 16;     // row offset within the batch
    WaveOut o;
    mlp_forward_wave(sm.m, G + (size_t)row0 * kSurrFeat, r0, nrows, o);
    float de[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int rr = r0 + 4 * (l >> 4) + i;
        const float res = rr < nrows ? o.e[i] - u[row0 + rr] : 0.f;
        de[i] = 2.0f * res / (float)nrows;
        if ((l & 15) == 0) sm.res[w * 16 + 4 * (l >> 4) + i] = res;
    }
    f32x4 d2[4], d1[4];
    mlp_backward_wave(sm.m, o, de, d2, d1);
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int n = 16 * t + (l & 15);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int m = w * 16 + 4 * (l >> 4) + i;
            const bool valid = r0 + 4 * (l >> 4) + i < nrows;
            sm.d2[m][n] = valid ? d2[t][i] : 0.f;
            sm.d1[m][n] = valid ? d1[t][i] : 0.f;
            sm.h1[m][n] = (__bf16)silu_f(o.a1[t][i]);
            sm.h2[m][n] = (__bf16)silu_f(o.a2[t][i]);
        }
    }
    {
        const int m = w * 16 + (l & 15), k0 = 8 * (l >> 4);
        const int rr = r0 + (l & 15);
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const float gv = rr < nrows ? G[(size_t)(row0 + rr) * kSurrFeat + k0 + i] : 0.f;
            sm.z[m][k0 + i] = (__bf16)((gv - sm.m.shift[k0 + i]) * sm.m.scale[k0 + i]);
        }
    }
    __syncthreads();
    const float two_n = 2.0f / (float)nrows;
    float* out = partial + (size_t)blockIdx.x * kSurrNGrad;
    for (int p = threadIdx.x; p < kSurrNGrad; p += blockDim.x) {
        float acc = 0.f;
        if (p < kOffB1) {                       // W1[o][i] += sum_m d1[m][o] z[m][i]
            const int oo = p / kSurrFeat, ii = p % kSurrFeat;
            for (int m = 0; m < kGradRows; m++) acc += sm.d1[m][oo] * (float)sm.z[m][ii];
        } else if (p < kOffW2) {                // b1
            const int oo = p - kOffB1;
            for (int m = 0; m < kGradRows; m++) acc += sm.d1[m][oo];
        } else if (p < kOffB2) {                // W2[o][i] += sum_m d2[m][o] h1[m][i]
            const int q = p - kOffW2, oo = q / kSurrHid, ii = q % kSurrHid;
            for (int m = 0; m < kGradRows; m++) acc += sm.d2[m][oo] * (float)sm.h1[m][ii];
        } else if (p < kOffW3) {                // b2
            const int oo = p - kOffB2;
            for (int m = 0; m < kGradRows; m++) acc += sm.d2[m][oo];
        } else if (p < kOffB3) {                // w3[n] += sum_m h2[m][n] de[m]
            const int nn = p - kOffW3;
            for (int m = 0; m < kGradRows; m++) acc += (float)sm.h2[m][nn] * (two_n * sm.res[m]);
        } else if (p == kOffB3) {               // b3
            for (int m = 0; m < kGradRows; m++) acc += two_n * sm.res[m];
        } else {                                // loss slot: sum of squared residuals
            for (int m = 0; m < kGradRows; m++) acc += sm.res[m] * sm.res[m];
        }
        out[p] = acc;
    }
}

// fixed-order sum of the workgroup partials (no update): gradient checks against the oracle
// (partials of ngrad floats: the trained parameters, then ngrad - kSurrNTrain loss sums)
__global__ __launch_bounds__(256) void k_surr_sum(const float* __restrict__ partial, int nparts, int ngrad,
                                                  float* __restrict__ out)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= ngrad) return;
    float gsum = 0.f;
    for (int b = 0; b < nparts; b++) gsum += partial[(size_t)b * ngrad + p];
    out[p] = gsum;
}

// the bfloat16 copies (both orientations) of W1 / W2 entry p of the parameter blob
__device__ __forceinline__ void store_bf16_weight(const SurrView& s, int p, float val)
{
    const __bf16 b = (__bf16)val;
    if (p < kOffB1) {
        s.w1b[p] = b;
        s.w1tb[(p % kSurrFeat) * kSurrHid + p / kSurrFeat] = b;
    } else if (p >= kOffW2 && p < kOffB2) {
        const int q = p - kOffW2;
        s.w2b[q] = b;
        s.w2tb[(q % kSurrHid) * kSurrHid + q / kSurrHid] = b;
    }
}

// one thread per trainable parameter: fixed-order sum over workgroup partials, then Adam
__global__ __launch_bounds__(256) void k_surr_adam(SurrView s, const float* __restrict__ partial, int nparts, int ngrad,
                                                   float lr, float beta1, float beta2, float eps, int step,
                                                   float* __restrict__ loss_out)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= ngrad) return;
    float gsum = 0.f;
    for (int b = 0; b < nparts; b++) gsum += partial[(size_t)b * ngrad + p];
    if (p >= kSurrNTrain) {                      // the loss slots
        if (loss_out) loss_out[p - kSurrNTrain] = gsum;
        return;
    }
    float m = s.adam_m[p] = beta1 * s.adam_m[p] + (1.f - beta1) * gsum;
    float vv = s.adam_v[p] = beta2 * s.adam_v[p] + (1.f - beta2) * gsum * gsum;
    const float mh = m / (1.f - powf(beta1, (float)step));
    const float vh = vv / (1.f - powf(beta2, (float)step));
    const float np = s.params[p] - lr * mh / (sqrtf(vh) + eps);
    s.params[p] = np;
    store_bf16_weight(s, p, np);
}

__global__ void k_surr_refresh(SurrView s)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p < kOffB2) store_bf16_weight(s, p, s.params[p]);
}

// per-atom LJ energies u_i = 2 eps sum_j sr6 (sr6 - 1) in sorted order, FP64 pairs from the surrogate's
// own pair list (identical pair set: same stencil and cutoff)
__global__ __launch_bounds__(256) void k_surr_targets(DevView v, SurrView s, float* __restrict__ u_out)
{
    const int a = blockIdx.x * 256 + threadIdx.x;
    if (a >= v.Ntot) return;
    const int atom = __float_as_int(v.xs[a].w);
    const ChainDev& C = v.chains[v.atom_chain[atom]];
    double acc = 0.0;
    const int n = s.nb_n[a];
    for (int q = 0; q < n; q++) {
        const double r = s.nb_d[(size_t)a * kSurrMaxNb + q].w;
        const double sr2 = C.sigma2 / (r * r);
        const double sr6 = sr2 * sr2 * sr2;
        acc += sr6 * (sr6 - 1.0);
    }
    u_out[a] = (float)(2.0 * C.eps * acc);
}

// D layout -> bf16 A-layout transpose buffer of this wave, then the two K = 32 fragments of a 64-wide row
__device__ __forceinline__ void act_store(MlpSmem& sm, const f32x4 x[4])
{
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int i = 0; i < 4; i++) sm.act[w][4 * (l >> 4) + i][16 * t + (l & 15)] = (__bf16)x[t][i];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// x[16 rows][64] (act buffer) times a 64 x 64 bf16 matrix given by rows B[n][k] (B = W2: x W2^T; B = W2T: x W2)
__device__ __forceinline__ void act_times(MlpSmem& sm, const __bf16* B, f32x4 out[4])
{
    const int w = threadIdx.x >> 6;
    const bf16x8 a0 = frag_rows(&sm.act[w][0][0], kSurrHid, 0);
    const bf16x8 a1 = frag_rows(&sm.act[w][0][0], kSurrHid, 32);
#pragma unroll
    for (int t = 0; t < 4; t++) {
        f32x4 acc = { 0.f, 0.f, 0.f, 0.f };
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, frag_rows(B + 16 * t * kSurrHid, kSurrHid, 0), acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, frag_rows(B + 16 * t * kSurrHid, kSurrHid, 32), acc, 0, 0, 0);
        out[t] = acc;
    }
}

// --------------------------------------------------------------------------------------------------
// surrogate on the neighbour lists: descriptors + MLP in one launch
// --------------------------------------------------------------------------------------------------
// 32 slots per workgroup, 8 lanes per slot.  Pass 1 (stage_pairs): the lanes test the slot's list
// entries (the pairs within rc + skin of the last rebuild, the first 16 per lane loaded together) and
// append the distances within rc to an LDS row in entry order; pass 2: they walk that row densely and
// accumulate the 32 Gaussian features in registers (two exponentials per block of 8 centres, the
// recurrence of the surrogate sweep).  The eight partial descriptors meet in a butterfly and go to LDS;
// waves 0 and 1 then run the MLP forward pass and its input gradient on 16 slots each (as k_surr_mlp).
// Slots follow the current list buffer of their chain; groups of 32 slots are mapped XCD-contiguously.
constexpr int kDescPairCap = 128;   // pairs within rc per atom staged in LDS
template <bool W16>
__global__ __launch_bounds__(256) void k_lsurr_desc_mlp(DevView v, SurrView s, int q, float* __restrict__ g_out,
                                                        float* __restrict__ e_out, float* __restrict__ G_out)
{
    __shared__ MlpSmem sm;
    __shared__ float sG[32][kSurrFeat];
    __shared__ float srd[32][kDescPairCap];
    __shared__ int4 seg_tab[W16 ? 32 : 1][kSegStride / 4];   // W16: segment starts of each slot's cell
    load_weights(sm, s);
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ngroups = (v.Ntot + 31) / 32;
    const int s0 = xcd_block(blockIdx.x, ngroups) * 32;               // first slot of this workgroup
    const int a = threadIdx.x >> 3, sub = threadIdx.x & 7;             // slot within the workgroup
    const int slot = s0 + a;
    const int N = v.Ntot;
    float G[kSurrFeat];
#pragma unroll
    for (int f = 0; f < kSurrFeat; f++) G[f] = 0.f;
    if (slot < N) {
        const int ch = v.atom_chain[slot];
        const ChainDev& C = v.chains[ch];
        const ListBuf L = list_buf(v, *list_cur(v, ch));
        const float4* __restrict__ xq = L.xsl + (size_t)q * N;
        const int n = L.nbn[slot];
        if constexpr (W16) {
            load_seg_row<8>(L, C, slot, sub, seg_tab[a]);      // the slot's 8 lanes share a wave
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        int np = stage_pairs<8, 16, false, W16>(C, xq, L.nb, v.nbcap, slot, reinterpret_cast<const int*>(seg_tab[W16 ? a : 0]),
                                                n, xq[slot], sub, nullptr, &srd[a][0], kDescPairCap);
        if (np > kDescPairCap) {
            if (sub == 0) atomicOr(s.flag, 1);
            np = kDescPairCap;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const float pirc = 3.14159265358979f / sqrtf(C.rc2_f);
        const float c2 = 1.0f / (2.0f * kSurrWidth * kSurrWidth);
        constexpr float kD = (kSurrMu1 - kSurrMu0) / (float)(kSurrFeat - 1);
        const float cq = __expf(-2.0f * c2 * kD * kD);
        for (int k = sub; k < np; k += 8) {
            const float r = srd[a][k];
            const float fc = 0.5f * (__cosf(pirc * r) + 1.0f);
#pragma unroll
            for (int b = 0; b < kSurrFeat / 8; b++) {
                const float tb = r - mu_k(8 * b);
                float ex = __expf(-tb * tb * c2) * fc;
                float rho = __expf(c2 * kD * (2.0f * tb - kD));
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    G[8 * b + u] += ex;
                    ex *= rho;
                    rho *= cq;
                }
            }
        }
    }
#pragma unroll
    for (int f = 0; f < kSurrFeat; f++) {
        G[f] += __shfl_xor(G[f], 1, 64);
        G[f] += __shfl_xor(G[f], 2, 64);
        G[f] += __shfl_xor(G[f], 4, 64);
    }
    if (sub == 0) {
#pragma unroll
        for (int f = 0; f < kSurrFeat; f++) sG[a][f] = G[f];
        if (G_out && slot < N) {
#pragma unroll
            for (int f = 0; f < kSurrFeat; f++) G_out[(size_t)slot * kSurrFeat + f] = G[f];
        }
    }
    __syncthreads();                                   // weights and descriptors in LDS
    const int m0 = s0 + 16 * w;                        // first slot of this wave's MLP rows
    if (w >= 2 || m0 >= N) return;                     // wave-uniform; no block barrier follows
    const int nval = min(16, N - m0);
    WaveOut o;
    mlp_forward_wave(sm, &sG[16 * w][0], 0, nval, o);
    if (e_out && (l & 15) == 0) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int m = 4 * (l >> 4) + i;
            if (m < nval) e_out[m0 + m] = o.e[i];
        }
    }
    const float one[4] = { 1.f, 1.f, 1.f, 1.f };
    f32x4 d2[4], d1[4];
    mlp_backward_wave(sm, o, one, d2, d1);
    act_store(sm, d1);
    const bf16x8 d1a0 = frag_rows(&sm.act[w][0][0], kSurrHid, 0);
    const bf16x8 d1a1 = frag_rows(&sm.act[w][0][0], kSurrHid, 32);
#pragma unroll
    for (int t = 0; t < 2; t++) {
        f32x4 acc = { 0.f, 0.f, 0.f, 0.f };
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(d1a0, frag_rows(&sm.W1T[16 * t][0], kSurrHid, 0), acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(d1a1, frag_rows(&sm.W1T[16 * t][0], kSurrHid, 32), acc, 0, 0, 0);
        const int n = 16 * t + (l & 15);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int m = 4 * (l >> 4) + i;
            if (m < nval) g_out[(size_t)(m0 + m) * kSurrFeat + n] = acc[i] * sm.scale[n];
        }
    }
}

// --------------------------------------------------------------------------------------------------
// force matching: loss = mean_i (e_i - u_i)^2 + w_f mean_{i,d} (dE/dx_id - dU/dx_id)^2 over whole
// configurations (all atoms of every chain at the current positions)
// --------------------------------------------------------------------------------------------------
__device__ __forceinline__ float d2silu_f(float a)
{
    const float s = 1.0f / (1.0f + __expf(-a));
    return s * (1.0f - s) * (2.0f + a * (1.0f - 2.0f * s));
}

// per slot: R = dE/dx - dU/dx (surrogate gradient from g = dE/dG, LJ gradient 48 eps sr6 (1/2 - sr6) / r^2 d
// in FP32 from the pair list) and the per-atom LJ energy u = 2 eps sum_j sr6 (sr6 - 1) (FP64 pairs);
// 32 lanes per atom: lane f takes feature f of every pair for dE/dx and pairs f, f + 32, ... for U
__global__ __launch_bounds__(256) void k_surr_fm_resid(DevView v, SurrView s, const float* __restrict__ g,
                                                       float4* __restrict__ R, float* __restrict__ u_out)
{
    const int a = blockIdx.x * 8 + (threadIdx.x >> 5);
    const int f = threadIdx.x & 31;
    if (a >= v.Ntot) return;
    const int atom = __float_as_int(v.xs[a].w);
    const ChainDev& C = v.chains[v.atom_chain[atom]];
    const float mu = mu_k(f), c2 = 1.0f / (2.0f * kSurrWidth * kSurrWidth), iw2 = 1.0f / (kSurrWidth * kSurrWidth);
    const float gi = g[(size_t)a * kSurrFeat + f];
    const float sig2 = (float)C.sigma2, c48 = (float)(48.0 * C.eps);
    const int n = s.nb_n[a];
    float ex_ = 0.f, ey = 0.f, ez = 0.f, lx = 0.f, ly = 0.f, lz = 0.f;
    double uacc = 0.0;
    for (int q = 0; q < n; q++) {
        const size_t e = (size_t)a * kSurrMaxNb + q;
        const float4 d = s.nb_d[e];
        const float2 fc = s.nb_f[e];
        const int j = s.nb_j[e];
        const float t = d.w - mu;
        const float ex = __expf(-t * t * c2);
        const float dphi = ex * (fc.y - t * iw2 * fc.x);
        const float term = (gi + g[(size_t)j * kSurrFeat + f]) * dphi / d.w;
        ex_ += term * d.x;
        ey += term * d.y;
        ez += term * d.z;
        if ((q & 31) == f) {
            const float sr2 = sig2 / (d.w * d.w), sr6 = sr2 * sr2 * sr2;
            const float cf = c48 * sr6 * (0.5f - sr6) / (d.w * d.w);
            lx += cf * d.x;
            ly += cf * d.y;
            lz += cf * d.z;
            const double r = d.w, dr2 = C.sigma2 / (r * r), dr6 = dr2 * dr2 * dr2;
            uacc += dr6 * (dr6 - 1.0);
        }
    }
#pragma unroll
    for (int m = 1; m < 32; m <<= 1) {
        ex_ += __shfl_xor(ex_, m, 64);
        ey += __shfl_xor(ey, m, 64);
        ez += __shfl_xor(ez, m, 64);
        lx += __shfl_xor(lx, m, 64);
        ly += __shfl_xor(ly, m, 64);
        lz += __shfl_xor(lz, m, 64);
        uacc += __shfl_xor(uacc, m, 64);
    }
    if (f == 0) {
        R[a] = make_float4(ex_ - lx, ey - ly, ez - lz, 0.f);
        u_out[a] = (float)(2.0 * C.eps * uacc);
    }
}

// per slot and feature: v_k(i) = c sum_j phi'_k(r_ij) (d_ij / r_ij) . (R_i - R_j), c = 2 w_f / (3 n): the
// derivative of the force term with respect to g(i) = dE/dG(i)
__global__ __launch_bounds__(256) void k_surr_fm_v(DevView v, SurrView s, const float4* __restrict__ R, float cw,
                                                   float* __restrict__ vout)
{
    const int a = blockIdx.x * 8 + (threadIdx.x >> 5);
    const int f = threadIdx.x & 31;
    if (a >= v.Ntot) return;
    const float mu = mu_k(f), c2 = 1.0f / (2.0f * kSurrWidth * kSurrWidth), iw2 = 1.0f / (kSurrWidth * kSurrWidth);
    const float4 Ri = R[a];
    const int n = s.nb_n[a];
    float acc = 0.f;
    for (int q = 0; q < n; q++) {
        const size_t e = (size_t)a * kSurrMaxNb + q;
        const float4 d = s.nb_d[e];
        const float2 fc = s.nb_f[e];
        const float4 Rj = R[s.nb_j[e]];
        const float t = d.w - mu;
        const float dphi = __expf(-t * t * c2) * (fc.y - t * iw2 * fc.x);
        acc += dphi * (d.x * (Ri.x - Rj.x) + d.y * (Ri.y - Rj.y) + d.z * (Ri.z - Rj.z)) / d.w;
    }
    vout[(size_t)a * kSurrFeat + f] = cw * acc;
}

constexpr int kFmRows = 32;            // rows per chunk of the force-matching gradient (2 waves x 16)

struct FmSmem {
    MlpSmem m;
    float D1[kFmRows][kSurrHid], D2[kFmRows][kSurrHid];      // de delta + tangent terms
    float dl1[kFmRows][kSurrHid], dl2[kFmRows][kSurrHid];    // delta1, delta2 (backward of e with de = 1)
    float q3[kFmRows][kSurrHid];                             // h2 de + silu'(a2) t2  (w3 gradient)
    __bf16 h1[kFmRows][kSurrHid], ta1[kFmRows][kSurrHid];
    __bf16 z[kFmRows][kSurrFeat], uu[kFmRows][kSurrFeat];
    float de[kFmRows];
    float lsum[2][2];                                        // per wave: sum (e - u)^2, sum |R|^2
};

// Gradient of the energy + force-matching loss over rows [0, nrows) (sorted slots of the current
// configuration).  With d = dL/dg = v (from k_surr_fm_v) the force term is differentiated by a forward
// tangent along u = v * scale through the input gradient, then reverse mode:
//   delta2 = w3 silu'(a2), gamma1 = delta2 W2, delta1 = gamma1 silu'(a1)           (de = 1 backward)
//   t1 = W1 u, ta1 = silu'(a1) t1, t2 = W2 ta1, zb2 = w3 silu''(a2) t2, ab1 = zb2 W2,
//   zb1 = silu''(a1) t1 gamma1 + silu'(a1) ab1, D2 = de delta2 + zb2, D1 = de delta1 + zb1
//   dW2 = D2^T h1 + delta2^T ta1, dW1 = D1^T z + delta1^T u, db = sum D, dw3 = h2^T de + sum silu'(a2) t2
// bfloat16 operands of the MFMA products (delta2, u, ta1, zb2) are rounded once and the rounded ta1 and
// u enter the weight gradients.  Each workgroup takes rows_per_wg consecutive rows in chunks of 32 and
// writes one partial of kSurrNTrain + 2 floats (the last two: sum (e - u)^2 and sum |R|^2).
__global__ __launch_bounds__(128) void k_surr_grad_fm(SurrView s, const float* __restrict__ G,
                                                      const float* __restrict__ u, const float* __restrict__ vin,
                                                      const float4* __restrict__ R, int nrows, int rows_per_wg,
                                                      float two_over_n, float* __restrict__ partial)
{
    __shared__ FmSmem sm;
    load_weights(sm.m, s);
    const int tid = threadIdx.x, l = tid & 63, w = tid >> 6;
    if (tid < 4) sm.lsum[tid >> 1][tid & 1] = 0.f;
    __syncthreads();
    // register-blocked weight gradients: blocks b = tid + 128 j of 4 x 4 (W2: b < 256, W1: 256 <= b < 384)
    float acc[3][4][4];
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
        for (int x = 0; x < 4; x++)
#pragma unroll
            for (int y = 0; y < 4; y++) acc[j][x][y] = 0.f;
    float accb = 0.f, accw3 = 0.f, accb3 = 0.f;
    const int wg0 = blockIdx.x * rows_per_wg;
    const int wg1 = min(nrows, wg0 + rows_per_wg);
    for (int base = wg0; base < wg1; base += kFmRows) {
        const int r0 = base + 16 * w;
        WaveOut o;
        mlp_forward_wave(sm.m, G, r0, wg1, o);
        float de[4];
        float le = 0.f;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int rr = r0 + 4 * (l >> 4) + i;
            const float res = rr < wg1 ? o.e[i] - u[rr] : 0.f;
            de[i] = two_over_n * res;
            le += res * res;
        }
        f32x4 s1[4], s2[4], dl2[4], gam[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int n = 16 * t + (l & 15);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                s1[t][i] = dsilu_f(o.a1[t][i]);
                s2[t][i] = dsilu_f(o.a2[t][i]);
                dl2[t][i] = sm.m.w3[n] * s2[t][i];
            }
        }
        act_store(sm.m, dl2);
        act_times(sm.m, &sm.m.W2T[0][0], gam);              // gamma1 = bf16(delta2) W2
        // u = bf16(v * scale) in A layout: row l % 16, features 8 (l / 16) .. + 7
        bf16x8 ua;
        {
            const int m = l & 15, k0 = 8 * (l >> 4);
            const int rr = r0 + m;
#pragma unroll
            for (int i = 0; i < 8; i++)
                ua[i] = (__bf16)(rr < wg1 ? vin[(size_t)rr * kSurrFeat + k0 + i] * sm.m.scale[k0 + i] : 0.f);
            const int mm = 16 * w + m;
#pragma unroll
            for (int i = 0; i < 8; i++) {
                sm.uu[mm][k0 + i] = ua[i];
                const float gv = rr < wg1 ? G[(size_t)rr * kSurrFeat + k0 + i] : 0.f;
                sm.z[mm][k0 + i] = (__bf16)((gv - sm.m.shift[k0 + i]) * sm.m.scale[k0 + i]);
            }
        }
        f32x4 t1[4], ta1[4], t2[4], zb2[4], ab1[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            f32x4 a = { 0.f, 0.f, 0.f, 0.f };
            t1[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ua, frag_rows(&sm.m.W1[16 * t][0], kSurrFeat, 0), a, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 4; i++) ta1[t][i] = s1[t][i] * t1[t][i];
        }
        act_store(sm.m, ta1);
        act_times(sm.m, &sm.m.W2[0][0], t2);                // t2 = bf16(ta1) W2^T
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int n = 16 * t + (l & 15);
#pragma unroll
            for (int i = 0; i < 4; i++) zb2[t][i] = sm.m.w3[n] * d2silu_f(o.a2[t][i]) * t2[t][i];
        }
        // rows of ta1 (rounded as the product saw them) into the gradient stage before act is reused
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int i = 0; i < 4; i++) sm.ta1[16 * w + 4 * (l >> 4) + i][16 * t + (l & 15)] = (__bf16)ta1[t][i];
        act_store(sm.m, zb2);
        act_times(sm.m, &sm.m.W2T[0][0], ab1);              // ab1 = bf16(zb2) W2
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int n = 16 * t + (l & 15);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int m = 16 * w + 4 * (l >> 4) + i;
                const bool valid = r0 + 4 * (l >> 4) + i < wg1;
                const float dl1 = gam[t][i] * s1[t][i];
                const float zb1 = d2silu_f(o.a1[t][i]) * t1[t][i] * gam[t][i] + s1[t][i] * ab1[t][i];
                const float h2 = (float)(__bf16)silu_f(o.a2[t][i]);
                sm.D1[m][n] = valid ? de[i] * dl1 + zb1 : 0.f;
                sm.D2[m][n] = valid ? de[i] * dl2[t][i] + zb2[t][i] : 0.f;
                sm.dl1[m][n] = valid ? dl1 : 0.f;
                sm.dl2[m][n] = valid ? dl2[t][i] : 0.f;
                sm.q3[m][n] = valid ? h2 * de[i] + s2[t][i] * t2[t][i] : 0.f;
                sm.h1[m][n] = (__bf16)silu_f(o.a1[t][i]);
            }
        }
        if ((l & 15) == 0) {
#pragma unroll
            for (int i = 0; i < 4; i++) sm.de[16 * w + 4 * (l >> 4) + i] = de[i];
        }
        // loss sums: (e - u)^2 over the lane's rows (lanes l % 16 == 0 hold them), |R|^2 over rows l < 16
        float lf = 0.f;
        if (l < 16 && r0 + l < wg1) {
            const float4 q = R[r0 + l];
            lf = q.x * q.x + q.y * q.y + q.z * q.z;
        }
        le = (l & 15) == 0 ? le : 0.f;
#pragma unroll
        for (int m = 1; m < 64; m <<= 1) {
            le += __shfl_xor(le, m, 64);
            lf += __shfl_xor(lf, m, 64);
        }
        if (l == 0) {
            sm.lsum[w][0] += le;
            sm.lsum[w][1] += lf;
        }
        __syncthreads();
        // weight gradients of this chunk
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int b = tid + 128 * j;
            if (b < 256) {
                const int o0 = 4 * (b >> 4), i0 = 4 * (b & 15);
                for (int m = 0; m < kFmRows; m++) {
                    const float4 dv = *reinterpret_cast<const float4*>(&sm.D2[m][o0]);
                    const float4 ev = *reinterpret_cast<const float4*>(&sm.dl2[m][o0]);
                    const float d4[4] = { dv.x, dv.y, dv.z, dv.w }, e4[4] = { ev.x, ev.y, ev.z, ev.w };
#pragma unroll
                    for (int y = 0; y < 4; y++) {
                        const float h = (float)sm.h1[m][i0 + y], ta = (float)sm.ta1[m][i0 + y];
#pragma unroll
                        for (int x = 0; x < 4; x++) acc[j][x][y] += d4[x] * h + e4[x] * ta;
                    }
                }
            } else {
                const int bb = b - 256, o0 = 4 * (bb >> 3), k0 = 4 * (bb & 7);
                for (int m = 0; m < kFmRows; m++) {
                    const float4 dv = *reinterpret_cast<const float4*>(&sm.D1[m][o0]);
                    const float4 ev = *reinterpret_cast<const float4*>(&sm.dl1[m][o0]);
                    const float d4[4] = { dv.x, dv.y, dv.z, dv.w }, e4[4] = { ev.x, ev.y, ev.z, ev.w };
#pragma unroll
                    for (int y = 0; y < 4; y++) {
                        const float zz = (float)sm.z[m][k0 + y], uv = (float)sm.uu[m][k0 + y];
#pragma unroll
                        for (int x = 0; x < 4; x++) acc[j][x][y] += d4[x] * zz + e4[x] * uv;
                    }
                }
            }
        }
        if (tid < 64) {
            for (int m = 0; m < kFmRows; m++) {
                accb += sm.D1[m][tid];
                accw3 += sm.q3[m][tid];
            }
        } else {
            for (int m = 0; m < kFmRows; m++) accb += sm.D2[m][tid - 64];
        }
        if (tid == 0)
            for (int m = 0; m < kFmRows; m++) accb3 += sm.de[m];
        __syncthreads();
    }
    float* out = partial + (size_t)blockIdx.x * (kSurrNTrain + 2);
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int b = tid + 128 * j;
        if (b < 256) {
            const int o0 = 4 * (b >> 4), i0 = 4 * (b & 15);
#pragma unroll
            for (int x = 0; x < 4; x++)
#pragma unroll
                for (int y = 0; y < 4; y++) out[kOffW2 + (o0 + x) * kSurrHid + i0 + y] = acc[j][x][y];
        } else {
            const int bb = b - 256, o0 = 4 * (bb >> 3), k0 = 4 * (bb & 7);
#pragma unroll
            for (int x = 0; x < 4; x++)
#pragma unroll
                for (int y = 0; y < 4; y++) out[kOffW1 + (o0 + x) * kSurrFeat + k0 + y] = acc[j][x][y];
        }
    }
    if (tid < 64) {
        out[kOffB1 + tid] = accb;
        out[kOffW3 + tid] = accw3;
    } else {
        out[kOffB2 + tid - 64] = accb;
    }
    if (tid == 0) {
        out[kOffB3] = accb3;
        out[kSurrNTrain] = sm.lsum[0][0] + sm.lsum[1][0];
        out[kSurrNTrain + 1] = sm.lsum[0][1] + sm.lsum[1][1];
    }
}

// ---- launch wrappers --------------------------------------------------------------------------------
hipError_t launch_surr_pairs(const DevView& v, const SurrView& s, hipStream_t st)
{
    hipLaunchKernelGGL(k_surr_pairs, dim3(v.Ctot), dim3(256), 0, st, v, s);
    return hipGetLastError();
}

hipError_t launch_surr_desc(const DevView& v, const SurrView& s, float* G, hipStream_t st)
{
    hipLaunchKernelGGL(k_surr_desc, dim3((v.Ntot + 7) / 8), dim3(256), 0, st, v, s, G);
    return hipGetLastError();
}

hipError_t launch_surr_mlp(const DevView& v, const SurrView& s, const float* G, float* g, float* e, hipStream_t st)
{
    hipLaunchKernelGGL(k_surr_mlp, dim3((v.Ntot + 63) / 64), dim3(256), 0, st, v, s, G, g, e);
    return hipGetLastError();
}

hipError_t launch_surr_force(const DevView& v, const SurrView& s, const float* g, double kick_dt, bool store, hipStream_t st)
{
    hipLaunchKernelGGL(k_surr_force, dim3((v.Ntot + 7) / 8), dim3(256), 0, st, v, s, g, kick_dt, store ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_surr_targets(const DevView& v, const SurrView& s, float* u, hipStream_t st)
{
    hipLaunchKernelGGL(k_surr_targets, dim3((v.Ntot + 255) / 256), dim3(256), 0, st, v, s, u);
    return hipGetLastError();
}

hipError_t launch_surr_train_step(const SurrView& s, const float* G, const float* u, long long row0, int nrows,
                                  float* partial, float lr, int step, float* loss_out, hipStream_t st)
{
    const int nb = (nrows + kGradRows - 1) / kGradRows;
    hipLaunchKernelGGL(k_surr_grad, dim3(nb), dim3(128), 0, st, s, G, u, row0, nrows, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_surr_adam, dim3((kSurrNGrad + 255) / 256), dim3(256), 0, st, s, partial, nb, kSurrNGrad, lr,
                       0.9f, 0.999f, 1e-8f, step, loss_out);
    return hipGetLastError();
}

hipError_t launch_surr_grad_only(const SurrView& s, const float* G, const float* u, long long row0, int nrows,
                                 float* partial, float* grad_out, hipStream_t st)
{
    const int nb = (nrows + kGradRows - 1) / kGradRows;
    hipLaunchKernelGGL(k_surr_grad, dim3(nb), dim3(128), 0, st, s, G, u, row0, nrows, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_surr_sum, dim3((kSurrNGrad + 255) / 256), dim3(256), 0, st, partial, nb, kSurrNGrad, grad_out);
    return hipGetLastError();
}

// force matching at the current configuration: residuals, v, per-workgroup partials; then either the
// summed gradient (grad_out, no update) or an Adam step (lr > 0) with the two loss sums in loss_out[2]
int surr_fm_nparts(int nrows)
{
    const int rows_per_wg = std::max(kFmRows, ((nrows + 255) / 256 + kFmRows - 1) / kFmRows * kFmRows);
    return (nrows + rows_per_wg - 1) / rows_per_wg;
}

hipError_t launch_surr_fm(const DevView& v, const SurrView& s, const float* G, const float* g, float4* R, float* u,
                          float* vbuf, float w_force, float* partial, float* grad_out, float lr, int step,
                          float* loss_out, hipStream_t st)
{
    const int nrows = v.Ntot;
    const int rows_per_wg = std::max(kFmRows, ((nrows + 255) / 256 + kFmRows - 1) / kFmRows * kFmRows);
    const int nb = (nrows + rows_per_wg - 1) / rows_per_wg;
    hipLaunchKernelGGL(k_surr_fm_resid, dim3((nrows + 7) / 8), dim3(256), 0, st, v, s, g, R, u);
    hipLaunchKernelGGL(k_surr_fm_v, dim3((nrows + 7) / 8), dim3(256), 0, st, v, s, R,
                       (float)(2.0 * w_force / (3.0 * nrows)), vbuf);
    hipLaunchKernelGGL(k_surr_grad_fm, dim3(nb), dim3(128), 0, st, s, G, u, vbuf, R, nrows, rows_per_wg,
                       (float)(2.0 / nrows), partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    constexpr int ng = kSurrNTrain + 2;
    if (grad_out)
        hipLaunchKernelGGL(k_surr_sum, dim3((ng + 255) / 256), dim3(256), 0, st, partial, nb, ng, grad_out);
    else
        hipLaunchKernelGGL(k_surr_adam, dim3((ng + 255) / 256), dim3(256), 0, st, s, partial, nb, ng, lr, 0.9f, 0.999f,
                           1e-8f, step, loss_out);
    return hipGetLastError();
}

hipError_t launch_lsurr_desc_mlp(const DevView& v, const SurrView& s, int q, float* g, float* e, float* G, hipStream_t st)
{
    if (v.entry16) hipLaunchKernelGGL(k_lsurr_desc_mlp<true>, dim3((v.Ntot + 31) / 32), dim3(256), 0, st, v, s, q, g, e, G);
    else hipLaunchKernelGGL(k_lsurr_desc_mlp<false>, dim3((v.Ntot + 31) / 32), dim3(256), 0, st, v, s, q, g, e, G);
    return hipGetLastError();
}

hipError_t launch_surr_refresh(const SurrView& s, hipStream_t st)
{
    hipLaunchKernelGGL(k_surr_refresh, dim3((kOffB2 + 255) / 256), dim3(256), 0, st, s);
    return hipGetLastError();
}


}  // namespace chmc
