// Constants and device view of the learned surrogate (descriptor + 32-64-64-1 MLP).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include "chmc_internal.hpp"

namespace chmc {

constexpr int kSurrFeat = 32, kSurrHid = 64, kSurrMaxNb = 128;
constexpr float kSurrMu0 = 0.8f, kSurrMu1 = 2.5f, kSurrWidth = 0.1f;
// parameter blob (FP32): trainable W1[64][32], b1, W2[64][64], b2, w3, b3, then the fixed input
// normalisation shift[32], scale[32]
constexpr int kOffW1 = 0, kOffB1 = 2048, kOffW2 = 2112, kOffB2 = 6208, kOffW3 = 6272, kOffB3 = 6336;
constexpr int kSurrNTrain = 6337, kOffShift = 6337, kOffScale = 6369, kSurrNParams = 6401;
constexpr int kSurrNGrad = kSurrNTrain + 1;     // + the loss slot of the training partials

// centre of descriptor k: uniform on [kSurrMu0, kSurrMu1]
__device__ __forceinline__ float mu_k(int k) { return kSurrMu0 + (kSurrMu1 - kSurrMu0) * (float)k / (float)(kSurrFeat - 1); }

struct SurrView {
    float* params;          // [kSurrNParams] FP32 master copy
    __bf16* w1b;            // [64*32] bfloat16 copy of W1
    __bf16* w2b;            // [64*64] bfloat16 copy of W2
    __bf16* w1tb;           // [32*64] bfloat16 copy of W1 transposed
    __bf16* w2tb;           // [64*64] bfloat16 copy of W2 transposed
    float *adam_m, *adam_v; // [kSurrNTrain]
    int* nb_n;              // [Ntot] pairs per atom (by sorted slot)
    int* nb_j;              // [Ntot * kSurrMaxNb] sorted slot of the neighbour
    float4* nb_d;           // (dx, dy, dz, r)
    float2* nb_f;           // (fc, dfc/dr)
    int* flag;              // bit 0: a list overflowed kSurrMaxNb; bit 1: a j-list overflowed the LDS stage
};

// The pairs within rc of one slot, from its neighbour-list row (surrogate kernels): the LP lanes of the
// slot read KB entries each without a count test (entries past the count read the row's first entry,
// which is a real entry or padding), gather the neighbours' coordinates together (32-bit entries: the
// slot and image code; 16-bit (W16): slot = segment start from the slot's LDS segment row segs + rank),
// then test them; a row longer than KB * LP entries continues in a loop.  The pairs within rc are
// appended in entry order to the slot's LDS row by ballot prefix counts: with GEOM, pd = (dx, dy, dz,
// neighbour slot as float bits), otherwise r only (rd).  Every lane of the slot returns the same count
// (which may exceed cap: the caller flags it).
template <int LP, int KB, bool GEOM, bool W16>
__device__ __forceinline__ int stage_pairs(const ChainDev& C, const float4* __restrict__ xq, const char* nb, int nbcap,
                                           int slot, const int* segs, int n, float4 xi, int sub, float4* pd, float* rd,
                                           int cap)
{
    using E = std::conditional_t<W16, uint16_t, int>;
    const E* __restrict__ nbs = reinterpret_cast<const E*>(nb) + (size_t)slot * nbcap;
    const int gbase = (threadIdx.x & 63) & ~(LP - 1);
    const float rc2 = C.rc2_f;
    int np = 0;
    auto slot_of = [&](unsigned e) -> int {
        if constexpr (W16) return entry_slot(segs, e);
        else return (int)(e >> kImageBits);
    };
    auto test = [&](unsigned e, float4 q, bool live) {
        float sx, sy, sz;
        if constexpr (W16) {
            const int k = (int)(e >> kRankBits);
            sx = seg_shift(k, 0, C.lsize_f[0]); sy = seg_shift(k, 1, C.lsize_f[1]); sz = seg_shift(k, 2, C.lsize_f[2]);
        } else {
            const int code = (int)(e & 63u);
            sx = image_shift(code, 0, C.lsize_f[0]); sy = image_shift(code, 1, C.lsize_f[1]);
            sz = image_shift(code, 2, C.lsize_f[2]);
        }
        const float dx = xi.x - (q.x + sx);
        const float dy = xi.y - (q.y + sy);
        const float dz = xi.z - (q.z + sz);
        const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
        const bool hit = live && r2 < rc2;
        const unsigned grp = (unsigned)((__ballot(hit) >> gbase) & ((1ull << LP) - 1ull));
        const int pos = np + __popc(grp & ((1u << sub) - 1u));
        if (hit && pos < cap) {
            if constexpr (GEOM) {
                pd[pos] = make_float4(dx, dy, dz, __int_as_float(slot_of(e)));
            } else {
                rd[pos] = sqrtf(r2);
            }
        }
        np += __popc(grp);
    };
    unsigned ent[KB];
    float4 q[KB];
#pragma unroll
    for (int u = 0; u < KB; u++) {
        const int k = sub + u * LP;
        ent[u] = (unsigned)nbs[k < n ? k : 0];
    }
#pragma unroll
    for (int u = 0; u < KB; u++) q[u] = xq[slot_of(ent[u])];
#pragma unroll
    for (int u = 0; u < KB; u++) test(ent[u], q[u], sub + u * LP < n);
    for (int k0 = KB * LP; k0 < n; k0 += LP) {
        const int k = k0 + sub;
        const bool live = k < n;
        const unsigned e = (unsigned)nbs[live ? k : k0];
        test(e, xq[slot_of(e)], live);
    }
    return np;
}

hipError_t launch_surr_pairs(const DevView& v, const SurrView& s, hipStream_t st);
hipError_t launch_surr_desc(const DevView& v, const SurrView& s, float* G, hipStream_t st);
hipError_t launch_surr_mlp(const DevView& v, const SurrView& s, const float* G, float* g, float* e, hipStream_t st);
hipError_t launch_surr_force(const DevView& v, const SurrView& s, const float* g, double kick_dt, bool store, hipStream_t st);
hipError_t launch_surr_targets(const DevView& v, const SurrView& s, float* u, hipStream_t st);
hipError_t launch_surr_train_step(const SurrView& s, const float* G, const float* u, long long row0, int nrows,
                                  float* partial, float lr, int step, float* loss_out, hipStream_t st);
hipError_t launch_surr_grad_only(const SurrView& s, const float* G, const float* u, long long row0, int nrows,
                                 float* partial, float* grad_out, hipStream_t st);
hipError_t launch_surr_refresh(const SurrView& s, hipStream_t st);
// surrogate MD on the neighbour lists (slot-ordered state of the current list buffer, FP32 coordinates of
// parity q): descriptors of the pairs within rc from each atom's list and the MLP with its input
// gradient in one launch (g[slot][32], optional e[slot] and G[slot][32]); then the force from g and the lists with the
// list sweep's kick / drift / rebuild-request tail (mode: LJ_* flags of launch_list_sweep)
hipError_t launch_lsurr_desc_mlp(const DevView& v, const SurrView& s, int q, float* g, float* e, float* G, hipStream_t st);
hipError_t launch_lsurr_sweep(const DevView& v, const float* g, int mode, int q, double kick_dt, double drift_dt,
                              bool checked, hipStream_t st);
int surr_fm_nparts(int nrows);
hipError_t launch_surr_fm(const DevView& v, const SurrView& s, const float* G, const float* g, float4* R, float* u,
                          float* vbuf, float w_force, float* partial, float* grad_out, float lr, int step,
                          float* loss_out, hipStream_t st);

}  // namespace chmc
