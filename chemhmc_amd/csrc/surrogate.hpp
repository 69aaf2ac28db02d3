// Constants and device view of the learned surrogate (descriptor + 32-64-64-1 MLP).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

namespace chmc {

constexpr int kSurrFeat = 32, kSurrHid = 64, kSurrMaxNb = 128;
constexpr float kSurrMu0 = 0.8f, kSurrMu1 = 2.5f, kSurrWidth = 0.1f;
// parameter blob (FP32): trainable W1[64][32], b1, W2[64][64], b2, w3, b3, then the fixed input
// normalisation shift[32], scale[32]
constexpr int kOffW1 = 0, kOffB1 = 2048, kOffW2 = 2112, kOffB2 = 6208, kOffW3 = 6272, kOffB3 = 6336;
constexpr int kSurrNTrain = 6337, kOffShift = 6337, kOffScale = 6369, kSurrNParams = 6401;
constexpr int kSurrNGrad = kSurrNTrain + 1;     // + the loss slot of the training partials

struct SurrView {
    float* params;          // [kSurrNParams] FP32 master copy
    __bf16* w1b;            // [64*32] bfloat16 copy of W1
    __bf16* w2b;            // [64*64] bfloat16 copy of W2
    float *adam_m, *adam_v; // [kSurrNTrain]
    int* nb_n;              // [Ntot] pairs per atom (by sorted slot)
    int* nb_j;              // [Ntot * kSurrMaxNb] sorted slot of the neighbour
    float4* nb_d;           // (dx, dy, dz, r)
    float2* nb_f;           // (fc, dfc/dr)
    int* flag;              // bit 0: a list overflowed kSurrMaxNb; bit 1: a j-list overflowed the LDS stage
};

struct DevView;
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
int surr_fm_nparts(int nrows);
hipError_t launch_surr_fm(const DevView& v, const SurrView& s, const float* G, const float* g, float4* R, float* u,
                          float* vbuf, float w_force, float* partial, float* grad_out, float lr, int step,
                          float* loss_out, hipStream_t st);

}  // namespace chmc
