// kt_learn.h -- device entry points of the learned count-kernel picker.
#pragma once

#include <cstdint>

#include <hip/hip_runtime_api.h>

#include "ktruss.h"

#define KT_FEAT_ACC 12     // u64 partials per workgroup of the phase-feature kernels
#define KT_FEAT_BLOCKS 256 // workgroups of the phase-feature kernels

namespace kt {
namespace dev {
uint32_t mlp_num_params(uint32_t n_out);
// acc: KT_FEAT_BLOCKS * KT_FEAT_ACC u64 partials.
// rowlen (optional): live length of every row, read instead of the stored
// row lengths ptr[r + 1] - ptr[r].  P (optional): the model (C outputs, input
// range R), whose pick on the features goes to *choice in the same launch.
hipError_t phase_features(const uint32_t* ptr, const uint32_t* rowlen, const uint32_t* dst,
                          const uint32_t* sup, uint32_t n, uint32_t E, uint32_t E0, uint32_t k,
                          bool sup_valid, unsigned long long* acc, float* feat, const float* P,
                          uint32_t C, const float* R, unsigned int* choice, hipStream_t s,
                          uint32_t* host_choice = nullptr);
// R: the input range, lo[KT_MLP_IN] then hi[KT_MLP_IN]; inputs are clamped to it.
hipError_t mlp_forward(const float* P, uint32_t C, const float* R, const float* X, uint32_t batch,
                       float* Y, hipStream_t s);
// Gradient of the weighted masked squared error into grad (deterministic);
// part needs nparts * mlp_num_params(C) floats, part_loss nparts floats.
hipError_t mlp_loss_grad(const float* P, uint32_t C, const float* X, const float* T, const float* M,
                         const float* W, uint32_t batch, float inv_wsum, float* part,
                         float* part_loss, uint32_t nparts, float* grad, hipStream_t s);
hipError_t adam_step(float* P, const float* G, float* m1, float* m2, uint32_t total, float lr,
                     uint32_t t, hipStream_t s);
}  // namespace dev
}  // namespace kt
