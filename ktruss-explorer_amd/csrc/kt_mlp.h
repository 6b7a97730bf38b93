// kt_mlp.h -- the count-kernel cost model (host object over device buffers).
#pragma once

#include <cstdint>

#include <hip/hip_runtime_api.h>

#include "ktruss.h"

namespace kt {

struct Mlp {
    static constexpr uint32_t kParts = 64;  // workgroups of the gradient kernel
    uint32_t C = 0, total = 0, steps = 0, cap = 0;
    float *P = nullptr, *G = nullptr, *m1 = nullptr, *m2 = nullptr;
    float *part = nullptr, *part_loss = nullptr;
    float *X = nullptr, *T = nullptr, *M = nullptr, *W = nullptr, *Y = nullptr;
    float* R = nullptr;  // input range: lo[KT_MLP_IN], hi[KT_MLP_IN]
    hipStream_t stream = nullptr;

    ~Mlp();
    int create(uint32_t n_out, uint64_t seed);
    int forward(const float* x, uint32_t batch, float* y);
    int loss_grad(const float* x, const float* t, const float* mask, const float* w, uint32_t batch,
                  float* loss, float* grad);
    int train(const float* x, const float* t, const float* mask, const float* w, uint32_t n,
              uint32_t epochs, uint32_t batch, float lr, uint64_t seed, float* loss_out);
    int get(float* out, uint32_t n) const;
    int set(const float* in, uint32_t n);
    int set_range(const float* lo, const float* hi);
    int get_range(float* lo, float* hi) const;

   private:
    int reset_adam();
    int ensure_batch(uint32_t n);
    int upload(const float* x, const float* t, const float* mask, const float* w, uint32_t n);
    int device_loss_grad(uint32_t batch, bool has_mask, bool has_w, float wsum, float* loss);
};

}  // namespace kt

struct kt_mlp {
    kt::Mlp impl;
};
