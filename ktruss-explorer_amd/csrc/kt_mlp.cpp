// kt_mlp.cpp -- host side of the count-kernel cost model: parameter storage
// on the device, seeded initialisation, mini-batch Adam training driven from
// the host, all arithmetic in the kt_learn.hip kernels.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "kt_internal.h"
#include "kt_learn.h"
#include "kt_mlp.h"

namespace kt {

Mlp::~Mlp() {
    for (void* p : {(void*)P, (void*)G, (void*)m1, (void*)m2, (void*)part, (void*)part_loss,
                    (void*)X, (void*)T, (void*)M, (void*)W, (void*)Y, (void*)R})
        if (p) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
}

int Mlp::create(uint32_t n_out, uint64_t seed) {
    if (n_out == 0 || n_out > KT_MLP_MAX_OUT) return fail(KT_ERR_ARG, "n_out must be in [1, 16]");
    C = n_out;
    total = dev::mlp_num_params(C);
    KT_HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    KT_HIP_TRY(hipMalloc((void**)&P, total * 4));
    KT_HIP_TRY(hipMalloc((void**)&G, total * 4));
    KT_HIP_TRY(hipMalloc((void**)&m1, total * 4));
    KT_HIP_TRY(hipMalloc((void**)&m2, total * 4));
    KT_HIP_TRY(hipMalloc((void**)&part, (size_t)kParts * total * 4));
    KT_HIP_TRY(hipMalloc((void**)&part_loss, kParts * 4));
    KT_HIP_TRY(hipMalloc((void**)&R, 2 * KT_MLP_IN * 4));
    {
        std::vector<float> r(2 * KT_MLP_IN);
        for (int i = 0; i < KT_MLP_IN; ++i) {
            r[i] = -FLT_MAX;
            r[KT_MLP_IN + i] = FLT_MAX;
        }
        KT_HIP_TRY(hipMemcpy(R, r.data(), r.size() * 4, hipMemcpyHostToDevice));
    }
    // He-uniform weights from a seeded mt19937_64, zero biases
    std::vector<float> h(total, 0.f);
    std::mt19937_64 rng(seed);
    auto fill = [&](size_t off, size_t count, int fan_in) {
        const float a = std::sqrt(6.f / (float)fan_in);
        std::uniform_real_distribution<float> u(-a, a);
        for (size_t i = 0; i < count; ++i) h[off + i] = u(rng);
    };
    const uint32_t w1 = 0, b1 = w1 + KT_MLP_HIDDEN * KT_MLP_IN, w2 = b1 + KT_MLP_HIDDEN;
    const uint32_t b2 = w2 + KT_MLP_HIDDEN * KT_MLP_HIDDEN, w3 = b2 + KT_MLP_HIDDEN;
    fill(w1, KT_MLP_HIDDEN * KT_MLP_IN, KT_MLP_IN);
    fill(w2, KT_MLP_HIDDEN * KT_MLP_HIDDEN, KT_MLP_HIDDEN);
    fill(w3, (size_t)C * KT_MLP_HIDDEN, KT_MLP_HIDDEN);
    KT_HIP_TRY(hipMemcpy(P, h.data(), total * 4, hipMemcpyHostToDevice));
    return reset_adam();
}

int Mlp::reset_adam() {
    KT_HIP_TRY(hipMemset(m1, 0, total * 4));
    KT_HIP_TRY(hipMemset(m2, 0, total * 4));
    steps = 0;
    return KT_OK;
}

int Mlp::ensure_batch(uint32_t n) {
    if (n <= cap) return KT_OK;
    for (float** p : {&X, &T, &M, &W, &Y})
        if (*p) {
            (void)hipFree(*p);
            *p = nullptr;
        }
    cap = std::max<uint32_t>(n, 1024);
    KT_HIP_TRY(hipMalloc((void**)&X, (size_t)cap * KT_MLP_IN * 4));
    KT_HIP_TRY(hipMalloc((void**)&T, (size_t)cap * C * 4));
    KT_HIP_TRY(hipMalloc((void**)&M, (size_t)cap * C * 4));
    KT_HIP_TRY(hipMalloc((void**)&W, (size_t)cap * 4));
    KT_HIP_TRY(hipMalloc((void**)&Y, (size_t)cap * C * 4));
    return KT_OK;
}

int Mlp::forward(const float* x, uint32_t batch, float* y) {
    int rc = ensure_batch(batch);
    if (rc) return rc;
    KT_HIP_TRY(hipMemcpyAsync(X, x, (size_t)batch * KT_MLP_IN * 4, hipMemcpyHostToDevice, stream));
    KT_HIP_TRY(dev::mlp_forward(P, C, R, X, batch, Y, stream));
    KT_HIP_TRY(hipMemcpyAsync(y, Y, (size_t)batch * C * 4, hipMemcpyDeviceToHost, stream));
    KT_HIP_TRY(hipStreamSynchronize(stream));
    return KT_OK;
}

// Loss and gradient on the device-resident batch [X, T, M, W] of `batch` rows.
int Mlp::device_loss_grad(uint32_t batch, bool has_mask, bool has_w, float wsum, float* loss) {
    const uint32_t nparts = std::min<uint32_t>(kParts, std::max<uint32_t>(batch, 1));
    KT_HIP_TRY(dev::mlp_loss_grad(P, C, X, T, has_mask ? M : nullptr, has_w ? W : nullptr, batch,
                                  1.f / wsum, part, part_loss, nparts, G, stream));
    if (loss) {
        std::vector<float> pl(nparts);
        KT_HIP_TRY(hipMemcpyAsync(pl.data(), part_loss, nparts * 4, hipMemcpyDeviceToHost, stream));
        KT_HIP_TRY(hipStreamSynchronize(stream));
        double s = 0;
        for (float v : pl) s += v;
        *loss = (float)s;
    }
    return KT_OK;
}

int Mlp::upload(const float* x, const float* t, const float* mask, const float* w, uint32_t n) {
    int rc = ensure_batch(n);
    if (rc) return rc;
    KT_HIP_TRY(hipMemcpyAsync(X, x, (size_t)n * KT_MLP_IN * 4, hipMemcpyHostToDevice, stream));
    KT_HIP_TRY(hipMemcpyAsync(T, t, (size_t)n * C * 4, hipMemcpyHostToDevice, stream));
    if (mask) KT_HIP_TRY(hipMemcpyAsync(M, mask, (size_t)n * C * 4, hipMemcpyHostToDevice, stream));
    if (w) KT_HIP_TRY(hipMemcpyAsync(W, w, (size_t)n * 4, hipMemcpyHostToDevice, stream));
    return KT_OK;
}

int Mlp::loss_grad(const float* x, const float* t, const float* mask, const float* w,
                   uint32_t batch, float* loss, float* grad) {
    if (!batch) return fail(KT_ERR_ARG, "empty batch");
    int rc = upload(x, t, mask, w, batch);
    if (rc) return rc;
    float wsum = (float)batch;
    if (w) {
        double s = 0;
        for (uint32_t i = 0; i < batch; ++i) s += w[i];
        wsum = (float)s;
    }
    if (!(wsum > 0)) return fail(KT_ERR_ARG, "sample weights sum to zero");
    rc = device_loss_grad(batch, mask != nullptr, w != nullptr, wsum, loss);
    if (rc) return rc;
    if (grad) {
        KT_HIP_TRY(hipMemcpyAsync(grad, G, total * 4, hipMemcpyDeviceToHost, stream));
        KT_HIP_TRY(hipStreamSynchronize(stream));
    }
    return KT_OK;
}

int Mlp::train(const float* x, const float* t, const float* mask, const float* w, uint32_t n,
               uint32_t epochs, uint32_t batch, float lr, uint64_t seed, float* loss_out) {
    if (!n || !batch) return fail(KT_ERR_ARG, "empty training set or batch");
    batch = std::min(batch, n);
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    std::mt19937_64 rng(seed);
    std::vector<float> bx((size_t)batch * KT_MLP_IN), bt((size_t)batch * C), bm((size_t)batch * C),
        bw(batch);
    for (uint32_t ep = 0; ep < epochs; ++ep) {
        std::shuffle(order.begin(), order.end(), rng);
        for (uint32_t b0 = 0; b0 + batch <= n; b0 += batch) {
            double wsum = 0;
            for (uint32_t i = 0; i < batch; ++i) {
                const uint32_t s = order[b0 + i];
                std::memcpy(&bx[(size_t)i * KT_MLP_IN], x + (size_t)s * KT_MLP_IN, KT_MLP_IN * 4);
                std::memcpy(&bt[(size_t)i * C], t + (size_t)s * C, C * 4);
                if (mask) std::memcpy(&bm[(size_t)i * C], mask + (size_t)s * C, C * 4);
                bw[i] = w ? w[s] : 1.f;
                wsum += bw[i];
            }
            if (!(wsum > 0)) continue;
            int rc = upload(bx.data(), bt.data(), mask ? bm.data() : nullptr, bw.data(), batch);
            if (rc) return rc;
            rc = device_loss_grad(batch, mask != nullptr, true, (float)wsum, nullptr);
            if (rc) return rc;
            KT_HIP_TRY(dev::adam_step(P, G, m1, m2, total, lr, ++steps, stream));
        }
    }
    KT_HIP_TRY(hipStreamSynchronize(stream));
    if (loss_out) return loss_grad(x, t, mask, w, n, loss_out, nullptr);
    return KT_OK;
}

int Mlp::get(float* out, uint32_t n) const {
    if (n != total) return fail(KT_ERR_CAPACITY, "model has " + std::to_string(total) + " parameters");
    KT_HIP_TRY(hipMemcpy(out, P, total * 4, hipMemcpyDeviceToHost));
    return KT_OK;
}

int Mlp::set(const float* in, uint32_t n) {
    if (n != total) return fail(KT_ERR_CAPACITY, "model has " + std::to_string(total) + " parameters");
    KT_HIP_TRY(hipMemcpy(P, in, total * 4, hipMemcpyHostToDevice));
    return reset_adam();
}

int Mlp::set_range(const float* lo, const float* hi) {
    std::vector<float> r(2 * KT_MLP_IN);
    for (int i = 0; i < KT_MLP_IN; ++i) {
        r[i] = lo ? lo[i] : -FLT_MAX;
        r[KT_MLP_IN + i] = hi ? hi[i] : FLT_MAX;
        if (!(r[i] <= r[KT_MLP_IN + i])) return fail(KT_ERR_ARG, "input range has lo > hi");
    }
    KT_HIP_TRY(hipMemcpy(R, r.data(), r.size() * 4, hipMemcpyHostToDevice));
    return KT_OK;
}

int Mlp::get_range(float* lo, float* hi) const {
    std::vector<float> r(2 * KT_MLP_IN);
    KT_HIP_TRY(hipMemcpy(r.data(), R, r.size() * 4, hipMemcpyDeviceToHost));
    for (int i = 0; i < KT_MLP_IN; ++i) {
        if (lo) lo[i] = r[i];
        if (hi) hi[i] = r[KT_MLP_IN + i];
    }
    return KT_OK;
}

}  // namespace kt
