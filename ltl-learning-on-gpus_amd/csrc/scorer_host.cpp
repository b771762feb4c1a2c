// scorer_host.cpp -- the scorer's C ABI: creation with seeded weights,
// parameter access, forward, gradient and Adam training on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/ltl_amd.h"
#include "scorer.hpp"

using namespace ltl;

static_assert(LTL_SCORER_PARAMS == P_N, "parameter layout");
static_assert(LTL_SCORER_FEATURES == kFeat, "feature count");

namespace {

int32_t fail(hipError_t e) { return e == hipErrorOutOfMemory ? LTL_E_OOM : LTL_E_HIP; }
#define SC_TRY(x)                                  \
    do {                                           \
        hipError_t e_ = (x);                       \
        if (e_ != hipSuccess) return fail(e_);     \
    } while (0)

uint64_t mix(uint64_t &s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
float unit(uint64_t &s) { return (float)((mix(s) >> 40) * (1.0 / 16777216.0)); }  // [0, 1)

// Uniform fan-in initialisation (He et al. 2015, uniform form): hidden
// weights U(-sqrt(6 / fan_in), +), output weights U(-sqrt(3 / 128), +),
// hidden biases 0, output bias 0.5 (labels are accuracies in [0, 1]).
void init_params(uint64_t seed, std::vector<float> &p) {
    p.assign(P_N, 0.f);
    uint64_t s = seed;
    const float a1 = std::sqrt(6.f / kFeat), a2 = std::sqrt(6.f / kHid), a3 = std::sqrt(3.f / kHid);
    for (int h = 0; h < kHid; ++h)
        for (int f = 0; f < kIn; ++f) p[P_W1 + h * kIn + f] = f < kFeat ? (2.f * unit(s) - 1.f) * a1 : 0.f;
    for (int i = 0; i < kHid * kHid; ++i) p[P_W2 + i] = (2.f * unit(s) - 1.f) * a2;
    for (int i = 0; i < kHid; ++i) p[P_W3 + i] = (2.f * unit(s) - 1.f) * a3;
    p[P_B3] = 0.5f;
}

int32_t upload_params(ltl_scorer *sc, const float *p) {
    SC_TRY(hipMemcpyAsync(sc->m.p, p, P_N * sizeof(float), hipMemcpyHostToDevice, sc->stream));
    SC_TRY(hipMemsetAsync(sc->m.g, 0, (size_t)3 * P_N * sizeof(float), sc->stream));  // g, m, v
    SC_TRY(launch_refresh(sc->m, sc->stream));
    SC_TRY(hipStreamSynchronize(sc->stream));
    sc->step = 0;
    return 0;
}

int32_t ensure(float *&buf, int64_t &cap, int64_t n) {
    if (n <= cap) return 0;
    if (buf) hipFree(buf);
    buf = nullptr;
    cap = 0;
    SC_TRY(hipMalloc((void **)&buf, (size_t)n * sizeof(float)));
    cap = n;
    return 0;
}

}  // namespace

extern "C" ltl_scorer *ltl_scorer_create(int32_t device, uint64_t seed, int32_t *status) {
    auto out = [&](int32_t s) { if (status) *status = s; };
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) { out(LTL_E_NODEVICE); return nullptr; }
    if (device < 0 || device >= n) { out(LTL_E_ARG); return nullptr; }
    if (hipSetDevice(device) != hipSuccess) { out(LTL_E_HIP); return nullptr; }
    auto *sc = new ltl_scorer;
    sc->device = device;
    const size_t nf = (size_t)4 * P_N + 64;
    const size_t nb = (size_t)kHid * kIn + 2 * (size_t)kHid * kHid;
    if (hipStreamCreateWithFlags(&sc->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void **)&sc->d_f, nf * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&sc->d_b, nb * sizeof(uint16_t)) != hipSuccess) {
        ltl_scorer_destroy(sc);
        out(LTL_E_OOM);
        return nullptr;
    }
    sc->m.p = sc->d_f;
    sc->m.g = sc->d_f + P_N;
    sc->m.m = sc->d_f + 2 * P_N;
    sc->m.v = sc->d_f + 3 * P_N;
    sc->m.loss = sc->d_f + 4 * P_N;
    sc->m.w1b = sc->d_b;
    sc->m.w2b = sc->d_b + kHid * kIn;
    sc->m.w2tb = sc->d_b + kHid * kIn + kHid * kHid;
    std::vector<float> p;
    init_params(seed, p);
    const int32_t r = upload_params(sc, p.data());
    if (r) {
        ltl_scorer_destroy(sc);
        out(r);
        return nullptr;
    }
    out(0);
    return sc;
}

extern "C" void ltl_scorer_destroy(ltl_scorer *sc) {
    if (!sc) return;
    hipSetDevice(sc->device);
    if (sc->stream) hipStreamSynchronize(sc->stream);
    if (sc->d_f) hipFree(sc->d_f);
    if (sc->d_b) hipFree(sc->d_b);
    if (sc->d_x) hipFree(sc->d_x);
    if (sc->d_y) hipFree(sc->d_y);
    if (sc->stream) hipStreamDestroy(sc->stream);
    delete sc;
}

extern "C" int32_t ltl_scorer_get_params(ltl_scorer *sc, float *params) {
    if (!sc || !params) return LTL_E_ARG;
    SC_TRY(hipSetDevice(sc->device));
    SC_TRY(hipMemcpyAsync(params, sc->m.p, P_N * sizeof(float), hipMemcpyDeviceToHost, sc->stream));
    SC_TRY(hipStreamSynchronize(sc->stream));
    return 0;
}

extern "C" int32_t ltl_scorer_set_params(ltl_scorer *sc, const float *params) {
    if (!sc || !params) return LTL_E_ARG;
    for (int i = 0; i < P_N; ++i)
        if (!std::isfinite(params[i])) return LTL_E_ARG;
    SC_TRY(hipSetDevice(sc->device));
    return upload_params(sc, params);
}

extern "C" int32_t ltl_scorer_forward(ltl_scorer *sc, const float *x, int64_t n, float *y) {
    if (!sc || n < 0 || (n && (!x || !y))) return LTL_E_ARG;
    if (!n) return 0;
    SC_TRY(hipSetDevice(sc->device));
    if (int32_t r = ensure(sc->d_x, sc->cap_x, n * kFeat)) return r;
    if (int32_t r = ensure(sc->d_y, sc->cap_y, n)) return r;
    SC_TRY(hipMemcpyAsync(sc->d_x, x, (size_t)n * kFeat * sizeof(float), hipMemcpyHostToDevice, sc->stream));
    SC_TRY(launch_forward(sc->m, sc->d_x, kFeat, n, sc->d_y, sc->stream));
    SC_TRY(hipMemcpyAsync(y, sc->d_y, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, sc->stream));
    SC_TRY(hipStreamSynchronize(sc->stream));
    return 0;
}

extern "C" int32_t ltl_scorer_grad(ltl_scorer *sc, const float *x, const float *t, int64_t n, float *grad,
                                   float *loss) {
    if (!sc || n <= 0 || !x || !t || !grad) return LTL_E_ARG;
    SC_TRY(hipSetDevice(sc->device));
    if (int32_t r = ensure(sc->d_x, sc->cap_x, n * kFeat)) return r;
    if (int32_t r = ensure(sc->d_y, sc->cap_y, n)) return r;
    hipStream_t s = sc->stream;
    SC_TRY(hipMemcpyAsync(sc->d_x, x, (size_t)n * kFeat * sizeof(float), hipMemcpyHostToDevice, s));
    SC_TRY(hipMemcpyAsync(sc->d_y, t, (size_t)n * sizeof(float), hipMemcpyHostToDevice, s));
    SC_TRY(launch_zero(sc->m.g, P_N, s));
    SC_TRY(launch_zero(sc->m.loss, 1, s));
    SC_TRY(launch_grad(sc->m, sc->d_x, kFeat, sc->d_y, 1, n, 1.f / (float)n, s));
    SC_TRY(hipMemcpyAsync(grad, sc->m.g, P_N * sizeof(float), hipMemcpyDeviceToHost, s));
    float l = 0.f;
    SC_TRY(hipMemcpyAsync(&l, sc->m.loss, sizeof(float), hipMemcpyDeviceToHost, s));
    SC_TRY(launch_zero(sc->m.g, P_N, s));
    SC_TRY(hipStreamSynchronize(s));
    if (loss) *loss = l;
    return 0;
}

extern "C" int32_t ltl_scorer_train(ltl_scorer *sc, const float *rows, int64_t n, int32_t epochs, int32_t batch,
                                    float lr, uint64_t seed, float *loss) {
    if (!sc || !rows || n <= 0 || epochs < 1 || batch < 1 || !(lr > 0.f)) return LTL_E_ARG;
    SC_TRY(hipSetDevice(sc->device));
    constexpr int R = kFeat + 1;
    // Seeded shuffle of the rows (Fisher-Yates), uploaded once.
    std::vector<int64_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    uint64_t s = seed;
    for (int64_t i = n - 1; i > 0; --i) std::swap(idx[i], idx[mix(s) % (uint64_t)(i + 1)]);
    std::vector<float> sh((size_t)n * R);
    for (int64_t i = 0; i < n; ++i) memcpy(&sh[(size_t)i * R], rows + (size_t)idx[i] * R, R * sizeof(float));
    if (int32_t r = ensure(sc->d_x, sc->cap_x, n * R)) return r;
    hipStream_t st = sc->stream;
    SC_TRY(hipMemcpyAsync(sc->d_x, sh.data(), sh.size() * sizeof(float), hipMemcpyHostToDevice, st));
    SC_TRY(launch_zero(sc->m.g, P_N, st));
    const int64_t nb = (n + batch - 1) / batch;
    float last = 0.f;
    for (int e = 0; e < epochs; ++e) {
        SC_TRY(launch_zero(sc->m.loss, 1, st));
        for (int64_t b = 0; b < nb; ++b) {
            const int64_t r0 = b * batch, cnt = std::min<int64_t>(batch, n - r0);
            const float *x = sc->d_x + (size_t)r0 * R;
            SC_TRY(launch_grad(sc->m, x, R, x + kFeat, R, cnt, 1.f / (float)cnt, st));
            SC_TRY(launch_adam(sc->m, lr, 0.9f, 0.999f, 1e-8f, ++sc->step, st));
        }
        if (e == epochs - 1) {
            SC_TRY(hipMemcpyAsync(&last, sc->m.loss, sizeof(float), hipMemcpyDeviceToHost, st));
        }
    }
    SC_TRY(hipStreamSynchronize(st));
    if (loss) *loss = last / (float)nb;
    return 0;
}
