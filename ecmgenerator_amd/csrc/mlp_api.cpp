// mlp_api.cpp -- host runtime behind include/ecm_mlp.h: parameters,
// dataset, the training loop (forward, backward and Adam on the GPU with
// the kernels of mlp_kernels.hip) and host-data entry points for checks.
#include "mlp_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <vector>

using namespace ecm;
using namespace ecm::mlp;

static uint16_t f32_to_bf16_rne(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;          // NaN
    const uint32_t r = ((u >> 16) & 1u) + 0x7fffu;
    return (uint16_t)((u + r) >> 16);
}
static float bf16_to_f32(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

extern "C" int ecm_mlp_n_params(void) { return kNParams; }

extern "C" int ecm_mlp_create(uint64_t seed, ecm_mlp** out) {
    if (!out) return fail(ECM_E_ARG, "null argument");
    *out = nullptr;
    if (!ecm_device_available()) return fail(ECM_E_NOGPU, "no gfx950 device available");
    ecm_mlp* m = new ecm_mlp();
    (void)hipGetDevice(&m->dev);
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, m->dev) == hipSuccess && pr.multiProcessorCount > 0) m->n_cu = pr.multiProcessorCount;
    int rc = 0;
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) rc = fail(ECM_E_HIP, "stream create failed");
    if (!rc) rc = m->params.alloc(kNParams);
    if (!rc) rc = m->grads.alloc(kNParams);
    if (!rc) rc = m->adam_m.alloc(kNParams);
    if (!rc) rc = m->adam_v.alloc(kNParams);
    if (!rc) rc = m->wbf.alloc(kNParams);
    if (!rc) rc = m->frags.alloc((size_t)kNFrags * 64);
    if (!rc) rc = m->pbias.alloc(kNBias);
    if (!rc) rc = m->ds_count.alloc(1);
    if (rc) { ecm_mlp_destroy(m); return rc; }
    // seeded uniform(+-1/sqrt(fan_in)) init, the padding inputs of W1 zero
    std::vector<float> p(kNParams, 0.0f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int rows, int cols, int fan_in, int used_cols) {
        std::uniform_real_distribution<float> U(-1.0f / std::sqrt((float)fan_in), 1.0f / std::sqrt((float)fan_in));
        for (int r = 0; r < rows; ++r)
            for (int c = 0; c < cols; ++c) p[off + r * cols + c] = c < used_cols ? U(rng) : 0.0f;
    };
    fill(kOffW1, kHid, kInPad, kIn, kIn);
    fill(kOffB1, 1, kHid, kIn, kHid);
    fill(kOffW2, kHid, kHid, kHid, kHid);
    fill(kOffB2, 1, kHid, kHid, kHid);
    fill(kOffW3, kHid, kHid, kHid, kHid);
    fill(kOffB3, 1, kHid, kHid, kHid);
    fill(kOffW4, kOut, kHid, kHid, kHid);
    fill(kOffB4, 1, kOut, kHid, kOut);
    rc = ecm_mlp_set_params(m, p.data(), kNParams);
    if (rc) { ecm_mlp_destroy(m); return rc; }
    *out = m;
    return 0;
}

// A simulator step returns before its kernels finish, and a learned step
// reads the packed weights (frags, pbias) on the simulator's stream.  The
// calls that free or rewrite them therefore wait for the whole device.
extern "C" int ecm_mlp_destroy(ecm_mlp* m) {
    if (!m) return 0;
    (void)hipSetDevice(m->dev);
    (void)hipDeviceSynchronize();
    for (auto* b : {&m->params, &m->grads, &m->adam_m, &m->adam_v, &m->pbias, &m->z1, &m->z2, &m->z3, &m->y, &m->dyf,
                    &m->dh, &m->dzf, &m->loss})
        b->release();
    for (auto* b : {&m->wbf, &m->h1, &m->h2, &m->h3, &m->dy, &m->dz}) b->release();
    m->frags.release(); m->ds_x.release(); m->xb.release();
    m->ds_label.release(); m->ds_count.release(); m->lb.release();
    m->ds_v.release(); m->vb.release(); m->perm.release();
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return 0;
}

extern "C" int ecm_mlp_get_params(ecm_mlp* m, float* out, int n) {
    if (!m || !out || n != kNParams) return fail(ECM_E_ARG, "expected %d parameters", kNParams);
    (void)hipSetDevice(m->dev);
    HIPCHK(hipMemcpyAsync(out, m->params.p, kNParams * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return 0;
}

extern "C" int ecm_mlp_set_params(ecm_mlp* m, const float* in, int n) {
    if (!m || !in || n != kNParams) return fail(ECM_E_ARG, "expected %d parameters", kNParams);
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(in[i])) return fail(ECM_E_ARG, "parameter %d is not finite", i);
    (void)hipSetDevice(m->dev);
    HIPCHK(hipDeviceSynchronize());          // a simulator step may still read the packed weights
    HIPCHK(hipMemcpyAsync(m->params.p, in, kNParams * sizeof(float), hipMemcpyHostToDevice, m->stream));
    HIPCHK(launch_to_bf16(m->params.p, m->wbf.p, kNParams, m->stream));
    HIPCHK(launch_fill_f32(m->adam_m.p, 0.0f, kNParams, m->stream));
    HIPCHK(launch_fill_f32(m->adam_v.p, 0.0f, kNParams, m->stream));
    m->adam_t = 0;
    int rc = ecm_mlp_pack(m);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(m->stream));
    return 0;
}

int ecm_mlp_pack(ecm_mlp* m) {
    HIPCHK(launch_pack(m->params.p, m->frags.p, m->pbias.p, m->stream));
    return 0;
}

// ----------------------------------------------------------------- dataset
extern "C" int ecm_mlp_dataset_reserve(ecm_mlp* m, int64_t capacity) {
    if (!m || capacity < 1 || capacity > (int64_t)0x7fffffff) return fail(ECM_E_ARG, "bad capacity");
    (void)hipSetDevice(m->dev);
    int64_t n = 0;
    int rc = ecm_mlp_dataset_size(m, &n);
    if (rc) return rc;
    if (capacity <= m->ds_cap) return 0;
    ecm::DBuf<bf16x8_t> x;
    ecm::DBuf<uint32_t> l;
    ecm::DBuf<float2> v;
    rc = x.alloc((size_t)capacity * 8);
    if (!rc) rc = l.alloc((size_t)capacity);
    if (!rc) rc = v.alloc((size_t)capacity);
    if (rc) { x.release(); l.release(); v.release(); return rc; }
    if (n > 0) {
        HIPCHK(hipMemcpyAsync(x.p, m->ds_x.p, (size_t)n * 8 * sizeof(bf16x8_t), hipMemcpyDeviceToDevice, m->stream));
        HIPCHK(hipMemcpyAsync(l.p, m->ds_label.p, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToDevice, m->stream));
        HIPCHK(hipMemcpyAsync(v.p, m->ds_v.p, (size_t)n * sizeof(float2), hipMemcpyDeviceToDevice, m->stream));
    }
    HIPCHK(hipStreamSynchronize(m->stream));
    m->ds_x.release(); m->ds_label.release(); m->ds_v.release();
    m->ds_x = x; m->ds_label = l; m->ds_v = v;
    x.p = nullptr; l.p = nullptr; v.p = nullptr;
    m->ds_cap = capacity;
    if (n == 0) return ecm_mlp_dataset_clear(m);
    return 0;
}

extern "C" int ecm_mlp_dataset_size(ecm_mlp* m, int64_t* n) {
    if (!m || !n) return fail(ECM_E_ARG, "null argument");
    (void)hipSetDevice(m->dev);
    uint32_t c = 0;
    HIPCHK(hipMemcpyAsync(&c, m->ds_count.p, sizeof(uint32_t), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    *n = std::min<int64_t>(c, m->ds_cap);
    return 0;
}

extern "C" int ecm_mlp_dataset_clear(ecm_mlp* m) {
    if (!m) return fail(ECM_E_ARG, "null argument");
    (void)hipSetDevice(m->dev);
    const uint32_t z = 0;
    HIPCHK(hipMemcpyAsync(m->ds_count.p, &z, sizeof(uint32_t), hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return 0;
}

extern "C" int ecm_mlp_dataset_read(ecm_mlp* m, int64_t first, int n, float* x, uint32_t* label, float* v) {
    if (!m || n < 0 || first < 0) return fail(ECM_E_ARG, "bad arguments");
    int64_t have = 0;
    int rc = ecm_mlp_dataset_size(m, &have);
    if (rc) return rc;
    if (first + n > have) return fail(ECM_E_ARG, "samples [%lld, %lld) beyond the %lld logged", (long long)first,
                                      (long long)(first + n), (long long)have);
    if (n == 0) return 0;
    if (x) {
        std::vector<uint16_t> h((size_t)n * 64);
        HIPCHK(hipMemcpyAsync(h.data(), m->ds_x.p + first * 8, (size_t)n * 128, hipMemcpyDeviceToHost, m->stream));
        HIPCHK(hipStreamSynchronize(m->stream));
        for (size_t i = 0; i < h.size(); ++i) x[i] = bf16_to_f32(h[i]);
    }
    if (label) HIPCHK(hipMemcpyAsync(label, m->ds_label.p + first, (size_t)n * 4, hipMemcpyDeviceToHost, m->stream));
    if (v) HIPCHK(hipMemcpyAsync(v, m->ds_v.p + first, (size_t)n * 8, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return 0;
}

// ---------------------------------------------------------------- training
static int ensure_ws(ecm_mlp* m, int B) {
    if (m->ws_batch >= B) return 0;
    int rc = m->xb.alloc((size_t)B * 8);
    if (!rc) rc = m->lb.alloc(B);
    if (!rc) rc = m->vb.alloc(B);
    for (auto* b : {&m->z1, &m->z2, &m->z3, &m->dh, &m->dzf})
        if (!rc) rc = b->alloc((size_t)B * kHid);
    if (!rc) rc = m->y.alloc((size_t)B * kOut);
    if (!rc) rc = m->dyf.alloc((size_t)B * kOut);
    if (!rc) rc = m->loss.alloc(1);
    for (auto* b : {&m->h1, &m->h2, &m->h3, &m->dz})
        if (!rc) rc = b->alloc((size_t)B * kHid);
    if (!rc) rc = m->dy.alloc((size_t)B * kOut);
    if (rc) return rc;
    m->ws_batch = B;
    return 0;
}

// Forward and backward of the batch staged in xb / lb / vb; adds the
// gradient into grads and the batch loss into loss[0].
static int fwd_bwd(ecm_mlp* m, int B, float pos_w, float vel_w) {
    hipStream_t st = m->stream;
    const __bf16* X = reinterpret_cast<const __bf16*>(m->xb.p);
    const __bf16* W = m->wbf.p;
    const float* P = m->params.p;
    float* G = m->grads.p;
    const int sk = std::max(1, B / 512);
    // forward: Z = X W^T, H = silu(Z + b)
    HIPCHK(launch_gemm(B, kHid, kInPad, X, kInPad, 1, W + kOffW1, kInPad, 1, m->z1.p, kHid, 1, 0, 1, st));
    HIPCHK(launch_bias_silu(m->z1.p, P + kOffB1, m->h1.p, B, kHid, st));
    HIPCHK(launch_gemm(B, kHid, kHid, m->h1.p, kHid, 1, W + kOffW2, kHid, 1, m->z2.p, kHid, 1, 0, 1, st));
    HIPCHK(launch_bias_silu(m->z2.p, P + kOffB2, m->h2.p, B, kHid, st));
    HIPCHK(launch_gemm(B, kHid, kHid, m->h2.p, kHid, 1, W + kOffW3, kHid, 1, m->z3.p, kHid, 1, 0, 1, st));
    HIPCHK(launch_bias_silu(m->z3.p, P + kOffB3, m->h3.p, B, kHid, st));
    HIPCHK(launch_gemm(B, kOut, kHid, m->h3.p, kHid, 1, W + kOffW4, kHid, 1, m->y.p, kOut, 1, 0, 1, st));
    HIPCHK(launch_head(m->y.p, P + kOffB4, m->lb.p, m->vb.p, m->dy.p, m->dyf.p, m->loss.p, B, pos_w, vel_w, st));
    // layer 4: dW4 += dY^T H3, db4 += sum dY, dH3 = dY W4
    HIPCHK(launch_gemm(kOut, kHid, B, m->dy.p, 1, kOut, m->h3.p, 1, kHid, G + kOffW4, kHid, 1, 1, sk, st));
    HIPCHK(launch_colsum(m->dyf.p, B, kOut, G + kOffB4, st));
    HIPCHK(launch_gemm(B, kHid, kOut, m->dy.p, kOut, 1, W + kOffW4, 1, kHid, m->dh.p, kHid, 1, 0, 1, st));
    // hidden layers 3, 2, 1
    const __bf16* Hin[3] = {X, m->h1.p, m->h2.p};
    const float* Zs[3] = {m->z1.p, m->z2.p, m->z3.p};
    const int offW[3] = {kOffW1, kOffW2, kOffW3}, offB[3] = {kOffB1, kOffB2, kOffB3};
    const int nin[3] = {kInPad, kHid, kHid};
    for (int l = 2; l >= 0; --l) {
        HIPCHK(launch_silu_bwd(m->dh.p, Zs[l], m->dz.p, m->dzf.p, B, kHid, st));
        HIPCHK(launch_colsum(m->dzf.p, B, kHid, G + offB[l], st));
        HIPCHK(launch_gemm(kHid, nin[l], B, m->dz.p, 1, kHid, Hin[l], 1, nin[l], G + offW[l], nin[l], 1, 1, sk, st));
        if (l > 0)
            HIPCHK(launch_gemm(B, nin[l], kHid, m->dz.p, kHid, 1, W + offW[l], 1, nin[l], m->dh.p, kHid, 1, 0, 1, st));
    }
    return 0;
}

static int stage_host_batch(ecm_mlp* m, const float* x, const uint32_t* label, const float* v, int n) {
    std::vector<uint16_t> xb((size_t)n * 64);
    for (size_t i = 0; i < xb.size(); ++i) xb[i] = f32_to_bf16_rne(x[i]);
    HIPCHK(hipMemcpyAsync(m->xb.p, xb.data(), xb.size() * 2, hipMemcpyHostToDevice, m->stream));
    if (label) HIPCHK(hipMemcpyAsync(m->lb.p, label, (size_t)n * 4, hipMemcpyHostToDevice, m->stream));
    if (v) HIPCHK(hipMemcpyAsync(m->vb.p, v, (size_t)n * 8, hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));      // xb (host vector) must outlive the copy
    return 0;
}

extern "C" int ecm_mlp_loss_grad(ecm_mlp* m, const float* x, const uint32_t* label, const float* v, int n,
                                 float pos_weight, float vel_weight, float* loss, float* grad) {
    if (!m || !x || !label || !v || n < 1 || !loss || !grad) return fail(ECM_E_ARG, "bad arguments");
    (void)hipSetDevice(m->dev);
    int rc = ensure_ws(m, n);
    if (!rc) rc = stage_host_batch(m, x, label, v, n);
    if (rc) return rc;
    HIPCHK(launch_fill_f32(m->grads.p, 0.0f, kNParams, m->stream));
    HIPCHK(launch_fill_f32(m->loss.p, 0.0f, 1, m->stream));
    rc = fwd_bwd(m, n, pos_weight, vel_weight);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(loss, m->loss.p, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipMemcpyAsync(grad, m->grads.p, kNParams * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return 0;
}

extern "C" int ecm_mlp_forward(ecm_mlp* m, const float* x, int n, float* y) {
    if (!m || !x || !y || n < 1) return fail(ECM_E_ARG, "bad arguments");
    (void)hipSetDevice(m->dev);
    int rc = ensure_ws(m, n);
    if (!rc) rc = stage_host_batch(m, x, nullptr, nullptr, n);
    if (rc) return rc;
    HIPCHK(launch_mlp_infer(m->xb.p, 0, nullptr, n, m->frags.p, m->pbias.p, 0.0f, m->lb.p, nullptr, m->y.p, m->n_cu,
                            m->stream));
    HIPCHK(hipMemcpyAsync(y, m->y.p, (size_t)n * kOut * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return 0;
}

// ------------------------------------------- forward without allocation
// Scratch of ecm_mlp_forward_noalloc: the BF16 features ([n][64]) and the
// proposal masks ([n]) of up to ECM_MLP_NOALLOC_MAX_SAMPLES samples.  Every
// GPU holds its own copy until the process ends.  Nothing in it has to
// start at zero: the conversion kernel writes every feature the network
// reads and the network writes every mask.
constexpr size_t kNoAllocFeatBytes = (size_t)ECM_MLP_NOALLOC_MAX_SAMPLES * kInPad * 2;
constexpr size_t kNoAllocBytes = kNoAllocFeatBytes + (size_t)ECM_MLP_NOALLOC_MAX_SAMPLES * 4;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// Device address of `scratch` on device `dev`, looked up once per device
// with that device current (each GPU holds the array at its own address).
static int scratch_address(int dev, unsigned char** out) {
    static std::mutex mu;
    static std::vector<unsigned char*> addr;
    std::lock_guard<std::mutex> lock(mu);
    if (dev < 0) return fail(ECM_E_ARG, "bad device %d", dev);
    if ((size_t)dev >= addr.size()) addr.resize((size_t)dev + 1, nullptr);
    if (!addr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(ECM_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        addr[dev] = (unsigned char*)p;
    }
    *out = addr[dev];
    return 0;
}

extern "C" int ecm_mlp_forward_noalloc(ecm_mlp* m, const float* x, int n, float* y) {
    if (!m || !x || !y || n < 1) return fail(ECM_E_ARG, "bad arguments");
    if (n > ECM_MLP_NOALLOC_MAX_SAMPLES)
        return fail(ECM_E_ARG, "%d samples exceed the %d of the static scratch", n, ECM_MLP_NOALLOC_MAX_SAMPLES);
    HIPCHK(hipSetDevice(m->dev));
    unsigned char* base = nullptr;
    int rc = scratch_address(m->dev, &base);
    if (rc) return rc;
    uint16_t* xb = reinterpret_cast<uint16_t*>(base);
    uint32_t* mask = reinterpret_cast<uint32_t*>(base + kNoAllocFeatBytes);
    HIPCHK(launch_f32_to_bf16_rne(x, xb, (long)n * kInPad, m->stream));
    HIPCHK(launch_mlp_infer(xb, 0, nullptr, n, m->frags.p, m->pbias.p, 0.0f, mask, nullptr, y, m->n_cu, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return 0;
}

extern "C" int ecm_mlp_train(ecm_mlp* m, int epochs, int batch, float lr, float pos_weight, float vel_weight,
                             uint64_t seed, float* losses) {
    if (!m || epochs < 0 || batch < 64 || !(lr > 0.0f)) return fail(ECM_E_ARG, "bad training arguments");
    (void)hipSetDevice(m->dev);
    HIPCHK(hipDeviceSynchronize());          // a simulator step may still read the packed weights
    int64_t n = 0;
    int rc = ecm_mlp_dataset_size(m, &n);
    if (rc) return rc;
    if (n < batch) return fail(ECM_E_STATE, "dataset holds %lld samples, fewer than one batch of %d", (long long)n, batch);
    rc = ensure_ws(m, batch);
    if (!rc && m->perm.n < (size_t)n) rc = m->perm.alloc((size_t)n);
    if (rc) return rc;
    std::vector<int64_t> perm((size_t)n);
    for (int64_t i = 0; i < n; ++i) perm[i] = i;
    std::mt19937_64 rng(seed);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const int64_t nb = n / batch;
    for (int e = 0; e < epochs; ++e) {
        // Fisher-Yates with the seeded generator (reproducible across hosts)
        for (int64_t i = n - 1; i > 0; --i) {
            const int64_t j = (int64_t)(rng() % (uint64_t)(i + 1));
            std::swap(perm[i], perm[j]);
        }
        HIPCHK(hipMemcpyAsync(m->perm.p, perm.data(), (size_t)n * sizeof(int64_t), hipMemcpyHostToDevice, m->stream));
        HIPCHK(launch_fill_f32(m->loss.p, 0.0f, 1, m->stream));
        for (int64_t b = 0; b < nb; ++b) {
            HIPCHK(launch_gather(m->perm.p + b * batch, batch, m->ds_x.p, m->ds_label.p, m->ds_v.p, m->xb.p, m->lb.p,
                                 m->vb.p, m->stream));
            HIPCHK(launch_fill_f32(m->grads.p, 0.0f, kNParams, m->stream));
            rc = fwd_bwd(m, batch, pos_weight, vel_weight);
            if (rc) return rc;
            ++m->adam_t;
            const float bc1 = 1.0f - std::pow(b1, (float)m->adam_t), bc2 = 1.0f - std::pow(b2, (float)m->adam_t);
            HIPCHK(launch_adam(m->params.p, m->grads.p, m->adam_m.p, m->adam_v.p, kNParams, lr, b1, b2, eps, bc1, bc2,
                               m->wbf.p, m->stream));
        }
        float L = 0.0f;
        HIPCHK(hipMemcpyAsync(&L, m->loss.p, sizeof(float), hipMemcpyDeviceToHost, m->stream));
        HIPCHK(hipStreamSynchronize(m->stream));      // also keeps `perm` alive for the upload
        if (losses) losses[e] = L / (float)nb;
    }
    rc = ecm_mlp_pack(m);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(m->stream));
    return 0;
}
