// mlp_host.h -- host-side state of an ecm_mlp handle, shared by the MLP
// runtime (mlp_api.cpp) and the step (sim_api.cpp), plus the kernel
// launchers of mlp_kernels.hip.
#pragma once

#include "../../include/ecm_mlp.h"
#include "host_util.h"
#include "mlp.h"

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ecm {
namespace mlp {

// feat: [n][8] chunks of 8 BF16 per agent, or with `tiled` chunk c of agent
// i at ((i / 32) * 8 + c) * 32 + i % 32.
hipError_t launch_mlp_infer(const void* feat, int tiled, const uint32_t* n_dev, int n_host, const void* frags,
                            const float* bias, float theta, uint32_t* mask, float2* vel, float* raw, int n_cu,
                            hipStream_t st);
hipError_t launch_gemm(int M, int N, int K, const __bf16* A, long sam, long sak, const __bf16* B, long sbn, long sbk,
                       float* C, long scm, long scn, int accumulate, int splitk, hipStream_t st);
hipError_t launch_bias_silu(float* Z, const float* bias, __bf16* H, int B, int N, hipStream_t st);
hipError_t launch_silu_bwd(const float* dH, const float* Z, __bf16* dZ, float* dZf, int B, int N, hipStream_t st);
hipError_t launch_head(float* Y, const float* bias, const uint32_t* lab, const float2* vt, __bf16* dY, float* dYf,
                       float* loss, int B, float pos_w, float vel_w, hipStream_t st);
hipError_t launch_colsum(const float* X, int B, int N, float* out, hipStream_t st);
hipError_t launch_adam(float* w, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                       float bc1, float bc2, __bf16* wb, hipStream_t st);
hipError_t launch_to_bf16(const float* w, __bf16* wb, int n, hipStream_t st);
hipError_t launch_f32_to_bf16_rne(const float* x, uint16_t* out, long n, hipStream_t st);
hipError_t launch_fill_f32(float* p, float v, long n, hipStream_t st);
hipError_t launch_gather(const int64_t* perm, int B, const void* X, const uint32_t* L, const float2* V, void* Xb,
                         uint32_t* Lb, float2* Vb, hipStream_t st);
hipError_t launch_pack(const float* P, void* frags, float* bias, hipStream_t st);

}  // namespace mlp
}  // namespace ecm

// 16 bytes = 8 BF16 values
struct bf16x8_t { uint32_t w[4]; };

struct ecm_mlp {
    int dev = 0;
    hipStream_t stream = nullptr;
    int n_cu = 256;
    // parameters: FP32 master, BF16 row-major copy (training GEMMs),
    // packed inference fragments and biases
    ecm::DBuf<float> params, grads, adam_m, adam_v;
    ecm::DBuf<__bf16> wbf;
    ecm::DBuf<bf16x8_t> frags;
    ecm::DBuf<float> pbias;
    int64_t adam_t = 0;
    // dataset
    ecm::DBuf<bf16x8_t> ds_x;      // [cap][8]
    ecm::DBuf<uint32_t> ds_label;  // [cap]
    ecm::DBuf<float2> ds_v;        // [cap]
    ecm::DBuf<uint32_t> ds_count;  // [1] device append counter
    int64_t ds_cap = 0;
    // training workspace for batch size ws_batch
    int ws_batch = 0;
    ecm::DBuf<int64_t> perm;
    ecm::DBuf<bf16x8_t> xb;
    ecm::DBuf<uint32_t> lb;
    ecm::DBuf<float2> vb;
    ecm::DBuf<float> z1, z2, z3, y, dyf, dh, dzf, loss;
    ecm::DBuf<__bf16> h1, h2, h3, dy, dz;
};

// Packs the current parameters into the inference layout (stream-ordered).
int ecm_mlp_pack(ecm_mlp* m);
