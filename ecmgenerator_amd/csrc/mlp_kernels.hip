// mlp_kernels.hip -- gfx950 kernels of the learned ORCA proposal network:
// inference (persistent, weights resident in LDS, one 32-agent MFMA tile
// per wavefront iteration) and training (a strided BF16 MFMA GEMM plus the
// elementwise, loss, Adam and packing kernels around it).
#include "mlp.h"

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

namespace ecm {
namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// SiLU with the hardware exp2 / reciprocal (both ~1 ulp); the training
// kernels use the same expression.
constexpr float kLog2e = 1.4426950408889634f;
__device__ __forceinline__ float sigmoid_f(float z) { return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-z * kLog2e)); }
__device__ __forceinline__ float silu_f(float z) { return z * sigmoid_f(z); }

// accumulator <- bias of rows 8 j + 4 g + r (register 4 j + r)
__device__ __forceinline__ void bias_init(f32x16& acc, const float* b, int g) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float4 v = *reinterpret_cast<const float4*>(b + 8 * j + 4 * g);
        acc[4 * j + 0] = v.x;
        acc[4 * j + 1] = v.y;
        acc[4 * j + 2] = v.z;
        acc[4 * j + 3] = v.w;
    }
}

// SiLU of one m-tile, packed as B-operand chunks 2 mi and 2 mi + 1 of the
// next layer
__device__ __forceinline__ void act_pack(const f32x16& acc, bf16x8& lo, bf16x8& hi) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        lo[e] = (__bf16)silu_f(acc[e]);
        hi[e] = (__bf16)silu_f(acc[8 + e]);
    }
}

// One 128-wide hidden layer for a 32-agent tile: m-tiles one after the
// other (one live accumulator), inputs `in`, outputs `out`.
template <int NKC>
__device__ __forceinline__ void hidden_layer(const bf16x8* sw, int frag0, const float* sb, int g, int lane,
                                             const bf16x8 (&in)[NKC], bf16x8 (&out)[8]) {
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
        f32x16 acc;
        bias_init(acc, sb + 32 * mi, g);
#pragma unroll
        for (int kc = 0; kc < NKC; ++kc) acc = mfma(sw[(frag0 + mi * NKC + kc) * 64 + lane], in[kc], acc);
        act_pack(acc, out[2 * mi], out[2 * mi + 1]);
    }
}

struct InferArgs {
    const bf16x8* feat;       // 64 bf16 features per agent: [n][8], or tiled (below)
    int tiled;                // chunk c of agent i at ((i / 32) * 8 + c) * 32 + i % 32
    const uint32_t* n_dev;    // agent count on the device, or null
    int n_host;               // agent count when n_dev is null
    const bf16x8* frags;      // [kNFrags][64] packed A fragments
    const float* bias;        // [kNBias]
    float theta;              // slot predicted active when its logit > theta
    uint32_t* mask;           // [n] predicted active slots
    float2* vel;              // [n] velocity head (m/s), or null
    float* raw;               // [n][16] raw outputs, or null
};

// 1024 threads = 16 wavefronts per workgroup, one workgroup per CU: the 88
// KiB of fragments are loaded into LDS once and every wavefront loops over
// 32-agent tiles.  All 64 lanes execute every MFMA (tile bounds only gate
// loads and stores).
__global__ __launch_bounds__(1024) void k_mlp_infer(InferArgs a) {
    __shared__ bf16x8 sw[kNFrags * 64];
    __shared__ float sb[kNBias];
    for (int i = threadIdx.x; i < kNFrags * 64; i += blockDim.x) sw[i] = a.frags[i];
    for (int i = threadIdx.x; i < kNBias; i += blockDim.x) sb[i] = a.bias[i];
    __syncthreads();
    const int n = a.n_dev ? (int)*a.n_dev : a.n_host;
    const int lane = threadIdx.x & 63, g = lane >> 5, col = lane & 31;
    const int waves = blockDim.x >> 6;
    const int ntiles = (n + 31) >> 5;
    bf16x8 zero;
#pragma unroll
    for (int e = 0; e < 8; ++e) zero[e] = (__bf16)0.0f;
    for (int tile = blockIdx.x * waves + (threadIdx.x >> 6); tile < ntiles; tile += gridDim.x * waves) {
        // the weights are re-read from LDS every tile: without this barrier
        // the compiler hoists all 88 fragments out of the loop and spills
        __asm__ volatile("" ::: "memory");
        const int agent = tile * 32 + col;
        const bool ok = agent < n;
        // layer 1: 64 inputs straight from HBM as B operands
        bf16x8 x[4], h[8], h2[8];
#pragma unroll
        for (int kc = 0; kc < 4; ++kc) {
            const size_t at = a.tiled ? ((size_t)tile * 8 + 2 * kc + g) * 32 + col : (size_t)agent * 8 + 2 * kc + g;
            x[kc] = ok ? a.feat[at] : zero;
        }
        hidden_layer<4>(sw, kFragL1, sb + kBiasL1, g, lane, x, h);
        hidden_layer<8>(sw, kFragL2, sb + kBiasL2, g, lane, h, h2);
        hidden_layer<8>(sw, kFragL3, sb + kBiasL3, g, lane, h2, h);
        // layer 4: 16 outputs in rows 0..15 of one padded tile
        f32x16 o;
        bias_init(o, sb + kBiasL4, g);
#pragma unroll
        for (int kc = 0; kc < 8; ++kc) o = mfma(sw[(kFragL4 + kc) * 64 + lane], h[kc], o);
        // lane pair (L, L + 32) holds rows {0-3, 8-11} and {4-7, 12-15}
        uint32_t m = 0;
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 8 * j + 4 * g + r;
                if (row < kSlots && o[4 * j + r] > a.theta) m |= 1u << row;
            }
        m |= __shfl_xor(m, 32);
        if (ok) {
            if (g == 0) a.mask[agent] = m;
            else if (a.vel) a.vel[agent] = make_float2(o[6] * (1.0f / kVelScale), o[7] * (1.0f / kVelScale));
            if (a.raw) {
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r) a.raw[(size_t)agent * 16 + 8 * j + 4 * g + r] = o[4 * j + r];
            }
        }
    }
}

// ------------------------------------------------------------- training

struct GemmArgs {
    int M, N, K, kchunk;      // kchunk: K range per blockIdx.z (multiple of 16)
    const __bf16* A; long sam, sak;
    const __bf16* B; long sbn, sbk;
    float* C; long scm, scn;
    int accumulate;           // C += (else C =); with gridDim.z > 1, atomic adds
};

// C[m][n] (+)= sum_k A[m sam + k sak] * B[n sbn + k sbk]: 64x64 output tile
// per 256-thread workgroup (four 32x32 MFMA tiles), K in steps of 16
// staged through LDS with arbitrary strides (training-size problems).
__global__ __launch_bounds__(256) void k_gemm(GemmArgs p) {
    __shared__ __attribute__((aligned(16))) __bf16 As[64 * 24];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[64 * 24];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wm = w >> 1, wn = w & 1;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int k0 = blockIdx.z * p.kchunk, k1 = min(p.K, k0 + p.kchunk);
    f32x16 acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
    const __bf16 z = (__bf16)0.0f;
    for (int k = k0; k < k1; k += 16) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int e = tid * 4 + i, r = e >> 4, kk = e & 15;
            const int gk = k + kk;
            const int gm = m0 + r, gn = n0 + r;
            As[r * 24 + kk] = (gm < p.M && gk < k1) ? p.A[(long)gm * p.sam + (long)gk * p.sak] : z;
            Bs[r * 24 + kk] = (gn < p.N && gk < k1) ? p.B[(long)gn * p.sbn + (long)gk * p.sbk] : z;
        }
        __syncthreads();
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[(32 * wm + (lane & 31)) * 24 + 8 * (lane >> 5)]);
        const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[(32 * wn + (lane & 31)) * 24 + 8 * (lane >> 5)]);
        acc = mfma(a, b, acc);
        __syncthreads();
    }
    const bool atomic = gridDim.z > 1;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int row = m0 + 32 * wm + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        const int col = n0 + 32 * wn + (lane & 31);
        if (row < p.M && col < p.N) {
            float* c = p.C + (long)row * p.scm + (long)col * p.scn;
            if (atomic) atomicAdd(c, acc[v]);
            else if (p.accumulate) *c += acc[v];
            else *c = acc[v];
        }
    }
}

// Z[b][n] += bias[n]; H = bf16(silu(Z))
__global__ void k_bias_silu(float* Z, const float* bias, __bf16* H, int B, int N) {
    const long n = (long)B * N;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float z = Z[i] + bias[i % N];
        Z[i] = z;
        H[i] = (__bf16)silu_f(z);
    }
}

// dZ = dH * silu'(Z), as bf16 (next GEMMs) and fp32 (bias gradient)
__global__ void k_silu_bwd(const float* dH, const float* Z, __bf16* dZ, float* dZf, int B, int N) {
    const long n = (long)B * N;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float z = Z[i];
        const float s = sigmoid_f(z);
        const float d = dH[i] * (s * (1.0f + z * (1.0f - s)));
        dZ[i] = (__bf16)d;
        dZf[i] = d;
    }
}

// Output head: Y[b][0..13] logits (BCE with a positive-class weight over
// the slots that exist), Y[b][14..15] velocity (squared error against the
// scaled LP solution).  Writes dY (mean over the batch) and adds the loss.
__global__ void k_head(float* Y, const float* bias, const uint32_t* lab, const float2* vt, __bf16* dY, float* dYf,
                       float* loss, int B, float pos_w, float vel_w) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (b < B) {
        const uint32_t L = lab[b];
        const float invB = 1.0f / (float)B;
        for (int c = 0; c < kOut; ++c) {
            const float y = Y[(long)b * kOut + c] + bias[c];
            Y[(long)b * kOut + c] = y;
            float d = 0.0f;
            if (c < kSlots) {
                if ((L >> (kLabelValidShift + c)) & 1u) {
                    const float t = (float)((L >> c) & 1u);
                    const float sp = fmaxf(y, 0.0f) + log1pf(expf(-fabsf(y)));      // softplus(y)
                    l += t * pos_w * (sp - y) + (1.0f - t) * sp;
                    const float s = sigmoid_f(y);
                    d = (s * (1.0f - t + pos_w * t) - pos_w * t) * invB;
                }
            } else {
                const float tv = (c == kSlots ? vt[b].x : vt[b].y) * kVelScale;
                const float e = y - tv;
                l += 0.5f * vel_w * e * e;
                d = vel_w * e * invB;
            }
            dY[(long)b * kOut + c] = (__bf16)d;
            dYf[(long)b * kOut + c] = d;
        }
        l *= invB;
    }
    // one atomic per wavefront
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, l);
}

// out[n] += sum_b X[b][n]
__global__ void k_colsum(const float* X, int B, int N, float* out) {
    const int n = blockIdx.x;
    float s = 0.0f;
    for (int b = threadIdx.x; b < B; b += blockDim.x) s += X[(long)b * N + n];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    __shared__ float part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.0f;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += part[i];
        out[n] += t;
    }
}

// Adam on the FP32 master parameters; refreshes the row-major BF16 copy.
__global__ void k_adam(float* w, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                       float bc1, float bc2, __bf16* wb) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
        w[i] = wi;
        wb[i] = (__bf16)wi;
    }
}

__global__ void k_to_bf16(const float* w, __bf16* wb, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) wb[i] = (__bf16)w[i];
}

// FP32 -> BF16, round to nearest even, NaN -> quiet NaN: the same bit
// arithmetic as the host staging of ecm_mlp_forward, so both paths feed
// the network identical operands.
__global__ void k_f32_to_bf16_rne(const float* x, uint16_t* out, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const uint32_t u = __float_as_uint(x[i]);
        out[i] = (u & 0x7fffffffu) > 0x7f800000u ? (uint16_t)0x7fc0 : (uint16_t)((u + (((u >> 16) & 1u) + 0x7fffu)) >> 16);
    }
}

__global__ void k_fill_f32(float* p, float v, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

// Batch gather through a permutation of the dataset.
__global__ void k_gather(const int64_t* perm, int B, const bf16x8* X, const uint32_t* L, const float2* V, bf16x8* Xb,
                         uint32_t* Lb, float2* Vb) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < B * 8; i += gridDim.x * blockDim.x) {
        const int b = i >> 3, c = i & 7;
        const int64_t src = perm[b];
        Xb[i] = X[src * 8 + c];
        if (c == 0) {
            Lb[b] = L[src];
            Vb[b] = V[src];
        }
    }
}

// Inference packing of the FP32 parameters (fragment layout in mlp.h).
__global__ void k_pack(const float* P, bf16x8* frags, float* bias) {
    const int f = blockIdx.x, lane = threadIdx.x;          // one fragment per block, 64 lanes
    int mi, kc, nin;
    const float* W;
    bool hidden_in;
    int rows;
    if (f < kFragL2) { const int q = f - kFragL1; mi = q / 4; kc = q % 4; W = P + kOffW1; nin = kInPad; hidden_in = false; rows = kHid; }
    else if (f < kFragL3) { const int q = f - kFragL2; mi = q / 8; kc = q % 8; W = P + kOffW2; nin = kHid; hidden_in = true; rows = kHid; }
    else if (f < kFragL4) { const int q = f - kFragL3; mi = q / 8; kc = q % 8; W = P + kOffW3; nin = kHid; hidden_in = true; rows = kHid; }
    else { const int q = f - kFragL4; mi = 0; kc = q; W = P + kOffW4; nin = kHid; hidden_in = true; rows = kOut; }
    const int row = 32 * mi + (lane & 31);
    bf16x8 v;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int k = 16 * kc + 8 * (lane >> 5) + e;
        const int col = hidden_in ? hidden_perm(k) : k;
        v[e] = (__bf16)(row < rows ? W[row * nin + col] : 0.0f);
    }
    frags[f * 64 + lane] = v;
    if (f == 0) {
        for (int i = lane; i < kNBias; i += 64) {
            float b = 0.0f;
            if (i < kBiasL2) b = P[kOffB1 + i];
            else if (i < kBiasL3) b = P[kOffB2 + i - kBiasL2];
            else if (i < kBiasL4) b = P[kOffB3 + i - kBiasL3];
            else if (i - kBiasL4 < kOut) b = P[kOffB4 + i - kBiasL4];
            bias[i] = b;
        }
    }
}

// ------------------------------------------------------------- launchers
static inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

hipError_t launch_mlp_infer(const void* feat, int tiled, const uint32_t* n_dev, int n_host, const void* frags,
                            const float* bias, float theta, uint32_t* mask, float2* vel, float* raw, int n_cu,
                            hipStream_t st) {
    InferArgs a{(const bf16x8*)feat, tiled, n_dev, n_host, (const bf16x8*)frags, bias, theta, mask, vel, raw};
    hipLaunchKernelGGL(k_mlp_infer, dim3(n_cu), dim3(1024), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_f32_to_bf16_rne(const float* x, uint16_t* out, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_f32_to_bf16_rne, dim3((unsigned)std::min<long>(cdiv(n, 256), 4096)), dim3(256), 0, st, x, out, n);
    return hipGetLastError();
}

hipError_t launch_gemm(int M, int N, int K, const __bf16* A, long sam, long sak, const __bf16* B, long sbn, long sbk,
                       float* C, long scm, long scn, int accumulate, int splitk, hipStream_t st) {
    if (splitk < 1) splitk = 1;
    int kchunk = ((cdiv(K, splitk) + 15) / 16) * 16;
    if (kchunk < 16) kchunk = 16;
    const int gz = cdiv(K, kchunk);
    GemmArgs p{M, N, K, kchunk, A, sam, sak, B, sbn, sbk, C, scm, scn, accumulate};
    hipLaunchKernelGGL(k_gemm, dim3(cdiv(N, 64), cdiv(M, 64), gz), dim3(256), 0, st, p);
    return hipGetLastError();
}

hipError_t launch_bias_silu(float* Z, const float* bias, __bf16* H, int B, int N, hipStream_t st) {
    hipLaunchKernelGGL(k_bias_silu, dim3(std::min(cdiv((long)B * N, 256), 4096)), dim3(256), 0, st, Z, bias, H, B, N);
    return hipGetLastError();
}

hipError_t launch_silu_bwd(const float* dH, const float* Z, __bf16* dZ, float* dZf, int B, int N, hipStream_t st) {
    hipLaunchKernelGGL(k_silu_bwd, dim3(std::min(cdiv((long)B * N, 256), 4096)), dim3(256), 0, st, dH, Z, dZ, dZf, B, N);
    return hipGetLastError();
}

hipError_t launch_head(float* Y, const float* bias, const uint32_t* lab, const float2* vt, __bf16* dY, float* dYf,
                       float* loss, int B, float pos_w, float vel_w, hipStream_t st) {
    hipLaunchKernelGGL(k_head, dim3(cdiv(B, 256)), dim3(256), 0, st, Y, bias, lab, vt, dY, dYf, loss, B, pos_w, vel_w);
    return hipGetLastError();
}

hipError_t launch_colsum(const float* X, int B, int N, float* out, hipStream_t st) {
    hipLaunchKernelGGL(k_colsum, dim3(N), dim3(256), 0, st, X, B, N, out);
    return hipGetLastError();
}

hipError_t launch_adam(float* w, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                       float bc1, float bc2, __bf16* wb, hipStream_t st) {
    hipLaunchKernelGGL(k_adam, dim3(cdiv(n, 256)), dim3(256), 0, st, w, g, m, v, n, lr, b1, b2, eps, bc1, bc2, wb);
    return hipGetLastError();
}

hipError_t launch_to_bf16(const float* w, __bf16* wb, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_to_bf16, dim3(cdiv(n, 256)), dim3(256), 0, st, w, wb, n);
    return hipGetLastError();
}

hipError_t launch_fill_f32(float* p, float v, long n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_f32, dim3(std::min(cdiv(n, 256), 4096)), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_gather(const int64_t* perm, int B, const void* X, const uint32_t* L, const float2* V, void* Xb,
                         uint32_t* Lb, float2* Vb, hipStream_t st) {
    hipLaunchKernelGGL(k_gather, dim3(cdiv((long)B * 8, 256)), dim3(256), 0, st, perm, B, (const bf16x8*)X, L, V,
                       (bf16x8*)Xb, Lb, Vb);
    return hipGetLastError();
}

hipError_t launch_pack(const float* P, void* frags, float* bias, hipStream_t st) {
    hipLaunchKernelGGL(k_pack, dim3(kNFrags), dim3(64), 0, st, P, (bf16x8*)frags, bias);
    return hipGetLastError();
}

}  // namespace mlp
}  // namespace ecm
