// mlp.h -- the learned ORCA proposal network: layout constants, the
// per-agent feature vector, and the MFMA fragment layout shared by the
// inference kernel, the training kernels and the host-side packing.
//
// Network: 61 inputs (padded to 64) -> 128 -> 128 -> 128 -> 16, SiLU after
// each hidden layer, BF16 weights and activations, FP32 accumulation.
// Outputs: 14 active-constraint logits (slots 0..9: the agent constraints
// of neighbours 0..9 in (d2, id) order; slots 10..13: the constraints of the
// 4 nearest obstacle edges in range) and a 2D velocity (scaled).
//
// Features (index: meaning, all relative to the agent, scaled, clamped to
// +-kFeatClamp):
//   [4j .. 4j+3], j < 10: neighbour j's (pos - p) * kPosScale and
//                          (vel - v) * kVelScale; zeros past the count
//   [40 + 4i .. 43 + 4i], i < 4: endpoints (L - p, R - p) * kPosScale of the
//                          i-th nearest obstacle edge in range; zeros past
//   [56, 57]: preferred velocity * kVelScale
//   [58, 59]: velocity * kVelScale
//   [60]: radius * kRadScale
//   [61..63]: 0
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ecm {
namespace mlp {

constexpr int kIn = 61, kInPad = 64, kHid = 128, kOut = 16, kSlots = 14;
constexpr int kAgentSlots = 10, kObstFeat = 4;
constexpr float kPosScale = 0.25f, kVelScale = 0.5f, kRadScale = 4.0f, kFeatClamp = 8.0f;

// Flat FP32 parameter vector (row-major W[out][in], then the bias).
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kHid * kInPad;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kHid * kHid;
constexpr int kOffW4 = kOffB3 + kHid;
constexpr int kOffB4 = kOffW4 + kOut * kHid;
constexpr int kNParams = kOffB4 + kOut;

// Packed inference weights: v_mfma_f32_32x32x16_bf16 A-operand fragments.
// Fragment (mi, kc) of a layer covers output rows [32 mi, 32 mi + 32) and
// input chunk kc (16 inputs); lane L holds 8 bf16: row 32 mi + L % 32,
// inputs 16 kc + 8 (L / 32) + e, e < 8, with hidden inputs permuted by
// hidden_perm so that the accumulator registers of one layer are the
// B operand of the next without any data movement.
constexpr int kFragElems = 512;                      // 64 lanes x 8 bf16
constexpr int kL1Frags = 4 * 4, kL2Frags = 4 * 8, kL3Frags = 4 * 8, kL4Frags = 1 * 8;
constexpr int kFragL1 = 0, kFragL2 = kFragL1 + kL1Frags, kFragL3 = kFragL2 + kL2Frags, kFragL4 = kFragL3 + kL3Frags;
constexpr int kNFrags = kFragL4 + kL4Frags;          // 88 fragments = 88 KiB
// Packed biases (FP32): b1, b2, b3 (128 each), b4 padded to 32.
constexpr int kBiasL1 = 0, kBiasL2 = 128, kBiasL3 = 256, kBiasL4 = 384, kNBias = 416;

// The accumulator of a 32x32 MFMA tile holds, in lane L, rows
// 8 j + 4 (L / 32) + r (register 4 j + r) of column L % 32.  The next
// layer's B operand for chunk kc = 2 mi + h takes registers 8 h .. 8 h + 7
// of m-tile mi; its input k_local = 8 (L / 32) + e is then hidden unit:
__host__ __device__ constexpr int hidden_perm(int k) {
    return 32 * ((k / 16) / 2) + 16 * ((k / 16) % 2) + 8 * (((k % 16) % 8) / 4) + 4 * ((k % 16) / 8) + ((k % 16) % 4);
}

// Label word of a logged sample: bits 0..13 active slots, 16..29 slots
// that exist (a constraint is behind the slot), bit 30 every active
// constraint is behind some slot.
constexpr uint32_t kLabelActive = 0x3fffu, kLabelValidShift = 16, kLabelRepresentable = 1u << 30;

}  // namespace mlp
}  // namespace ecm
