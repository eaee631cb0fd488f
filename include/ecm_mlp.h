/* ecm_mlp.h -- C ABI of the learned ORCA proposal: a small MLP that
 * predicts which ORCA constraints are active at the LP optimum, trained on
 * the GPU from samples the simulator logs, and used by the step as a
 * proposal that the exact LP verifies.
 *
 * The reference has no learned component; the plain incremental LP it
 * proposes for (ORCA::RandomizedLP, ORCA.cpp:428-587, and
 * ORCA::RandomizedLP3D, ORCA.cpp:592-669) has the last word: an agent's
 * proposal is accepted only when the LP restricted to the predicted
 * constraints returns a velocity that every other constraint contains
 * (Constraint::Contains, ORCA.h:49-60), which makes it the optimum of the
 * full LP; otherwise the agent takes the full 2D / 3D solve.
 *
 * Network: 61 per-agent features (neighbour and obstacle geometry relative
 * to the agent, preferred and current velocity, radius) -> 128 -> 128 ->
 * 128 -> 16, SiLU, BF16 weights and activations with FP32 accumulation on
 * MFMA.  Outputs: 14 active-constraint logits
 * and a 2D velocity.  All functions return 0 on success or a negative
 * ECM_E_* code (include/ecm_sim.h) with ecm_last_error() set. */
#ifndef ECM_MLP_H
#define ECM_MLP_H

#include <stdint.h>

#include "ecm_sim.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ecm_mlp ecm_mlp;

#define ECM_MLP_N_IN 61
#define ECM_MLP_N_IN_PAD 64
#define ECM_MLP_N_OUT 16
#define ECM_MLP_N_SLOTS 14

/* Number of FP32 parameters: W1[128][64], b1[128], W2[128][128], b2[128],
 * W3[128][128], b3[128], W4[16][128], b4[16] (row-major W[out][in]). */
int ecm_mlp_n_params(void);

/* Network on the current device with seeded uniform(+-1/sqrt(fan_in))
 * initial weights (columns 61..63 of W1 are zero). */
int ecm_mlp_create(uint64_t seed, ecm_mlp** out);
int ecm_mlp_destroy(ecm_mlp* m);
int ecm_mlp_get_params(ecm_mlp* m, float* out, int n);
int ecm_mlp_set_params(ecm_mlp* m, const float* in, int n);

/* Training samples live in device memory owned by the handle.  Each is
 * 64 BF16 features, a label word (bits 0..13: active slots; 16..29: slots
 * that hold a constraint; 30: every active constraint is behind a slot)
 * and the LP velocity. */
int ecm_mlp_dataset_reserve(ecm_mlp* m, int64_t capacity);
int ecm_mlp_dataset_size(ecm_mlp* m, int64_t* n);
int ecm_mlp_dataset_clear(ecm_mlp* m);
/* Copies samples [first, first + n) to the host (x: n*64 float, the BF16
 * features widened; label: n; v: 2n). */
int ecm_mlp_dataset_read(ecm_mlp* m, int64_t first, int n, float* x, uint32_t* label, float* v);

/* One plain step of `sim` that also appends a sample for every agent
 * updated by the main ORCA pass whose 2D LP is feasible.  Stops appending
 * at capacity; *n_logged (optional) receives the number appended. */
int ecm_sim_log_step(ecm_sim* sim, ecm_mlp* m, int* n_logged);

/* Adam training over the dataset: `epochs` passes of seeded shuffled
 * batches of `batch` samples.  Loss = BCE over existing slots (positive
 * class weighted by pos_weight) + vel_weight * 0.5 |v - v_LP|^2 (scaled
 * velocities), mean over the batch.  losses (optional, [epochs]) receives
 * the mean batch loss of each epoch. */
int ecm_mlp_train(ecm_mlp* m, int epochs, int batch, float lr, float pos_weight, float vel_weight, uint64_t seed,
                  float* losses);

/* Forward pass of n host samples (x: n*64 FP32, cast to BF16 on the
 * device) with the inference kernel: y (n*16) raw outputs. */
int ecm_mlp_forward(ecm_mlp* m, const float* x, int n, float* y);

/* ecm_mlp_forward for device buffers, allocating nothing: x (n*64 FP32)
 * and y (n*16 raw outputs) are device memory of the handle's GPU, and the
 * BF16 features and proposal masks go to a static scratch array of the
 * library (one per GPU, ECM_MLP_NOALLOC_MAX_SAMPLES samples = 4.125 MiB,
 * held until the process ends).  No hipMalloc, hipFree or other allocating
 * call; n above the maximum is rejected with ECM_E_ARG.  Calls that
 * overlap on one GPU share the scratch and race.  Returns after y is
 * written.  No counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops). */
#define ECM_MLP_NOALLOC_MAX_SAMPLES 32768
int ecm_mlp_forward_noalloc(ecm_mlp* m, const float* x, int n, float* y);

/* Loss and gradient (n_params FP32, same order as the parameters) of n
 * host samples, with the training kernels (one batch, no update). */
int ecm_mlp_loss_grad(ecm_mlp* m, const float* x, const uint32_t* label, const float* v, int n, float pos_weight,
                      float vel_weight, float* loss, float* grad);

/* Learned proposal in the step of `sim` (null: the plain LP).  A slot is
 * proposed active when its logit exceeds theta.  The handle must outlive
 * its use by `sim`; parameters changed later are picked up at the next
 * call of this function. */
int ecm_sim_set_learned(ecm_sim* sim, ecm_mlp* m, float theta);

#ifdef __cplusplus
}
#endif

#endif
