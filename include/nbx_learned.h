/*
 * nbx_learned.h - learned Barnes-Hut residual correction (C ABI, libnbx.so).
 *
 * A 3-layer MLP (16 -> 64 -> 64 -> 3, SiLU after the first two layers) maps
 * per-particle walk features to the part of the walk's error that its
 * quadrupole force q does not remove: (a_direct - a_BH - q) / |q| (q below;
 * the residual of a theta = 1 walk lies almost along q, so q is added as is
 * and the network predicts the bounded, N-independent remainder), expressed
 * in the particle's local frame: e1 along a_BH, e2 along the
 * quadrupole force of the accepted cells perpendicular to e1 (or the
 * coordinate axis least aligned with e1 when that part vanishes), e3 = e1 x e2. Weights and
 * activations are BF16 on the MFMA units with FP32 accumulation; master
 * weights, gradients and Adam moments are FP32. The reference has no learned
 * component; it only counts accepted cells during its walk
 * (src/c/barnes_hut/barnes_hut.c:611), which is one of the features.
 *
 * The learned path's walk is the group walk with its member test, which stays
 * on under NBX_FLAG_BH_BOX_TEST.
 *
 * Features (raw, before standardisation; all frame-free), per particle, with
 * q = sum over accepted cells of the quadrupole force 5/2 (d.Q.d) d / r^7 -
 * Q.d / r^5 (d from the particle to the cell's centre of mass, r^2 = d.d +
 * eps^2, Q = sum m (3 x x^T - |x|^2 I) about the cell's centre of mass, the
 * traceless moment the build computes for the learned path) and a = the
 * walk's acceleration:
 *   0  rms accepted distance / cell edge     1  minimum accepted distance / edge
 *   2  log2 local density: own mass / (leaf cell edge)^3
 *   3  log2(1 + accepted cells)              4  log2(1 + exact leaf terms)
 *   5  q along e1 / |a|                      6  |q perpendicular to e1| / |a|
 *   7  q along e1 / |q|                      8  |q perpendicular to e1| / |q|
 *   9  log2(|q| / |a| + 2^-40)               10-15 zero
 * The counts (3, 4) grow with N and are left out of the committed model,
 * as is the density (2).
 * A feature whose std is +inf standardises to 0, i.e. the model ignores it.
 * Parameter layout (float, nbx_mlp_param_count() values):
 *   W1[64][16], b1[64], W2[64][64], b2[64], W3[16][64] (rows 3..15 unused,
 *   kept zero), b3[16] (entries 3..15 unused).
 * The model standardises features as (f - mean) * float32(1 / std); its output y is
 * scaled: (a_direct - a_BH - q) / |q| = (y0 e1 + y1 e2 + y2 e3) / out_scale,
 * so the corrected acceleration is a_BH + q + |q| (y0 e1 + y1 e2 + y2 e3) /
 * out_scale (no accepted cell: q = 0, no correction; a zero model gives the
 * quadrupole-corrected walk a_BH + q).
 */
#ifndef NBX_LEARNED_H
#define NBX_LEARNED_H

#include "nbx_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { NBX_MLP_FEATURES = 16, NBX_MLP_HIDDEN = 64, NBX_MLP_OUTPUTS = 3 };

typedef struct NBXMlp NBXMlp;

size_t nbx_mlp_param_count(void);
NBXMlp* nbx_mlp_create(int device, NBodySimProError* error);
void nbx_mlp_destroy(NBXMlp* mlp);

/* FP32 parameters, feature standardisation (16 means, 16 stds) and output scale. */
NBodySimProStatus nbx_mlp_set_params(NBXMlp* mlp, const float* params, const float* mean,
                                     const float* std, float out_scale, NBodySimProError* error);
NBodySimProStatus nbx_mlp_get_params(NBXMlp* mlp, float* params, float* mean, float* std,
                                     float* out_scale, NBodySimProError* error);

/* Inference on host arrays: raw features (n x 16) -> outputs y (n x 3). */
NBodySimProStatus nbx_mlp_forward(NBXMlp* mlp, const float* features, size_t n, float* out,
                                  NBodySimProError* error);
/* Loss L = mean_i sum_c (y_ic - target_ic)^2 and dL/dparams for host arrays,
 * without updating the parameters (gradient checks). */
NBodySimProStatus nbx_mlp_gradients(NBXMlp* mlp, const float* features, const float* targets,
                                    size_t n, float* grads, float* loss, NBodySimProError* error);
/* Training set kept on the device (targets are already scaled, n x 3). */
NBodySimProStatus nbx_mlp_load_dataset(NBXMlp* mlp, const float* features, const float* targets,
                                       size_t n, NBodySimProError* error);
/* Adam epochs over the device training set (shuffled per epoch from `seed`);
 * loss_history (nullable) receives the mean loss of every epoch. */
NBodySimProStatus nbx_mlp_train(NBXMlp* mlp, int epochs, int batch, float learning_rate,
                                uint64_t seed, float* loss_history, NBodySimProError* error);

/* Attach a trained model to an engine (NULL detaches): hip_bh_learned then
 * corrects its theta walk with it; without a model it is plain hip_barnes_hut. */
NBodySimProStatus nbx_engine_set_mlp(NBXEngine* engine, NBXMlp* mlp, NBodySimProError* error);
/* Build + walk (the learned path's walk) on the device-resident state,
 * returning the raw features (n x 16, original particle order) and, when
 * `frames` is not NULL, the local frames (n x 9: e1, e2, e3); the walk's
 * accelerations are left in the engine (as by hip_barnes_hut's default group
 * walk; with NBX_FLAG_BH_REFERENCE_LISTS, its per-particle walk). */
NBodySimProStatus nbx_engine_bh_features(NBXEngine* engine, const NBodySimProGravity* gravity,
                                         float* features, float* frames, NBodySimProError* error);
/* FP64 direct sum (the reference's scalar order) for selected targets only:
 * out[3k..3k+2] = acceleration of particle targets[k]. */
NBodySimProStatus nbx_engine_direct_fp64_targets(NBXEngine* engine,
                                                 const NBodySimProGravity* gravity,
                                                 const int64_t* targets, size_t count,
                                                 double* out, NBodySimProError* error);

/* Lane-layout probe of one v_mfma_f32_16x16x16bf16_1k: d = a * b for
 * row-major 16x16 FP32 inputs (rounded to BF16). */
NBodySimProStatus nbx_mlp_probe_mfma(int device, const float* a, const float* b, float* d,
                                     NBodySimProError* error);

#ifdef __cplusplus
}
#endif

#endif /* NBX_LEARNED_H */
