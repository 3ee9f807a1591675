/*
 * lbm_closure.h -- C ABI of the learned sub-grid closure (liblbm_core_hip.so).
 *
 * No reference counterpart: the reference's only turbulence closures are
 * OmegaNo and OmegaLES (3D/simulation/src/solvers/turbulence/OmegaLES.h:25-137).
 * Here a network 12 -> 64 -> 64 -> 1 (SiLU, BF16 weights on matrix cores,
 * FP32 accumulation) maps local features of a cell (features.hpp) to a
 * coefficient C(x) >= 0 that replaces OmegaLES's constant in its rate formula:
 *   tau(x) = tau/2 + sqrt(tau^2 + 18 C(x) |Pi_neq| / rho) / 2.
 * Output head: C = c_les * softplus(z) / ln 2, so z = 0 gives the
 * Smagorinsky constant.  Cells whose normalised features leave the training
 * envelope use OmegaLES with c_les unchanged.
 *
 * Parameter vector (FP32, NPARAM = 5057): W1[64][12], b1[64], W2[64][64],
 * b2[64], w3[64], b3.  Features are passed normalised: (x - mu) / sd.
 */
#ifndef LBM_CLOSURE_H
#define LBM_CLOSURE_H

#include <stddef.h>
#include <stdint.h>

#include "lbm_core.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_MLP_FEATURES 12
#define LBM_MLP_PARAMS 5057

typedef struct lbm_mlp lbm_mlp;

/* A network on the current device.  params == NULL: seeded initialisation
 * (W1, W2 ~ N(0, 1/fan_in), biases and w3 zero, so C = c_les everywhere). */
lbm_status lbm_mlp_create(const float* params, uint64_t seed, lbm_mlp** out);

/* Training set: x [n][12] normalised features, y [n] target coefficients,
 * w [n] sample weights (NULL = 1).  Copied to the device. */
lbm_status lbm_mlp_set_data(lbm_mlp* m, const float* x, const float* y, const float* w, int64_t n);

/* Loss and gradient over samples first .. first+count-1 of the training set
 * (count a positive multiple of 16), without updating the parameters.
 * loss_sum = sum w (C - y)^2 (FP64); grad = d(loss_sum / count) / d params.
 * use_bf16 = 1: the matrix-core kernels (BF16 operands); 0: the same kernels
 * with FP32 products (a check of the data flow). */
lbm_status lbm_mlp_gradient(lbm_mlp* m, int64_t first, int32_t count, float c_les, int32_t use_bf16,
                            double* loss_sum, float* grad);

/* steps Adam steps (beta 0.9 / 0.999, eps 1e-8) on random batches of
 * `batch` samples (multiple of 16) drawn with replacement from a hash of
 * (seed, step, position); loss[k] = mean w (C - y)^2 of batch k (may be NULL). */
lbm_status lbm_mlp_train(lbm_mlp* m, int32_t steps, int32_t batch, float lr, float c_les, uint64_t seed,
                         double* loss);

/* Coefficient C for n normalised feature rows: the inference path (BF16
 * weights on matrix cores) that the fused collision kernel runs. */
lbm_status lbm_mlp_forward(lbm_mlp* m, const float* x, int64_t n, float c_les, float* c_out);

lbm_status lbm_mlp_get_params(lbm_mlp* m, float* params);
lbm_status lbm_mlp_set_params(lbm_mlp* m, const float* params);
void lbm_mlp_destroy(lbm_mlp* m);

/* Install a trained network in a lattice created with LBM_CLOSURE_MLP.
 * norm = {mu[12], sd[12], lo[12], hi[12]}: normalisation and the training
 * envelope in normalised units.  Until this is called lbm_step returns
 * LBM_E_STATE. */
lbm_status lbm_set_closure(lbm_ctx* ctx, const float* params, const float* norm);

/* Evaluate the network every `every` steps (default 1) and hold each cell's
 * coefficient C(x) in between: the steps in between still apply the rate
 * formula to the cell's current |Pi_neq| / rho, only C is held.  The
 * network's coefficient moves on the turbulence time scale, many steps long.
 * Restarts the cycle: the next step evaluates the network. */
lbm_status lbm_set_closure_refresh(lbm_ctx* ctx, int32_t every);

/* Per cell of the current state (caller layout): normalised features
 * [12][N] (may be NULL), envelope flag [N] (may be NULL) and the coefficient
 * the next step uses [N] (c_les outside the envelope, the held value when the
 * next step does not evaluate the network; may be NULL).
 * Requires LBM_CLOSURE_MLP and an installed network.  Synchronises. */
lbm_status lbm_get_closure_features(lbm_ctx* ctx, float* features, int32_t* in_envelope, double* coefficient);

/* Box filter of the current state over factor^3 blocks of cells
 * (nx, ny, nz divisible by factor): out = 10 planes of (nx/f)(ny/f)(nz/f)
 * FP64 values, plain averages of rho, ux, uy, uz, ux ux, uy uy, uz uz,
 * ux uy, ux uz, uy uz over the non-solid cells of each block.  Synchronises. */
lbm_status lbm_filter(lbm_ctx* ctx, int32_t factor, double* out);

#ifdef __cplusplus
}
#endif

#endif /* LBM_CLOSURE_H */
