// Learned smoothing-length proposal (gf_mlp.hip).  Device arrays; x holds
// NIN = 4 BF16 features per sample, row-major.
#pragma once
#include "gf_common.h"

#define GF_MLP_PARAMS 1249

// Floats of scratch gf_mlp_grad / gf_mlp_train_step need for n samples.
size_t gf_mlp_partial_floats(int n);
int gf_mlp_forward(int n, const uint16_t* x, const float* params, float* out, hipStream_t s);
// Gradient (GF_MLP_PARAMS floats) of the mean loss over n_total samples --
// squared error 0.5 (y - t)^2 for tau <= 0, the pinball loss of the
// tau-quantile for 0 < tau < 1 -- and the summed loss (1 float, squared errors
// for tau <= 0); either output may be null.
int gf_mlp_grad(int n, int n_total, const uint16_t* x, const float* target, const float* params,
                float tau, float* grad, float* loss, float* partial, hipStream_t s);
// One Adam step (step >= 1) on the mean loss of the n samples.
int gf_mlp_train_step(int n, const uint16_t* x, const float* target, float* params, float* m,
                      float* v, int step, float lr, float b1, float b2, float eps, float tau,
                      float* loss, float* partial, hipStream_t s);
// Samples per call of gf_mlp_train_step_noalloc (its scratch is static).
#define GF_MLP_NOALLOC_MAX 65536
// gf_mlp_train_step with its partials in the library's static device array
// instead of a caller's buffer; calls no allocator.  n <= GF_MLP_NOALLOC_MAX.
int gf_mlp_train_step_noalloc(int n, const uint16_t* x, const float* target, float* params,
                              float* m, float* v, int step, float lr, float b1, float b2,
                              float eps, float tau, float* loss, hipStream_t s);
int gf_mlp_features(int n, int k, const double* dk_prev, const double* cell_rho,
                    const double* cell_count, const double* divv, const double* dtc_prev,
                    float dt, const float* norm, uint16_t* x, float* raw, hipStream_t s);
int gf_mlp_propose(int n, const uint16_t* x, const float* params, const double* dk_prev,
                   float margin, double* r, hipStream_t s);
