/*
 * symcost.h - learned cost model of the autotuner: a small FP32 multilayer perceptron
 * (num_features -> hidden -> hidden -> 1, ReLU) that predicts the log of a launch's time from
 * features of its tiles and code.  Training (mini-batch Adam on the squared error) and
 * inference run as hand-written HIP kernels on the GPU, compiled in process by hipRTC.
 *
 * No counterpart in the reference: SymbolicLib picks no launch parameters (its HIP path always
 * launches ceil(n / ThreadsPerBlock) blocks per group, src/ComputeUnit.cpp:440-447).  The
 * autotuner that uses this model keeps exhaustive search as its fallback and as the measure of
 * the model's regret.
 *
 * Features are standardised with the mean and deviation of the training set, the target with
 * its own mean and deviation; both are stored with the weights.  Parameter layout (FP32):
 * W1 [hidden x features], b1 [hidden], W2 [hidden x hidden], b2 [hidden], w3 [hidden], b3.
 *
 * All functions return 0 on success; symcost_last_error() holds the message otherwise.
 */
#ifndef SYMCOST_H
#define SYMCOST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct symcost_model* symcost_t;

/* New model with He-initialised weights from `seed` (needs a HIP device). */
int symcost_create(symcost_t* out, int num_features, int hidden, uint64_t seed);

/* Train on n samples: X row-major [n x num_features], y = log(time).  `epochs` passes of
 * shuffled mini-batches of `batch` samples, Adam with learning rate `lr`.  The first call fixes
 * the standardisation.  loss_out (may be NULL) receives the mean squared error on the
 * standardised target after the last epoch. */
int symcost_train(symcost_t m, const float* X, const float* y, size_t n, int epochs, int batch, float lr,
                  uint64_t seed, float* loss_out);

/* Predicted log(time) for n samples. */
int symcost_predict(symcost_t m, const float* X, size_t n, float* y_out);

/* Parameters (FP32, layout above) and standardisation (feature means, feature deviations,
 * target mean, target deviation), for inspection and checks. */
int symcost_num_params(symcost_t m, size_t* n);
int symcost_get_params(symcost_t m, float* params, size_t n);
int symcost_set_params(symcost_t m, const float* params, size_t n);
int symcost_get_scaling(symcost_t m, float* mean, float* dev, float* y_mean_dev);
int symcost_set_scaling(symcost_t m, const float* mean, const float* dev, const float* y_mean_dev);

/* One Adam step on one batch without shuffling (for checks against a host computation);
 * uses the stored standardisation. */
int symcost_step(symcost_t m, const float* X, const float* y, size_t n, float lr);

/* Gradient of the last step's mean squared error (standardised target) per parameter. */
int symcost_get_gradient(symcost_t m, float* grad, size_t n);

/* Text file: header line, scaling, parameters. */
int symcost_save(symcost_t m, const char* path);
int symcost_load(symcost_t* out, const char* path);

void symcost_destroy(symcost_t m);
const char* symcost_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SYMCOST_H */
