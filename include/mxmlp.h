/* The warm-start network: a multilayer perceptron (SiLU activations) whose products run in BF16 on
 * the gfx950 matrix cores with FP32 accumulation, FP32 master weights trained with Adam.
 * No counterpart in the reference: it maps a bump-channel cell (centroid, wall distance) and the
 * case's bump height, Mach number and flow angle to an initial state (rho, u, v, w, p).
 * Samples are rows (host arrays, row-major); inputs and outputs are normalised on the device with
 * the means and scales given. Every function returns 0, or 1 with a message in err. */
#ifndef MXMLP_H
#define MXMLP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mxmlp mxmlp;

/* inputs <= 32, outputs <= 16, hidden a multiple of 64, layers hidden layers. */
int mxmlp_create(int32_t inputs, int32_t hidden, int32_t layers, int32_t outputs, uint64_t seed, mxmlp** out, char* err,
                 size_t err_len);
void mxmlp_destroy(mxmlp* h);
/* Parameters in order: per layer its weights (outputs x inputs, padded to 32 inputs for the first
 * layer and 16 outputs for the last) then its biases. */
int64_t mxmlp_param_count(const mxmlp* h);
int mxmlp_set_normalization(mxmlp* h, const float* in_mean, const float* in_scale, const float* out_mean,
                            const float* out_scale, char* err, size_t err_len);
int mxmlp_get_params(mxmlp* h, float* params, char* err, size_t err_len);
int mxmlp_set_params(mxmlp* h, const float* params, char* err, size_t err_len);
/* y (n x outputs, physical units) for x (n x inputs). */
int mxmlp_predict(mxmlp* h, const float* x, int64_t n, float* y, char* err, size_t err_len);
/* The mean squared error of the normalised outputs over one batch and its gradient (grads may be
 * NULL), in parameter order. */
int mxmlp_gradients(mxmlp* h, const float* x, const float* t, int32_t n, float* grads, float* loss, char* err,
                    size_t err_len);
/* One batch's forward and backward (as mxmlp_gradients), then one of its tensors, row-major n x
 * width in FP32: what = 0 the input activations of `layer` (BF16 values; layer 0: the normalised,
 * padded inputs), 1 its pre-activations Z, 2 the gradient of the loss by Z (BF16 values). For tests
 * that check each product and epilogue on its own inputs. */
int mxmlp_trace(mxmlp* h, const float* x, const float* t, int32_t n, int32_t what, int32_t layer, float* out,
                char* err, size_t err_len);
/* One Adam step with the gradients of the last mxmlp_gradients call. */
int mxmlp_adam_step(mxmlp* h, float lr, char* err, size_t err_len);
/* Training data kept on the device, then epochs of shuffled batches, one Adam step per batch. */
int mxmlp_set_data(mxmlp* h, const float* x, const float* t, int64_t n, char* err, size_t err_len);
int mxmlp_train(mxmlp* h, int32_t epochs, int32_t batch, float lr, uint64_t seed, float* loss, double* seconds,
                char* err, size_t err_len);
/* Milliseconds of one inference of n device-resident samples (average of repeats). */
int mxmlp_time_forward(mxmlp* h, int32_t n, int32_t repeats, double* ms, char* err, size_t err_len);
/* Inference (mxmlp_predict, mxmlp_time_forward) by the fused forward, one kernel over all layers (1,
 * the default; hidden widths up to 256), or by the training path's layer-by-layer products (0). The
 * outputs are the same bits. */
int mxmlp_set_fused(mxmlp* h, int32_t on, char* err, size_t err_len);
/* Samples per workgroup of the fused forward (64, 96 or 128; default 128): each workgroup reads
 * every weight once for its samples, so more samples per workgroup read fewer L2 bytes per product
 * (128 with eight waves when the hidden width is a multiple of 128). The outputs are the same bits
 * whatever the tile. */
int mxmlp_set_infer_rows(mxmlp* h, int32_t rows, char* err, size_t err_len);

#ifdef __cplusplus
}
#endif

#endif
