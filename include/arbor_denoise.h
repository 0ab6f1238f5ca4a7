/* The learned denoiser (no counterpart in the reference, which denoises with
 * filterLighting only, Arbor.cu:292-302): a 3x3-convolution encoder-decoder
 * in BF16 on the matrix cores with FP32 accumulation, its training kernels
 * and a network runner for training and inference. filterLighting stays the
 * fallback and the yardstick.
 *
 * Tensors are NHWC bf16 (uint16_t bit patterns), channel counts multiples of
 * 8. Convolution weights are [9 taps][cout][cin] bf16, tap = 3*(dy+1)+(dx+1);
 * biases fp32 [cout]. All calls return ARBOR_OK or a negative status
 * (arbor.h). */
#ifndef ARBOR_DENOISE_H
#define ARBOR_DENOISE_H

#include <stdint.h>

#include "arbor.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Pixels per partial sum of the weight gradient. */
#define ARBOR_DN_WGRAD_CHUNK 4096
/* Widest tensor the convolution takes (input or output channels). */
#define ARBOR_DN_MAX_CHANNELS 256

/* Y = act(conv3x3(X, Wt) + bias), zero padding, stride 1; cout a multiple of
 * 16, cin a multiple of 8, both up to ARBOR_DN_MAX_CHANNELS; act = ReLU when
 * relu != 0; bias may be NULL. */
int arbor_dn_conv3x3(const uint16_t* X, int n, int h, int w, int cin, const uint16_t* Wt, const float* bias, int relu,
                     uint16_t* Y, int cout, hipStream_t stream);
/* Weight and bias gradients of that convolution: dW [9][cout][cin] and dB
 * [cout] (fp32) from X and dY; `partial` is scratch of at least
 * arbor_dn_wgrad_partial_floats(...) floats. */
long long arbor_dn_wgrad_partial_floats(int n, int h, int w, int cin, int cout);
int arbor_dn_conv3x3_wgrad(const uint16_t* X, const uint16_t* dY, int n, int h, int w, int cin, int cout,
                           float* partial, long long partial_floats, float* dW, float* dB, hipStream_t stream);
/* Wflip[tap][ci][co] = Wt[8 - tap][co][ci]: conv3x3(dY, Wflip) is the gradient
 * of the convolution with respect to X. */
int arbor_dn_flip_weights(const uint16_t* Wt, int cin, int cout, uint16_t* Wflip, hipStream_t stream);
/* 2x2 average pooling (h, w even) and its gradient (h, w: the fine size). */
int arbor_dn_pool2(const uint16_t* X, int n, int h, int w, int c, uint16_t* Y, hipStream_t stream);
int arbor_dn_pool2_bwd(const uint16_t* dY, int n, int h, int w, int c, uint16_t* dX, hipStream_t stream);
/* Nearest 2x upsampling and its gradient (h, w: the coarse size). */
int arbor_dn_up2(const uint16_t* X, int n, int h, int w, int c, uint16_t* Y, hipStream_t stream);
int arbor_dn_up2_bwd(const uint16_t* dY, int n, int h, int w, int c, uint16_t* dX, hipStream_t stream);
/* Channel concatenation Y = [A, B] per pixel and its split. */
int arbor_dn_concat(const uint16_t* A, int ca, const uint16_t* B, int cb, long long pixels, uint16_t* Y,
                    hipStream_t stream);
int arbor_dn_split(const uint16_t* Y, int ca, int cb, long long pixels, uint16_t* A, uint16_t* B, hipStream_t stream);
/* G *= (Yact > 0); A += B. */
int arbor_dn_relu_bwd(uint16_t* G, const uint16_t* Yact, long long count, hipStream_t stream);
int arbor_dn_add(uint16_t* A, const uint16_t* B, long long count, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* ARBOR_DENOISE_H */
