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
/* Y = act(X W1^T + bias): the 1x1 convolution of `pixels` pixels of cin = 16
 * or 32 channels with W1 [cout][cin] bf16 (cout a multiple of 16). */
int arbor_dn_conv1x1(const uint16_t* X, long long pixels, int cin, const uint16_t* W1, const float* bias, int relu,
                     uint16_t* Y, int cout, hipStream_t stream);
/* The same convolution of the channel concatenation [up2(Xc), X]: Xc is
 * n x (h/2) x (w/2) x cc (nearest 2x upsampling), X is n x h x w x cs; Wt has
 * cin = cc + cs input channels. Equals conv3x3 of the materialised tensor. */
int arbor_dn_conv3x3_cat(const uint16_t* Xc, int cc, const uint16_t* X, int cs, int n, int h, int w, const uint16_t* Wt,
                         const float* bias, int relu, uint16_t* Y, int cout, hipStream_t stream);
/* Weight and bias gradients of that convolution: dW [9][cout][cin] and dB
 * [cout] (fp32) from X and dY; `partial` is scratch of at least
 * arbor_dn_wgrad_partial_floats(...) floats. */
long long arbor_dn_wgrad_partial_floats(int n, int h, int w, int cin, int cout);
int arbor_dn_conv3x3_wgrad(const uint16_t* X, const uint16_t* dY, int n, int h, int w, int cin, int cout,
                           float* partial, long long partial_floats, float* dW, float* dB, hipStream_t stream);
/* Wflip[tap][ci][co] = Wt[8 - tap][co][ci]: conv3x3(dY, Wflip) is the gradient
 * of the convolution with respect to X. */
int arbor_dn_flip_weights(const uint16_t* Wt, int cin, int cout, uint16_t* Wflip, hipStream_t stream);
/* The elementwise operations below take channel counts (and counts) that are
 * multiples of 8, and tensors below 2^31 elements. */
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

/* ---- the network ----
 * Either the lean network (levels = 0) or a U-Net (levels 1..5).
 *
 * The lean network: one 3x3 convolution of widths[0] = 16 or 32 channels
 * with ReLU, then a linear 1x1 output layer. Its input is the compact
 * feature set, 8 bf16 per pixel (16 bytes): log(1 + radiance) rgb of one
 * sample, albedo rgb (Guide), depth log(1 + min(t, 1000))/7 (1 for sky) and
 * a zero channel. arbor_shadePixelsFeatures writes it directly, and the
 * network runs as one kernel from it (arbor_denoiser_denoise_features).
 * Its 1x1 layer is stored as a 3x3 layer whose taps other than the centre
 * are zero.
 *
 * A U-Net: L levels, 3x3 convolutions (ReLU), 2x2 average pooling,
 * nearest upsampling and skip concatenation, a linear 3x3 output layer. The
 * default is 5 levels of 32/48/64/96/128 channels: 1,021,904 parameters stored
 * (1,016,419 excluding the zero padding of the input and output channels).
 * Input per pixel: log(1 + radiance) rgb
 * of one sample, albedo rgb (Guide), shading normal (arbor_shadePixelsAOV)
 * and depth log(1 + min(t, 1000))/7 (1 for sky); output log(1 + radiance).
 * The runner holds a batch of max_batch slots of height x width (multiples
 * of 16); images smaller than a slot are zero-padded. */
typedef struct arbor_denoiser arbor_denoiser;
/* train != 0 also allocates gradients, Adam state and targets. Weights are
 * He-initialised from seed. */
int arbor_denoiser_create(arbor_denoiser** out, int max_batch, int height, int width, int train, int seed);
/* The same with `levels` levels of widths[0..levels-1] channels (multiples of
 * 16; a width plus the next one at most ARBOR_DN_MAX_CHANNELS); levels = 0:
 * the lean network of widths[0] (16 or 32) channels. */
int arbor_denoiser_create_net(arbor_denoiser** out, int max_batch, int height, int width, int train, int seed,
                              int levels, const int* widths);
/* The network's levels and widths (widths: 5 ints, unused levels 0). */
int arbor_denoiser_shape(const arbor_denoiser* d, int* levels, int* widths);
/* The levels and widths a checkpoint file was saved from (widths: 5 ints;
 * levels 0 and widths[0]: the lean network). */
int arbor_denoiser_checkpoint_shape(const char* path, int* levels, int* widths);
/* The runner's slot count and slot size (height x width). */
int arbor_denoiser_slot(const arbor_denoiser* d, int* max_batch, int* height, int* width);
int arbor_denoiser_destroy(arbor_denoiser* d);
long long arbor_denoiser_parameters(const arbor_denoiser* d);
/* Input features of slot `index` from one frame's buffers (float4 per pixel,
 * w x h): Linear, Guide, Normal, Hit (the lean network does not read Normal,
 * which may then be NULL). */
int arbor_denoiser_set_input(arbor_denoiser* d, int index, const float* Linear, const float* Guide,
                             const float* Normal, const float* Hit, int w, int h, hipStream_t stream);
/* Training target of slot `index`: radiance float4 per pixel (w x h). */
int arbor_denoiser_set_target(arbor_denoiser* d, int index, const float* Radiance, int w, int h, hipStream_t stream);
/* Forward pass over slots 0..n-1. */
int arbor_denoiser_forward(arbor_denoiser* d, int n, hipStream_t stream);
/* Radiance of slot `index` (float4, alpha 1, w x h): a drop-in for Filtered
 * ahead of resolvePixels. */
int arbor_denoiser_get_output(arbor_denoiser* d, int index, float* Out, int w, int h, hipStream_t stream);
/* One frame through the network: input features, forward and output on
 * slot 0 into Out (float4, alpha 1, w x h; a drop-in for Filtered). A
 * 1-level network of 16 channels runs as one kernel with its intermediate
 * tensors on chip (the same bits as the three calls); other shapes run the
 * three calls. */
int arbor_denoiser_denoise(arbor_denoiser* d, const float* Linear, const float* Guide, const float* Normal,
                           const float* Hit, int w, int h, float* Out, hipStream_t stream);
/* shadePixels (arbor_shadePixels, Arbor.cu:277-289) writing the compact
 * features of each pixel (8 bf16, 16 bytes per pixel, w x h) in place of
 * Linear and Guide: the bits arbor_denoiser_set_input computes for the lean
 * network from arbor_shadePixels' Linear and Guide and the Hit. */
int arbor_shadePixelsFeatures(const float* P, const float* Nodes, const unsigned* Order, const float* C,
                              const float* Hit, void* Feat, int width, int height, hipStream_t stream);
/* The lean network on one frame's compact features (w x h, within the slot)
 * into Out (float4, alpha 1; a drop-in for Filtered), in one kernel: the
 * bits of arbor_denoiser_denoise on the frame the features came from.
 * ARBOR_ERR_STATE for a U-Net. */
int arbor_denoiser_denoise_features(arbor_denoiser* d, const void* Feat, int w, int h, float* Out,
                                    hipStream_t stream);
/* Forward, L1 loss on log(1 + radiance) against the targets, backward, one
 * Adam step (beta 0.9 / 0.999) over slots 0..n-1. */
int arbor_denoiser_train_step(arbor_denoiser* d, int n, float lr, hipStream_t stream);
/* Mean L1 loss of the last train step (synchronises). */
int arbor_denoiser_loss(arbor_denoiser* d, float* host_loss);
/* Weights as "ARBORDN1" (a U-Net) or "ARBORDN2" and the int32 kind 1 (the
 * lean network), layer count, then per layer cin, cout (int32) and fp32
 * weights [9][cout][cin] and biases [cout]. */
int arbor_denoiser_save(const arbor_denoiser* d, const char* path);
int arbor_denoiser_load(arbor_denoiser* d, const char* path);
/* Device pointer of "input" (bf16 NHWC, 16 channels; 8 for the lean
 * network), "output" (bf16 NHWC,
 * 16 channels, 3 used), "target" (float4: log(1 + radiance), mask), or of
 * layer i's fp32 "weight:i" / "bias:i" and, after a train step, their
 * gradients "wgrad:i" / "bgrad:i". */
void* arbor_denoiser_tensor(arbor_denoiser* d, const char* name);

#ifdef __cplusplus
}
#endif
#endif /* ARBOR_DENOISE_H */
