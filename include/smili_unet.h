/* C ABI of the warm-start network (libsmili_unet.so).
 *
 * None of these entry points has a counterpart in the reference: SMILI
 * starts L-BFGS-B from a prior image (smili/imaging/lbfgs.py:202-489,
 * src/fortlib/fftim2d.f90:286-351).  This library proposes a better start
 * from the dirty image and dirty beam; the host imager then runs the same
 * L-BFGS-B from that proposal.
 *
 * Network (fixed architecture): 4-level U-Net, 32/64/128/256 channels, two
 * 3x3 convolutions (+ bias, ReLU) per level, 2x2 average pooling down,
 * nearest x2 upsampling concatenated with the skip connection up, linear
 * 1x1 head z.  Input [n][h][w][2] float32 (dirty image, dirty beam),
 * output image max(z, 0) [n][h][w] float32; training fits z itself.
 * Activations and 3x3 weights are bfloat16 in HBM, products accumulate in
 * float32 (MFMA); the first convolution (2 input channels) and the head run
 * in float32.
 *
 * Parameters: one flat float32 vector of smili_unet_nparams() values; for
 * each 3x3 convolution (conv_in, enc0b, enc1a, enc1b, enc2a, enc2b, enc3a,
 * enc3b, dec2a, dec2b, dec1a, dec1b, dec0a, dec0b) W[co][tap][ci]
 * (tap = 3 dy + dx) then bias[co]; then head w[32], b[1].
 *
 * Status codes: 0 ok, -1 invalid argument, -2 device error; the message is
 * in smili_unet_last_error().  Host arrays are copied; a handle's calls
 * must not overlap (one HIP stream per handle).
 */
#ifndef SMILI_UNET_H
#define SMILI_UNET_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct smili_unet smili_unet;

/* A network for batches of up to `batch` images of height x width
 * (height % 64 == 0, width % 128 == 0), parameters zero until set. */
int smili_unet_create(int32_t batch, int32_t height, int32_t width, smili_unet** out);
void smili_unet_destroy(smili_unet* net);
int64_t smili_unet_nparams(void);
int smili_unet_set_params(smili_unet* net, const float* params);
int smili_unet_get_params(smili_unet* net, float* params);

/* output[n][h][w] = network(input[n][h][w][2]) for any n >= 1 (processed
 * in batches). */
int smili_unet_forward(smili_unet* net, const float* input, int32_t n, float* output);

/* Mean squared error (1/(n h w)) sum (z - target)^2 of the linear head
 * over n <= batch images and its gradient grad[nparams] (float32; NULL to
 * skip). */
int smili_unet_loss_grad(smili_unet* net, const float* input, const float* target, int32_t n,
                         double* loss, float* grad);

/* Uploads a training set (inputs [count][h][w][2], targets [count][h][w])
 * into device memory, replacing any earlier one. */
int smili_unet_set_data(smili_unet* net, const float* inputs, const float* targets, int64_t count);

/* nsteps Adam steps (decoupled weight decay); step s trains on the batch
 * index[s*batch .. s*batch+batch) of the uploaded set with learning rate
 * lr[s]; hyper = {beta1, beta2, eps, weight_decay}; the optimiser's own
 * step counter continues across calls.  losses[s] receives each step's
 * loss (before its update).  An index is 8 * sample + k, k a symmetry of
 * the square applied to input and target alike (k = 0 identity; k >= 4
 * transposes, then k % 4 quarter turns out[i][j] = in[j][n-1-i]; k != 0
 * needs height == width). */
int smili_unet_train(smili_unet* net, const int64_t* index, int64_t nsteps, const float* lr,
                     const float* hyper, double* losses);
int smili_unet_reset_optimizer(smili_unet* net);

/* Times `reps` training steps on the first `batch` uploaded samples with
 * HIP events on the handle's stream: ms[0] = mean step, ms[1] = mean
 * forward pass.  The steps run with learning rate 0, and the Adam moments
 * and step count are restored afterwards: the handle's training state is
 * what it was before the call. */
int smili_unet_time_steps(smili_unet* net, int32_t reps, float* ms);

/* Times `reps` launches of one 3x3 convolution kernel of layer `layer`
 * (1..13, the parameter order above; 0 is the float32 first convolution)
 * on the activations of the last forward pass: pass 0 forward, 1 input
 * gradient (for the odd layers, whose input has no other consumer, fused
 * with that input's ReLU mask and bias sums as in training), 2 weight
 * gradient (this one adds into the gradient buffer).
 * ms = mean launch time (HIP events on the handle's stream), flop = FLOP
 * of one launch for a full batch. */
int smili_unet_time_layer(smili_unet* net, int32_t layer, int32_t pass, int32_t reps, float* ms, double* flop);

/* Work of one training step / forward pass of a full batch:
 * flop[0] = 3x3-convolution FLOP of a training step, flop[1] = of a
 * forward pass. */
int smili_unet_flops(smili_unet* net, double* flop);

const char* smili_unet_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
