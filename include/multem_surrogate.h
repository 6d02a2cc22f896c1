/*
 * multem_surrogate.h -- C ABI of the frozen-phonon surrogate network.
 *
 * No counterpart in the reference.  The network maps a few-configuration
 * frozen-phonon ADF image plus the specimen's projected-potential map (two
 * input channels) to the many-configuration ADF image (one channel):
 *
 *   h0 = relu(conv(x, W0) + b0)
 *   n_blocks times:  u = relu(conv(h, Wa) + ba);  h = h + conv(u, Wb) + bb
 *   y  = x[channel 0] + conv(h, Wh) + bh
 *
 * with 3x3 convolutions under periodic padding.  Convolution operands are
 * rounded to bfloat16 and multiplied on the matrix cores with FP32
 * accumulation; biases, activations, residual sums, the loss and the Adam
 * state are FP32.
 *
 * Images are NHWC float32 arrays: x [batch][h][w][2], y and targets
 * [batch][h][w].  Parameters are one flat float32 vector, per layer the
 * weights [cout][3][3][cin] then the biases [cout], layers in the order
 * W0, Wa_0, Wb_0, ..., Wa_{n-1}, Wb_{n-1}, Wh.  All arrays are host memory
 * owned by the caller.  h*w must be a multiple of 16.
 */
#ifndef MULTEM_SURROGATE_H
#define MULTEM_SURROGATE_H

#include <stddef.h>

#include "multem_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mt_sg_net mt_sg_net;

/* New network on GPU `device` with n_blocks residual blocks of `channels`
   channels (1..64, a multiple of 4); parameters are zero until set. */
mt_status mt_hip_sg_create(int n_blocks, int channels, int device, mt_sg_net **net);
void mt_hip_sg_destroy(mt_sg_net *net);
size_t mt_hip_sg_param_count(const mt_sg_net *net);

/* Copies parameters in / out; setting them also resets the Adam state. */
mt_status mt_hip_sg_set_params(mt_sg_net *net, const float *theta, size_t n);
mt_status mt_hip_sg_get_params(const mt_sg_net *net, float *theta, size_t n);

/* y = net(x) for `batch` images of h x w pixels. */
mt_status mt_hip_sg_forward(mt_sg_net *net, const float *x, int batch, int h, int w, float *y);

/* Mean squared error against t and its gradient w.r.t. the flat parameters
   (no update). */
mt_status mt_hip_sg_loss_grad(mt_sg_net *net, const float *x, const float *t, int batch, int h, int w,
                              double *loss, float *grad);

/* Adam training on a resident data set of n_images input/target pairs:
   `steps` steps of minibatches of `batch` images drawn with replacement by a
   generator seeded with `seed`, each image randomly translated (periodic)
   and transformed by one of the 8 symmetries of the square.  Step k uses
   learning rate lr * 0.5 * (1 + cos(pi k / steps)).  losses (optional,
   `steps` values) receives the minibatch loss before each update. */
mt_status mt_hip_sg_train(mt_sg_net *net, const float *x, const float *t, int n_images, int h, int w, int batch,
                          int steps, float lr, unsigned long long seed, float *losses);

/* Device time (ms) of the last forward call (HIP events around its kernels). */
double mt_hip_sg_last_forward_ms(const mt_sg_net *net);

#ifdef __cplusplus
}
#endif

#endif /* MULTEM_SURROGATE_H */
