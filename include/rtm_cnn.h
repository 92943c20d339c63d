/* rtm_cnn.h -- C ABI of the image-domain deblurring network (librtm_hip.so).
 *
 * Not in the reference, which stops at the migrated image (kernel.cu:935-1059).
 * A small convolutional network approximates the inverse Hessian of Born
 * modelling in the image domain: it maps a migrated image (input channel 0,
 * plus optional auxiliary channels such as depth) to a reflectivity estimate,
 *   y = x[0] + net(x),
 * net = nlayer 3x3 convolutions with bias (zero padding, stride 1), ReLU after
 * every layer but the last.  Convolutions run on the matrix cores with BF16
 * operands and FP32 accumulation; parameters, Adam state and reductions are
 * FP32 / FP64.  Training pairs come from the engine itself (random
 * reflectivity -> Born data -> adjoint image), see rtm_hip.h.
 *
 * Parameters are one FP32 vector: per layer W[cout][cin][3][3] then b[cout].
 * Images are [n][channels][h][w] FP32 on the host. */
#ifndef RTM_CNN_H
#define RTM_CNN_H

#ifdef __cplusplus
extern "C" {
#endif

enum { RTM_CNN_OK = 0, RTM_CNN_E_ARG = -1, RTM_CNN_E_HIP = -2, RTM_CNN_E_NOMEM = -3,
       RTM_CNN_E_STATE = -4 };

typedef struct rtm_cnn rtm_cnn;

/* channels[0..nlayer]: input channels, hidden widths, 1 output channel; every
 * width 1..32, nlayer 1..16.  dilation[0..nlayer-1]: per-layer dilation 1..64
 * of the 3x3 taps (NULL: all 1; zero padding of the same width).  max_pixels
 * (at most 2^25) bounds n*h*w of a batch, max_rows (at most max_pixels)
 * bounds n*h.  Parameters start at zero (the network then passes x[0]). */
int rtm_cnn_create(int nlayer, const int* channels, const int* dilation, long long max_pixels,
                   long long max_rows, rtm_cnn** out);
int rtm_cnn_destroy(rtm_cnn* net);
const char* rtm_cnn_last_error(const rtm_cnn* net);
int rtm_cnn_num_params(const rtm_cnn* net);
/* Set / read the FP32 parameters; setting resets the Adam state. */
int rtm_cnn_set_params(rtm_cnn* net, const float* params);
int rtm_cnn_get_params(rtm_cnn* net, float* params);
/* y[n][h][w] = x[:,0] + net(x) for x[n][c0][h][w]. */
int rtm_cnn_forward(rtm_cnn* net, const float* x, int n, int h, int w, float* y);
/* *loss = mean((y - t)^2) over the batch and, when grad is not NULL, its
 * gradient with respect to the parameters (kept for rtm_cnn_adam_step). */
int rtm_cnn_loss_grad(rtm_cnn* net, const float* x, const float* t, int n, int h, int w,
                      double* loss, float* grad);
/* One Adam step (bias-corrected) with the last gradient. */
int rtm_cnn_adam_step(rtm_cnn* net, float lr, float beta1, float beta2, float eps);
/* Mini-batch training on a device-resident copy of x[n][c0][h][w] and
 * targets t[n][h][w]: epochs passes over a seeded shuffle, Adam (0.9, 0.999,
 * 1e-8) at learning rate lr; the mean batch loss of each epoch goes to
 * loss_per_epoch[epochs] when not NULL. */
int rtm_cnn_fit(rtm_cnn* net, const float* x, const float* t, int n, int h, int w, int epochs,
                int batch, float lr, unsigned seed, double* loss_per_epoch);

#ifdef __cplusplus
}
#endif
#endif
