/* mfista_learn.h -- kernels of the learned warm start (a 4-level U-Net with
 * BF16 weights and FP32 accumulation) in libmfista_amd.so.
 *
 * No counterpart in the reference: the reference has no learned component;
 * these serve the warm start of mfista_imaging_core_nufft (cold-start MFISTA
 * stays the fallback and decides which start is kept).  Every array is a device pointer; tensors
 * are NHWC FP32, 3x3 weights are [ky][kx][ci][co] in BF16 (uint16_t), the FP32
 * master copy lives with the optimiser.  `stream` is a hipStream_t (NULL = the
 * default stream).  Functions return 0 or a negative status. */
#ifndef MFISTA_LEARN_H
#define MFISTA_LEARN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UNET_E_ARG (-5)
#define UNET_E_HIP (-100)

/* out = relu?(bias + conv3x3(concat(in1[C1], in2[C2]))), zero padding; in2 may be NULL (C2 = 0) */
int unet_conv3x3_fwd(const float *in1, int C1, const float *in2, int C2, const uint16_t *w,
                     const float *bias, float *out, int B, int H, int W, int Cout, int relu,
                     void *stream);
/* gradient of the input (split into the two concatenated parts; either may be NULL) */
int unet_conv3x3_bwd_data(const float *dout, int Cout, const uint16_t *w, float *din1, int C1,
                          float *din2, int C2, int B, int H, int W, void *stream);
/* accumulates the weight gradient dw[9][Cin][Cout] (FP32) and, if db != NULL, the bias gradient */
int unet_conv3x3_bwd_weight(const float *in1, int C1, const float *in2, int C2, const float *dout,
                            int Cout, float *dw, float *db, int B, int H, int W, void *stream);
/* dout[i] = 0 where out[i] <= 0 */
int unet_relu_bwd(const float *out, float *dout, long n, void *stream);
/* 2x2 average pooling and its gradient (added to din if accumulate != 0);
 * H, W are the input (fine) sizes */
int unet_avgpool2_fwd(const float *in, float *out, int B, int H, int W, int C, void *stream);
int unet_avgpool2_bwd(const float *dout, float *din, int B, int H, int W, int C, int accumulate,
                      void *stream);
/* nearest-neighbour 2x upsampling and its gradient; H, W are the input (coarse) sizes */
int unet_upsample2_fwd(const float *in, float *out, int B, int H, int W, int C, void *stream);
int unet_upsample2_bwd(const float *dout, float *din, int B, int H, int W, int C, void *stream);
/* 1x1 convolution to one channel (C <= 32): out[p] = relu?(bias + sum_c in[p][c] w[c]) */
int unet_head_fwd(const float *in, int C, const uint16_t *w, const float *bias, float *out, long P,
                  int relu, void *stream);
/* its gradients: din (may be NULL), accumulated dw[C] and db[1] */
int unet_head_bwd(const float *in, int C, const uint16_t *w, const float *dout, float *din,
                  float *dw, float *db, long P, void *stream);
/* *loss = mean (y - t)^2 (FP64), dy = 2 (y - t) / n (dy may be NULL) */
int unet_mse(const float *y, const float *t, float *dy, double *loss, long n, void *stream);
/* per-sample weighted squared error: loss = sum_b wt[b] sum_i (y - t)^2 / n over
 * samples of `per` elements (n a multiple of per), and its gradient (dy may be NULL) */
int unet_mse_weighted(const float *y, const float *t, const float *wt, long per, float *dy, double *loss,
                      long n, void *stream);
/* one Adam step on the FP32 master weights w; refreshes the BF16 copy w16 */
int unet_adam(float *w, uint16_t *w16, const float *g, float *m, float *v, long n, float lr,
              float beta1, float beta2, float eps, int step, void *stream);
/* w16 = bf16(w), round to nearest even */
int unet_to_bf16(const float *w, uint16_t *w16, long n, void *stream);

#ifdef __cplusplus
}
#endif

#endif
