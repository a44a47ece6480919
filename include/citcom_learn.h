/*
 * citcom_learn.h -- C ABI of the learned warm start's GPU kernels (in
 * libcitcom_hip.so): a 3-D convolutional network, 3x3x3 kernels, 16
 * channels, zero padding, trained on the device.  No counterpart in the
 * reference: the network's output is only an initial guess for the Uzawa
 * solve (before Stokes_flow_Incomp.c:354), which still iterates to the same
 * criterion.
 *
 * Activations are bf16 [nb][nx*ny*nz][16] (node order k + nz*(i + nx*j)),
 * weights bf16 [16 out][27 taps][16 in], tap = (dz+1) + 3(dx+1) + 9(dy+1),
 * with FP32 master copies.  All products are v_mfma_f32_16x16x16_bf16 with
 * FP32 accumulation.  Pointers are device pointers; stream is a hipStream_t
 * (NULL: the default stream).  Every function returns 0 or -1.
 */
#ifndef CITCOM_LEARN_H
#define CITCOM_LEARN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* y = act(conv(x) + bias); act = ReLU when relu != 0; yf (optional) gets the
 * FP32 values */
int cl_conv3d_fwd(const uint16_t *x, int nb, int nx, int ny, int nz, const uint16_t *w, const float *bias, int relu,
                  uint16_t *y, float *yf, void *stream);
/* dy[i] = 0 where y[i] <= 0 (backward through ReLU), n entries */
int cl_relu_bwd(const uint16_t *y, uint16_t *dy, long n, void *stream);
/* dx = conv^T(dy) */
int cl_conv3d_bwd_data(const uint16_t *dy, int nb, int nx, int ny, int nz, const uint16_t *w, uint16_t *dx,
                       void *stream);
/* grad = [dW (16*27*16), db (16)] FP32, summed in a fixed order; workspace
 * of cl_conv3d_bwd_weight_workspace bytes */
long cl_conv3d_bwd_weight_workspace(int nb, int nx, int ny, int nz);
int cl_conv3d_bwd_weight(const uint16_t *x, const uint16_t *dy, int nb, int nx, int ny, int nz, float *grad,
                         float *workspace, void *stream);
/* The same weight gradient with no workspace argument and no allocation of
 * any kind: the partial sums live in a 16 MiB file-scope device array
 * (cl_scratch_bytes()), one copy per GPU, held until the process ends.  It
 * has no counterpart in the reference; it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight training or
 * serving loops, harnesses that forbid hipMalloc).  Returns -1, with the
 * reason in cl_last_error(), when nb volumes of nx*ny*nz voxels need more
 * than cl_scratch_bytes() of partial sums (cl_conv3d_bwd_weight_workspace).
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (different streams or threads) race. */
int cl_conv3d_bwd_weight_noalloc(const uint16_t *x, const uint16_t *dy, int nb, int nx, int ny, int nz, float *grad,
                                 void *stream);
long cl_scratch_bytes(void);
const char *cl_last_error(void);
/* mean squared error of y (FP32, 16 channels) against target over the first
 * nout channels: dy (bf16) = d loss / d y; partial[block] = block sums of
 * squared errors, ceil(nvox*16/256) blocks */
int cl_mse(const float *y, const float *target, long nvox, int nout, uint16_t *dy, float *partial, void *stream);
/* Adam step `step` (1-based) on FP32 weights w with gradient g, moments m, v;
 * wbf (optional) receives the bf16 copy; mask (optional) freezes entries */
int cl_adam(float *w, const float *g, float *m, float *v, long n, float lr, float b1, float b2, float eps, int step,
            uint16_t *wbf, const unsigned char *mask, void *stream);
int cl_to_bf16(const float *a, uint16_t *b, long n, void *stream);

#ifdef __cplusplus
}
#endif
#endif
