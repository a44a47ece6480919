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

/* ---- the warm start's U-Net: convolutions with any channel counts (ci, co
 * multiples of 16 up to 256), activations bf16 [nb][nx*ny*nz][C] (channels
 * fastest), weights bf16 [co][27 taps][ci].
 * y = act(conv(x) + bias + res): bias (FP32 [co]) and res (bf16
 * [nb][nv][co], the U-Net's skip connection) optional; relu != 0: ReLU;
 * mirror != 0 reads tap t at offset -off_t (used by the data gradient); yf
 * (optional) gets the FP32 values. */
int cl_conv3d_fwd_c(const uint16_t *x, int nb, int nx, int ny, int nz, int ci, int co, const uint16_t *w,
                    const float *bias, const uint16_t *res, int relu, int mirror, uint16_t *y, float *yf,
                    void *stream);
/* wt[c][t][o] = w[o][t][c] */
int cl_conv3d_transpose_w(const uint16_t *w, int ci, int co, uint16_t *wt, void *stream);
/* dx = conv^T(dy) with the transposed weights wt of cl_conv3d_transpose_w */
int cl_conv3d_bwd_data_c(const uint16_t *dy, int nb, int nx, int ny, int nz, int ci, int co, const uint16_t *wt,
                         uint16_t *dx, void *stream);
/* grad = [dW (co*27*ci), db (co)] FP32, summed over voxels and the batch in
 * a fixed order; workspace of cl_conv3d_bwd_weight_c_workspace bytes */
long cl_conv3d_bwd_weight_c_workspace(int nb, int nx, int ny, int nz, int ci, int co);
int cl_conv3d_bwd_weight_c(const uint16_t *x, const uint16_t *dy, int nb, int nx, int ny, int nz, int ci, int co,
                           float *grad, float *workspace, void *stream);
/* Grid transfers between node-coincident levels of a grid of nx, ny, nz
 * nodes (all odd, >= 3) and the coarse grid of (n-1)/2+1 nodes per axis,
 * ch channels: full-weighting restriction coarse[c] = scale * sum_d w(d)
 * fine[2c+d] (w = product over axes of 1/2 at d = 0, 1/4 at d = +-1;
 * neighbours outside left out) and trilinear interpolation fine[f] = scale *
 * sum_c p(f-2c) coarse[c] (p = product of 1 at 0, 1/2 at +-1).  The
 * restriction's transpose is the interpolation / 8. */
int cl_restrict(const uint16_t *fine, int nb, int nx, int ny, int nz, int ch, float scale, uint16_t *coarse,
                void *stream);
int cl_interp(const uint16_t *coarse, int nb, int nx, int ny, int nz, int ch, float scale, uint16_t *fine,
              void *stream);
/* a += b, n bf16 values */
int cl_add_bf16(uint16_t *a, const uint16_t *b, long n, void *stream);
/* weighted mean squared error of y (FP32, ystride channels) against target
 * (FP32 [nvox][nout]) over the first nout channels, channel weights wch
 * (device FP32 [nout]): dy (bf16, ystride channels) = d loss / d y;
 * partial[block] = block sums of the weighted squared errors,
 * ceil(nvox*ystride/256) blocks */
int cl_wmse(const float *y, int ystride, const float *target, long nvox, int nout, const float *wch, uint16_t *dy,
            float *partial, void *stream);

/* The warm start's U-Net (fixed architecture): 4 levels of (n-1)/2^l + 1
 * nodes per axis (n - 1 divisible by 8, n >= 17), 16/32/48/64 channels, 15
 * convolutions with weights concatenated in layer order, channel pairs
 * (ci, co) = (16,16) (16,16) | (16,32) (32,32) | (32,48) (48,48) |
 * (48,64) (64,64) | (64,48)+skip (48,48) | (48,32)+skip (32,32) |
 * (32,16)+skip (16,16) (16,16); ReLU after all but the last.  Encoder:
 * two convolutions per level, the second's output kept as that level's
 * skip, then cl_restrict (scale 1); decoder: cl_interp (scale 1), a
 * convolution plus the level's skip, a second convolution.  Input x: bf16
 * [nb][nv][16]; out: FP32 [nb][nv][16] (the last convolution's values).
 * cl_unet_params gives the weight and bias counts (594432, 528). */
int cl_unet_params(long *n_weights, long *n_biases);
long cl_unet_workspace(int nb, int nx, int ny, int nz);
int cl_unet_forward(const uint16_t *x, int nb, int nx, int ny, int nz, const uint16_t *w, const float *bias,
                    float *out, void *workspace, void *stream);

#ifdef __cplusplus
}
#endif
#endif
