/*
 * vmbd_filters.h -- C ABI of the image filters of libvmbd_hip.so (the
 * kernels of kernel.cu, edgetaper.cu and resize.cu of the reference).
 * float32 device images in C order, asynchronous on `stream`; return 0 or a
 * negative VMBD_E* code.  Stencils write a separate output (out != in).
 */
#ifndef VMBD_FILTERS_H
#define VMBD_FILTERS_H

#include "vmbd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One shock-filter step out = in - dt*sign(I_nn)*|grad I|/h (minmod
 * gradients; top/left border replicated, bottom/right neighbours 0).
 * Replaces shock (kernel.cu:9-88); shock_filter_gpu (gputools.py:1282-1314)
 * runs it `iter` times. */
int vmbd_shock_step(float* out, const float* in, int32_t H, int32_t W, float dt, float h, vmbd_stream_t stream);

/* Bilateral filter of half size w (w <= 24), spatial sigma_d, range sigma_r;
 * neighbours outside the image are skipped.  Replaces bfilter
 * (kernel.cu:90-136, gputools.py:251-285). */
int vmbd_bilateral(float* out, const float* in, int32_t H, int32_t W, int32_t w, float sigma_d, float sigma_r,
                   vmbd_stream_t stream);

/* Forward differences scaled by 1/sqrt(2): gx = (y[i,j]-y[i,j+1])/sqrt2,
 * gy = (y[i,j]-y[i+1,j])/sqrt2.  same=0: gx H x (W-1), gy (H-1) x W;
 * same=1: H x W with a zero last column/row.  Replaces gradient_valid/same
 * (kernel.cu:180-242, gputools.py:854-881). */
int vmbd_gradient(float* gx, float* gy, const float* y, int32_t H, int32_t W, int32_t same, vmbd_stream_t stream);

/* Replicate padding to (H+padh) x (W+padw), top/left = floor(pad/2).
 * Replaces impad (kernel.cu:244-266, gputools.py:884-907). */
int vmbd_impad(float* out, const float* in, int32_t H, int32_t W, int32_t padh, int32_t padw, vmbd_stream_t stream);

/* 5-point Laplacian (4c - sum of neighbours)/sqrt(20) of each of K slices;
 * same=0: (H-2) x (W-2) valid output, same=1: H x W with a zero border;
 * zero_end_slices=1 also zeroes the first and last slice (laplace_stack_same).
 * Replaces laplace_valid/same/stack_same (kernel.cu:268-375). */
int vmbd_laplace(float* out, const float* in, int32_t K, int32_t H, int32_t W, int32_t same, int32_t zero_end_slices,
                 vmbd_stream_t stream);

/* 7-point 3-D Laplacian 6c - sum of neighbours, zero on the faces.
 * Replaces laplace3d_same (kernel.cu:377-418). */
int vmbd_laplace3d(float* out, const float* in, int32_t D, int32_t H, int32_t W, vmbd_stream_t stream);

/* Hyper-Laplacian shrinkage, in place: argmin_x |x|^alpha + beta/2 (x-v)^2
 * by 4 Newton steps, 0 when not better than x = 0.  alpha23: alpha = 2/3
 * (modify_alpha23, kernel.cu:555-578); alpha: general (modify_alpha,
 * kernel.cu:529-553). */
int vmbd_shrink_alpha23(float* y, int64_t n, float beta, vmbd_stream_t stream);
int vmbd_shrink_alpha(float* y, int64_t n, float beta, float alpha, vmbd_stream_t stream);

/* In-place border taper: rows i < hy and i > H-1-hy are multiplied by
 * wy[i] / wy[H-1-i], columns likewise by wx (device tables of hx, hy floats).
 * Replaces edgetaper (edgetaper.cu:6-31, gputools.py:643-673). */
int vmbd_edgetaper(float* im, int32_t H, int32_t W, const float* wx, int32_t hx, const float* wy, int32_t hy,
                   vmbd_stream_t stream);

/* Bilinear resampling of K slices H x W to OH x OW at normalised coordinates
 * (j/OW, i/OH), texel centres at +0.5, clamped addressing.  Replaces resize
 * (resize.cu:5-15, gputools.py:1205-1266). */
int vmbd_resize(float* out, int32_t OH, int32_t OW, const float* in, int32_t K, int32_t H, int32_t W,
                vmbd_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* VMBD_FILTERS_H */
