/*
 * vmbd_plain.h -- C ABI of the reference-decomposition ("plain") kernels of
 * libvmbd_hip.so: one entry point per kernel that gputools.py launches on
 * the OLA path, with the same index semantics.  All sizes are element
 * counts, all pointers device pointers, all calls asynchronous on `stream`;
 * return 0 or a negative VMBD_E* code (message in vmbd_last_error()).
 * `cplx` selects complex64 (interleaved re,im) instead of float32 for the
 * buffer the entry point's comment names.
 */
#ifndef VMBD_PLAIN_H
#define VMBD_PLAIN_H

#include "vmbd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Complex product modes of vmbd_cmul. */
#define VMBD_CMUL_AB 0        /* a * b        */
#define VMBD_CMUL_A_CONJB 1   /* a * conj(b)  */
#define VMBD_CMUL_CONJA_B 2   /* conj(a) * b  */

/* out (m x n, complex if cplx) = in (p x q, float) placed at (s, t), zero
 * elsewhere.  Replaces pad_cpu2gpu / zeroPad(Complex)Kernel
 * (gputools.py:1099-1137, gputools.cu:154-182). */
int vmbd_pad(void* out, int32_t m, int32_t n, const float* in, int32_t p, int32_t q, int32_t s, int32_t t,
             int32_t cplx, vmbd_stream_t stream);

/* out (csf0*csf1 x sz0 x sz1, complex if cplx): patch k of x (H x W, zero
 * padded to sxp = nhop*csf + sw - nhop) at tile offset (off0, off1), times
 * win[win_idx[k]] (sw0 x sw1 windows) when win != NULL (win_idx == NULL:
 * window k).  Replaces chop_pad_GPU / chop_mod_pad_GPU (gputools.py:322-401,
 * gputools.cu:16-126). */
int vmbd_chop_pad(void* out, int32_t sz0, int32_t sz1, const float* x, int32_t H, int32_t W, int32_t csf0,
                  int32_t csf1, int32_t sw0, int32_t sw1, int32_t nhop0, int32_t nhop1, int32_t off0, int32_t off1,
                  const float* win, const int32_t* win_idx, int32_t cplx, vmbd_stream_t stream);

/* out (K x sz0 x sz1, complex if cplx) = each slice of in (K x xm x xn) at
 * (off0, off1).  Replaces pad_stack_GPU (gputools.py:1140-1181,
 * gputools.cu:184-222). */
int vmbd_pad_stack(void* out, int32_t K, int32_t sz0, int32_t sz1, const float* in, int32_t xm, int32_t xn,
                   int32_t off0, int32_t off1, int32_t cplx, vmbd_stream_t stream);

/* out (K x sz0 x sz1, float) = in (K x inm x inn, real part if cplx) cropped at
 * (off0, off1).  Replaces crop_stack_GPU (gputools.py:599-640,
 * gputools.cu:225-251). */
int vmbd_crop_stack(float* out, int32_t K, int32_t sz0, int32_t sz1, const void* in, int32_t inm, int32_t inn,
                    int32_t off0, int32_t off1, int32_t cplx, vmbd_stream_t stream);

/* out (m x n, float) = in (p x q, real part if cplx) cropped at (s, t).
 * Replaces crop_gpu2cpu (gputools.py:546-569, gputools.cu:254-274). */
int vmbd_crop(float* out, int32_t m, int32_t n, const void* in, int32_t p, int32_t q, int32_t s, int32_t t,
              int32_t cplx, vmbd_stream_t stream);

/* Overlap-add gather: out ((nhop*csf + w - nhop) per dim) sums, at every
 * pixel, element (om, on) + (pixel - origin) of each patch whose extent
 * (wm, wn) covers it; in is (csf0*csf1 x inm x inn), real part if cplx.
 * Replaces ola_GPU_test (gputools.py:1068-1096, gputools.cu:276-314); the
 * extent must fit in the slices (the reference reads past them, see
 * olaGPU.py:307-309). */
int vmbd_ola(float* out, const void* in, int32_t inm, int32_t inn, int32_t csf0, int32_t csf1, int32_t wm,
             int32_t wn, int32_t nhop0, int32_t nhop1, int32_t om, int32_t on, int32_t cplx, vmbd_stream_t stream);

/* In-place batched 2-D complex FFT of `batch` Th x Tw complex64 slices;
 * dir -1 forward (e^{-i...}), +1 inverse (unnormalised), result scaled by
 * `scale`.  twiddles: device table filled by vmbd_twiddles (1024 complex64).
 * Replaces the pyfft plan (olaGPU.py:155-157, 813-824). */
int vmbd_fft2(void* data, int32_t batch, int32_t Th, int32_t Tw, int32_t dir, float scale, const void* twiddles,
              vmbd_stream_t stream);
int vmbd_twiddles(void* twiddles, vmbd_stream_t stream);

/* out[e] = op(a[e], b[e % bper]) * scale over n complex64 elements
 * (the GPUArray products of olaGPU.py:217, 228, 295). */
int vmbd_cmul(void* out, const void* a, const void* b, int64_t n, int64_t bper, int32_t mode, float scale,
              vmbd_stream_t stream);

/* Fourier-domain divisions of the closed-form deconvolutions; L, gx, gy are
 * broadcast with period `per` (one tile).
 *   vmbd_fourier_div  : conj(f) y / (|f|^2 + alpha L + beta)
 *                       (comp_ola_deconv, gputools.cu:316-341)
 *   vmbd_fourier_div_g: (conj(xx) yx + conj(xy) yy) / (|xx|^2 + |xy|^2 + alpha L + beta)
 *                       (comp_ola_gdeconv, gputools.cu:343-374)
 *   vmbd_fourier_div_s: (conj(gx) wx + conj(gy) wy + ftpy/(alpha beta)) /
 *                       (|f|^2/(alpha beta) + L + gamma)
 *                       (comp_ola_sdeconv, gputools.cu:376-406) */
int vmbd_fourier_div(void* out, const void* f, const void* y, const void* L, int64_t n, int64_t per, float alpha,
                     float beta, vmbd_stream_t stream);
int vmbd_fourier_div_g(void* out, const void* xx, const void* xy, const void* yx, const void* yy, const void* L,
                       int64_t n, int64_t per, float alpha, float beta, vmbd_stream_t stream);
int vmbd_fourier_div_s(void* out, const void* gx, const void* gy, const void* wx, const void* wy, const void* ftpy,
                       const void* f, const void* L, int64_t n, int64_t per, float alpha, float beta, float gamma,
                       vmbd_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* VMBD_PLAIN_H */
