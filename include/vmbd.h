/*
 * vmbd.h -- C ABI of the MI355X (gfx950) EFF/OLA operator library
 * (libvmbd_hip.so).
 *
 * Every entry point takes raw device pointers (float32, C order), sizes and a
 * HIP stream (hipStream_t passed as void*), enqueues its work asynchronously
 * on that stream and returns 0 on success or a negative VMBD_E* code; the
 * message of the last failure on the calling thread is in vmbd_last_error().
 * No entry point synchronises the device.
 *
 * A plan belongs to the device that was current when it was created; calls
 * made while another device is current fail with VMBD_E_ARG.  A plan's
 * workspaces serve one call at a time: calls on one plan must be ordered
 * (one stream, or streams the caller synchronises); separate plans are
 * independent.
 *
 * Reference interfaces replaced (files of the reference repository):
 *   plan + windows      imagetools.py:398-454 (win2winaux), olaGPU.py:95-166
 *   operator            olaGPU.py:170-311 (OlaGPU.cnv / OlaGPU.cnvtp) and the
 *                       kernels they launch in gputools.cu:16-314 plus the
 *                       pyfft plan (olaGPU.py:155-157, 813-824)
 */
#ifndef VMBD_H
#define VMBD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VMBD_OK 0
#define VMBD_E_ARG (-1)     /* bad argument / shape                       */
#define VMBD_E_HIP (-2)     /* HIP runtime error                          */
#define VMBD_E_NOMEM (-3)   /* device allocation failed                   */
#define VMBD_E_UNSUP (-4)   /* geometry outside what the kernels support  */

/* Output-size modes of olaGPU.py:127-136 (circ is not supported). */
#define VMBD_MODE_VALID 0
#define VMBD_MODE_SAME 1
#define VMBD_MODE_FULL 2

/* Flags of vmbd_ola_adjoint_xf. */
#define VMBD_REUSE_ROWS 1 /* x and f unchanged since the last vmbd_ola_forward on
                             this plan with the same nframes: reuse its row
                             transforms of x and f                            */

/* Which operand an OlaGPU instance caches (olaGPU.py:104-118). */
#define VMBD_CACHE_PSF 0   /* 'F' instance: spectra of the PSF stack          */
#define VMBD_CACHE_IMAGE 1 /* 'X' instance: spectra of the windowed patches    */

typedef void* vmbd_stream_t; /* hipStream_t */

/* Patch-grid geometry (imagetools.py:423-436, olaGPU.py:127-141).
 * Image x is H x W; PSF stack f is (csf0*csf1) x sfh x sfw; patch
 * k = m*csf1 + n covers rows [m*nhop0, m*nhop0+sw0) and columns
 * [n*nhop1, n*nhop1+sw1); every patch is transformed on a Th x Tw tile with
 * Th >= sw0+sfh-1, Tw >= sw1+sfw-1, both powers of two in [16, 512].
 * max_frames (>= 1) bounds the frames of one batched call: the operators take
 * nframes frames of the same geometry stored back to back (x: nframes*H*W,
 * f: nframes*P*sfh*sfw, y: nframes*sy0*sy1) and process all their patches in
 * the same launches. */
typedef struct vmbd_geom {
  int32_t H, W;
  int32_t sfh, sfw;
  int32_t csf0, csf1;
  int32_t sw0, sw1;
  int32_t nhop0, nhop1;
  int32_t Th, Tw;
  int32_t mode;
  int32_t max_frames;
} vmbd_geom;

typedef struct vmbd_plan vmbd_plan;

/* Launch shapes of a plan's passes (no counterpart in the reference, whose
 * kernels have fixed 16x16 / 32x32 blocks, gputools.py:372-401); 0 =
 * automatic.  Any setting gives the same results (bit for bit): it only
 * moves work between workgroups.
 *   ola_threads  threads per workgroup of the overlap-add passes: 128, 192
 *                or 256 (one output row per workgroup; 128 and 192 need
 *                every row's transforms in one round of that many threads)
 *   col_threads  threads per workgroup of the column passes: 128 or 256
 *                (a workgroup carries 16 * threads / Th columns of one
 *                patch, i.e. threads / 64 wavefronts per column strip)
 *   image_rows   with separable windows (vmbd_plan_set_window_factors):
 *                1 transforms each image row once per patch column and
 *                applies the patch rows' factors in the column pass, 2
 *                once per patch, 0 chooses (see
 *                vmbd_plan_image_rows_shared).  The two forms round
 *                differently (each within the operators' accuracy). */
typedef struct vmbd_tuning {
  int32_t ola_threads;
  int32_t col_threads;
  int32_t image_rows;
} vmbd_tuning;

/* Library version (major*10000 + minor*100 + patch). */
int vmbd_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* vmbd_last_error(void);
/* Status code (VMBD_E*) of the last failed call on this thread, VMBD_OK if
 * the last call succeeded; for vmbd_plan_create, which returns NULL on
 * failure (no counterpart in the reference, whose errors are Python
 * exceptions, olaGPU.py:104-136). */
int vmbd_last_status(void);

/* Create a plan on the current HIP device.
 * win_table: host array (n_win, sw0, sw1) of float32 windows ws_k;
 * win_idx:   host array (csf0*csf1) mapping patch k to its window.
 * Replaces win2winaux's device windows (imagetools.py:439-454) and the
 * OlaGPU constructor's set-up (olaGPU.py:95-166).  Returns NULL on error. */
vmbd_plan* vmbd_plan_create(const vmbd_geom* geom, const float* win_table, int32_t n_win,
                            const int32_t* win_idx);
void vmbd_plan_destroy(vmbd_plan* plan);

/* Set / read the plan's launch shapes (VMBD_E_UNSUP for a shape the
 * plan's geometry cannot use, or image_rows = 1 without separable
 * windows).  Setting them drops the row transforms a forward left for
 * VMBD_REUSE_ROWS. */
int vmbd_plan_set_tuning(vmbd_plan* plan, const vmbd_tuning* tuning);
int vmbd_plan_get_tuning(const vmbd_plan* plan, vmbd_tuning* tuning);

/* Declare the plan's windows separable: every window class equals
 * nom / wimg, with nom(r, c) = max(a[r] * b[c], floor) (a: sw0 values, b:
 * sw1 values) and wimg the overlap-add of nom over the patch grid, as
 * win2winaux builds them (imagetools.py:314-344, 438-445).  The plan checks
 * that against its tables (relative 1e-5) and then runs the x-adjoint's
 * overlap-add in summed form: sum_n b(c) IFFT(sum_m a(r) Y[m, n][r]) / wimg,
 * with the pixels whose covering windows are all at the floor (the outer
 * rows and columns of the patch grid) taken as the reference takes them,
 * the mean of the covering patches.  The plan may then also transform each
 * image row once per patch column instead of once per covering patch,
 * with the patch rows' factors applied in the column pass (see
 * vmbd_plan_image_rows_shared).  VMBD_E_UNSUP when the tables do not factor (the
 * plan keeps its two-dimensional windows).  Allocates; call once after
 * vmbd_plan_create.  No counterpart in the reference's API. */
int vmbd_plan_set_window_factors(vmbd_plan* plan, const float* a, const float* b, float floor);

/* 1 when the forward and the two-adjoint call of nframes frames transform
 * the image rows once per patch column, 0 when once per patch, < 0 on a
 * bad argument.  By default (image_rows 0) they share them only when the
 * launches fill the GPU (a single 1024^2 frame does not); the factor load
 * of the column pass costs more than it saves in launches that leave most
 * of it idle.  vmbd_ola_adjoint_f follows them.  The
 * calls that transform the image rows alone (the image spectra, the forward
 * of cached PSF spectra) share them whenever the windows are separable and
 * image_rows is not 2.  No counterpart in the reference's API. */
int vmbd_plan_image_rows_shared(const vmbd_plan* plan, int32_t nframes);

/* Device bytes held by the plan (workspaces, windows, twiddles). */
size_t vmbd_plan_workspace_bytes(const vmbd_plan* plan);
/* Output size of the forward operator for the plan's mode (olaGPU.py:127-136). */
int vmbd_plan_out_shape(const vmbd_plan* plan, int32_t* sy0, int32_t* sy1);
/* Number of complex64 elements of the cached spectrum stack of one frame,
 * (csf0*csf1) * (Tw/2+1) * Th, for vmbd_spectra. */
size_t vmbd_plan_spectrum_elems(const vmbd_plan* plan);

/* y = B(x, f) = F x = X f for nframes frames.  x: H x W, f: P x sfh x sfw,
 * y: sy0 x sy1 per frame.  Replaces OlaGPU.cnv (olaGPU.py:170-257) with
 * neither operand cached. */
int vmbd_ola_forward(vmbd_plan* plan, int32_t nframes, const float* x, const float* f, float* y,
                     vmbd_stream_t stream);

/* xbar = F' g (transpose with respect to the image).  g: sy0 x sy1,
 * xbar: H x W.  exact=1: true adjoint of vmbd_ola_forward; exact=0: the
 * reference's patch-local transpose (olaGPU.py:276-311, residual chopped with
 * patch size sw; OLA extent sw). */
int vmbd_ola_adjoint_x(vmbd_plan* plan, int32_t nframes, const float* g, const float* f, float* xbar,
                       int32_t exact, vmbd_stream_t stream);

/* fbar = X' g (transpose with respect to the PSFs), fbar: P x sfh x sfw.
 * Replaces OlaGPU.cnvtp of an 'X' instance (olaGPU.py:276-302). */
int vmbd_ola_adjoint_f(vmbd_plan* plan, int32_t nframes, const float* g, const float* x, float* fbar,
                       int32_t exact, vmbd_stream_t stream);

/* Both transposes from one residual (shares the residual spectra).  flags:
 * VMBD_REUSE_ROWS skips the row transforms of x and f (see above). */
int vmbd_ola_adjoint_xf(vmbd_plan* plan, int32_t nframes, const float* g, const float* x, const float* f,
                        float* xbar, float* fbar, int32_t exact, int32_t flags, vmbd_stream_t stream);

/* Cached spectra, the OlaGPU constructor (olaGPU.py:142-157):
 * kind VMBD_CACHE_PSF  : z = f stack  -> spec = FFT(pad_stack(f))
 * kind VMBD_CACHE_IMAGE: z = image x  -> spec = FFT(chop_mod_pad(x, ws))
 * spec: complex64 device buffer of vmbd_plan_spectrum_elems() elements, in
 * the operator's layout [P][Tw/2+1][Th] (row bin, then column frequency):
 * bin 0 holds the column transform of the two real row bins 0 and Tw/2
 * packed as one complex column (bin0 + i*bin(Tw/2)); bin Tw/2 is zero. */
int vmbd_spectra(vmbd_plan* plan, int32_t nframes, int32_t kind, const float* z, void* spec,
                 vmbd_stream_t stream);

/* The two OlaGPU constructors by name: vmbd_spectra with kind PSF ('F'
 * instance, olaGPU.py:142-146) or IMAGE ('X' instance, olaGPU.py:148-157). */
int vmbd_cache_psf_spectra(vmbd_plan* plan, int32_t nframes, const float* f, void* spec, vmbd_stream_t stream);
int vmbd_cache_image_spectra(vmbd_plan* plan, int32_t nframes, const float* x, void* spec, vmbd_stream_t stream);

/* The normalised per-patch windows, ws (csf0*csf1) x sw0 x sw1 float32 on the
 * device (win2winaux.ws_gpu, imagetools.py:439-454), expanded from the plan's
 * window classes. */
int vmbd_windows(const vmbd_plan* plan, float* ws, vmbd_stream_t stream);

/* OlaGPU.cnv with the cached operand (olaGPU.py:170-257):
 * kind PSF: u = image x; kind IMAGE: u = PSF stack.  y: sy0 x sy1. */
int vmbd_ola_forward_cached(vmbd_plan* plan, int32_t nframes, int32_t kind, const void* spec, const float* u,
                            float* y, vmbd_stream_t stream);

/* OlaGPU.cnvtp with the cached operand (olaGPU.py:261-311):
 * kind PSF -> out = xbar (H x W); kind IMAGE -> out = fbar (P x sfh x sfw). */
int vmbd_ola_adjoint_cached(vmbd_plan* plan, int32_t nframes, int32_t kind, const void* spec, const float* g,
                            float* out, int32_t exact, vmbd_stream_t stream);

/* Solver forms of the cached operator (no counterpart in the reference,
 * whose objectives form the residual and the Tikhonov term in separate
 * PyCUDA passes, optGPU.py:33-46, olaGPU.py:756-811); the extra arithmetic
 * is done in the final store of the operator:
 *   forward_cached_ex: y = (B u - sub) * mul     (sub, mul may be NULL)
 *   adjoint_cached_ex: out = B' g + beta * z     (z may be NULL)
 * with sub, mul shaped like y and z like out. */
int vmbd_ola_forward_cached_ex(vmbd_plan* plan, int32_t nframes, int32_t kind, const void* spec, const float* u,
                               const float* sub, const float* mul, float* y, vmbd_stream_t stream);
int vmbd_ola_adjoint_cached_ex(vmbd_plan* plan, int32_t nframes, int32_t kind, const void* spec, const float* g,
                               const float* z, float beta, float* out, int32_t exact, vmbd_stream_t stream);

/* Closed-form deconvolutions of OlaGPU.deconv on the plan (one frame): the
 * observation's patches are transformed by the row pass, the per-patch
 * Fourier division (gputools.cu:316-406) is formed on the half spectra in the
 * column pass with the filter energies in closed form, and the inverse rows
 * go straight into the output: three launches per division instead of the
 * reference's chop / pad / FFT / divide / IFFT / crop / OLA sequence.
 * spec is the plan's cached spectrum (vmbd_spectra) of the instance's
 * operand; work holds vmbd_deconv_workspace_bytes() bytes of device memory.
 *
 * vmbd_deconv_direct  ('xdirect', olaGPU.py:550-609): per patch
 *   conj(C) Y / (|C|^2 + alpha L + beta), L the energy of the [-1 1] and
 *   [-1 1]' filters, Y the observation's patch at the mode offset times the
 *   window at tile offset 0 (the reference's placement).  kind PSF ('F'
 *   instance): out = image H x W (patches cropped to sw, overlap-added);
 *   kind IMAGE ('X' instance): out = PSFs P x sfh x sfw.
 * vmbd_deconv_gdirect ('gdirect', olaGPU.py:416-511, 'X' instance): PSFs from
 *   the valid-mode gradients xx (H x W-1), xy (H-1 x W) of the latent image
 *   and yx (sy0 x sy1-1), yy (sy0-1 x sy1) of the (tapered) observation:
 *   (conj(Xx) Yx + conj(Xy) Yy) / (|Xx|^2 + |Xy|^2 + alpha L + beta), L of
 *   the 3x2 / 2x3 difference filters.
 * vmbd_deconv_sparse_setup / _step ('sparse', olaGPU.py:613-733, 'F'
 *   instance), the half-quadratic splitting: setup forms ftpy = conj(F) Y (a
 *   spectrum of vmbd_plan_spectrum_elems() complex64) from the tapered
 *   observation y and cfac = 0.01 / max(xdirect-like deconvolution of the
 *   tapered constant image c, 0.01) (H x W); each step takes the shrunk
 *   gradients wx (wx_rows x wx_cols), wy of the current estimate and writes
 *   x = max(cfac * OLA(windows * IFFT((conj(Gx) Wx + conj(Gy) Wy +
 *   ftpy/(alpha beta)) / (|F|^2/(alpha beta) + Ls + gamma))), 0). */
size_t vmbd_deconv_workspace_bytes(const vmbd_plan* plan);
int vmbd_deconv_direct(vmbd_plan* plan, int32_t kind, const void* spec, const float* y, float alpha, float beta,
                       float* out, vmbd_stream_t stream);
int vmbd_deconv_gdirect(vmbd_plan* plan, const float* xx, const float* xy, const float* yx, const float* yy,
                        float alpha, float beta, float* fout, void* work, vmbd_stream_t stream);
int vmbd_deconv_sparse_setup(vmbd_plan* plan, const void* spec, const float* y, const float* c, float alpha,
                             float gamma, void* ftpy, float* cfac, void* work, vmbd_stream_t stream);
int vmbd_deconv_sparse_step(vmbd_plan* plan, const void* spec, const void* ftpy, const float* wx, int32_t wx_rows,
                            int32_t wx_cols, const float* wy, int32_t wy_rows, int32_t wy_cols, float alpha,
                            float beta, float gamma, const float* cfac, float* x, void* work, vmbd_stream_t stream);

/* Batches of frames of different geometry (no counterpart in the reference,
 * which runs one OlaGPU instance at a time; its batched pyfft call over all
 * patches of one instance is olaGPU.py:813-824).  A batch binds n plans
 * (each with its own geometry, tile and PSF size) and, per plan, nframes[i]
 * resident frames: x[i], f[i], g[i] in and y[i], xbar[i], fbar[i] out, laid
 * out as for the batched operators of that plan.  vmbd_batch_step runs the
 * forward and both transposes (exact or patch-local) of every frame of every
 * plan, each pass of all plans as one launch over a device table of groups
 * (one launch per pass and distinct tile: the row, overlap-add and stack
 * passes are specialised to Tw, the column passes to Th), on `stream`.  The
 * plans and buffers must outlive the batch; the batch uses the plans'
 * workspaces, so calls on those plans and steps of the batch must be
 * ordered.  At most 16 plans; all on the current device. */
typedef struct vmbd_batch vmbd_batch;
vmbd_batch* vmbd_batch_create(vmbd_plan* const* plans, int32_t n, const int32_t* nframes, const float* const* x,
                              const float* const* f, const float* const* g, float* const* y, float* const* xbar,
                              float* const* fbar, int32_t exact);
void vmbd_batch_destroy(vmbd_batch* batch);
/* launches per step (passes x distinct tiles) */
int vmbd_batch_launches(const vmbd_batch* batch);
int vmbd_batch_step(vmbd_batch* batch, vmbd_stream_t stream);

/* Opt-in per-kernel timing (no counterpart in the reference, whose only
 * timing is wall-clock prints, stopwatch.py:5-22).  When enabled, every kernel
 * launch is bracketed by HIP events on its stream; vmbd_prof_report()
 * synchronises them and writes "name count total_us avg_us" lines into buf
 * and returns the bytes needed.  vmbd_prof_reset() drops the records. */
int vmbd_prof_enable(int32_t on);
int vmbd_prof_reset(void);
int vmbd_prof_report(char* buf, int32_t len);

#ifdef __cplusplus
}
#endif
#endif /* VMBD_H */
