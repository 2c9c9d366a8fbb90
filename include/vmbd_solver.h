/*
 * vmbd_solver.h -- C ABI of the solver kernels of libvmbd_hip.so: clipping,
 * PSF sum-to-one projection, deterministic reductions, the projected
 * Barzilai-Borwein step and the multiplicative x-step of the online loop.
 * Device pointers, element counts, asynchronous on `stream`; return 0 or a
 * negative VMBD_E* code.  No entry point synchronises: step lengths and
 * reduction results stay on the device, so a fixed number of iterations can
 * be captured in a HIP graph.
 *
 * Reduction scratch: `partials` must hold 3 * vmbd_reduce_blocks(n) doubles.
 */
#ifndef VMBD_SOLVER_H
#define VMBD_SOLVER_H

#include "vmbd.h"

#ifdef __cplusplus
extern "C" {
#endif

#define VMBD_MAX_PARTIALS 512
#define VMBD_CLIP_LOWER 1
#define VMBD_CLIP_UPPER 2

/* Number of partial sums a reduction over n elements uses. */
int vmbd_reduce_blocks(int64_t n);

/* x = max(x, lb) and/or min(x, ub) in place.  Replaces cliplower_GPU,
 * clipupper_GPU, clip_GPU (gputools.py:472-495). */
int vmbd_clip(float* x, int64_t n, float lb, float ub, int32_t flags, vmbd_stream_t stream);

/* f_k /= sum(f_k) for each of K slices of len elements whose sum exceeds eps
 * (K = 1: global).  Replaces normalize (gputools.py:1025-1044) without the
 * device-host round trip. */
int vmbd_psf_normalize(float* f, int32_t K, int32_t len, float eps, vmbd_stream_t stream);

/* *out = sum a*b (double accumulation, fixed reduction order).  Replaces
 * cua.dot (olaGPU.py:768-784, optGPU.py:27, 42). */
int vmbd_dot(const float* a, const float* b, int64_t n, double* partials, double* out, vmbd_stream_t stream);

/* Projected Barzilai-Borwein step (optGPU.py:198-241):
 * vmbd_pbb_stats : partials of sum dx^2, sum dx dg, sum dg^2 with
 *                  dx = x - oldx, dg = g - oldg zeroed where x == 0 and g > 0
 *                  (clip2bound, optGPU.py:129-132) unless unconstrained;
 *                  also the three sums in `sums` when non-NULL.
 * vmbd_pbb_update: xnew = P(x - step g), step = BB1 (iter even) or BB2 (odd)
 *                  with the 1e-5 guards of optGPU.py:222-225, P = projection
 *                  onto x >= 0 unless unconstrained; `step_out` (optional)
 *                  receives the step.  Uses the partials of the matching
 *                  vmbd_pbb_stats call (same n). */
int vmbd_pbb_stats(const float* x, const float* oldx, const float* g, const float* oldg, int64_t n,
                   int32_t unconstrained, double* partials, double* sums, vmbd_stream_t stream);
int vmbd_pbb_update(float* xnew, const float* x, const float* g, int64_t n, const double* partials, int32_t iter,
                    int32_t unconstrained, float* step_out, vmbd_stream_t stream);

/* One projected BB iteration in one call (vmbd_pbb_stats + vmbd_pbb_update,
 * optGPU.py:200-245): xnew = P(x - step g) with the step of iteration `iter`
 * computed on the device from (x - oldx, g - oldg); partials:
 * 3 * vmbd_reduce_blocks(n) doubles. */
int vmbd_pbb_step(float* xnew, const float* x, const float* oldx, const float* g, const float* oldg, int64_t n,
                  int32_t iter, int32_t unconstrained, double* partials, float* step_out, vmbd_stream_t stream);

/* vmbd_pbb_step with its reduction scratch in a static device array of the
 * library instead of a caller's `partials` buffer: it allocates nothing
 * (no hipMalloc, no caller workspace).  No counterpart in the reference; it
 * is for callers that cannot allocate: code under HIP graph capture, tight
 * solver loops, harnesses that forbid allocation.  Limit: the scratch holds
 * 3 * VMBD_MAX_PARTIALS doubles, the partial sums of a step over any n (the
 * reductions use at most VMBD_MAX_PARTIALS blocks); larger needs are
 * refused with VMBD_E_UNSUP.  Each GPU holds one copy until the process
 * ends, and every call on that GPU shares it: calls that overlap on one GPU
 * (other streams, other host threads) race and must be ordered by the
 * caller. */
int vmbd_pbb_step_noalloc(float* xnew, const float* x, const float* oldx, const float* g, const float* oldg,
                          int64_t n, int32_t iter, int32_t unconstrained, float* step_out, vmbd_stream_t stream);

/* Start pair of PBB (optGPU.py:256-297, replaces the host comparison of
 * the two start objectives): dots = device [r.r, z.z] at the first point and
 * at the second, objective r.r/2 + beta z.z/2; when the second point is not
 * worse, (x, oldx) and (g, oldg) exchange contents.  No host synchronisation. */
int vmbd_pbb_select(float* x, float* oldx, float* g, float* oldg, int64_t n, const double* dots, double beta,
                    vmbd_stream_t stream);

/* out = a*s + b*t (b may be NULL). */
int vmbd_axpby(float* out, const float* a, float s, const float* b, float t, int64_t n, vmbd_stream_t stream);
/* Elementwise family: out = a*b, a/b, s/a, a-b, a + s*b or s*a (op
 * VMBD_EW_*); out may alias a or b. */
#define VMBD_EW_MUL 0
#define VMBD_EW_DIV 1
#define VMBD_EW_SDIV 2
#define VMBD_EW_SUB 3
#define VMBD_EW_AXPY 4
#define VMBD_EW_SCALE 5
#define VMBD_EW_LT 6   /* out = (a < s) ? 1 : 0 (saturation mask, vmbd_chimney.py:117) */
#define VMBD_EW_ADDS 7 /* out = a + s */
int vmbd_ew(float* out, const float* a, const float* b, float s, int64_t n, int32_t op, vmbd_stream_t stream);

/* out = a * b elementwise. */
int vmbd_mul(float* out, const float* a, const float* b, int64_t n, vmbd_stream_t stream);

/* Isotropic eps-smoothed total variation of K images H x W (the TV term of
 * the autograd loss; no counterpart in the reference, which has no TV prior):
 * TV = sum sqrt(dx^2 + dy^2 + eps^2) with forward differences, zero past the
 * last row/column.  value_out: one device double; partials:
 * vmbd_reduce_blocks(K*H*W) doubles. */
int vmbd_tv(const float* x, int32_t K, int32_t H, int32_t W, float eps, double* partials, double* value_out,
            vmbd_stream_t stream);

/* grad = (accumulate ? grad : 0) + weight * scale[0] * dTV/dx; scale is a
 * device float (the upstream gradient) or NULL for 1. */
int vmbd_tv_grad(float* grad, const float* x, int32_t K, int32_t H, int32_t W, float eps, const float* scale,
                 float weight, int32_t accumulate, vmbd_stream_t stream);

/* Multiplicative latent-image step (vmbd_chimney.py:143-147):
 * x *= max(num / (den + tol), tol); then x = min(x, max of x over rows
 * [border0, H - border0) and columns [border1, W - border1)).  partial_max:
 * vmbd_reduce_blocks(H*W) floats; xmax_out (optional) receives the clip level. */
int vmbd_rl_update(float* x, const float* num, const float* den, int32_t H, int32_t W, float tol, int32_t border0,
                   int32_t border1, float* partial_max, float* xmax_out, vmbd_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* VMBD_SOLVER_H */
