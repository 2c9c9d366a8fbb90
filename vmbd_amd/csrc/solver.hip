// Solver kernels for gfx950: projections, deterministic reductions and the
// projected Barzilai-Borwein step with its step length computed on the device.
//
// Reductions are two-stage with a fixed grid: every block writes one partial
// per quantity, and the consumer reduces the partials in a fixed order, so
// results are bitwise reproducible and no host synchronisation is needed
// (the iteration can be captured in a HIP graph).  Block sums use wavefront
// shuffles (DPP/permute through __shfl_xor) and one LDS exchange per wave.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <initializer_list>
#include <mutex>

#include "../../include/vmbd_solver.h"
#include "capi_util.h"
#include "prof.h"

namespace vmbd {
namespace {

constexpr int TB = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// block sum of up to 3 values (double accumulation), result valid in thread 0
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K]) {
  __shared__ double sh[K][TB / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    v[k] = wave_sum(v[k]);
    if (lane == 0) sh[k][w] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double s = 0.0;
      for (int i = 0; i < TB / 64; ++i) s += sh[k][i];
      v[k] = s;
    }
  }
  __syncthreads();
}

// sum of K partial arrays (layout [K][nb]) in fixed order, broadcast to the block
template <int K>
__device__ __forceinline__ void reduce_partials(const double* partials, int nb, double (&out)[K]) {
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += TB) s += partials[k * nb + i];
    v[k] = s;
  }
  block_sum<K>(v);
  __shared__ double bc[K];
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) bc[k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) out[k] = bc[k];
}

__global__ __launch_bounds__(TB) void k_clip(float* x, long long n, float lb, float ub, int flags) {
  for (long long e = (long long)blockIdx.x * TB + threadIdx.x; e < n; e += (long long)gridDim.x * TB) {
    float v = x[e];
    if (flags & VMBD_CLIP_LOWER) v = v < lb ? lb : v;
    if (flags & VMBD_CLIP_UPPER) v = v > ub ? ub : v;
    x[e] = v;
  }
}

// one block per slice: f_k /= sum(f_k) when the sum exceeds eps
__global__ __launch_bounds__(TB) void k_normalize(float* f, int len, float eps) {
  float* s = f + (long long)blockIdx.x * len;
  double v[1] = {0.0};
  for (int i = threadIdx.x; i < len; i += TB) v[0] += (double)s[i];
  block_sum<1>(v);
  __shared__ float tot;
  if (threadIdx.x == 0) tot = (float)v[0];
  __syncthreads();
  const float t = tot;
  if (t > eps)
    for (int i = threadIdx.x; i < len; i += TB) s[i] /= t;
}

// partials[0][b] = sum a*b over the block's share; (a == b allowed)
__global__ __launch_bounds__(TB) void k_dot_partials(const float* a, const float* b, long long n, double* partials) {
  double v[1] = {0.0};
  for (long long e = (long long)blockIdx.x * TB + threadIdx.x; e < n; e += (long long)gridDim.x * TB)
    v[0] += (double)a[e] * (double)b[e];
  block_sum<1>(v);
  if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
}

__global__ __launch_bounds__(TB) void k_reduce_out(const double* partials, int nb, int K, double* out) {
  if (K == 1) {
    double r[1];
    reduce_partials<1>(partials, nb, r);
    if (threadIdx.x == 0) out[0] = r[0];
  } else {
    double r[3];
    reduce_partials<3>(partials, nb, r);
    if (threadIdx.x == 0) {
      out[0] = r[0];
      out[1] = r[1];
      out[2] = r[2];
    }
  }
}

__device__ __forceinline__ void pbb_acc(double (&v)[3], float xv, float ox, float gv, float og, int unconstrained) {
  float dx = xv - ox;
  float dg = gv - og;
  if (!unconstrained && xv == 0.f && gv > 0.f) {
    dx = 0.f;
    dg = 0.f;
  }
  v[0] += (double)dx * dx;
  v[1] += (double)dx * dg;
  v[2] += (double)dg * dg;
}

__device__ __forceinline__ void pbb_acc4(double (&v)[3], float4 a, float4 b, float4 c, float4 d, int unc) {
  pbb_acc(v, a.x, b.x, c.x, d.x, unc);
  pbb_acc(v, a.y, b.y, c.y, d.y, unc);
  pbb_acc(v, a.z, b.z, c.z, d.z, unc);
  pbb_acc(v, a.w, b.w, c.w, d.w, unc);
}

// PBB, stage 1 (optGPU.py:202-207 + 222-225): dx = x - oldx, dg = g - oldg,
// zeroed where x == 0 and g > 0 (clip2bound, optGPU.py:129-132) unless
// unconstrained; partials of sum dx^2, sum dx dg, sum dg^2.  n4 > 0: the
// first 4*n4 elements are read as float4 (all operands 16-byte aligned), two
// chunks per thread in flight; the rest element by element.
__global__ __launch_bounds__(TB) void k_pbb_stats(const float* x, const float* oldx, const float* g,
                                                  const float* oldg, long long n, long long n4, int unconstrained,
                                                  double* partials) {
  double v[3] = {0.0, 0.0, 0.0};
  const long long stride = (long long)gridDim.x * TB;
  const long long tid = (long long)blockIdx.x * TB + threadIdx.x;
  const float4 *x4 = (const float4*)x, *ox4 = (const float4*)oldx, *g4 = (const float4*)g, *og4 = (const float4*)oldg;
  for (long long e = tid; e < n4; e += 2 * stride) {
    const long long e1 = e + stride;
    const float4 a = x4[e], b = ox4[e], c = g4[e], d = og4[e];
    float4 a1 = make_float4(0.f, 0.f, 0.f, 0.f), b1 = a1, c1 = a1, d1 = a1;
    if (e1 < n4) {
      a1 = x4[e1];
      b1 = ox4[e1];
      c1 = g4[e1];
      d1 = og4[e1];
    }
    pbb_acc4(v, a, b, c, d, unconstrained);
    if (e1 < n4) pbb_acc4(v, a1, b1, c1, d1, unconstrained);
  }
  for (long long e = 4 * n4 + tid; e < n; e += stride) pbb_acc(v, x[e], oldx[e], g[e], oldg[e], unconstrained);
  block_sum<3>(v);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = v[0];
    partials[gridDim.x + blockIdx.x] = v[1];
    partials[2 * gridDim.x + blockIdx.x] = v[2];
  }
}

// PBB, stage 2 (optGPU.py:222-229): BB1 step on even iterations,
// sum dx^2/(1e-5 + sum dx dg); BB2 on odd, sum dx dg/(1e-5 + sum dg^2);
// xnew = x - step g, projected onto x >= 0 unless unconstrained.  Every block
// reduces the stage-1 partials itself (same order -> same step everywhere).
__device__ __forceinline__ float pbb_new(float xv, float gv, float step, int unconstrained) {
  const float v = xv - gv * step;
  return (!unconstrained && v < 0.f) ? 0.f : v;
}

// n4 as in k_pbb_stats; the first two chunks of each thread are loaded
// before the partials are reduced, so their latency overlaps the reduction.
__global__ __launch_bounds__(TB) void k_pbb_update(float* xnew, const float* x, const float* g, long long n,
                                                   long long n4, const double* partials, int nb, int iter,
                                                   int unconstrained, float* step_out) {
  const long long stride = (long long)gridDim.x * TB;
  const long long tid = (long long)blockIdx.x * TB + threadIdx.x;
  const float4 *x4 = (const float4*)x, *g4 = (const float4*)g;
  const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
  float4 a0 = z4, c0 = z4, a1 = z4, c1 = z4;
  if (tid < n4) {
    a0 = x4[tid];
    c0 = g4[tid];
  }
  if (tid + stride < n4) {
    a1 = x4[tid + stride];
    c1 = g4[tid + stride];
  }
  double s[3];
  reduce_partials<3>(partials, nb, s);
  const float sxx = (float)s[0], sxg = (float)s[1], sgg = (float)s[2];
  const float step = (iter % 2 == 0) ? sxx / (1e-5f + sxg) : sxg / (1e-5f + sgg);
  if (blockIdx.x == 0 && threadIdx.x == 0 && step_out) *step_out = step;
  float4* o4 = (float4*)xnew;
  for (long long e = tid; e < n4; e += stride) {
    // the first two chunks were loaded before the reduction
    const float4 a = e == tid ? a0 : e == tid + stride ? a1 : x4[e];
    const float4 c = e == tid ? c0 : e == tid + stride ? c1 : g4[e];
    o4[e] = make_float4(pbb_new(a.x, c.x, step, unconstrained), pbb_new(a.y, c.y, step, unconstrained),
                        pbb_new(a.z, c.z, step, unconstrained), pbb_new(a.w, c.w, step, unconstrained));
  }
  for (long long e = 4 * n4 + tid; e < n; e += stride) xnew[e] = pbb_new(x[e], g[e], step, unconstrained);
}

// PBB start (optGPU.py:256-297): the iterate is the start point with the
// lower objective.  dots = [r.r, z.z] at the first point, then at the second;
// objective = r.r/2 + beta z.z/2.  When the second point is not worse the
// pairs (x, oldx) and (g, oldg) exchange contents.
__global__ __launch_bounds__(TB) void k_pbb_select(float* x, float* oldx, float* g, float* oldg, long long n,
                                                   const double* dots, double beta) {
  const double fx = 0.5 * dots[0] + 0.5 * beta * dots[1];
  const double fy = 0.5 * dots[2] + 0.5 * beta * dots[3];
  if (fx < fy) return;
  for (long long e = (long long)blockIdx.x * TB + threadIdx.x; e < n; e += (long long)gridDim.x * TB) {
    const float a = x[e], b = g[e];
    x[e] = oldx[e];
    oldx[e] = a;
    g[e] = oldg[e];
    oldg[e] = b;
  }
}

// out = a*s + b*t (residuals, Tikhonov terms)
__global__ __launch_bounds__(TB) void k_axpby(float* out, const float* a, float s, const float* b, float t,
                                              long long n) {
  for (long long e = (long long)blockIdx.x * TB + threadIdx.x; e < n; e += (long long)gridDim.x * TB)
    out[e] = a[e] * s + (b ? b[e] * t : 0.f);
}

// elementwise family of vmbd_ew
__global__ __launch_bounds__(TB) void k_ew(float* out, const float* a, const float* b, float s, long long n, int op) {
  for (long long e = (long long)blockIdx.x * TB + threadIdx.x; e < n; e += (long long)gridDim.x * TB) {
    const float x = a[e];
    float r;
    switch (op) {
      case VMBD_EW_MUL: r = x * b[e]; break;
      case VMBD_EW_DIV: r = x / b[e]; break;
      case VMBD_EW_SDIV: r = s / x; break;
      case VMBD_EW_SUB: r = x - b[e]; break;
      case VMBD_EW_AXPY: r = x + s * b[e]; break;
      case VMBD_EW_LT: r = x < s ? 1.f : 0.f; break;
      case VMBD_EW_ADDS: r = x + s; break;
      default: r = s * x; break;
    }
    out[e] = r;
  }
}

// out = a * b (masks)
__global__ __launch_bounds__(TB) void k_mul(float* out, const float* a, const float* b, long long n) {
  for (long long e = (long long)blockIdx.x * TB + threadIdx.x; e < n; e += (long long)gridDim.x * TB)
    out[e] = a[e] * b[e];
}

// Multiplicative x-step of the online loop (vmbd_chimney.py:143-147), stage 1:
// x *= max(num/(den + tol), tol); block maxima of x over the interior
// rows [b0, H-b0) and columns [b1, W-b1).
__global__ __launch_bounds__(TB) void k_rl_step(float* x, const float* num, const float* den, int H, int W, float tol,
                                                int b0, int b1, float* partial_max) {
  float mx = -INFINITY;
  const long long n = (long long)H * W;
  for (long long e = (long long)blockIdx.x * TB + threadIdx.x; e < n; e += (long long)gridDim.x * TB) {
    float fac = num[e] / (den[e] + tol);
    fac = fac < tol ? tol : fac;
    const float v = x[e] * fac;
    x[e] = v;
    const int i = (int)(e / W), j = (int)(e - (long long)(e / W) * W);
    if (i >= b0 && i < H - b0 && j >= b1 && j < W - b1) mx = fmaxf(mx, v);
  }
  mx = wave_max(mx);
  __shared__ float sh[TB / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = sh[0];
    for (int i = 1; i < TB / 64; ++i) m = fmaxf(m, sh[i]);
    partial_max[blockIdx.x] = m;
  }
}

// stage 2: x = min(x, max over partials); the maximum is also stored
__global__ __launch_bounds__(TB) void k_rl_clip(float* x, long long n, const float* partial_max, int nb,
                                                float* xmax_out) {
  float m = -INFINITY;
  for (int i = threadIdx.x; i < nb; i += TB) m = fmaxf(m, partial_max[i]);
  m = wave_max(m);
  __shared__ float sh[TB / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  m = sh[0];
  for (int i = 1; i < TB / 64; ++i) m = fmaxf(m, sh[i]);
  if (blockIdx.x == 0 && threadIdx.x == 0 && xmax_out) *xmax_out = m;
  for (long long e = (long long)blockIdx.x * TB + threadIdx.x; e < n; e += (long long)gridDim.x * TB)
    x[e] = x[e] > m ? m : x[e];
}

// Isotropic, eps-smoothed total variation of K images (H x W):
//   TV = sum_ij n_ij,  n_ij = sqrt(dx_ij^2 + dy_ij^2 + eps^2),
//   dx_ij = x[i, j+1] - x[i, j] (0 in the last column),
//   dy_ij = x[i+1, j] - x[i, j] (0 in the last row).
__device__ __forceinline__ float tv_norm(const float* s, int H, int W, int i, int j, float eps2, float* dx,
                                         float* dy) {
  const float c = s[(long long)i * W + j];
  const float a = (j + 1 < W) ? s[(long long)i * W + j + 1] - c : 0.f;
  const float b = (i + 1 < H) ? s[(long long)(i + 1) * W + j] - c : 0.f;
  *dx = a;
  *dy = b;
  return sqrtf(a * a + b * b + eps2);
}

__global__ __launch_bounds__(TB) void k_tv_value(const float* x, long long n, int H, int W, float eps2,
                                                 double* partials) {
  double v[1] = {0.0};
  const long long HW = (long long)H * W;
  for (long long e = (long long)blockIdx.x * TB + threadIdx.x; e < n; e += (long long)gridDim.x * TB) {
    const long long k = e / HW;
    const int r = (int)(e - k * HW);
    const int i = r / W, j = r - (r / W) * W;
    float a, b;
    v[0] += (double)tv_norm(x + k * HW, H, W, i, j, eps2, &a, &b);
  }
  block_sum<1>(v);
  if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
}

// grad (+)= weight * scale[0] * dTV/dx:
//   dTV/dx_ij = -(dx_ij + dy_ij)/n_ij + dx_{i,j-1}/n_{i,j-1} + dy_{i-1,j}/n_{i-1,j}
__global__ __launch_bounds__(TB) void k_tv_grad(float* grad, const float* x, long long n, int H, int W, float eps2,
                                                const float* scale, float weight, int accumulate) {
  const float w = weight * (scale ? scale[0] : 1.f);
  const long long HW = (long long)H * W;
  for (long long e = (long long)blockIdx.x * TB + threadIdx.x; e < n; e += (long long)gridDim.x * TB) {
    const long long k = e / HW;
    const int r = (int)(e - k * HW);
    const int i = r / W, j = r - (r / W) * W;
    const float* s = x + k * HW;
    float a, b;
    const float nc = tv_norm(s, H, W, i, j, eps2, &a, &b);
    float gv = -(a + b) / nc;
    if (j > 0) {
      const float nl = tv_norm(s, H, W, i, j - 1, eps2, &a, &b);
      gv += a / nl;
    }
    if (i > 0) {
      const float nu = tv_norm(s, H, W, i - 1, j, eps2, &a, &b);
      gv += b / nu;
    }
    gv *= w;
    grad[e] = accumulate ? grad[e] + gv : gv;
  }
}

}  // namespace

// Scratch of vmbd_pbb_step_noalloc: the partial sums of one PBB step
// (3 per reduction block, at most VMBD_MAX_PARTIALS blocks), held by every
// GPU the library runs on until the process ends.  Host code never uses the
// array's name as an address: its device address comes from
// hipGetSymbolAddress, once per device (noalloc_scratch).
constexpr int NOALLOC_DOUBLES = 3 * VMBD_MAX_PARTIALS;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// device address of this device's scratch (looked up on its first use)
int noalloc_scratch(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return vmbd_fail_hip("vmbd_pbb_step_noalloc: hipGetDevice", e);
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
    return vmbd_fail(VMBD_E_UNSUP, "vmbd_pbb_step_noalloc: device %d beyond the %d supported", dev,
                     NOALLOC_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return vmbd_fail_hip("vmbd_pbb_step_noalloc: hipGetSymbolAddress", e);
    if (!p) return vmbd_fail(VMBD_E_HIP, "vmbd_pbb_step_noalloc: no device address for the scratch");
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return VMBD_OK;
}

}  // namespace
}  // namespace vmbd

using namespace vmbd;

namespace {
// float4 chunks of an n-element operation whose operands are all 16-byte
// aligned (0 otherwise: element loop only)
inline long long vec4_chunks(long long n, std::initializer_list<const void*> ps) {
  for (const void* p : ps)
    if (reinterpret_cast<uintptr_t>(p) & 15) return 0;
  return n / 4;
}

inline int grid_for(long long n) {
  long long b = (n + TB - 1) / TB;
  if (b > VMBD_MAX_PARTIALS) b = VMBD_MAX_PARTIALS;
  return b < 1 ? 1 : (int)b;
}

// blocks of the PBB statistics and update passes: two 4-float chunks per
// thread (so a small operand, such as a PSF stack, runs few blocks and its
// update reduces few partials), at most VMBD_MAX_PARTIALS; never more than
// grid_for(n), the partials callers allocate
inline int grid_pbb(long long n) {
  long long b = (n + 8 * TB - 1) / (8 * TB);
  if (b > VMBD_MAX_PARTIALS) b = VMBD_MAX_PARTIALS;
  return b < 1 ? 1 : (int)b;
}
#define SOLVER_DONE(what)                                            \
  do {                                                               \
    hipError_t e__ = hipGetLastError();                              \
    if (e__ != hipSuccess) return vmbd_fail_hip(what, e__);          \
    return VMBD_OK;                                                  \
  } while (0)
}  // namespace

extern "C" {

int vmbd_reduce_blocks(int64_t n) { return grid_for(n); }

int vmbd_clip(float* x, int64_t n, float lb, float ub, int32_t flags, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!x || n < 0) return vmbd_fail(VMBD_E_ARG, "vmbd_clip: bad argument");
  if (n == 0) return VMBD_OK;
  hipStream_t st = (hipStream_t)stream;
  ProfScope ps("clip", st);
  hipLaunchKernelGGL(k_clip, dim3(grid_for(n)), dim3(TB), 0, st, x, (long long)n, lb, ub, flags);
  SOLVER_DONE("vmbd_clip");
}

int vmbd_psf_normalize(float* f, int32_t K, int32_t len, float eps, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!f || K <= 0 || len <= 0) return vmbd_fail(VMBD_E_ARG, "vmbd_psf_normalize: bad argument");
  hipStream_t st = (hipStream_t)stream;
  ProfScope ps("psf_normalize", st);
  hipLaunchKernelGGL(k_normalize, dim3(K), dim3(TB), 0, st, f, len, eps);
  SOLVER_DONE("vmbd_psf_normalize");
}

int vmbd_dot(const float* a, const float* b, int64_t n, double* partials, double* out, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!a || !b || !partials || !out || n < 0) return vmbd_fail(VMBD_E_ARG, "vmbd_dot: bad argument");
  hipStream_t st = (hipStream_t)stream;
  const int nb = grid_for(n);
  ProfScope ps("dot", st);
  hipLaunchKernelGGL(k_dot_partials, dim3(nb), dim3(TB), 0, st, a, b, (long long)n, partials);
  hipLaunchKernelGGL(k_reduce_out, dim3(1), dim3(TB), 0, st, partials, nb, 1, out);
  SOLVER_DONE("vmbd_dot");
}

int vmbd_pbb_stats(const float* x, const float* oldx, const float* g, const float* oldg, int64_t n,
                   int32_t unconstrained, double* partials, double* sums, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!x || !oldx || !g || !oldg || !partials || n <= 0) return vmbd_fail(VMBD_E_ARG, "vmbd_pbb_stats: bad argument");
  hipStream_t st = (hipStream_t)stream;
  const int nb = grid_pbb(n);
  ProfScope ps("pbb_stats", st);
  const long long n4 = vec4_chunks(n, {x, oldx, g, oldg});
  hipLaunchKernelGGL(k_pbb_stats, dim3(nb), dim3(TB), 0, st, x, oldx, g, oldg, (long long)n, n4, unconstrained,
                     partials);
  if (sums) hipLaunchKernelGGL(k_reduce_out, dim3(1), dim3(TB), 0, st, partials, nb, 3, sums);
  SOLVER_DONE("vmbd_pbb_stats");
}

int vmbd_pbb_update(float* xnew, const float* x, const float* g, int64_t n, const double* partials, int32_t iter,
                    int32_t unconstrained, float* step_out, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!xnew || !x || !g || !partials || n <= 0) return vmbd_fail(VMBD_E_ARG, "vmbd_pbb_update: bad argument");
  hipStream_t st = (hipStream_t)stream;
  const int nb = grid_pbb(n);
  ProfScope ps("pbb_update", st);
  const long long n4 = vec4_chunks(n, {xnew, x, g});
  hipLaunchKernelGGL(k_pbb_update, dim3(nb), dim3(TB), 0, st, xnew, x, g, (long long)n, n4, partials, nb, iter,
                     unconstrained, step_out);
  SOLVER_DONE("vmbd_pbb_update");
}

int vmbd_pbb_step(float* xnew, const float* x, const float* oldx, const float* g, const float* oldg, int64_t n,
                  int32_t iter, int32_t unconstrained, double* partials, float* step_out, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!xnew || !x || !oldx || !g || !oldg || !partials || n <= 0)
    return vmbd_fail(VMBD_E_ARG, "vmbd_pbb_step: bad argument");
  hipStream_t st = (hipStream_t)stream;
  const int nb = grid_pbb(n);
  ProfScope ps("pbb_step", st);
  const long long n4s = vec4_chunks(n, {x, oldx, g, oldg});
  const long long n4u = vec4_chunks(n, {xnew, x, g});
  hipLaunchKernelGGL(k_pbb_stats, dim3(nb), dim3(TB), 0, st, x, oldx, g, oldg, (long long)n, n4s, unconstrained,
                     partials);
  hipLaunchKernelGGL(k_pbb_update, dim3(nb), dim3(TB), 0, st, xnew, x, g, (long long)n, n4u, partials, nb, iter,
                     unconstrained, step_out);
  SOLVER_DONE("vmbd_pbb_step");
}

int vmbd_pbb_select(float* x, float* oldx, float* g, float* oldg, int64_t n, const double* dots, double beta,
                    vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!x || !oldx || !g || !oldg || !dots || n <= 0) return vmbd_fail(VMBD_E_ARG, "vmbd_pbb_select: bad argument");
  hipStream_t st = (hipStream_t)stream;
  ProfScope ps("pbb_select", st);
  hipLaunchKernelGGL(k_pbb_select, dim3(grid_for(n)), dim3(TB), 0, st, x, oldx, g, oldg, (long long)n, dots, beta);
  SOLVER_DONE("vmbd_pbb_select");
}

int vmbd_axpby(float* out, const float* a, float s, const float* b, float t, int64_t n, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!out || !a || n < 0) return vmbd_fail(VMBD_E_ARG, "vmbd_axpby: bad argument");
  if (n == 0) return VMBD_OK;
  hipStream_t st = (hipStream_t)stream;
  ProfScope ps("axpby", st);
  hipLaunchKernelGGL(k_axpby, dim3(grid_for(n)), dim3(TB), 0, st, out, a, s, b, t, (long long)n);
  SOLVER_DONE("vmbd_axpby");
}

int vmbd_mul(float* out, const float* a, const float* b, int64_t n, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!out || !a || !b || n < 0) return vmbd_fail(VMBD_E_ARG, "vmbd_mul: bad argument");
  if (n == 0) return VMBD_OK;
  hipStream_t st = (hipStream_t)stream;
  ProfScope ps("mul", st);
  hipLaunchKernelGGL(k_mul, dim3(grid_for(n)), dim3(TB), 0, st, out, a, b, (long long)n);
  SOLVER_DONE("vmbd_mul");
}

int vmbd_ew(float* out, const float* a, const float* b, float s, int64_t n, int32_t op, vmbd_stream_t stream) {
  vmbd_clear_error();
  const bool needb = op == VMBD_EW_MUL || op == VMBD_EW_DIV || op == VMBD_EW_SUB || op == VMBD_EW_AXPY;
  if (!out || !a || (needb && !b) || n < 0 || op < 0 || op > VMBD_EW_ADDS)
    return vmbd_fail(VMBD_E_ARG, "vmbd_ew: bad argument");
  if (n == 0) return VMBD_OK;
  hipStream_t st = (hipStream_t)stream;
  ProfScope ps("ew", st);
  hipLaunchKernelGGL(k_ew, dim3(grid_for(n)), dim3(TB), 0, st, out, a, b, s, (long long)n, op);
  SOLVER_DONE("vmbd_ew");
}

int vmbd_tv(const float* x, int32_t K, int32_t H, int32_t W, float eps, double* partials, double* value_out,
            vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!x || !partials || !value_out || K <= 0 || H <= 0 || W <= 0) return vmbd_fail(VMBD_E_ARG, "vmbd_tv: bad argument");
  hipStream_t st = (hipStream_t)stream;
  const long long n = (long long)K * H * W;
  const int nb = grid_for(n);
  ProfScope ps("tv", st);
  hipLaunchKernelGGL(k_tv_value, dim3(nb), dim3(TB), 0, st, x, n, H, W, eps * eps, partials);
  hipLaunchKernelGGL(k_reduce_out, dim3(1), dim3(TB), 0, st, partials, nb, 1, value_out);
  SOLVER_DONE("vmbd_tv");
}

int vmbd_tv_grad(float* grad, const float* x, int32_t K, int32_t H, int32_t W, float eps, const float* scale,
                 float weight, int32_t accumulate, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!grad || !x || K <= 0 || H <= 0 || W <= 0) return vmbd_fail(VMBD_E_ARG, "vmbd_tv_grad: bad argument");
  hipStream_t st = (hipStream_t)stream;
  const long long n = (long long)K * H * W;
  ProfScope ps("tv_grad", st);
  hipLaunchKernelGGL(k_tv_grad, dim3(grid_for(n)), dim3(TB), 0, st, grad, x, n, H, W, eps * eps, scale, weight,
                     accumulate);
  SOLVER_DONE("vmbd_tv_grad");
}

int vmbd_rl_update(float* x, const float* num, const float* den, int32_t H, int32_t W, float tol, int32_t border0,
                   int32_t border1, float* partial_max, float* xmax_out, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!x || !num || !den || !partial_max || H <= 0 || W <= 0 || border0 < 0 || border1 < 0 || 2 * border0 >= H ||
      2 * border1 >= W)
    return vmbd_fail(VMBD_E_ARG, "vmbd_rl_update: bad argument");
  hipStream_t st = (hipStream_t)stream;
  const long long n = (long long)H * W;
  const int nb = grid_for(n);
  ProfScope ps("rl_update", st);
  hipLaunchKernelGGL(k_rl_step, dim3(nb), dim3(TB), 0, st, x, num, den, H, W, tol, border0, border1,
                     partial_max);
  hipLaunchKernelGGL(k_rl_clip, dim3(nb), dim3(TB), 0, st, x, n, partial_max, nb, xmax_out);
  SOLVER_DONE("vmbd_rl_update");
}

}  // extern "C"

extern "C" int vmbd_pbb_step_noalloc(float* xnew, const float* x, const float* oldx, const float* g,
                                     const float* oldg, int64_t n, int32_t iter, int32_t unconstrained,
                                     float* step_out, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!xnew || !x || !oldx || !g || !oldg || n <= 0)
    return vmbd_fail(VMBD_E_ARG, "vmbd_pbb_step_noalloc: bad argument");
  // the partials of any n fit: the reductions use at most VMBD_MAX_PARTIALS blocks
  if (3 * grid_for(n) > NOALLOC_DOUBLES)
    return vmbd_fail(VMBD_E_UNSUP, "vmbd_pbb_step_noalloc: %d partial sums exceed the %d-double scratch",
                     3 * grid_for(n), NOALLOC_DOUBLES);
  double* partials = nullptr;
  if (int r = noalloc_scratch(&partials)) return r;
  // every partial the update reads is written by the statistics kernel first
  return vmbd_pbb_step(xnew, x, oldx, g, oldg, n, iter, unconstrained, partials, step_out, stream);
}
