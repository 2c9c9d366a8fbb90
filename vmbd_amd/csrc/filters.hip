// Image stencils and pointwise filters of the deconvolution path for gfx950
// (kernel.cu, edgetaper.cu, resize.cu of the reference).
//
// Every stencil reads its input and writes a separate output (Jacobi
// semantics): the reference's shock and bilateral kernels update the image
// in place while neighbouring blocks still read their halos (kernel.cu:86,
// 134).  Border handling follows the reference kernels exactly:
//   shock      top/left rows and columns replicate, bottom/right read 0 (kernel.cu:22-30)
//   bilateral  out-of-image neighbours are skipped (kernel.cu:124)
//   laplace    'valid' or zero border ('same')
// 16x16 workgroups tile the image; the bilateral filter stages its tile and
// halo in LDS (121 taps per pixel).
#include <hip/hip_runtime.h>

#include "../../include/vmbd_filters.h"
#include "capi_util.h"
#include "prof.h"

namespace vmbd {
namespace {

constexpr int BX = 16, BY = 16;

inline dim3 grid2(int rows, int cols) { return dim3((cols + BX - 1) / BX, (rows + BY - 1) / BY); }

__device__ __forceinline__ float at_shock(const float* a, int H, int W, int i, int j) {
  // kernel.cu:20-30: negative indices clamp to 0, indices past the end read 0
  if (i < 0) i = 0;
  if (j < 0) j = 0;
  if (i >= H || j >= W) return 0.f;
  return a[(long long)i * W + j];
}

// one Osher-Rudin shock step (kernel.cu:36-86)
__global__ __launch_bounds__(BX* BY) void k_shock(float* out, const float* a, int H, int W, float dt, float h) {
  const int i = blockIdx.y * BY + threadIdx.y, j = blockIdx.x * BX + threadIdx.x;
  if (i >= H || j >= W) return;
  const float c = at_shock(a, H, W, i, j);
  const float l = at_shock(a, H, W, i, j - 1), r = at_shock(a, H, W, i, j + 1);
  const float u = at_shock(a, H, W, i - 1, j), d = at_shock(a, H, W, i + 1, j);
  float t1 = -c + r, t2 = c - l;
  const float Ix = (t1 + t2) / 2.f;
  float mx = (t1 * t2 < 0.f) ? 0.f : fminf(fabsf(t1), fabsf(t2));
  t1 = -c + d;
  t2 = c - u;
  const float Iy = (t1 + t2) / 2.f;
  float my = (t1 * t2 < 0.f) ? 0.f : fminf(fabsf(t1), fabsf(t2));
  const float flow = sqrtf(mx * mx + my * my);
  const float Ixx = r + l - 2.f * c;
  const float Iyy = d + u - 2.f * c;
  const float Ixy = (at_shock(a, H, W, i + 1, j + 1) - at_shock(a, H, W, i + 1, j - 1) -
                     at_shock(a, H, W, i - 1, j + 1) + at_shock(a, H, W, i - 1, j - 1)) /
                    4.f;
  float Inn = (Iyy * Iy * Iy + 2.f * Ixy * Iy * Ix + Ixx * Ix * Ix) / (Ix * Ix + Iy * Iy + 0.00000001f);
  if (Ix == 0.f && Iy == 0.f) Inn = Ixx;
  out[(long long)i * W + j] = c - dt * copysignf(flow, Inn) / h;
}

// bilateral filter (kernel.cu:90-136), tile + halo in LDS (dynamic size)
__global__ __launch_bounds__(BX* BY) void k_bfilter(float* out, const float* a, int H, int W, int w, float sd,
                                                    float sr) {
  extern __shared__ float tile[];
  const int TW = BX + 2 * w, TH = BY + 2 * w;
  const int i0 = blockIdx.y * BY - w, j0 = blockIdx.x * BX - w;
  for (int idx = threadIdx.y * BX + threadIdx.x; idx < TW * TH; idx += BX * BY) {
    const int ti = idx / TW, tj = idx - (idx / TW) * TW;
    const int ii = i0 + ti, jj = j0 + tj;
    tile[idx] = (ii >= 0 && ii < H && jj >= 0 && jj < W) ? a[(long long)ii * W + jj] : 0.f;
  }
  __syncthreads();
  const int i = blockIdx.y * BY + threadIdx.y, j = blockIdx.x * BX + threadIdx.x;
  if (i >= H || j >= W) return;
  const float c = tile[(threadIdx.y + w) * TW + threadIdx.x + w];
  const float inv_r = 1.f / (2.f * sr * sr), inv_d = 1.f / (2.f * sd * sd);
  float sw = 0.f, acc = 0.f;
  for (int x = -w; x <= w; ++x) {
    for (int y = -w; y <= w; ++y) {
      if (i + y >= 0 && i + y < H && j + x >= 0 && j + x < W) {
        const float v = tile[(threadIdx.y + w + y) * TW + threadIdx.x + w + x];
        const float dv = v - c;
        const float wt = expf(-(dv * dv * inv_r + (float)(x * x + y * y) * inv_d));
        sw += wt;
        acc += wt * v;
      }
    }
  }
  out[(long long)i * W + j] = acc / sw;
}

// The same filter for a half size WR known at compile time (the MFBD prior's
// w = 5): loops unrolled, the spatial factor folded into one exponent,
// exp(-t) evaluated as the hardware 2^(-t log2 e), and the in-image test
// only in workgroups whose window crosses the image border.  Same sums in
// the same order as k_bfilter.
template <int WR>
__global__ __launch_bounds__(BX* BY) void k_bfilter_w(float* out, const float* a, int H, int W, float sd, float sr) {
  extern __shared__ float tile[];
  constexpr int TW = BX + 2 * WR, TH = BY + 2 * WR;
  const int i0 = blockIdx.y * BY - WR, j0 = blockIdx.x * BX - WR;
  for (int idx = threadIdx.y * BX + threadIdx.x; idx < TW * TH; idx += BX * BY) {
    const int ti = idx / TW, tj = idx - (idx / TW) * TW;
    const int ii = i0 + ti, jj = j0 + tj;
    tile[idx] = (ii >= 0 && ii < H && jj >= 0 && jj < W) ? a[(long long)ii * W + jj] : 0.f;
  }
  __syncthreads();
  const int i = blockIdx.y * BY + threadIdx.y, j = blockIdx.x * BX + threadIdx.x;
  if (i >= H || j >= W) return;
  const bool inner = i0 >= 0 && j0 >= 0 && i0 + TH <= H && j0 + TW <= W;  // uniform per workgroup
  const float* tc = tile + (threadIdx.y + WR) * TW + threadIdx.x + WR;
  const float c = tc[0];
  constexpr float LOG2E = 1.44269504088896340736f;
  const float kr = -LOG2E / (2.f * sr * sr), kd = -LOG2E / (2.f * sd * sd);
  float sw = 0.f, acc = 0.f;
#pragma unroll
  for (int x = -WR; x <= WR; ++x) {
#pragma unroll
    for (int y = -WR; y <= WR; ++y) {
      const float v = tc[y * TW + x];
      const float dv = v - c;
      float wt = __builtin_amdgcn_exp2f(fmaf(dv * dv, kr, (float)(x * x + y * y) * kd));
      if (!inner && !(i + y >= 0 && i + y < H && j + x >= 0 && j + x < W)) wt = 0.f;
      sw += wt;
      acc += wt * v;
    }
  }
  out[(long long)i * W + j] = acc / sw;
}

#define SQRT2H 0.70710678118654752440f

// kernel.cu:180-242
__global__ __launch_bounds__(BX* BY) void k_gradient(float* gx, float* gy, const float* y, int H, int W, int same) {
  const int i = blockIdx.y * BY + threadIdx.y, j = blockIdx.x * BX + threadIdx.x;
  if (i >= H || j >= W) return;
  const float c = y[(long long)i * W + j];
  if (same) {
    gy[(long long)i * W + j] = (i < H - 1) ? SQRT2H * (c - y[(long long)(i + 1) * W + j]) : 0.f;
    gx[(long long)i * W + j] = (j < W - 1) ? SQRT2H * (c - y[(long long)i * W + j + 1]) : 0.f;
  } else {
    if (i < H - 1) gy[(long long)i * W + j] = SQRT2H * (c - y[(long long)(i + 1) * W + j]);
    if (j < W - 1) gx[(long long)i * (W - 1) + j] = SQRT2H * (c - y[(long long)i * W + j + 1]);
  }
}

// kernel.cu:244-266: replicate padding, (top, left) = floor(sf/2)
__global__ __launch_bounds__(BX* BY) void k_impad(float* out, int OH, int OW, const float* im, int H, int W, int top,
                                                  int left) {
  const int i = blockIdx.y * BY + threadIdx.y, j = blockIdx.x * BX + threadIdx.x;
  if (i >= OH || j >= OW) return;
  int si = i - top, sj = j - left;
  si = si < 0 ? 0 : (si >= H ? H - 1 : si);
  sj = sj < 0 ? 0 : (sj >= W ? W - 1 : sj);
  out[(long long)i * OW + j] = im[(long long)si * W + sj];
}

#define LAP_N 0.22360679774997896964f  // 1/sqrt(20)
#define LAP_C 0.89442719099991587856f  // 4/sqrt(20)

// kernel.cu:268-333 (per slice of a stack of K images); zero_k: the first and
// last slice are zero (laplace_stack_same, kernel.cu:364)
__global__ __launch_bounds__(BX* BY) void k_laplace(float* out, const float* im, int H, int W, int same, int K,
                                                    int zero_k) {
  const int i = blockIdx.y * BY + threadIdx.y, j = blockIdx.x * BX + threadIdx.x;
  const int k = blockIdx.z;
  const float* s = im + (long long)k * H * W;
  if (same) {
    if (i >= H || j >= W) return;
    float v = 0.f;
    const bool kin = !zero_k || (k > 0 && k < K - 1);
    if (kin && i > 0 && i < H - 1 && j > 0 && j < W - 1) {
      const long long c = (long long)i * W + j;
      v = -LAP_N * (s[c + W] + s[c - W] + s[c + 1] + s[c - 1]) + LAP_C * s[c];
    }
    out[(long long)k * H * W + (long long)i * W + j] = v;
  } else {
    if (i >= H - 2 || j >= W - 2) return;
    const long long c = (long long)(i + 1) * W + j + 1;
    out[(long long)k * (H - 2) * (W - 2) + (long long)i * (W - 2) + j] =
        -LAP_N * (s[c + W] + s[c - W] + s[c + 1] + s[c - 1]) + LAP_C * s[c];
  }
}

// kernel.cu:377-418: 7-point 3-D Laplacian, zero on every face
__global__ __launch_bounds__(BX* BY) void k_laplace3d(float* out, const float* im, int D, int H, int W) {
  const int i = blockIdx.y * BY + threadIdx.y, j = blockIdx.x * BX + threadIdx.x;
  const int k = blockIdx.z;
  if (i >= H || j >= W) return;
  const long long HW = (long long)H * W;
  const long long c = k * HW + (long long)i * W + j;
  float v = 0.f;
  if (i > 0 && i < H - 1 && j > 0 && j < W - 1 && k > 0 && k < D - 1)
    v = -(im[c + W] + im[c - W] + im[c + 1] + im[c - 1] + im[c + HW] + im[c - HW]) + 6.f * im[c];
  out[c] = v;
}

// hyper-Laplacian w-step, alpha = 2/3 (kernel.cu:555-578)
__global__ __launch_bounds__(256) void k_alpha23(float* y, long long n, float beta) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= n) return;
  const float v = y[e];
  float x = v;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float t1 = (2.f / 3.f) / cbrtf(fabsf(x));
    x -= (copysignf(t1, x) + beta * (x - v)) / (t1 / (-3.f * fabsf(x)) + beta);
  }
  if (!isfinite(x)) x = 0.f;
  float t = cbrtf(fabsf(x));
  t *= t;
  if (t + beta * 0.5f * (x - v) * (x - v) > beta * 0.5f * v * v) x = 0.f;
  y[e] = x;
}

// general-alpha w-step (kernel.cu:529-553)
__global__ __launch_bounds__(256) void k_alpha(float* y, long long n, float beta, float alpha) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= n) return;
  const float v = y[e];
  float x = v;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float fd = alpha * (x > 0.f ? 1.f : -1.f) * powf(fabsf(x), alpha - 1.f) + beta * (x - v);
    const float fdd = alpha * (alpha - 1.f) * powf(fabsf(x), alpha - 2.f) + beta;
    x -= fd / fdd;
  }
  if (!isfinite(x)) x = 0.f;
  if (powf(fabsf(x), alpha) + beta * 0.5f * (x - v) * (x - v) > beta * 0.5f * v * v) x = 0.f;
  y[e] = x;
}

// edgetaper.cu:6-31 (tables already divided by max(wx)*max(wy) on the host)
__global__ __launch_bounds__(BX* BY) void k_edgetaper(float* im, int H, int W, const float* wx, int hx,
                                                      const float* wy, int hy) {
  const int i = blockIdx.y * BY + threadIdx.y, j = blockIdx.x * BX + threadIdx.x;
  if (i >= H || j >= W) return;
  if (!(i < hy || i > H - 1 - hy || j < hx || j > W - 1 - hx)) return;
  float v = im[(long long)i * W + j];
  if (i < hy) v *= wy[i];
  else if (i > H - 1 - hy) v *= wy[H - 1 - i];
  if (j < hx) v *= wx[j];
  else if (j > W - 1 - hx) v *= wx[W - 1 - j];
  im[(long long)i * W + j] = v;
}

// resize.cu:5-15: sample at normalised (j/OW, i/OH) with bilinear filtering
// and clamp addressing (texel centres at +0.5), per slice
__global__ __launch_bounds__(BX* BY) void k_resize(float* out, int OH, int OW, const float* in, int H, int W) {
  const int i = blockIdx.y * BY + threadIdx.y, j = blockIdx.x * BX + threadIdx.x;
  const int k = blockIdx.z;
  if (i >= OH || j >= OW) return;
  const float* s = in + (long long)k * H * W;
  const float xs = (float)j / (float)OW * (float)W - 0.5f;
  const float ys = (float)i / (float)OH * (float)H - 0.5f;
  const float fx = floorf(xs), fy = floorf(ys);
  const float ax = xs - fx, ay = ys - fy;
  auto cl = [](int v, int n) { return v < 0 ? 0 : (v >= n ? n - 1 : v); };
  const int x0 = cl((int)fx, W), x1 = cl((int)fx + 1, W);
  const int y0 = cl((int)fy, H), y1 = cl((int)fy + 1, H);
  const float t0 = s[(long long)y0 * W + x0] * (1.f - ax) + s[(long long)y0 * W + x1] * ax;
  const float t1 = s[(long long)y1 * W + x0] * (1.f - ax) + s[(long long)y1 * W + x1] * ax;
  out[(long long)k * OH * OW + (long long)i * OW + j] = t0 * (1.f - ay) + t1 * ay;
}

}  // namespace
}  // namespace vmbd

using namespace vmbd;

#define FILT_DONE(what)                                              \
  do {                                                               \
    hipError_t e__ = hipGetLastError();                              \
    if (e__ != hipSuccess) return vmbd_fail_hip(what, e__);          \
    return VMBD_OK;                                                  \
  } while (0)

extern "C" {

int vmbd_shock_step(float* out, const float* in, int32_t H, int32_t W, float dt, float h, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!out || !in || out == in || H <= 0 || W <= 0) return vmbd_fail(VMBD_E_ARG, "vmbd_shock_step: bad argument");
  hipStream_t st = (hipStream_t)stream;
  ProfScope ps("shock", st);
  hipLaunchKernelGGL(k_shock, grid2(H, W), dim3(BX, BY), 0, st, out, in, H, W, dt, h);
  FILT_DONE("vmbd_shock_step");
}

int vmbd_bilateral(float* out, const float* in, int32_t H, int32_t W, int32_t w, float sigma_d, float sigma_r,
                   vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!out || !in || out == in || H <= 0 || W <= 0 || w < 0 || w > 24)
    return vmbd_fail(VMBD_E_ARG, "vmbd_bilateral: bad argument (half size in [0,24], out != in)");
  hipStream_t st = (hipStream_t)stream;
  const size_t shm = (size_t)(BX + 2 * w) * (BY + 2 * w) * sizeof(float);
  ProfScope ps("bilateral", st);
  if (w == 5)
    hipLaunchKernelGGL(k_bfilter_w<5>, grid2(H, W), dim3(BX, BY), shm, st, out, in, H, W, sigma_d, sigma_r);
  else
    hipLaunchKernelGGL(k_bfilter, grid2(H, W), dim3(BX, BY), shm, st, out, in, H, W, w, sigma_d, sigma_r);
  FILT_DONE("vmbd_bilateral");
}

int vmbd_gradient(float* gx, float* gy, const float* y, int32_t H, int32_t W, int32_t same, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!gx || !gy || !y || H <= 1 || W <= 1) return vmbd_fail(VMBD_E_ARG, "vmbd_gradient: bad argument");
  hipStream_t st = (hipStream_t)stream;
  ProfScope ps("gradient", st);
  hipLaunchKernelGGL(k_gradient, grid2(H, W), dim3(BX, BY), 0, st, gx, gy, y, H, W, same);
  FILT_DONE("vmbd_gradient");
}

int vmbd_impad(float* out, const float* in, int32_t H, int32_t W, int32_t padh, int32_t padw, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!out || !in || H <= 0 || W <= 0 || padh < 0 || padw < 0) return vmbd_fail(VMBD_E_ARG, "vmbd_impad: bad argument");
  hipStream_t st = (hipStream_t)stream;
  const int OH = H + padh, OW = W + padw;
  ProfScope ps("impad", st);
  hipLaunchKernelGGL(k_impad, grid2(OH, OW), dim3(BX, BY), 0, st, out, OH, OW, in, H, W, padh / 2, padw / 2);
  FILT_DONE("vmbd_impad");
}

int vmbd_laplace(float* out, const float* in, int32_t K, int32_t H, int32_t W, int32_t same, int32_t zero_end_slices,
                 vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!out || !in || K <= 0 || H <= 2 || W <= 2) return vmbd_fail(VMBD_E_ARG, "vmbd_laplace: bad argument");
  hipStream_t st = (hipStream_t)stream;
  dim3 g = grid2(H, W);
  g.z = K;
  ProfScope ps("laplace", st);
  hipLaunchKernelGGL(k_laplace, g, dim3(BX, BY), 0, st, out, in, H, W, same, K, zero_end_slices);
  FILT_DONE("vmbd_laplace");
}

int vmbd_laplace3d(float* out, const float* in, int32_t D, int32_t H, int32_t W, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!out || !in || D <= 0 || H <= 0 || W <= 0) return vmbd_fail(VMBD_E_ARG, "vmbd_laplace3d: bad argument");
  hipStream_t st = (hipStream_t)stream;
  dim3 g = grid2(H, W);
  g.z = D;
  ProfScope ps("laplace3d", st);
  hipLaunchKernelGGL(k_laplace3d, g, dim3(BX, BY), 0, st, out, in, D, H, W);
  FILT_DONE("vmbd_laplace3d");
}

int vmbd_shrink_alpha23(float* y, int64_t n, float beta, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!y || n < 0) return vmbd_fail(VMBD_E_ARG, "vmbd_shrink_alpha23: bad argument");
  if (n == 0) return VMBD_OK;
  hipStream_t st = (hipStream_t)stream;
  ProfScope ps("shrink23", st);
  hipLaunchKernelGGL(k_alpha23, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, y, (long long)n, beta);
  FILT_DONE("vmbd_shrink_alpha23");
}

int vmbd_shrink_alpha(float* y, int64_t n, float beta, float alpha, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!y || n < 0) return vmbd_fail(VMBD_E_ARG, "vmbd_shrink_alpha: bad argument");
  if (n == 0) return VMBD_OK;
  hipStream_t st = (hipStream_t)stream;
  ProfScope ps("shrink", st);
  hipLaunchKernelGGL(k_alpha, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, y, (long long)n, beta, alpha);
  FILT_DONE("vmbd_shrink_alpha");
}

int vmbd_edgetaper(float* im, int32_t H, int32_t W, const float* wx, int32_t hx, const float* wy, int32_t hy,
                   vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!im || H <= 0 || W <= 0 || hx < 0 || hy < 0 || (hx && !wx) || (hy && !wy))
    return vmbd_fail(VMBD_E_ARG, "vmbd_edgetaper: bad argument");
  hipStream_t st = (hipStream_t)stream;
  ProfScope ps("edgetaper", st);
  hipLaunchKernelGGL(k_edgetaper, grid2(H, W), dim3(BX, BY), 0, st, im, H, W, wx, hx, wy, hy);
  FILT_DONE("vmbd_edgetaper");
}

int vmbd_resize(float* out, int32_t OH, int32_t OW, const float* in, int32_t K, int32_t H, int32_t W,
                vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!out || !in || OH <= 0 || OW <= 0 || K <= 0 || H <= 0 || W <= 0)
    return vmbd_fail(VMBD_E_ARG, "vmbd_resize: bad argument");
  hipStream_t st = (hipStream_t)stream;
  dim3 g = grid2(OH, OW);
  g.z = K;
  ProfScope ps("resize", st);
  hipLaunchKernelGGL(k_resize, g, dim3(BX, BY), 0, st, out, OH, OW, in, H, W);
  FILT_DONE("vmbd_resize");
}

}  // extern "C"
