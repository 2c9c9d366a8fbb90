// Reference-decomposition kernels (the "plain" per-operator form) for gfx950.
//
// One kernel per operation of gputools.cu / the pyfft plan, each with the same
// index semantics but a flat, coalesced thread mapping (consecutive threads
// own consecutive output elements of the innermost dimension):
//   pad            gputools.cu:154-182  zeroPad(Complex)Kernel
//   chop_pad       gputools.cu:16-126   chop_(mod_)pad_(Complex)Kernel
//   pad_stack      gputools.cu:184-222  pad_stack_(Complex)Kernel
//   crop_stack     gputools.cu:225-251  crop_stack_(Complex)Kernel
//   crop           gputools.cu:254-274  crop_(Complex)Kernel
//   ola            gputools.cu:276-314  ola_(Complex)Kernel_test
//   fft2           pyfft Plan((Th,Tw)).execute, batched, in place
//   cmul           the complex elementwise products of olaGPU.py:217-231, 294-297
//   fourier_div    gputools.cu:316-406  comp_ola_{,g,s}deconv_Kernel
#include <hip/hip_runtime.h>

#include "../../include/vmbd_plain.h"
#include "capi_util.h"
#include "fft_core.h"
#include "prof.h"

namespace vmbd {
namespace {

constexpr int TB = 256;

inline unsigned blocks_for(long long n) { return (unsigned)((n + TB - 1) / TB); }

// complex (float2) or real (float) element store
__device__ __forceinline__ void put(void* out, long long idx, float v, int cplx) {
  if (cplx) ((float2*)out)[idx] = make_float2(v, 0.f);
  else ((float*)out)[idx] = v;
}
__device__ __forceinline__ float get(const void* in, long long idx, int cplx) {
  return cplx ? ((const float2*)in)[idx].x : ((const float*)in)[idx];
}

__global__ __launch_bounds__(TB) void k_pad(void* out, int m, int n, const float* in, int p, int q, int s, int t,
                                            int cplx) {
  const long long e = (long long)blockIdx.x * TB + threadIdx.x;
  if (e >= (long long)m * n) return;
  const int i = (int)(e / n), j = (int)(e - (long long)i * n);
  const int ii = i - s, jj = j - t;
  const float v = (ii >= 0 && ii < p && jj >= 0 && jj < q) ? in[(long long)ii * q + jj] : 0.f;
  put(out, e, v, cplx);
}

__global__ __launch_bounds__(TB) void k_chop_pad(void* out, int P, int sz0, int sz1, const float* x, int H, int W,
                                                 int sxp0, int sxp1, int sw0, int sw1, int hop0, int hop1, int csf1,
                                                 int off0, int off1, const float* win, const int* win_idx, int cplx) {
  const long long e = (long long)blockIdx.x * TB + threadIdx.x;
  const long long per = (long long)sz0 * sz1;
  if (e >= per * P) return;
  const int k = (int)(e / per);
  const int rem = (int)(e - (long long)k * per);
  const int a = rem / sz1, b = rem - (rem / sz1) * sz1;
  const int la = a - off0, lb = b - off1;
  float v = 0.f;
  if (la >= 0 && la < sw0 && lb >= 0 && lb < sw1) {
    const int m = k / csf1, n = k - (k / csf1) * csf1;
    const int i = m * hop0 + la, j = n * hop1 + lb;
    // x is zero-padded to sxp (and cut there) before the chop
    if (i < H && j < W && i < sxp0 && j < sxp1) v = x[(long long)i * W + j];
    if (win) v *= win[((long long)(win_idx ? win_idx[k] : k) * sw0 + la) * sw1 + lb];
  }
  put(out, e, v, cplx);
}

__global__ __launch_bounds__(TB) void k_pad_stack(void* out, int K, int sz0, int sz1, const float* in, int xm,
                                                  int xn, int off0, int off1, int cplx) {
  const long long e = (long long)blockIdx.x * TB + threadIdx.x;
  const long long per = (long long)sz0 * sz1;
  if (e >= per * K) return;
  const int k = (int)(e / per);
  const int rem = (int)(e - (long long)k * per);
  const int a = rem / sz1 - off0, b = rem - (rem / sz1) * sz1 - off1;
  const float v = (a >= 0 && a < xm && b >= 0 && b < xn) ? in[((long long)k * xm + a) * xn + b] : 0.f;
  put(out, e, v, cplx);
}

__global__ __launch_bounds__(TB) void k_crop_stack(float* out, int K, int sz0, int sz1, const void* in, int inm,
                                                   int inn, int off0, int off1, int cplx) {
  const long long e = (long long)blockIdx.x * TB + threadIdx.x;
  const long long per = (long long)sz0 * sz1;
  if (e >= per * K) return;
  const int k = (int)(e / per);
  const int rem = (int)(e - (long long)k * per);
  const int a = rem / sz1, b = rem - (rem / sz1) * sz1;
  out[e] = get(in, ((long long)k * inm + a + off0) * inn + b + off1, cplx);
}

__global__ __launch_bounds__(TB) void k_crop(float* out, int m, int n, const void* in, int p, int q, int s, int t,
                                             int cplx) {
  const long long e = (long long)blockIdx.x * TB + threadIdx.x;
  if (e >= (long long)m * n) return;
  const int i = (int)(e / n), j = (int)(e - (long long)(e / n) * n);
  out[e] = get(in, (long long)(i + s) * q + j + t, cplx);
}

// gather OLA: out (outm x outn) = sum over patches whose extent covers (i,j)
__global__ __launch_bounds__(TB) void k_ola(float* out, int outm, int outn, const void* in, int inm, int inn,
                                            int wm, int wn, int om, int on, int csf0, int csf1, int hop0, int hop1,
                                            int cplx) {
  const long long e = (long long)blockIdx.x * TB + threadIdx.x;
  if (e >= (long long)outm * outn) return;
  const int i = (int)(e / outn), j = (int)(e - (long long)(e / outn) * outn);
  const int ci = min(i / hop0, csf0 - 1);
  const int cj = min(j / hop1, csf1 - 1);
  float acc = 0.f;
  for (int mi = ci; mi >= 0 && i - mi * hop0 < wm; --mi)
    for (int nj = cj; nj >= 0 && j - nj * hop1 < wn; --nj) {
      const long long k = (long long)mi * csf1 + nj;
      acc += get(in, (k * inm + om + i - mi * hop0) * inn + on + j - nj * hop1, cplx);
    }
  out[e] = acc;
}

// batched complex FFT along the contiguous dimension (rows of length N)
template <int N, int S>
__global__ __launch_bounds__(TB) void k_fft_rows(float2* data, long long rows, float scale, const float2* tw) {
  constexpr int TPT = N / 16;
  constexpr int TPW = TB / TPT;
  constexpr int SL = FftLds<N>::floats2;
  __shared__ float2 sm[TPW * SL];
  const int tr = threadIdx.x / TPT, t = threadIdx.x - (threadIdx.x / TPT) * TPT;
  const long long row = (long long)blockIdx.x * TPW + tr;
  const bool act = row < rows;
  float2* rp = data + (act ? row : 0) * N;
  Twiddles<N> T;
  twiddles_init<N>(T, t, tw);
  float2 v[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r] = act ? rp[t + TPT * r] : make_float2(0.f, 0.f);
  fft<N, S>(v, sm + tr * SL, t, T, tw);
  if (act) {
#pragma unroll
    for (int m = 0; m < 16; ++m) rp[t + TPT * m] = cscale(v[m], scale);
  }
}

// batched complex FFT along the strided dimension: a workgroup transforms CS
// adjacent columns of one (Th x Tw) slice, staged through LDS so that every
// global access is a contiguous run of CS elements
template <int N, int S>
__global__ __launch_bounds__(TB) void k_fft_cols(float2* data, int Tw, int strips, float scale, const float2* tw) {
  constexpr int TPT = N / 16;
  constexpr int CS = TB / TPT;
  constexpr int SL = FftLds<N>::floats2;
  constexpr int TILE = N * (CS + 1);
  __shared__ float2 sm[TILE > CS * SL ? TILE : CS * SL];
  const long long slice = blockIdx.x / strips;
  const int strip = blockIdx.x - (int)(slice * strips);
  const int c0 = strip * CS;
  float2* base = data + slice * (long long)N * Tw + c0;
  const int ci = threadIdx.x & (CS - 1);
  const bool colok = c0 + ci < Tw;
  for (int r = threadIdx.x / CS; r < N; r += TB / CS)
    sm[r * (CS + 1) + ci] = colok ? base[(long long)r * Tw + ci] : make_float2(0.f, 0.f);
  __syncthreads();
  const int cc = threadIdx.x / TPT, t = threadIdx.x - (threadIdx.x / TPT) * TPT;
  float2 v[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r] = sm[(t + TPT * r) * (CS + 1) + cc];
  __syncthreads();
  Twiddles<N> T;
  twiddles_init<N>(T, t, tw);
  fft<N, S>(v, sm + cc * SL, t, T, tw);
  __syncthreads();
#pragma unroll
  for (int m = 0; m < 16; ++m) sm[(t + TPT * m) * (CS + 1) + cc] = cscale(v[m], scale);
  __syncthreads();
  if (colok)
    for (int r = threadIdx.x / CS; r < N; r += TB / CS) base[(long long)r * Tw + ci] = sm[r * (CS + 1) + ci];
}

template <int N>
hipError_t fft_rows_n(float2* d, long long rows, int dir, float scale, const float2* tw, hipStream_t st) {
  constexpr int TPW = TB / (N / 16);
  const unsigned grid = (unsigned)((rows + TPW - 1) / TPW);
  if (dir < 0) hipLaunchKernelGGL((k_fft_rows<N, -1>), dim3(grid), dim3(TB), 0, st, d, rows, scale, tw);
  else hipLaunchKernelGGL((k_fft_rows<N, 1>), dim3(grid), dim3(TB), 0, st, d, rows, scale, tw);
  return hipGetLastError();
}

template <int N>
hipError_t fft_cols_n(float2* d, long long slices, int Tw, int dir, float scale, const float2* tw, hipStream_t st) {
  constexpr int CS = TB / (N / 16);
  const int strips = (Tw + CS - 1) / CS;
  const unsigned grid = (unsigned)(slices * strips);
  if (dir < 0) hipLaunchKernelGGL((k_fft_cols<N, -1>), dim3(grid), dim3(TB), 0, st, d, Tw, strips, scale, tw);
  else hipLaunchKernelGGL((k_fft_cols<N, 1>), dim3(grid), dim3(TB), 0, st, d, Tw, strips, scale, tw);
  return hipGetLastError();
}

hipError_t fft_rows(float2* d, long long rows, int N, int dir, float scale, const float2* tw, hipStream_t st) {
  switch (N) {
    case 16: return fft_rows_n<16>(d, rows, dir, scale, tw, st);
    case 32: return fft_rows_n<32>(d, rows, dir, scale, tw, st);
    case 64: return fft_rows_n<64>(d, rows, dir, scale, tw, st);
    case 128: return fft_rows_n<128>(d, rows, dir, scale, tw, st);
    case 256: return fft_rows_n<256>(d, rows, dir, scale, tw, st);
    case 512: return fft_rows_n<512>(d, rows, dir, scale, tw, st);
    case 1024: return fft_rows_n<1024>(d, rows, dir, scale, tw, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t fft_cols(float2* d, long long slices, int Th, int Tw, int dir, float scale, const float2* tw,
                    hipStream_t st) {
  switch (Th) {
    case 16: return fft_cols_n<16>(d, slices, Tw, dir, scale, tw, st);
    case 32: return fft_cols_n<32>(d, slices, Tw, dir, scale, tw, st);
    case 64: return fft_cols_n<64>(d, slices, Tw, dir, scale, tw, st);
    case 128: return fft_cols_n<128>(d, slices, Tw, dir, scale, tw, st);
    case 256: return fft_cols_n<256>(d, slices, Tw, dir, scale, tw, st);
    case 512: return fft_cols_n<512>(d, slices, Tw, dir, scale, tw, st);
    case 1024: return fft_cols_n<1024>(d, slices, Tw, dir, scale, tw, st);
    default: return hipErrorInvalidValue;
  }
}

// out = op(a, b) * scale; b broadcast with period bper (== n for none)
__global__ __launch_bounds__(TB) void k_cmul(float2* out, const float2* a, const float2* b, long long n,
                                             long long bper, int mode, float scale) {
  const long long e = (long long)blockIdx.x * TB + threadIdx.x;
  if (e >= n) return;
  const float2 x = a[e];
  const float2 y = b[e % bper];
  float2 r;
  if (mode == VMBD_CMUL_AB) r = cmul(x, y);
  else if (mode == VMBD_CMUL_A_CONJB) r = cmulc(x, y);
  else r = cmulc(y, x);  // conj(a) * b
  out[e] = cscale(r, scale);
}

__device__ __forceinline__ float2 cdiv(float2 num, float2 den) {
  const float d = den.x * den.x + den.y * den.y;
  return make_float2((num.x * den.x + num.y * den.y) / d, (num.y * den.x - num.x * den.y) / d);
}

// comp_ola_deconv: conj(f) y / (|f|^2 + alpha L + beta)
__global__ __launch_bounds__(TB) void k_fdiv(float2* out, const float2* f, const float2* y, const float2* L,
                                             long long n, long long per, float alpha, float beta) {
  const long long e = (long long)blockIdx.x * TB + threadIdx.x;
  if (e >= n) return;
  const float2 fv = f[e], yv = y[e], lv = L[e % per];
  const float2 num = cmulc(yv, fv);
  const float2 den = make_float2(fv.x * fv.x + fv.y * fv.y + alpha * lv.x + beta, alpha * lv.y);
  out[e] = cdiv(num, den);
}

// comp_ola_gdeconv: (conj(xx) yx + conj(xy) yy) / (|xx|^2 + |xy|^2 + alpha L + beta)
__global__ __launch_bounds__(TB) void k_fdiv_g(float2* out, const float2* xx, const float2* xy, const float2* yx,
                                               const float2* yy, const float2* L, long long n, long long per,
                                               float alpha, float beta) {
  const long long e = (long long)blockIdx.x * TB + threadIdx.x;
  if (e >= n) return;
  const float2 a = xx[e], b = xy[e], lv = L[e % per];
  const float2 num = cadd(cmulc(yx[e], a), cmulc(yy[e], b));
  const float2 den =
      make_float2(a.x * a.x + a.y * a.y + b.x * b.x + b.y * b.y + alpha * lv.x + beta, alpha * lv.y);
  out[e] = cdiv(num, den);
}

// comp_ola_sdeconv: (conj(gx) wx + conj(gy) wy + Ftpy/(alpha beta)) /
//                   (|f|^2/(alpha beta) + L + gamma); gx, gy, L broadcast
__global__ __launch_bounds__(TB) void k_fdiv_s(float2* out, const float2* gx, const float2* gy, const float2* wx,
                                               const float2* wy, const float2* ftpy, const float2* f,
                                               const float2* L, long long n, long long per, float alpha, float beta,
                                               float gamma) {
  const long long e = (long long)blockIdx.x * TB + threadIdx.x;
  if (e >= n) return;
  const long long q = e % per;
  const float ab = alpha * beta;
  const float2 fy = ftpy[e], fv = f[e], lv = L[q];
  float2 num = cadd(cmulc(wx[e], gx[q]), cmulc(wy[e], gy[q]));
  num = make_float2(num.x + fy.x / ab, num.y + fy.y / ab);
  const float2 den = make_float2((fv.x * fv.x + fv.y * fv.y) / ab + lv.x + gamma, lv.y);
  out[e] = cdiv(num, den);
}

}  // namespace
}  // namespace vmbd

using namespace vmbd;

#define PLAIN_LAUNCH(name, st)                                            \
  ProfScope ps__(name, st);                                               \
  do {                                                                    \
  } while (0)

#define PLAIN_DONE()                                                      \
  do {                                                                    \
    hipError_t e__ = hipGetLastError();                                   \
    if (e__ != hipSuccess) return vmbd_fail_hip(__func__, e__);           \
    return VMBD_OK;                                                       \
  } while (0)

extern "C" {

int vmbd_pad(void* out, int32_t m, int32_t n, const float* in, int32_t p, int32_t q, int32_t s, int32_t t,
             int32_t cplx, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!out || !in || m < 0 || n < 0 || p < 0 || q < 0) return vmbd_fail(VMBD_E_ARG, "vmbd_pad: bad argument");
  hipStream_t st = (hipStream_t)stream;
  PLAIN_LAUNCH("pad", st);
  if ((long long)m * n > 0)
    hipLaunchKernelGGL(k_pad, dim3(blocks_for((long long)m * n)), dim3(TB), 0, st, out, m, n, in, p, q, s, t, cplx);
  PLAIN_DONE();
}

int vmbd_chop_pad(void* out, int32_t sz0, int32_t sz1, const float* x, int32_t H, int32_t W, int32_t csf0,
                  int32_t csf1, int32_t sw0, int32_t sw1, int32_t nhop0, int32_t nhop1, int32_t off0, int32_t off1,
                  const float* win, const int32_t* win_idx, int32_t cplx, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!out || !x || sz0 <= 0 || sz1 <= 0 || csf0 <= 0 || csf1 <= 0 || sw0 <= 0 || sw1 <= 0 || nhop0 <= 0 ||
      nhop1 <= 0 || off0 < 0 || off1 < 0)
    return vmbd_fail(VMBD_E_ARG, "vmbd_chop_pad: bad argument");
  hipStream_t st = (hipStream_t)stream;
  const int P = csf0 * csf1;
  const int sxp0 = nhop0 * csf0 + sw0 - nhop0, sxp1 = nhop1 * csf1 + sw1 - nhop1;
  PLAIN_LAUNCH(win ? "chop_mod_pad" : "chop_pad", st);
  hipLaunchKernelGGL(k_chop_pad, dim3(blocks_for((long long)P * sz0 * sz1)), dim3(TB), 0, st, out, P, sz0, sz1, x, H,
                     W, sxp0, sxp1, sw0, sw1, nhop0, nhop1, csf1, off0, off1, win, win_idx, cplx);
  PLAIN_DONE();
}

int vmbd_pad_stack(void* out, int32_t K, int32_t sz0, int32_t sz1, const float* in, int32_t xm, int32_t xn,
                   int32_t off0, int32_t off1, int32_t cplx, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!out || !in || K <= 0 || sz0 <= 0 || sz1 <= 0 || off0 < 0 || off1 < 0)
    return vmbd_fail(VMBD_E_ARG, "vmbd_pad_stack: bad argument");
  hipStream_t st = (hipStream_t)stream;
  PLAIN_LAUNCH("pad_stack", st);
  hipLaunchKernelGGL(k_pad_stack, dim3(blocks_for((long long)K * sz0 * sz1)), dim3(TB), 0, st, out, K, sz0, sz1, in,
                     xm, xn, off0, off1, cplx);
  PLAIN_DONE();
}

int vmbd_crop_stack(float* out, int32_t K, int32_t sz0, int32_t sz1, const void* in, int32_t inm, int32_t inn,
                    int32_t off0, int32_t off1, int32_t cplx, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!out || !in || K <= 0 || sz0 <= 0 || sz1 <= 0 || off0 < 0 || off1 < 0 || off0 + sz0 > inm ||
      off1 + sz1 > inn)
    return vmbd_fail(VMBD_E_ARG, "vmbd_crop_stack: Size missmatch: Size after - size before < offset");
  hipStream_t st = (hipStream_t)stream;
  PLAIN_LAUNCH("crop_stack", st);
  hipLaunchKernelGGL(k_crop_stack, dim3(blocks_for((long long)K * sz0 * sz1)), dim3(TB), 0, st, out, K, sz0, sz1, in,
                     inm, inn, off0, off1, cplx);
  PLAIN_DONE();
}

int vmbd_crop(float* out, int32_t m, int32_t n, const void* in, int32_t p, int32_t q, int32_t s, int32_t t,
              int32_t cplx, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!out || !in || m <= 0 || n <= 0 || s < 0 || t < 0 || s + m > p || t + n > q)
    return vmbd_fail(VMBD_E_ARG, "vmbd_crop: crop window %dx%d at (%d,%d) outside %dx%d", m, n, s, t, p, q);
  hipStream_t st = (hipStream_t)stream;
  PLAIN_LAUNCH("crop", st);
  hipLaunchKernelGGL(k_crop, dim3(blocks_for((long long)m * n)), dim3(TB), 0, st, out, m, n, in, p, q, s, t, cplx);
  PLAIN_DONE();
}

int vmbd_ola(float* out, const void* in, int32_t inm, int32_t inn, int32_t csf0, int32_t csf1, int32_t wm,
             int32_t wn, int32_t nhop0, int32_t nhop1, int32_t om, int32_t on, int32_t cplx, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!out || !in || csf0 <= 0 || csf1 <= 0 || wm <= 0 || wn <= 0 || nhop0 <= 0 || nhop1 <= 0 || om < 0 || on < 0)
    return vmbd_fail(VMBD_E_ARG, "vmbd_ola: bad argument");
  if (om + wm > inm || on + wn > inn)
    return vmbd_fail(VMBD_E_ARG, "vmbd_ola: extent %dx%d at offset (%d,%d) exceeds the %dx%d slices", wm, wn, om, on,
                     inm, inn);
  hipStream_t st = (hipStream_t)stream;
  const int outm = nhop0 * csf0 + wm - nhop0, outn = nhop1 * csf1 + wn - nhop1;
  PLAIN_LAUNCH("ola", st);
  hipLaunchKernelGGL(k_ola, dim3(blocks_for((long long)outm * outn)), dim3(TB), 0, st, out, outm, outn, in, inm, inn,
                     wm, wn, om, on, csf0, csf1, nhop0, nhop1, cplx);
  PLAIN_DONE();
}

int vmbd_fft2(void* data, int32_t batch, int32_t Th, int32_t Tw, int32_t dir, float scale, const void* twiddles,
              vmbd_stream_t stream) {
  vmbd_clear_error();
  auto ok = [](int v) { return v >= 16 && v <= 1024 && (v & (v - 1)) == 0; };
  if (!data || !twiddles || batch <= 0 || !ok(Th) || !ok(Tw) || (dir != -1 && dir != 1))
    return vmbd_fail(VMBD_E_ARG, "vmbd_fft2: bad argument (sizes must be powers of two in [16,1024])");
  hipStream_t st = (hipStream_t)stream;
  {
    ProfScope ps("fft2_rows", st);
    hipError_t e = fft_rows((float2*)data, (long long)batch * Th, Tw, dir, 1.0f, (const float2*)twiddles, st);
    if (e != hipSuccess) return vmbd_fail_hip("vmbd_fft2 rows", e);
  }
  {
    ProfScope ps("fft2_cols", st);
    hipError_t e = fft_cols((float2*)data, batch, Th, Tw, dir, scale, (const float2*)twiddles, st);
    if (e != hipSuccess) return vmbd_fail_hip("vmbd_fft2 cols", e);
  }
  return VMBD_OK;
}

namespace {
// exp(-2 pi i m / TW_LEN) in FP64, rounded once; built on first use and never
// written again, so asynchronous copies from it need no wait
struct TwiddleTable {
  float2 v[TW_LEN];
  TwiddleTable() {
    for (int m = 0; m < TW_LEN; ++m) {
      const double a = -2.0 * 3.14159265358979323846 * (double)m / (double)TW_LEN;
      v[m] = make_float2((float)cos(a), (float)sin(a));
    }
  }
};
const TwiddleTable& twiddle_table() {
  static const TwiddleTable t;
  return t;
}
}  // namespace

int vmbd_twiddles(void* tw, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!tw) return vmbd_fail(VMBD_E_ARG, "vmbd_twiddles: null");
  const TwiddleTable& t = twiddle_table();
  hipError_t e = hipMemcpyAsync(tw, t.v, sizeof(t.v), hipMemcpyHostToDevice, (hipStream_t)stream);
  if (e != hipSuccess) return vmbd_fail_hip("vmbd_twiddles", e);
  return VMBD_OK;
}

int vmbd_cmul(void* out, const void* a, const void* b, int64_t n, int64_t bper, int32_t mode, float scale,
              vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!out || !a || !b || n < 0 || bper <= 0 || mode < 0 || mode > 2)
    return vmbd_fail(VMBD_E_ARG, "vmbd_cmul: bad argument");
  hipStream_t st = (hipStream_t)stream;
  PLAIN_LAUNCH("cmul", st);
  if (n > 0)
    hipLaunchKernelGGL(k_cmul, dim3(blocks_for(n)), dim3(TB), 0, st, (float2*)out, (const float2*)a,
                       (const float2*)b, (long long)n, (long long)bper, mode, scale);
  PLAIN_DONE();
}

int vmbd_fourier_div(void* out, const void* f, const void* y, const void* L, int64_t n, int64_t per, float alpha,
                     float beta, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!out || !f || !y || !L || n < 0 || per <= 0) return vmbd_fail(VMBD_E_ARG, "vmbd_fourier_div: bad argument");
  hipStream_t st = (hipStream_t)stream;
  PLAIN_LAUNCH("fourier_div", st);
  if (n > 0)
    hipLaunchKernelGGL(k_fdiv, dim3(blocks_for(n)), dim3(TB), 0, st, (float2*)out, (const float2*)f,
                       (const float2*)y, (const float2*)L, (long long)n, (long long)per, alpha, beta);
  PLAIN_DONE();
}

int vmbd_fourier_div_g(void* out, const void* xx, const void* xy, const void* yx, const void* yy, const void* L,
                       int64_t n, int64_t per, float alpha, float beta, vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!out || !xx || !xy || !yx || !yy || !L || n < 0 || per <= 0)
    return vmbd_fail(VMBD_E_ARG, "vmbd_fourier_div_g: bad argument");
  hipStream_t st = (hipStream_t)stream;
  PLAIN_LAUNCH("fourier_div_g", st);
  if (n > 0)
    hipLaunchKernelGGL(k_fdiv_g, dim3(blocks_for(n)), dim3(TB), 0, st, (float2*)out, (const float2*)xx,
                       (const float2*)xy, (const float2*)yx, (const float2*)yy, (const float2*)L, (long long)n,
                       (long long)per, alpha, beta);
  PLAIN_DONE();
}

int vmbd_fourier_div_s(void* out, const void* gx, const void* gy, const void* wx, const void* wy, const void* ftpy,
                       const void* f, const void* L, int64_t n, int64_t per, float alpha, float beta, float gamma,
                       vmbd_stream_t stream) {
  vmbd_clear_error();
  if (!out || !gx || !gy || !wx || !wy || !ftpy || !f || !L || n < 0 || per <= 0)
    return vmbd_fail(VMBD_E_ARG, "vmbd_fourier_div_s: bad argument");
  hipStream_t st = (hipStream_t)stream;
  PLAIN_LAUNCH("fourier_div_s", st);
  if (n > 0)
    hipLaunchKernelGGL(k_fdiv_s, dim3(blocks_for(n)), dim3(TB), 0, st, (float2*)out, (const float2*)gx,
                       (const float2*)gy, (const float2*)wx, (const float2*)wy, (const float2*)ftpy,
                       (const float2*)f, (const float2*)L, (long long)n, (long long)per, alpha, beta, gamma);
  PLAIN_DONE();
}

}  // extern "C"
