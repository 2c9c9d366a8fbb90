// Column pass of the closed-form deconvolutions (olaGPU.py:416-733; the
// per-patch Fourier divisions of gputools.cu:316-406) for gfx950.
//
// Same decomposition as the operator's column pass (ola_kernels.hip): one
// workgroup per (patch, strip of 16 row bins), one 64-lane wavefront holding
// four 256-point column transforms, the row bins of real rows in half-spectrum
// form with bins 0 and Tw/2 packed into column 0.  Per column the kernel
// transforms its row-pass operands, reads its cached spectra, forms the
// division of DivOp (ola_kernels.h) bin by bin with the filter energies in
// closed form, and either inverse-transforms the quotient and stores the rows
// the caller needs or (SPROD) stores the product as a spectrum.
//
// The packed column carries two Hermitian spectra D (row bin 0) and N (row
// bin Tw/2) as P = D + iN.  A division is not bilinear, so its operands are
// taken apart first: each transform of that column is staged in LDS and
// D(k) = (P(k) + conj P(-k))/2, N(k) = (P(k) - conj P(-k))/(2i) are formed from
// it; the quotients of both parts (Hermitian again) are put back as D' + iN'.
#include "col_device.h"
#include "fft_core.h"
#include "ola_kernels.h"
#include "prof.h"

namespace vmbd {

namespace {

constexpr float SQRT1_2 = 0.70710678118654752440f;

// 4 sin^2(pi q / T) = 2 - 2 cos(2 pi q / T), without the cancellation near 0
__device__ __forceinline__ float four_sin2(int q, int T) {
  const float s = sinpif((float)q / (float)T);
  return 4.f * s * s;
}
// (1 + 2 cos(2 pi q / T))^2
__device__ __forceinline__ float one_2cos_sq(int q, int T) {
  const float c = 1.f + 2.f * cospif(2.f * (float)q / (float)T);
  return c * c;
}
// r (e^{-2 pi i q / T} - 1), r = sqrt(2)/2: the transform of the
// difference filter r [-1 1] placed at the tile origin
__device__ __forceinline__ float2 diff_filter(int q, int T) {
  float s, c;
  sincospif(2.f * (float)q / (float)T, &s, &c);
  return make_float2(SQRT1_2 * (c - 1.f), -SQRT1_2 * s);
}

__device__ __forceinline__ float cabs2(float2 a) { return fmaf(a.x, a.x, a.y * a.y); }
__device__ __forceinline__ float2 cdivr(float2 a, float d) { return make_float2(a.x / d, a.y / d); }

// Hermitian parts (D, N) of a packed transform P at frequency k, given P(-k)
__device__ __forceinline__ void split(float2 p, float2 q, float2& d, float2& n) {
  d = make_float2(0.5f * (p.x + q.x), 0.5f * (p.y - q.y));
  n = make_float2(0.5f * (p.y + q.y), 0.5f * (q.x - p.x));
}
__device__ __forceinline__ float2 join(float2 d, float2 n) { return make_float2(d.x - n.y, d.y + n.x); }

// the division of one (row bin, column frequency) from its operand values
// (order as in DivOp); cx = row bin, k = column frequency
template <int OP>
__device__ __forceinline__ float2 div_value(const float2 (&o)[4], int cx, int k, int Tw, int Th, DivParams p) {
  if constexpr (OP == DIV_X) {
    const float L = four_sin2(cx, Tw) + four_sin2(k, Th);
    return cdivr(cmulc(o[0], o[1]), cabs2(o[1]) + p.alpha * L + p.beta);
  } else if constexpr (OP == DIV_G) {
    const float L = one_2cos_sq(k, Th) * four_sin2(cx, Tw) + four_sin2(k, Th) * one_2cos_sq(cx, Tw);
    const float2 num = cadd(cmulc(o[1], o[0]), cmulc(o[3], o[2]));
    return cdivr(num, cabs2(o[0]) + cabs2(o[2]) + p.alpha * L + p.beta);
  } else if constexpr (OP == DIV_SPROD) {
    return cmulc(o[0], o[1]);
  } else {
    const float2 gx = diff_filter(cx, Tw), gy = diff_filter(k, Th);
    const float ab = p.alpha * p.beta;
    const float2 num = cadd(cadd(cmulc(o[0], gx), cmulc(o[1], gy)), cdivr(o[2], ab));
    return cdivr(num, cabs2(o[3]) / ab + cabs2(gx) + cabs2(gy) + p.gamma);
  }
}

template <int OP>
struct DivArity {
  static constexpr int n = (OP == DIV_X || OP == DIV_SPROD) ? 2 : 4;
};

template <int TH, int OP>
__global__ __launch_bounds__(WG, 1) void k_cols_div(Geo g, ColIn a0, ColIn a1, ColIn a2, ColIn a3, DivParams prm,
                                                   ColOut o, float2* __restrict__ spec_out,
                                                   const float2* __restrict__ tw) {
  constexpr int TPT = TH / 16;
  constexpr int CS = WG / TPT;  // columns per workgroup
  constexpr int SL = FftLds<TH>::floats2;
  constexpr int NA = DivArity<OP>::n;
  // FFT slices (aliased by the row staging of the store), then NA slices in
  // which the packed column's operand transforms are staged
  __shared__ float2 sm[cmax<CS * SL, TH * (CS + 1)>::v + NA * SL];
  float2* pk = sm + cmax<CS * SL, TH * (CS + 1)>::v;

  const int NB = g.NB;
  const int strip = blockIdx.x;
  const int p = (int)blockIdx.z * g.Pf + (int)blockIdx.y;
  const int cc = threadIdx.x / TPT;
  const int t = threadIdx.x - cc * TPT;
  const int c = strip * CS + cc;
  const bool active = c < NB - 1;  // bins [0, NB-1); NB-1 travels packed in bin 0
  const bool packed = c == 0;
  const bool wact = strip * CS + (int)(threadIdx.x >> 6) * (64 / TPT) < NB - 1;
  const int pfr = p / g.Pf;
  const int pk_ = p - pfr * g.Pf;
  const int pm = pk_ / g.csf1;
  const int grp = pfr * g.csf1 + (pk_ - pm * g.csf1);
  const int row0 = pm * g.hop0;
  const ColCtx cx{pm, pk_ - pm * g.csf1, pfr, g.hop0, g.Tw, tw};
  float2* sl = sm + cc * SL;

  Twiddles<TH> T;
  twiddles_init<TH, true>(T, t, tw);

  const ColIn ins[4] = {a0, a1, a2, a3};
  // Operands are consumed one at a time: transformed (row-pass outputs) or
  // loaded (spectra), staged when the column is the packed one, and folded
  // into the running numerator / denominator of the plain columns.
  float2 num[16], acc[16];
#pragma unroll
  for (int m = 0; m < 16; ++m) num[m] = acc[m] = make_float2(0.f, 0.f);
  float2 prev[16];
#pragma unroll
  for (int j = 0; j < NA; ++j) {
    float2 v[16];
    col_load<TH>(v, ins[j], p, c, NB, t, active, grp, row0, cx);
    if (!ins[j].spec && wact) fft<TH, -1>(v, sl, t, T, tw);
    if (packed) {
#pragma unroll
      for (int m = 0; m < 16; ++m) pk[j * SL + lds_pad(t + TPT * m)] = v[m];
    }
    // plain columns: keep what the formula needs (pairs for DIV_X / SPROD /
    // DIV_G, running sums for DIV_S)
    if constexpr (OP == DIV_S) {
      const float ab = prm.alpha * prm.beta;
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        const int k = t + TPT * m;
        if (j == 0) num[m] = cmulc(v[m], diff_filter(c, g.Tw));
        if (j == 1) num[m] = cadd(num[m], cmulc(v[m], diff_filter(k, g.Th)));
        if (j == 2) num[m] = cadd(num[m], cdivr(v[m], ab));
        if (j == 3) acc[m] = make_float2(cabs2(v[m]) / ab, 0.f);
      }
    } else if constexpr (OP == DIV_G) {
      if (j == 0 || j == 2) {
#pragma unroll
        for (int m = 0; m < 16; ++m) prev[m] = v[m];
      } else {
#pragma unroll
        for (int m = 0; m < 16; ++m) {
          num[m] = cadd(num[m], cmulc(v[m], prev[m]));
          acc[m].x += cabs2(prev[m]);
        }
      }
    } else {
      if (j == 0) {
#pragma unroll
        for (int m = 0; m < 16; ++m) prev[m] = v[m];
      } else {
#pragma unroll
        for (int m = 0; m < 16; ++m) {
          float2 ov[4] = {prev[m], v[m], make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
          num[m] = div_value<OP>(ov, c, t + TPT * m, g.Tw, g.Th, prm);
        }
      }
    }
  }
  float2 Z[16];
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    const int k = t + TPT * m;
    if constexpr (OP == DIV_S) {
      const float2 gx = diff_filter(c, g.Tw), gy = diff_filter(k, g.Th);
      Z[m] = cdivr(num[m], acc[m].x + cabs2(gx) + cabs2(gy) + prm.gamma);
    } else if constexpr (OP == DIV_G) {
      const float L = one_2cos_sq(k, g.Th) * four_sin2(c, g.Tw) + four_sin2(k, g.Th) * one_2cos_sq(c, g.Tw);
      Z[m] = cdivr(num[m], acc[m].x + prm.alpha * L + prm.beta);
    } else {
      Z[m] = num[m];
    }
  }
  if (packed) {
    // both Hermitian parts of every staged operand, the division of each
    // part (row bins 0 and Tw/2), the quotient packed again
    wave_sync();
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const int k = t + TPT * m;
      const int kr = (TH - k) & (TH - 1);
      float2 od[4], on[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        od[j] = on[j] = make_float2(0.f, 0.f);
        if (j < NA) split(pk[j * SL + lds_pad(k)], pk[j * SL + lds_pad(kr)], od[j], on[j]);
      }
      Z[m] = join(div_value<OP>(od, 0, k, g.Tw, g.Th, prm), div_value<OP>(on, g.Tw / 2, k, g.Tw, g.Th, prm));
    }
  }
  if constexpr (OP == DIV_SPROD) {
    if (active) {
      float2* col = spec_out + ((size_t)p * NB + c) * TH;
#pragma unroll
      for (int m = 0; m < 16; ++m) col[t + TPT * m] = Z[m];
      if (packed) {  // bin NB-1 travels in bin 0: its column is written as zeros
        float2* coln = spec_out + ((size_t)p * NB + NB - 1) * TH;
#pragma unroll
        for (int m = 0; m < 16; ++m) coln[t + TPT * m] = make_float2(0.f, 0.f);
      }
    }
  } else {
    if (wact) fft<TH, +1>(Z, sl, t, T, tw);
    col_store_rows<TH, CS, true>(Z, sm, o, p, strip, NB, g.NBP, cc, t);
  }
}

template <int TH>
hipError_t cols_div_t(const Geo& g, int op, const ColIn* a, DivParams prm, const ColOut& o, float2* spec_out,
                      const float2* tw, hipStream_t st) {
  constexpr int CS = WG / (TH / 16);
  const int strips = (g.NB - 1 + CS - 1) / CS;
  const dim3 grid((unsigned)strips, (unsigned)g.Pf, (unsigned)(g.P / g.Pf));
  static const char* names[4] = {"cols_div_x", "cols_div_g", "cols_sprod", "cols_div_s"};
  ProfScope ps(names[op & 3], st);
  switch (op) {
    case DIV_X:
      hipLaunchKernelGGL((k_cols_div<TH, DIV_X>), grid, dim3(WG), 0, st, g, a[0], a[1], a[2], a[3], prm, o,
                         spec_out, tw);
      break;
    case DIV_G:
      hipLaunchKernelGGL((k_cols_div<TH, DIV_G>), grid, dim3(WG), 0, st, g, a[0], a[1], a[2], a[3], prm, o,
                         spec_out, tw);
      break;
    case DIV_SPROD:
      hipLaunchKernelGGL((k_cols_div<TH, DIV_SPROD>), grid, dim3(WG), 0, st, g, a[0], a[1], a[2], a[3], prm, o,
                         spec_out, tw);
      break;
    case DIV_S:
      hipLaunchKernelGGL((k_cols_div<TH, DIV_S>), grid, dim3(WG), 0, st, g, a[0], a[1], a[2], a[3], prm, o,
                         spec_out, tw);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_cols_div(const Geo& g, int op, const ColIn* a, int na, DivParams prm, const ColOut& o,
                           float2* spec_out, const float2* tw, hipStream_t st) {
  const int need = (op == DIV_X || op == DIV_SPROD) ? 2 : 4;
  if (na != need || g.P <= 0) return hipErrorInvalidValue;
  ColIn in[4] = {a[0], a[1], ColIn{}, ColIn{}};
  if (na == 4) {
    in[2] = a[2];
    in[3] = a[3];
  }
  switch (g.Th) {
    case 16: return cols_div_t<16>(g, op, in, prm, o, spec_out, tw, st);
    case 32: return cols_div_t<32>(g, op, in, prm, o, spec_out, tw, st);
    case 64: return cols_div_t<64>(g, op, in, prm, o, spec_out, tw, st);
    case 128: return cols_div_t<128>(g, op, in, prm, o, spec_out, tw, st);
    case 256: return cols_div_t<256>(g, op, in, prm, o, spec_out, tw, st);
    case 512: return cols_div_t<512>(g, op, in, prm, o, spec_out, tw, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace vmbd
