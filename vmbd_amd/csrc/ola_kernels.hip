// EFF/OLA spectral pipeline kernels for gfx950.
//
// The operator of olaGPU.py:170-311 is evaluated per patch as a 2-D real FFT
// convolution on a Th x Tw tile, split in three kernels so that every launch
// has thousands of workgroups and all global traffic is coalesced:
//
//   rows_r2c      chop + window + zero pad fused into the load, real row FFTs
//                 (two rows packed in one complex transform), half spectrum
//                 written bin-major [P][NB][rows] so the column pass reads
//                 contiguous runs.  Rows that are all zero are never computed.
//   cols          per (patch, strip of bins): column FFTs of 1..3 inputs, the
//                 spectrum product (conj for the transposes), inverse column
//                 FFT, only the output rows that are needed are stored,
//                 row-major [P][rows][NB].
//   rows_c2r_ola  one workgroup per output row: inverse real row FFTs of every
//                 patch row that lands on it (two patches per transform), then
//                 a fixed-order overlap-add gather and crop (deterministic,
//                 no atomics), optional window multiply (x-adjoint).
//   rows_c2r_stack inverse real row FFTs written as a patch stack (f-adjoint).
#include "fft_core.h"
#include "ola_kernels.h"

namespace vmbd {

constexpr int WG = 256;

// ----------------------------------------------------------------------------
// rows_r2c
// ----------------------------------------------------------------------------
template <int TW>
__global__ __launch_bounds__(WG) void k_rows_r2c(Geo g, RowSrc s, float2* __restrict__ out,
                                                 const float2* __restrict__ tw) {
  constexpr int TPT = TW / 16;
  constexpr int TPW = WG / TPT;
  constexpr int NBN = TW / 2 + 1;
  constexpr int SL = FftLds<TW>::floats2;
  __shared__ float2 fl[TPW * SL];
  __shared__ float2 st[2 * TPW * NBN];

  const int tid = threadIdx.x;
  const int tr = tid / TPT;
  const int t = tid - tr * TPT;
  const int nr = s.rhi - s.rlo;
  const int NQ = (nr + 1) >> 1;
  const long long nitems = (long long)g.P * NQ;
  const long long item = (long long)blockIdx.x * TPW + tr;
  const bool active = item < nitems;
  const int p = active ? (int)(item / NQ) : 0;
  const int q = active ? (int)(item - (long long)p * NQ) : 0;
  const int r0 = s.rlo + 2 * q;
  const int m = p / g.csf1;
  const int n = p - m * g.csf1;
  const int o0 = (s.origin ? m * g.hop0 : 0) - s.sh0;
  const int o1 = (s.origin ? n * g.hop1 : 0) - s.sh1;
  const float* base = s.src + (long long)p * s.pstride;
  const float* wb = s.win ? s.win + (long long)s.win_idx[p] * g.sw0 * g.sw1 : nullptr;

  float2 v[16];
#pragma unroll
  for (int rr = 0; rr < 16; ++rr) {
    const int c = t + TPT * rr;
    float a = 0.f, b = 0.f;
    if (active && c >= s.clo && c < s.chi) {
      const int j = o1 + c;
      if (j >= 0 && j < s.srcW) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int r = r0 + h;
          const int i = o0 + r;
          if (r < s.rhi && i >= 0 && i < s.srcH) {
            float val = base[(long long)i * s.ld + j];
            if (wb) val *= wb[r * g.sw1 + c];
            if (h == 0) a = val; else b = val;
          }
        }
      }
    }
    v[rr] = make_float2(a, b);
  }

  float2* sl = fl + tr * SL;
  fft<TW, -1>(v, sl, t, tw);
  // natural-order spectrum of z = a + i b into the slice, then split
#pragma unroll
  for (int mm = 0; mm < 16; ++mm) sl[lds_pad(t + TPT * mm)] = v[mm];
  wave_sync();
#pragma unroll
  for (int mm = 0; mm <= 8; ++mm) {
    const int c = t + TPT * mm;
    if (c < NBN) {
      const float2 zc = sl[lds_pad(c)];
      const float2 zm = sl[lds_pad((TW - c) & (TW - 1))];
      // A = (Z[c] + conj Z[-c]) / 2,  B = (Z[c] - conj Z[-c]) / (2i)
      st[(2 * tr) * NBN + c] = make_float2(0.5f * (zc.x + zm.x), 0.5f * (zc.y - zm.y));
      st[(2 * tr + 1) * NBN + c] = make_float2(0.5f * (zc.y + zm.y), 0.5f * (zm.x - zc.x));
    }
  }
  __syncthreads();
  // bin-major store: consecutive lanes -> consecutive rows of one bin
  const long long item0 = (long long)blockIdx.x * TPW;
  for (int e = tid; e < 2 * TPW * NBN; e += WG) {
    const int c = e / (2 * TPW);
    const int rl = e - c * (2 * TPW);
    const long long it = item0 + (rl >> 1);
    if (it < nitems) {
      const int pp = (int)(it / NQ);
      const int qq = (int)(it - (long long)pp * NQ);
      const int rr = 2 * qq + (rl & 1);
      if (rr < nr) out[((long long)pp * NBN + c) * nr + rr] = st[rl * NBN + c];
    }
  }
}

template <int TW>
static hipError_t rows_r2c_t(const Geo& g, const RowSrc& s, float2* out, const float2* tw, hipStream_t st) {
  constexpr int TPW = WG / (TW / 16);
  const int nr = s.rhi - s.rlo;
  const long long items = (long long)g.P * ((nr + 1) / 2);
  const long long blocks = (items + TPW - 1) / TPW;
  if (blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rows_r2c<TW>, dim3((unsigned)blocks), dim3(WG), 0, st, g, s, out, tw);
  return hipGetLastError();
}

hipError_t launch_rows_r2c(const Geo& g, const RowSrc& s, float2* out, const float2* tw, hipStream_t st) {
  switch (g.Tw) {
    case 16: return rows_r2c_t<16>(g, s, out, tw, st);
    case 32: return rows_r2c_t<32>(g, s, out, tw, st);
    case 64: return rows_r2c_t<64>(g, s, out, tw, st);
    case 128: return rows_r2c_t<128>(g, s, out, tw, st);
    case 256: return rows_r2c_t<256>(g, s, out, tw, st);
    case 512: return rows_r2c_t<512>(g, s, out, tw, st);
    default: return hipErrorInvalidValue;
  }
}

// ----------------------------------------------------------------------------
// cols
// ----------------------------------------------------------------------------
template <int TH>
__device__ __forceinline__ void col_load(float2 (&v)[16], const ColIn& in, int p, int c, int NB, int t, bool active) {
  constexpr int TPT = TH / 16;
  if (in.spec) {
    const float2* col = in.p + ((long long)p * NB + c) * TH;
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = active ? col[t + TPT * r] : make_float2(0.f, 0.f);
  } else {
    const float2* col = in.p + ((long long)p * NB + c) * in.nr;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int rr = t + TPT * r - in.rlo;
      v[r] = (active && rr >= 0 && rr < in.nr) ? col[rr] : make_float2(0.f, 0.f);
    }
  }
}

// spectrum of one input in registers (output mapping: slot m = row t+TPT*m)
template <int TH>
__device__ __forceinline__ void col_spectrum(float2 (&v)[16], const ColIn& in, int p, int c, int NB, int t,
                                             bool active, float2* sl, const float2* tw) {
  col_load<TH>(v, in, p, c, NB, t, active);
  if (!in.spec) fft<TH, -1>(v, sl, t, tw);
}

// stage rows [0,nrows) of the strip in LDS and store row-major [P][nrows][NB]
template <int TH, int CS>
__device__ __forceinline__ void col_store_rows(const float2 (&v)[16], float2* st, const ColOut& o, int p, int strip,
                                               int NB, int cc, int t) {
  constexpr int TPT = TH / 16;
  __syncthreads();  // staging buffer free
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    const int r = t + TPT * m;
    if (r < o.nrows) st[r * (CS + 1) + cc] = v[m];
  }
  __syncthreads();
  const int c0 = strip * CS;
  const int ncol = min(CS, NB - c0);
  for (int e = threadIdx.x; e < o.nrows * CS; e += WG) {
    const int r = e / CS;
    const int ci = e - r * CS;
    if (ci < ncol) o.p[((long long)p * o.nrows + r) * NB + c0 + ci] = st[r * (CS + 1) + ci];
  }
}

template <int TH, int OP>
__global__ __launch_bounds__(WG) void k_cols(Geo g, ColIn a, ColIn b, ColIn c3, ColOut o1, ColOut o2, float scale,
                                             float2* __restrict__ spec_out, const float2* __restrict__ tw) {
  constexpr int TPT = TH / 16;
  constexpr int CS = WG / TPT;  // columns per workgroup
  constexpr int SL = FftLds<TH>::floats2;
  __shared__ float2 fl[CS * SL];
  __shared__ float2 st[(OP == COL_SPEC) ? 1 : TH * (CS + 1)];

  const int NB = g.NB;
  const int strips = (NB + CS - 1) / CS;
  const int p = blockIdx.x / strips;
  const int strip = blockIdx.x - p * strips;
  const int cc = threadIdx.x / TPT;
  const int t = threadIdx.x - cc * TPT;
  const int c = strip * CS + cc;
  const bool active = c < NB;
  float2* sl = fl + cc * SL;

  float2 A[16];
  col_spectrum<TH>(A, a, p, c, NB, t, active, sl, tw);
  if constexpr (OP == COL_SPEC) {
    if (active) {
      float2* col = spec_out + ((long long)p * NB + c) * TH;
#pragma unroll
      for (int m = 0; m < 16; ++m) col[t + TPT * m] = A[m];
    }
  } else {
    float2 B[16];
    col_spectrum<TH>(B, b, p, c, NB, t, active, sl, tw);
#pragma unroll
    for (int m = 0; m < 16; ++m)
      B[m] = cscale((OP == COL_FWD) ? cmul(A[m], B[m]) : cmulc(A[m], B[m]), scale);
    fft<TH, +1>(B, sl, t, tw);
    col_store_rows<TH, CS>(B, st, o1, p, strip, NB, cc, t);
    if constexpr (OP == COL_ADJ2) {
      col_spectrum<TH>(B, c3, p, c, NB, t, active, sl, tw);
#pragma unroll
      for (int m = 0; m < 16; ++m) B[m] = cscale(cmulc(A[m], B[m]), scale);
      fft<TH, +1>(B, sl, t, tw);
      col_store_rows<TH, CS>(B, st, o2, p, strip, NB, cc, t);
    }
  }
}

template <int TH>
static hipError_t cols_t(const Geo& g, int op, const ColIn& a, const ColIn& b, const ColIn& c, const ColOut& o1,
                         const ColOut& o2, float scale, float2* spec_out, const float2* tw, hipStream_t st) {
  constexpr int CS = WG / (TH / 16);
  const int strips = (g.NB + CS - 1) / CS;
  const dim3 grid((unsigned)(g.P * strips));
  switch (op) {
    case COL_SPEC:
      hipLaunchKernelGGL((k_cols<TH, COL_SPEC>), grid, dim3(WG), 0, st, g, a, b, c, o1, o2, scale, spec_out, tw);
      break;
    case COL_FWD:
      hipLaunchKernelGGL((k_cols<TH, COL_FWD>), grid, dim3(WG), 0, st, g, a, b, c, o1, o2, scale, spec_out, tw);
      break;
    case COL_ADJ1:
      hipLaunchKernelGGL((k_cols<TH, COL_ADJ1>), grid, dim3(WG), 0, st, g, a, b, c, o1, o2, scale, spec_out, tw);
      break;
    case COL_ADJ2:
      hipLaunchKernelGGL((k_cols<TH, COL_ADJ2>), grid, dim3(WG), 0, st, g, a, b, c, o1, o2, scale, spec_out, tw);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_cols(const Geo& g, int op, const ColIn& a, const ColIn& b, const ColIn& c, const ColOut& o1,
                       const ColOut& o2, float scale, float2* spec_out, const float2* tw, hipStream_t st) {
  switch (g.Th) {
    case 16: return cols_t<16>(g, op, a, b, c, o1, o2, scale, spec_out, tw, st);
    case 32: return cols_t<32>(g, op, a, b, c, o1, o2, scale, spec_out, tw, st);
    case 64: return cols_t<64>(g, op, a, b, c, o1, o2, scale, spec_out, tw, st);
    case 128: return cols_t<128>(g, op, a, b, c, o1, o2, scale, spec_out, tw, st);
    case 256: return cols_t<256>(g, op, a, b, c, o1, o2, scale, spec_out, tw, st);
    case 512: return cols_t<512>(g, op, a, b, c, o1, o2, scale, spec_out, tw, st);
    default: return hipErrorInvalidValue;
  }
}

// ----------------------------------------------------------------------------
// inverse real row transform helpers
// ----------------------------------------------------------------------------
// Z[c] = A[c] + i B[c] over the full length from two half spectra (Hermitian
// extension), in the input mapping of the transform.
template <int TW>
__device__ __forceinline__ void c2r_load(float2 (&v)[16], const float2* ra, const float2* rb, int t) {
  constexpr int TPT = TW / 16;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int c = t + TPT * r;
    float2 z = make_float2(0.f, 0.f);
    if (c <= TW / 2) {
      const float2 A = ra ? ra[c] : make_float2(0.f, 0.f);
      const float2 B = rb ? rb[c] : make_float2(0.f, 0.f);
      z = make_float2(A.x - B.y, A.y + B.x);
    } else {
      const int cm = TW - c;
      const float2 A = ra ? ra[cm] : make_float2(0.f, 0.f);
      const float2 B = rb ? rb[cm] : make_float2(0.f, 0.f);
      z = make_float2(A.x + B.y, B.x - A.y);
    }
    v[r] = z;
  }
}

// ----------------------------------------------------------------------------
// rows_c2r_ola: one workgroup per output row
// ----------------------------------------------------------------------------
template <int TW>
__global__ __launch_bounds__(WG) void k_rows_c2r_ola(Geo g, OlaDst d, const float2* __restrict__ tw) {
  constexpr int TPT = TW / 16;
  constexpr int TPW = WG / TPT;
  constexpr int SL = FftLds<TW>::floats2;
  __shared__ float2 fl[TPW * SL];
  extern __shared__ float rowbuf[];  // [mc][csf1][ext1]

  const int i = blockIdx.x;
  const int u = i + d.oc0;
  // covering patch rows m: m*hop0 <= u < m*hop0 + ext0
  int mlo = 0, mhi = -1;
  if (u >= 0) {
    mhi = min(g.csf0 - 1, u / g.hop0);
    const int lo = u - d.ext0 + 1;
    mlo = lo <= 0 ? 0 : (lo + g.hop0 - 1) / g.hop0;
  }
  const int Mc = mhi >= mlo ? (mhi - mlo + 1) : 0;
  const int NP2 = (g.csf1 + 1) >> 1;
  const int NT = Mc * NP2;
  const int tr = threadIdx.x / TPT;
  const int t = threadIdx.x - tr * TPT;
  float2* sl = fl + tr * SL;

  for (int base = 0; base < NT; base += TPW) {
    const int tix = base + tr;
    const bool act = tix < NT;
    const int mi = act ? tix / NP2 : 0;
    const int jp = act ? tix - mi * NP2 : 0;
    const int m = mlo + mi;
    const int r = u - m * g.hop0;
    const int na = 2 * jp, nb = 2 * jp + 1;
    const int ka = m * g.csf1 + na;
    const bool hasb = nb < g.csf1;
    const bool rok = act && r >= 0 && r < d.srow;
    const float2* ra = rok ? d.src + ((long long)ka * d.srow + r) * g.NB : nullptr;
    const float2* rb = (rok && hasb) ? d.src + ((long long)(ka + 1) * d.srow + r) * g.NB : nullptr;
    float2 v[16];
    c2r_load<TW>(v, ra, rb, t);
    fft<TW, +1>(v, sl, t, tw);
    if (act) {
      const float* wa = (d.win && r < g.sw0) ? d.win + ((long long)d.win_idx[ka] * g.sw0 + r) * g.sw1 : nullptr;
      const float* wbp =
          (d.win && hasb && r < g.sw0) ? d.win + ((long long)d.win_idx[ka + 1] * g.sw0 + r) * g.sw1 : nullptr;
      float* ba = rowbuf + ((long long)mi * g.csf1 + na) * d.ext1;
      float* bb = rowbuf + ((long long)mi * g.csf1 + nb) * d.ext1;
#pragma unroll
      for (int mm = 0; mm < 16; ++mm) {
        const int col = t + TPT * mm;
        if (col < d.ext1) {
          float xa = v[mm].x, xb = v[mm].y;
          if (d.win) {
            xa = (wa && col < g.sw1) ? xa * wa[col] : 0.f;
            xb = (wbp && col < g.sw1) ? xb * wbp[col] : 0.f;
          }
          ba[col] = xa;
          if (hasb) bb[col] = xb;
        }
      }
    }
  }
  __syncthreads();
  // fixed-order gather over covering patches
  for (int j = threadIdx.x; j < d.O1; j += WG) {
    const int vcol = j + d.oc1;
    float acc = 0.f;
    if (vcol >= 0) {
      const int nhi = min(g.csf1 - 1, vcol / g.hop1);
      const int lo = vcol - d.ext1 + 1;
      const int nlo = lo <= 0 ? 0 : (lo + g.hop1 - 1) / g.hop1;
      for (int mi = 0; mi < Mc; ++mi)
        for (int nn = nlo; nn <= nhi; ++nn) acc += rowbuf[((long long)mi * g.csf1 + nn) * d.ext1 + (vcol - nn * g.hop1)];
    }
    d.out[(long long)i * d.O1 + j] = acc;
  }
}

size_t ola_rowbuf_bytes(const Geo& g, const OlaDst& d) {
  return (size_t)d.mc * g.csf1 * d.ext1 * sizeof(float);
}

template <int TW>
static hipError_t c2r_ola_t(const Geo& g, const OlaDst& d, const float2* tw, hipStream_t st) {
  const size_t shm = ola_rowbuf_bytes(g, d);
  hipLaunchKernelGGL(k_rows_c2r_ola<TW>, dim3((unsigned)d.O0), dim3(WG), shm, st, g, d, tw);
  return hipGetLastError();
}

hipError_t launch_rows_c2r_ola(const Geo& g, const OlaDst& d, const float2* tw, hipStream_t st) {
  if (d.O0 <= 0 || d.O1 <= 0) return hipSuccess;
  switch (g.Tw) {
    case 16: return c2r_ola_t<16>(g, d, tw, st);
    case 32: return c2r_ola_t<32>(g, d, tw, st);
    case 64: return c2r_ola_t<64>(g, d, tw, st);
    case 128: return c2r_ola_t<128>(g, d, tw, st);
    case 256: return c2r_ola_t<256>(g, d, tw, st);
    case 512: return c2r_ola_t<512>(g, d, tw, st);
    default: return hipErrorInvalidValue;
  }
}

// ----------------------------------------------------------------------------
// rows_c2r_stack: (P, nrows, ncols) stack from the first rows/cols of each tile
// ----------------------------------------------------------------------------
template <int TW>
__global__ __launch_bounds__(WG) void k_rows_c2r_stack(Geo g, const float2* __restrict__ src, int srow,
                                                       float* __restrict__ out, int nrows, int ncols,
                                                       const float2* __restrict__ tw) {
  constexpr int TPT = TW / 16;
  constexpr int TPW = WG / TPT;
  constexpr int SL = FftLds<TW>::floats2;
  __shared__ float2 fl[TPW * SL];
  const int tr = threadIdx.x / TPT;
  const int t = threadIdx.x - tr * TPT;
  const int NQ = (nrows + 1) >> 1;
  const long long item = (long long)blockIdx.x * TPW + tr;
  const bool act = item < (long long)g.P * NQ;
  const int p = act ? (int)(item / NQ) : 0;
  const int q = act ? (int)(item - (long long)p * NQ) : 0;
  const int r0 = 2 * q, r1 = 2 * q + 1;
  const float2* ra = act ? src + ((long long)p * srow + r0) * g.NB : nullptr;
  const float2* rb = (act && r1 < nrows) ? src + ((long long)p * srow + r1) * g.NB : nullptr;
  float2 v[16];
  c2r_load<TW>(v, ra, rb, t);
  float2* sl = fl + tr * SL;
  fft<TW, +1>(v, sl, t, tw);
  if (act) {
#pragma unroll
    for (int mm = 0; mm < 16; ++mm) {
      const int col = t + TPT * mm;
      if (col < ncols) {
        out[((long long)p * nrows + r0) * ncols + col] = v[mm].x;
        if (r1 < nrows) out[((long long)p * nrows + r1) * ncols + col] = v[mm].y;
      }
    }
  }
}

template <int TW>
static hipError_t c2r_stack_t(const Geo& g, const float2* src, int srow, float* out, int nrows, int ncols,
                              const float2* tw, hipStream_t st) {
  constexpr int TPW = WG / (TW / 16);
  const long long items = (long long)g.P * ((nrows + 1) / 2);
  const long long blocks = (items + TPW - 1) / TPW;
  if (blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rows_c2r_stack<TW>, dim3((unsigned)blocks), dim3(WG), 0, st, g, src, srow, out, nrows,
                     ncols, tw);
  return hipGetLastError();
}

hipError_t launch_rows_c2r_stack(const Geo& g, const float2* src, int srow, float* out, int nrows, int ncols,
                                 const float2* tw, hipStream_t st) {
  switch (g.Tw) {
    case 16: return c2r_stack_t<16>(g, src, srow, out, nrows, ncols, tw, st);
    case 32: return c2r_stack_t<32>(g, src, srow, out, nrows, ncols, tw, st);
    case 64: return c2r_stack_t<64>(g, src, srow, out, nrows, ncols, tw, st);
    case 128: return c2r_stack_t<128>(g, src, srow, out, nrows, ncols, tw, st);
    case 256: return c2r_stack_t<256>(g, src, srow, out, nrows, ncols, tw, st);
    case 512: return c2r_stack_t<512>(g, src, srow, out, nrows, ncols, tw, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace vmbd
