// EFF/OLA spectral pipeline kernels for gfx950.
//
// The operator of olaGPU.py:170-311 is evaluated per patch as a 2-D real FFT
// convolution on a Th x Tw tile, split in three kernels so that every launch
// has hundreds to thousands of workgroups and all global traffic is coalesced:
//
//   rows_r2c      chop + window + zero pad fused into the load (16-byte
//                 loads, redistributed to the FFT mapping through LDS), real
//                 row FFTs (two rows packed in one complex transform), half
//                 spectrum written bin-major [P][NB][row_pitch(rows)] so the
//                 column pass reads contiguous runs; the real bins 0 and Tw/2
//                 of a row share bin 0, (X[0], X[Tw/2]), so Tw/2 columns
//                 remain.  Rows that are all zero are never computed.
//   cols          per (patch, strip of 16 columns): column FFTs of 1..3
//                 inputs, the spectrum product (conj for the transposes; the
//                 packed column 0 is taken apart and put back), inverse column
//                 FFT; only the output rows that are needed are stored,
//                 row-major [P][rows][NBP] (NBP = NB rounded up to 16 bins,
//                 so a 16-bin strip of a row is one 128-byte line).
//   rows_c2r_ola  one workgroup per output row: inverse real row FFTs of every
//                 patch row that lands on it (two patches per transform), then
//                 a fixed-order overlap-add gather and crop (deterministic,
//                 no atomics), optional window multiply (x-adjoint).
//   rows_c2r_stack inverse real row FFTs written as a patch stack (f-adjoint).
//
// LDS: each workgroup owns one buffer; the FFT slices and the staging area of
// the transposed store alias it (separated by workgroup barriers), which keeps
// the footprint at 35-40 KB and allows four workgroups per CU.
#include <algorithm>
#include <climits>

#include "col_device.h"
#include "fft_core.h"
#include "ola_kernels.h"
#include "prof.h"

namespace vmbd {



// ----------------------------------------------------------------------------
// rows_r2c
// ----------------------------------------------------------------------------
// LDS (float2) of one rows_r2c workgroup: the transforms' slices, aliased by
// the staging of the bin-major store
template <int TW>
struct RowsLds {
  static constexpr int TPW = WG / (TW / 16);
  static constexpr int v = cmax<TPW * FftLds<TW>::floats2, 2 * TPW * (TW / 2 + 1)>::v;
};

// One workgroup of the row pass: row pairs [q0, q0 + TPW) of patch pu of
// source s into out (sm: RowsLds<TW>::v float2 of LDS).
template <int TW>
__device__ __forceinline__ void rows_r2c_body(Geo g, RowSrc s, float2* __restrict__ out, int pu, int q0, float2* sm,
                                              const float2* __restrict__ tw) {
  constexpr int TPT = TW / 16;
  constexpr int TPW = WG / TPT;
  constexpr int NBN = TW / 2 + 1;
  constexpr int SL = FftLds<TW>::floats2;

  const int tid = threadIdx.x;
  const int tr = tid / TPT;
  const int t = tid - tr * TPT;
  const int nr = s.rhi - s.rlo;
  const int NQ = (nr + 1) >> 1;
  const int q = q0 + tr;
  const bool active = q < NQ;
  const int p = pu;
  const int r0 = s.rlo + 2 * q;
  const int fr = p / g.Pf;
  const int k = p - fr * g.Pf;
  const int m = k / g.csf1;
  const int n = k - m * g.csf1;
  const int o0 = (s.origin ? m * g.hop0 : 0) - s.sh0;
  const int o1 = (s.origin ? n * g.hop1 : 0) - s.sh1;
  const float* base = s.src + (size_t)p * s.pstride + (size_t)fr * s.fstride;
  const float* wb = s.win ? s.win + (size_t)s.win_idx[k] * g.sw0 * g.sw1 : nullptr;

  Twiddles<TW> T;
  twiddles_init<TW>(T, t, tw);
  // a wavefront whose row pairs all lie past the end skips its transforms
  const bool wact = q0 + (tid >> 6) * (64 / TPT) < NQ;

  // rows r0 (real part) and r0+1 (imaginary part)
  const int i0 = o0 + r0, i1 = i0 + 1;
  const bool ok0 = active && r0 < s.rhi && i0 >= 0 && i0 < s.srcH;
  const bool ok1 = active && r0 + 1 < s.rhi && i1 >= 0 && i1 < s.srcH;
  const float* row0 = base + (size_t)(ok0 ? i0 : 0) * s.ld;
  const float* row1 = base + (size_t)(ok1 ? i1 : 0) * s.ld;
  // window rows of r0 and r0+1, column c - wsh1 (only read inside
  // [clo,chi)): rows r0 - wsh0 and r0 + 1 - wsh0 of the patch's window, or
  // the patch column's weights of each source row's type (cw)
  const bool hw = wb || s.cw;
  const float* w0 = wb ? wb + ((long long)(r0 - s.wsh0) * g.sw1 - s.wsh1) : nullptr;
  const float* w1 = wb ? w0 + g.sw1 : nullptr;
  if (s.cw) {
    w0 = s.cw + ((long long)(k * 2 + (ok0 ? s.czr[i0] : 0)) * g.sw1 - s.wsh1);
    w1 = s.cw + ((long long)(k * 2 + (ok1 ? s.czr[i1] : 0)) * g.sw1 - s.wsh1);
  }

  float2* sl = sm + tr * SL;
  float2 za[9], zb[9];
  if (wact) {
    // Thread t reads columns 4m..4m+3, m = t + TPT*i, of both rows (and of
    // their window rows) with 16-byte loads, all issued before any is used;
    // chunks that cross the column range or the source edge take element
    // loads.  The transform's LDS slice then hands column t + TPT*r to slot r
    // of the FFT input mapping.
    float a[4][4], b[4][4], wa[4][4], wb4[4][4];
    bool full[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int c0 = 4 * (t + TPT * i);
      const int j0 = o1 + c0;
      full[i] = c0 >= s.clo && c0 + 4 <= s.chi && j0 >= 0 && j0 + 4 <= s.srcW;
#pragma unroll
      for (int e = 0; e < 4; ++e) a[i][e] = b[i][e] = wa[i][e] = wb4[i][e] = 0.f;
      if (full[i] && ok0) ld4(a[i], row0 + j0);
      if (full[i] && ok1) ld4(b[i], row1 + j0);
      if (hw && full[i] && ok0) ld4(wa[i], w0 + c0);
      if (hw && full[i] && ok1) ld4(wb4[i], w1 + c0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int c0 = 4 * (t + TPT * i);
      if (!full[i] && c0 + 4 > s.clo && c0 < s.chi) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int c = c0 + e;
          const int j = o1 + c;
          const bool okc = c >= s.clo && c < s.chi && j >= 0 && j < s.srcW;
          if (ok0 && okc) a[i][e] = row0[j];
          if (ok1 && okc) b[i][e] = row1[j];
          if (hw && ok0 && okc) wa[i][e] = w0[c];
          if (hw && ok1 && okc) wb4[i][e] = w1[c];
        }
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int c0 = 4 * (t + TPT * i);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float x0 = hw ? a[i][e] * wa[i][e] : a[i][e];
        const float x1 = hw ? b[i][e] * wb4[i][e] : b[i][e];
        sl[lds_pad(c0 + e)] = make_float2(x0, x1);
      }
    }
    wave_sync();
    float2 v[16];
    // columns at or past chi are zero: a job whose rows end within 3*TPT
    // columns (PSF rows) takes a transform whose first pass knows the other
    // slots are zero (uniform per job; not at 512 points, where the second
    // transform takes the kernel past 128 VGPRs)
    if (TW <= 256 && s.chi <= 3 * TPT) {
#pragma unroll
      for (int r = 0; r < 3; ++r) v[r] = sl[lds_pad(t + TPT * r)];
      wave_sync();
      fft<TW, -1, (TW <= 256 ? 3 : 16)>(v, sl, t, T, tw);
    } else {
#pragma unroll
      for (int r = 0; r < 16; ++r) v[r] = sl[lds_pad(t + TPT * r)];
      wave_sync();
      fft<TW, -1>(v, sl, t, T, tw);
    }
    // natural-order spectrum of z = a + i b into the slice, then split
#pragma unroll
    for (int mm = 0; mm < 16; ++mm) sl[lds_pad(t + TPT * mm)] = v[mm];
    wave_sync();
#pragma unroll
    for (int mm = 0; mm <= 8; ++mm) {
      const int c = t + TPT * mm;
      const float2 zc = sl[lds_pad(c < NBN ? c : 0)];
      const float2 zm = sl[lds_pad((TW - c) & (TW - 1))];
      // A = (Z[c] + conj Z[-c]) / 2,  B = (Z[c] - conj Z[-c]) / (2i)
      za[mm] = make_float2(0.5f * (zc.x + zm.x), 0.5f * (zc.y - zm.y));
      zb[mm] = make_float2(0.5f * (zc.y + zm.y), 0.5f * (zm.x - zc.x));
    }
    // bins 0 and TW/2 of a real row are real: bin 0 carries both,
    // (X[0], X[TW/2]), so the column pass has TW/2 columns (thread 0 holds
    // both bins, slots 0 and 8)
    if (t == 0) {
      za[0] = make_float2(za[0].x, za[8].x);
      zb[0] = make_float2(zb[0].x, zb[8].x);
    }
  } else {
#pragma unroll
    for (int mm = 0; mm <= 8; ++mm) za[mm] = zb[mm] = make_float2(0.f, 0.f);
  }
  // bin-major store staged through LDS: lane pairs of rows (2rp, 2rp+1) of
  // one bin are one 16-byte write, so the workgroup's 32 rows of a bin (row
  // blocks start on 32-row, 256-byte boundaries of the 16-row pitch) are two
  // whole lines written by 16 consecutive lanes
  __syncthreads();  // every slice has been read: the buffer becomes staging
#pragma unroll
  for (int mm = 0; mm <= 8; ++mm) {
    const int c = t + TPT * mm;
    if (c < NBN - 1) {  // bin TW/2 travels in bin 0
      sm[(2 * tr) * NBN + c] = za[mm];
      sm[(2 * tr + 1) * NBN + c] = zb[mm];
    }
  }
  __syncthreads();
  const int nrp = row_pitch(nr);
  float2* dst = out + (size_t)p * NBN * nrp + 2 * q0;
  const int nrow = min(2 * TPW, nr - 2 * q0);
  // (an odd row count writes its last pair into the padding row of the pitch)
  const int rp = tid & (TPW - 1);
  if (2 * rp < nrow)
    for (int c = tid / TPW; c < NBN - 1; c += WG / TPW) {
      const float2 a = sm[(2 * rp) * NBN + c], b = sm[(2 * rp + 1) * NBN + c];
      *(float4*)(dst + (size_t)c * nrp + 2 * rp) = make_float4(a.x, a.y, b.x, b.y);
    }
}

// 1-D grid over work units (patch, row block), row blocks of all jobs
// job-major.  Workgroup ids are dealt round-robin to the NXCD XCDs; XCD j
// runs the contiguous units [start_j, start_j + count_j).  Units are ordered
// in tiles of TP = ceil(P / NXCD) consecutive patches (one XCD's share),
// row-block-major inside a tile, so neighbouring patches, which read the
// same pixels, run on one XCD at the same time and share its L2.  The map
// id -> (patch, row block) is a bijection whatever the hardware placement.
// blk: the workgroup's id among the launch's g.P * (row blocks) (its XCD is
// blk mod NXCD).  Returns the patch, the job and its first row pair.
template <int TW>
__device__ __forceinline__ void rows_unit(const Geo& g, const RowJobs& jobs, int blk, int& pu, int& ji, int& q0) {
  constexpr int TPW = WG / (TW / 16);
  const int nblk = jobs.blk_end[jobs.n - 1];
  const int total = g.P * nblk;
  const int xcd = blk % NXCD;
  const int u = xcd * (total / NXCD) + min(xcd, total % NXCD) + blk / NXCD;
  const int TP = (g.P + NXCD - 1) / NXCD;
  const int tile = u / (TP * nblk);
  const int ur = u - tile * TP * nblk;
  const int Tt = min(TP, g.P - tile * TP);
  const int by = ur / Tt;
  pu = tile * TP + (ur - by * Tt);
  ji = 0;
  int blk0 = 0;
#pragma unroll
  for (int jj = 0; jj < ROW_JOBS_MAX - 1; ++jj)
    if (jobs.n > jj + 1 && by >= jobs.blk_end[jj]) {
      ji = jj + 1;
      blk0 = jobs.blk_end[jj];
    }
  q0 = (by - blk0) * TPW;
}

template <int TW>
__global__ __launch_bounds__(WG) void k_rows_r2c(Geo g, RowJobs jobs, const float2* __restrict__ tw) {
  __shared__ float2 sm[RowsLds<TW>::v];

  int pu, ji, q0;
  rows_unit<TW>(g, jobs, (int)blockIdx.x, pu, ji, q0);
  const RowSrc s = jobs.s[ji];
  rows_r2c_body<TW>(g, s, jobs.out[ji], pu, q0, sm, tw);
}

// Row jobs with a patch count of their own: the units of job 0, then of
// job 1, each in the tile order of rows_unit over its own patches; the XCD
// of workgroup b runs a contiguous range of them.
template <int TW>
__global__ __launch_bounds__(WG) void k_rows_r2c_geo(Geo g, RowJobs jobs, RowJobGeo jg,
                                                     const float2* __restrict__ tw) {
  __shared__ float2 sm[RowsLds<TW>::v];
  constexpr int TPW = WG / (TW / 16);
  const int blk = (int)blockIdx.x;
  const int total = jobs.n > 1 ? jg.uend[1] : jg.uend[0];
  const int xcd = blk % NXCD;
  const int u = xcd * (total / NXCD) + min(xcd, total % NXCD) + blk / NXCD;
  const bool j1 = jobs.n > 1 && u >= jg.uend[0];
  const int ul = j1 ? u - jg.uend[0] : u;
  const int P = j1 ? jg.P[1] : jg.P[0];
  const int nb = j1 ? jg.nb[1] : jg.nb[0];
  const int TP = (P + NXCD - 1) / NXCD;
  const int tile = ul / (TP * nb);
  const int ur = ul - tile * TP * nb;
  const int Tt = min(TP, P - tile * TP);
  const int by = ur / Tt;
  const int pu = tile * TP + (ur - by * Tt);
  Geo gj = g;
  gj.P = P;
  gj.Pf = j1 ? jg.Pf[1] : jg.Pf[0];
  rows_r2c_body<TW>(gj, j1 ? jobs.s[1] : jobs.s[0], j1 ? jobs.out[1] : jobs.out[0], pu, by * TPW, sm, tw);
}

template <int TW>
static hipError_t rows_r2c_geo_t(const Geo& g, const RowSrc* s, float2* const* out, const int* P, const int* Pf, int n,
                                 const float2* tw, hipStream_t st, const char* name) {
  constexpr int TPW = WG / (TW / 16);
  RowJobs jobs{};
  RowJobGeo jg{};
  jobs.n = n;
  long long units = 0;
  for (int i = 0; i < n; ++i) {
    jobs.s[i] = s[i];
    jobs.out[i] = out[i];
    jg.P[i] = P[i];
    jg.Pf[i] = Pf[i];
    jg.nb[i] = ((s[i].rhi - s[i].rlo + 1) / 2 + TPW - 1) / TPW;
    units += (long long)P[i] * jg.nb[i];
    if (jg.nb[i] <= 0 || P[i] <= 0 || units > INT_MAX) return hipErrorInvalidValue;
    jg.uend[i] = (int)units;
  }
  ProfScope ps(name ? name : "rows_r2c_multi", st);
  hipLaunchKernelGGL(k_rows_r2c_geo<TW>, dim3((unsigned)units), dim3(WG), 0, st, g, jobs, jg, tw);
  return hipGetLastError();
}

hipError_t launch_rows_r2c_geo(const Geo& g, const RowSrc* s, float2* const* out, const int* P, const int* Pf, int n,
                               const float2* tw, hipStream_t st, const char* name) {
  if (n < 1 || n > 2) return hipErrorInvalidValue;
  switch (g.Tw) {
    case 16: return rows_r2c_geo_t<16>(g, s, out, P, Pf, n, tw, st, name);
    case 32: return rows_r2c_geo_t<32>(g, s, out, P, Pf, n, tw, st, name);
    case 64: return rows_r2c_geo_t<64>(g, s, out, P, Pf, n, tw, st, name);
    case 128: return rows_r2c_geo_t<128>(g, s, out, P, Pf, n, tw, st, name);
    case 256: return rows_r2c_geo_t<256>(g, s, out, P, Pf, n, tw, st, name);
    case 512: return rows_r2c_geo_t<512>(g, s, out, P, Pf, n, tw, st, name);
    default: return hipErrorInvalidValue;
  }
}

template <int TW>
static hipError_t rows_r2c_t(const Geo& g, const RowSrc* s, float2* const* out, int n, const float2* tw,
                             hipStream_t st, const char* name) {
  constexpr int TPW = WG / (TW / 16);
  RowJobs jobs{};
  jobs.n = n;
  int blocks = 0;
  for (int i = 0; i < n; ++i) {
    jobs.s[i] = s[i];
    jobs.out[i] = out[i];
    const int nq = (s[i].rhi - s[i].rlo + 1) / 2;
    blocks += (nq + TPW - 1) / TPW;
    jobs.blk_end[i] = blocks;
  }
  if (blocks <= 0 || g.P <= 0) return hipSuccess;
  ProfScope ps(name ? name : n == 1 ? "rows_r2c" : "rows_r2c_multi", st);
  hipLaunchKernelGGL(k_rows_r2c<TW>, dim3((unsigned)(g.P * blocks)), dim3(WG), 0, st, g, jobs, tw);
  return hipGetLastError();
}

static hipError_t rows_r2c_any(const Geo& g, const RowSrc* s, float2* const* out, int n, const float2* tw,
                               hipStream_t st, const char* name) {
  if (n < 1 || n > ROW_JOBS_MAX) return hipErrorInvalidValue;
  switch (g.Tw) {
    case 16: return rows_r2c_t<16>(g, s, out, n, tw, st, name);
    case 32: return rows_r2c_t<32>(g, s, out, n, tw, st, name);
    case 64: return rows_r2c_t<64>(g, s, out, n, tw, st, name);
    case 128: return rows_r2c_t<128>(g, s, out, n, tw, st, name);
    case 256: return rows_r2c_t<256>(g, s, out, n, tw, st, name);
    case 512: return rows_r2c_t<512>(g, s, out, n, tw, st, name);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_rows_r2c_multi(const Geo& g, const RowSrc* s, float2* const* out, int n, const float2* tw,
                                 hipStream_t st) {
  return rows_r2c_any(g, s, out, n, tw, st, nullptr);
}

hipError_t launch_rows_r2c(const Geo& g, const RowSrc& s, float2* out, const float2* tw, hipStream_t st,
                           const char* name) {
  return rows_r2c_any(g, &s, &out, 1, tw, st, name);
}

// ----------------------------------------------------------------------------
// cols
// ----------------------------------------------------------------------------
// Spectrum product of one column: B = A * B (forward) or A * conj(B)
// (adjoints); the inverse transforms' 1/(Th*Tw) is applied to the summed
// output by the final stores (a power of two: exact).  The column of bin 0 (packed; the first TPT
// lanes of wavefront 0 of strip 0) carries two real row bins, 0 and TW/2: its
// transforms are P = D + i N of two Hermitian spectra D, N, taken apart with
// P(-k) (D = (P(k) + conj P(-k)) / 2, N = (P(k) - conj P(-k)) / 2i),
// multiplied bin by bin and put together as D' + i N'.  Its operands go to
// LDS (its own slice and the next column's, which belongs to the same
// wavefront and is idle between the transforms); after the plain product of
// every column, all 64 lanes of that wavefront form the packed product in
// place, each taking pairs of frequencies (k, TH-k), and its lanes read their
// slots back (SPREAD; otherwise the column's own lanes form it, 16 slots each).  Written
// as a macro: the same code behind a function boundary takes 32 more VGPRs
// (four wavefronts per SIMD become three).
#define VMBD_COL_PRODUCT(CONJ, A, B)                                                                \
  do {                                                                                              \
    if (packed) {                                                                                   \
      _Pragma("unroll") for (int m = 0; m < 16; ++m) {                                              \
        sm[lds_pad(t + TPT * m)] = A[m];                                                            \
        sm[SL + lds_pad(t + TPT * m)] = B[m];                                                       \
      }                                                                                             \
    }                                                                                               \
    _Pragma("unroll") for (int m = 0; m < 16; ++m) B[m] =                                           \
        (CONJ) ? cmulc(A[m], B[m]) : cmul(A[m], B[m]);                                              \
    if (SPREAD && pwave) {                                                                          \
      /* frequencies k <= TH/2 with their mirror TH-k: one lane reads both */                      \
      /* pairs of both operands and writes both products in place         */                       \
      constexpr int NP_ = TH / 128 + 1;                                                             \
      const int lane_ = threadIdx.x & 63;                                                           \
      wave_sync();                                                                                  \
      _Pragma("unroll 1") for (int j = 0; j < NP_; ++j) {                                           \
        const int k_ = lane_ + 64 * j;                                                              \
        if (k_ <= TH / 2) {                                                                         \
          const int kr_ = (TH - k_) & (TH - 1);                                                     \
          const float2 a_ = sm[lds_pad(k_)], ar_ = sm[lds_pad(kr_)];                                \
          const float2 b_ = sm[SL + lds_pad(k_)], br_ = sm[SL + lds_pad(kr_)];                      \
          const float2 ad_ = make_float2(0.5f * (a_.x + ar_.x), 0.5f * (a_.y - ar_.y));             \
          const float2 an_ = make_float2(0.5f * (a_.y + ar_.y), 0.5f * (ar_.x - a_.x));             \
          const float2 bd_ = make_float2(0.5f * (b_.x + br_.x), 0.5f * (b_.y - br_.y));             \
          const float2 bn_ = make_float2(0.5f * (b_.y + br_.y), 0.5f * (br_.x - b_.x));             \
          const float2 pd_ = (CONJ) ? cmulc(ad_, bd_) : cmul(ad_, bd_);                            \
          const float2 pn_ = (CONJ) ? cmulc(an_, bn_) : cmul(an_, bn_);                            \
          /* at -k: D(-k) = conj D(k), N(-k) = conj N(k) */                                         \
          const float2 pdr_ = cconj(pd_), pnr_ = cconj(pn_);                                        \
          sm[lds_pad(k_)] = make_float2(pd_.x - pn_.y, pd_.y + pn_.x);                              \
          if (kr_ != k_) sm[lds_pad(kr_)] = make_float2(pdr_.x - pnr_.y, pdr_.y + pnr_.x);          \
        }                                                                                           \
      }                                                                                             \
      wave_sync();                                                                                  \
      if (packed) {                                                                                 \
        _Pragma("unroll") for (int m = 0; m < 16; ++m) B[m] = sm[lds_pad(t + TPT * m)];             \
      }                                                                                             \
      wave_sync(); /* the slices are written next by the inverse transform */                      \
    }                                                                                               \
    if (!SPREAD && packed) {                                                                        \
      wave_sync();                                                                                  \
      _Pragma("unroll") for (int m = 0; m < 16; ++m) {                                              \
        const int k_ = t + TPT * m;                                                                 \
        const int kr_ = (TH - k_) & (TH - 1);                                                       \
        const float2 a_ = sm[lds_pad(k_)], ar_ = sm[lds_pad(kr_)];                                  \
        const float2 b_ = sm[SL + lds_pad(k_)], br_ = sm[SL + lds_pad(kr_)];                        \
        const float2 ad_ = make_float2(0.5f * (a_.x + ar_.x), 0.5f * (a_.y - ar_.y));               \
        const float2 an_ = make_float2(0.5f * (a_.y + ar_.y), 0.5f * (ar_.x - a_.x));               \
        const float2 bd_ = make_float2(0.5f * (b_.x + br_.x), 0.5f * (b_.y - br_.y));               \
        const float2 bn_ = make_float2(0.5f * (b_.y + br_.y), 0.5f * (br_.x - b_.x));               \
        const float2 pd_ = (CONJ) ? cmulc(ad_, bd_) : cmul(ad_, bd_);                              \
        const float2 pn_ = (CONJ) ? cmulc(an_, bn_) : cmul(an_, bn_);                              \
        B[m] = make_float2(pd_.x - pn_.y, pd_.y + pn_.x);                                           \
      }                                                                                             \
      wave_sync(); /* the slices are written next by the inverse transform */                      \
    }                                                                                               \
  } while (0)

// LDS (float2) of one column-pass workgroup: the transforms' slices, aliased
// by the row staging of the store
template <int TH, int OP, int WGC = WG>
struct ColsLds {
  static constexpr int CS = WGC / (TH / 16);
  static constexpr int v = (OP == COL_SPEC) ? CS * FftLds<TH>::floats2 : cmax<CS * FftLds<TH>::floats2, TH * (CS + 1)>::v;
};

// One workgroup of WGC threads of the column pass: strip `strip` of patch p
// (sm: ColsLds<TH, OP, WGC>::v float2 of LDS).
template <int TH, int OP, int WGC = WG>
// (parameters by value: with references the compiler keeps the column
// descriptors in vector registers and spills at 256 points)
__device__ __forceinline__ void cols_body(Geo g, ColIn a, ColIn b, ColIn c3, ColOut o1, ColOut o2,
                                          float2* __restrict__ spec_out, int strip, int p, float2* sm,
                                          const float2* __restrict__ tw) {
  constexpr int TPT = TH / 16;
  constexpr int CS = WGC / TPT;  // columns per workgroup
  constexpr int SL = FftLds<TH>::floats2;

  const int NB = g.NB;
  const int cc = threadIdx.x / TPT;
  const int t = threadIdx.x - cc * TPT;
  const int c = strip * CS + cc;
  // columns: bins [0, NB-1); bin 0 carries the real bins 0 and NB-1 packed
  const bool active = c < NB - 1;
  const bool packed = c == 0;
  // wavefront 0 of strip 0 holds the packed column (lanes 0..TPT-1) and
  // forms its product (uniform per wavefront)
  const bool pwave = strip == 0 && threadIdx.x < 64;
  // the first product spreads the packed column over the wavefront; the
  // two-adjoint pass forms its second one on the column's own lanes (spread
  // twice, its registers spill)
  constexpr bool SPREAD1 = true;
  constexpr bool SPREAD2 = false;
  // a wavefront whose columns all lie past the last bin skips its transforms
  const bool wact = strip * CS + (int)(threadIdx.x >> 6) * (64 / TPT) < NB - 1;
  // group and row offset of this patch in shared row transforms
  const int pfr = p / g.Pf;
  const int pk = p - pfr * g.Pf;
  const int pm = pk / g.csf1;
  const int grp = pfr * g.csf1 + (pk - pm * g.csf1);
  const int row0 = pm * g.hop0;
  const ColCtx cx{pm, pk - pm * g.csf1, pfr, g.hop0, g.Tw, tw};
  // operands that may be shared image rows: the forward's and the spectra's
  // first, the f-adjoint's second, the two-adjoint pass's third
  constexpr bool XA = OP == COL_FWD || OP == COL_SPEC;
  constexpr bool XB = OP == COL_ADJ1;
  float2* sl = sm + cc * SL;

  Twiddles<TH> T;
  twiddles_init<TH, true>(T, t, tw);

  float2 A[16];
  // operands of at most 3*TPT stored rows (PSF rows) take a transform whose
  // first pass knows the other slots are zero; so do stores of at most
  // 3*TPT rows (f-adjoint) for the inverse transform's last pass.  Not in
  // the 512-point two-adjoint pass, whose registers the extra paths spill.
  constexpr bool PRUNE = !(TH >= 512 && OP == COL_ADJ2);
  if (PRUNE && (OP == COL_FWD || OP == COL_SPEC) && !a.spec && a.rlo + a.nr <= 3 * TPT) {
    col_load<TH, 3, XA>(A, a, p, c, NB, t, active, grp, row0, cx);
    if (wact) fft<TH, -1, 3>(A, sl, t, T, tw);
  } else if (PRUNE && XA && !a.spec && a.rlo + a.nr <= 12 * TPT) {
    // image rows of a patch of at most 3/4 of the tile height
    col_load<TH, 12, XA>(A, a, p, c, NB, t, active, grp, row0, cx);
    if (wact) fft<TH, -1, 12>(A, sl, t, T, tw);
  } else {
    col_load<TH, 16, XA>(A, a, p, c, NB, t, active, grp, row0, cx);
    if (!a.spec && wact) fft<TH, -1>(A, sl, t, T, tw);
  }
  if constexpr (OP == COL_SPEC) {
    if (active) {
      float2* col = spec_out + ((size_t)p * NB + c) * TH;
#pragma unroll
      for (int m = 0; m < 16; ++m) col[t + TPT * m] = A[m];
      if (packed) {  // bin NB-1 travels in bin 0: its column is written as zeros
        float2* coln = spec_out + ((size_t)p * NB + NB - 1) * TH;
#pragma unroll
        for (int m = 0; m < 16; ++m) coln[t + TPT * m] = make_float2(0.f, 0.f);
      }
    }
  } else {
    float2 B[16];
    if (PRUNE && !b.spec && b.rlo + b.nr <= 3 * TPT) {
      col_load<TH, 3, XB>(B, b, p, c, NB, t, active, grp, row0, cx);
      if (wact) fft<TH, -1, 3>(B, sl, t, T, tw);
    } else {
      col_load<TH, 16, XB>(B, b, p, c, NB, t, active, grp, row0, cx);
      if (!b.spec && wact) fft<TH, -1>(B, sl, t, T, tw);
    }
    {
      constexpr bool SPREAD = SPREAD1;
      VMBD_COL_PRODUCT(OP != COL_FWD, A, B);
    }
    if (OP == COL_ADJ1 && o1.nrows <= 3 * TPT) {
      if (wact) fft<TH, +1>(B, sl, t, T, tw);
      col_store_rows<TH, CS, true, 3, WGC>(B, sm, o1, p, strip, NB, g.NBP, cc, t);
    } else {
      if (wact) fft<TH, +1>(B, sl, t, T, tw);
      col_store_rows<TH, CS, OP != COL_ADJ2, 16, WGC>(B, sm, o1, p, strip, NB, g.NBP, cc, t);
    }
    if constexpr (OP == COL_ADJ2) {
      if (PRUNE && !c3.spec && c3.rlo + c3.nr <= 12 * TPT) {
        col_load<TH, 12, true>(B, c3, p, c, NB, t, active, grp, row0, cx);
        if (wact) fft<TH, -1, 12>(B, sl, t, T, tw);
      } else {
        col_load<TH, 16, true>(B, c3, p, c, NB, t, active, grp, row0, cx);
        if (!c3.spec && wact) fft<TH, -1>(B, sl, t, T, tw);
      }
      {
        constexpr bool SPREAD = SPREAD2;
        VMBD_COL_PRODUCT(true, A, B);
      }
      if (wact) fft<TH, +1>(B, sl, t, T, tw);
      if (PRUNE && o2.nrows <= 3 * TPT)
        col_store_rows<TH, CS, true, 3, WGC>(B, sm, o2, p, strip, NB, g.NBP, cc, t);
      else
        col_store_rows<TH, CS, true, 16, WGC>(B, sm, o2, p, strip, NB, g.NBP, cc, t);
    }
  }
}

// 256- and 512-point columns: four wavefronts per SIMD (at most 128 VGPRs,
// a few spilled at 512) beat three without spills; the LDS allows four
// workgroups per CU.  Strips of one patch are consecutive workgroups; each
// writes whole 128-byte lines of its rows (16 bins, row pitch NBP a multiple
// of 16).
template <int TH, int OP, int WGC>
__global__ __launch_bounds__(WGC, (TH >= 256 ? 4 * WG / WGC : 1)) void k_cols(Geo g, ColIn a, ColIn b, ColIn c3,
                                                                             ColOut o1, ColOut o2,
                                                                             float2* __restrict__ spec_out,
                                                                             const float2* __restrict__ tw) {
  __shared__ float2 sm[ColsLds<TH, OP, WGC>::v];
  // frame blockIdx.z, patch blockIdx.y
  cols_body<TH, OP, WGC>(g, a, b, c3, o1, o2, spec_out, blockIdx.x, (int)blockIdx.z * g.Pf + (int)blockIdx.y, sm,
                         tw);
}

#undef VMBD_COL_PRODUCT

template <int TH, int WGC>
static hipError_t cols_wg(const Geo& g, int op, const ColIn& a, const ColIn& b, const ColIn& c, const ColOut& o1,
                          const ColOut& o2, float2* spec_out, const float2* tw, hipStream_t st) {
  constexpr int CS = WGC / (TH / 16);
  const int strips = (g.NB - 1 + CS - 1) / CS;  // bins [0, NB-1): NB-1 is packed into bin 0
  const dim3 grid((unsigned)strips, (unsigned)g.Pf, (unsigned)(g.P / g.Pf));
  static const char* names[4] = {"cols_spec", "cols_fwd", "cols_adj1", "cols_adj2"};
  ProfScope ps(names[op & 3], st);
  switch (op) {
    case COL_SPEC:
      hipLaunchKernelGGL((k_cols<TH, COL_SPEC, WGC>), grid, dim3(WGC), 0, st, g, a, b, c, o1, o2, spec_out, tw);
      break;
    case COL_FWD:
      hipLaunchKernelGGL((k_cols<TH, COL_FWD, WGC>), grid, dim3(WGC), 0, st, g, a, b, c, o1, o2, spec_out, tw);
      break;
    case COL_ADJ1:
      hipLaunchKernelGGL((k_cols<TH, COL_ADJ1, WGC>), grid, dim3(WGC), 0, st, g, a, b, c, o1, o2, spec_out, tw);
      break;
    case COL_ADJ2:
      hipLaunchKernelGGL((k_cols<TH, COL_ADJ2, WGC>), grid, dim3(WGC), 0, st, g, a, b, c, o1, o2, spec_out, tw);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int TH>
static hipError_t cols_t(const Geo& g, int op, const ColIn& a, const ColIn& b, const ColIn& c, const ColOut& o1,
                         const ColOut& o2, float2* spec_out, const float2* tw, hipStream_t st, int wgc) {
  if (wgc == 128) return cols_wg<TH, 128>(g, op, a, b, c, o1, o2, spec_out, tw, st);
  return cols_wg<TH, WG>(g, op, a, b, c, o1, o2, spec_out, tw, st);
}

hipError_t launch_cols(const Geo& g, int op, const ColIn& a, const ColIn& b, const ColIn& c, const ColOut& o1,
                       const ColOut& o2, float2* spec_out, const float2* tw, hipStream_t st, int wgc) {
  if (wgc != 0 && wgc != 128 && wgc != WG) return hipErrorInvalidValue;
  switch (g.Th) {
    case 16: return cols_t<16>(g, op, a, b, c, o1, o2, spec_out, tw, st, wgc);
    case 32: return cols_t<32>(g, op, a, b, c, o1, o2, spec_out, tw, st, wgc);
    case 64: return cols_t<64>(g, op, a, b, c, o1, o2, spec_out, tw, st, wgc);
    case 128: return cols_t<128>(g, op, a, b, c, o1, o2, spec_out, tw, st, wgc);
    case 256: return cols_t<256>(g, op, a, b, c, o1, o2, spec_out, tw, st, wgc);
    case 512: return cols_t<512>(g, op, a, b, c, o1, o2, spec_out, tw, st, wgc);
    default: return hipErrorInvalidValue;
  }
}

// ----------------------------------------------------------------------------
// inverse real row transform helpers
// ----------------------------------------------------------------------------
// Bins [0, TW/2] of each row are read with 16-byte loads of two bins (chunk k
// holds bins 2k, 2k+1; thread t reads chunks t + TPT*i), all issued before
// any is used.  Rows are stored with a pitch of at least TW/2 + 2 bins, so the
// last chunk stays inside its row; its bins are not used: bin 0 holds the
// real bins 0 and TW/2 of the row, (X[0], X[TW/2]).  The full
// spectrum Z[c] = A[c] + i B[c] (conjugate-symmetric halves mirrored to
// TW - c) is written into the transform's LDS slice and slot r takes
// Z[t + TPT*r].  A missing row (hasa / hasb false) reads as zero.
template <int TW>
struct C2r {
  static constexpr int TPT = TW / 16;
  static constexpr int NCH = TW / 4;                    // two-bin chunks covering [0, TW/2)
  static constexpr int NI = (NCH + TPT - 1) / TPT;      // chunks per thread
  static_assert(NI == 4, "chunk count per thread");
};

// Bins of an absent row: a transform without one of its rows reads these
// zeros, so every lane issues the same loads without a per-lane test.
__device__ const float4 c2r_zero_bins[512 / 4] = {};

// global loads of both rows' chunks (all issued before any is used)
template <int TW>
__device__ __forceinline__ void c2r_fetch(float4 (&A)[4], float4 (&B)[4], const float2* ra, const float2* rb,
                                          bool hasa, bool hasb, int t) {
  static_assert(C2r<TW>::NCH <= 512 / 4, "zero row shorter than a row's chunks");
  const float2* pa = hasa ? ra : (const float2*)c2r_zero_bins;
  const float2* pb = hasb ? rb : (const float2*)c2r_zero_bins;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int k = t + C2r<TW>::TPT * i;
    A[i] = B[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (k < C2r<TW>::NCH) {
      __builtin_memcpy(&A[i], pa + 2 * k, 16);
      __builtin_memcpy(&B[i], pb + 2 * k, 16);
    }
  }
}

// Hermitian extension in the slice and read-back in the input mapping
template <int TW>
__device__ __forceinline__ void c2r_stage(float2 (&v)[16], const float4 (&A)[4], const float4 (&B)[4], float2* sl,
                                          int t) {
  constexpr int TPT = C2r<TW>::TPT;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int k = t + TPT * i;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int c = 2 * k + e;
      const float2 a = e ? make_float2(A[i].z, A[i].w) : make_float2(A[i].x, A[i].y);
      const float2 b = e ? make_float2(B[i].z, B[i].w) : make_float2(B[i].x, B[i].y);
      if (c == 0) {
        // bin 0 holds the real bins 0 and TW/2 of each row: (X[0], X[TW/2])
        sl[lds_pad(0)] = make_float2(a.x, b.x);
        sl[lds_pad(TW / 2)] = make_float2(a.y, b.y);
      } else if (c < TW / 2) {
        sl[lds_pad(c)] = make_float2(a.x - b.y, a.y + b.x);
        sl[lds_pad(TW - c)] = make_float2(a.x + b.y, b.x - a.y);
      }
    }
  }
  wave_sync();
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r] = sl[lds_pad(t + TPT * r)];
  wave_sync();
}

template <int TW>
__device__ __forceinline__ void c2r_load(float2 (&v)[16], const float2* ra, const float2* rb, bool hasa, bool hasb,
                                         float2* sl, int t) {
  float4 A[4], B[4];
  c2r_fetch<TW>(A, B, ra, rb, hasa, hasb, t);
  c2r_stage<TW>(v, A, B, sl, t);
}

// ----------------------------------------------------------------------------
// rows_c2r_ola: one workgroup per output row
// ----------------------------------------------------------------------------
__host__ __device__ int ola_window_offset_floats(const Geo& g, const OlaDst& d);

// Window rows of output row u: row u - m*hop0 of every window class of every
// covering patch row m (pair pr = mi*nslot + slot), zero where the row lies
// outside the window or the slot is unused.  A staged row keeps the columns
// t + TPT*mm that thread t of a transform weights at floats t*MW + mm
// (ola_win_mw), so the thread reads them with 16-byte LDS loads.  The slot
// of every covering patch, [mi][n], is staged beside them.  Wavefront w
// stages pairs pr = w + 4k by direct global-to-LDS loads (lane l of a load
// writes LDS float l past the wavefront-uniform base, from any address): no
// registers hold the rows, so the spectrum rows are fetched while they are in
// flight.  The caller waits for the loads (s_waitcnt) and a workgroup
// barrier before reading them.
typedef __attribute__((address_space(3))) void lds_void_t;

template <int TW, int WGS>
__device__ __forceinline__ void ola_stage_windows(float* wl, int* wsl, const Geo& g, const OlaDst& d, int Mc, int mlo,
                                                  int u) {
  constexpr int TPT = TW / 16;
  constexpr int NW = WGS / 64;  // wavefronts of the workgroup
  constexpr int JU = (TW + 63) / 64;  // staged row width <= TW
  const int MW = ola_win_mw(g);
  const int RW = TPT * MW;
  const int npair = Mc * d.nslot;
  const int lane = threadIdx.x & 63;
  const int w = (int)(threadIdx.x >> 6);
  for (int pr = w; pr < npair; pr += NW) {
    const int mi = pr / d.nslot;
    const int slot = pr - mi * d.nslot;
    const int m = mlo + mi;
    const int r = u - m * g.hop0;
    const int cls = d.row_class[m * d.nslot + slot];
    float* dst = wl + pr * RW;
    if (cls >= 0 && r >= 0 && r < g.sw0) {
      const float* src = d.win + ((size_t)cls * g.sw0 + r) * g.sw1;
#pragma unroll
      for (int j = 0; j < JU; ++j) {
        const int q = lane + 64 * j;  // staged position: thread tt, slot mm
        const int tt = q / MW;
        const int c = tt + TPT * (q - tt * MW);
        if (q < RW && c < g.sw1)
          __builtin_amdgcn_global_load_lds((const void*)(src + c), (lds_void_t*)(dst + 64 * j), 4, 0, 0);
      }
    } else {
#pragma unroll
      for (int j = 0; j < JU; ++j) {
        const int q = lane + 64 * j;
        if (q < RW) dst[q] = 0.f;
      }
    }
  }
  for (int mi = w; mi < Mc; mi += NW)
    for (int n0 = 0; n0 < g.csf1; n0 += 64)
      if (n0 + lane < g.csf1)
        __builtin_amdgcn_global_load_lds((const void*)(d.win_slot + (mlo + mi) * g.csf1 + n0 + lane),
                                         (lds_void_t*)(wsl + mi * g.csf1 + n0), 4, 0, 0);
}


constexpr int OLA_KJ = 8;  // output columns per thread of streamed launches (O1 <= OLA_KJ * WG)

// STREAM: several rounds of TPW transforms; each round's rows go to a buffer
// that aliases the FFT slices and are added at once into per-column
// accumulators in LDS (fixed round and patch order), so the LDS holds one
// round instead of every covering patch row.  Otherwise all rows of the
// output row are gathered after the last round (single-round launches alias
// that buffer with the slices).
// WIN: x-adjoint launches (d.win set), instantiated apart so that the
// forward's code carries no window path.
// WGS: threads per workgroup (ola_wg): single-round launches of a full
// grid take only the threads their transforms need.
// One workgroup of the OLA pass: output row i of frame fr (dyn: the
// launch's ola_shared_bytes of LDS).
template <int TW, bool STREAM, bool WIN, int WGS>
__device__ __forceinline__ void c2r_ola_body(Geo g, OlaDst d, int i, int fr, float2* dyn,
                                             const float2* __restrict__ tw) {
  constexpr int TPT = TW / 16;
  constexpr int wgs = WGS;
  constexpr int TPW = WGS / TPT;
  constexpr int SL = FftLds<TW>::floats2;
  float2* fl = dyn;
  float* rowbuf = (STREAM || d.alias) ? (float*)dyn : (float*)(dyn + TPW * SL);

  const int u = i + d.oc0;
  const int kf = fr * g.Pf;  // first patch of this frame
  // covering patch rows m: m*hop0 <= u < m*hop0 + ext0
  int mlo = 0, mhi = -1;
  if (u >= 0) {
    mhi = min(g.csf0 - 1, u / g.hop0);
    const int lo = u - d.ext0 + 1;
    mlo = lo <= 0 ? 0 : (lo + g.hop0 - 1) / g.hop0;
  }
  const int Mc = mhi >= mlo ? (mhi - mlo + 1) : 0;
  const int NP2 = (g.csf1 + 1) >> 1;
  // class layout of the gather buffer (single-round launches)
  const int NC = STREAM ? 0 : d.ncls;
  const int RLc = ola_class_len(g, d);
  const int NT = Mc * NP2;
  const int tr = threadIdx.x / TPT;
  const int t = threadIdx.x - tr * TPT;
  float2* sl = fl + tr * SL;

  Twiddles<TW> T;
  twiddles_init<TW>(T, t, tw);

  // x-adjoint: the window rows this output row needs, one per (covering patch
  // row, window class of that row), staged in LDS past the gather buffer
  float* wl = (float*)(dyn) + ola_window_offset_floats(g, d);
  int* wsl = (int*)(wl + (WIN ? d.mc * d.nslot * ola_win_rw(g) : 0));
  if constexpr (WIN) ola_stage_windows<TW, WGS>(wl, wsl, g, d, Mc, mlo, u);  // waited for before the first rowbuf write

  const float inv_hop1 = 1.0f / (float)g.hop1;
  // streamed launches: this row's accumulators in LDS past the staged
  // windows; column j belongs to thread j mod wgs, so no barrier guards them
  float* accl = (float*)wsl + (WIN ? d.mc * g.csf1 : 0);
  if constexpr (STREAM)
    for (int j = threadIdx.x; j < d.O1; j += wgs) accl[j] = 0.f;

  for (int base = 0; base < NT; base += TPW) {
    const int tix = base + tr;
    const bool act = tix < NT;
    const int mi = act ? tix / NP2 : 0;
    const int jp = act ? tix - mi * NP2 : 0;
    const int m = mlo + mi;
    const int r = u - m * g.hop0;
    const int na = 2 * jp, nb = 2 * jp + 1;
    const int ka = kf + m * g.csf1 + na;
    const bool hasb = nb < g.csf1;
    const bool rok = act && r >= 0 && r < d.srow;
    const float2* ra = rok ? d.src + ((size_t)ka * d.srow + r) * g.NBP : d.src;
    const float2* rb = (rok && hasb) ? ra + (size_t)d.srow * g.NBP : ra;
    // the patches' rows lie inside their windows
    const bool wrow = WIN && rok && r < g.sw0;
    float2 v[16];
    // transforms without an item skip their work and never touch LDS, so a
    // single-round launch holds slices for its items only
    if (act) {
      c2r_load<TW>(v, ra, rb, rok, rok && hasb, sl, t);
      fft<TW, +1>(v, sl, t, T, tw);
    } else {
#pragma unroll
      for (int mm = 0; mm < 16; ++mm) v[mm] = make_float2(0.f, 0.f);
    }
    if (WIN && base == 0) __builtin_amdgcn_s_waitcnt(0);  // this wavefront's window loads have landed
    if (STREAM || d.alias || (WIN && base == 0)) __syncthreads();  // slices about to become rowbuf; windows staged
    if (!STREAM && NC) {
      // class rows start as zeros: the patches of a class leave gaps
      float4* z4 = (float4*)rowbuf;
      const int n4 = (Mc * NC * RLc) >> 2;
      for (int e = threadIdx.x; e < n4; e += wgs) z4[e] = make_float4(0.f, 0.f, 0.f, 0.f);
      __syncthreads();
    }
    if (act) {
      float* ba;
      float* bb;
      if (STREAM) {
        ba = rowbuf + 2 * tr * d.ext1;
        bb = ba + d.ext1;
      } else if (NC) {
        // patch n's row at column n*hop1 of class row n mod NC
        const int qa = na % NC;
        const int qb = qa + 1 == NC ? 0 : qa + 1;
        ba = rowbuf + (mi * NC + qa) * RLc + na * g.hop1;
        bb = rowbuf + (mi * NC + qb) * RLc + nb * g.hop1;
      } else {
        ba = rowbuf + (mi * g.csf1 + na) * d.ext1;
        bb = ba + d.ext1;
      }
      // staged window values of this thread's columns (x-adjoint: ext1 == sw1)
      const float* wa = wl;
      const float* wbp = wl;
      if constexpr (WIN) {
        const int MW = ola_win_mw(g);
        const int RW = TPT * MW;
        if (wrow) {
          wa = wl + (mi * d.nslot + wsl[mi * g.csf1 + na]) * RW + t * MW;
          wbp = hasb ? wl + (mi * d.nslot + wsl[mi * g.csf1 + nb]) * RW + t * MW : wa;
        }
      }
      float4 qa = make_float4(0.f, 0.f, 0.f, 0.f), qb = qa;
#pragma unroll
      for (int mm = 0; mm < 16; ++mm) {
        const int col = t + TPT * mm;
        if (col < d.ext1) {
          float xa = v[mm].x, xb = v[mm].y;
          if constexpr (WIN) {
            if ((mm & 3) == 0 && wrow) {
              qa = *(const float4*)(wa + mm);
              qb = *(const float4*)(wbp + mm);
            }
            const float fa = (mm & 3) == 0 ? qa.x : (mm & 3) == 1 ? qa.y : (mm & 3) == 2 ? qa.z : qa.w;
            const float fb = (mm & 3) == 0 ? qb.x : (mm & 3) == 1 ? qb.y : (mm & 3) == 2 ? qb.z : qb.w;
            xa = wrow ? xa * fa : 0.f;
            xb = wrow ? xb * fb : 0.f;
          }
          ba[col] = xa;
          if (hasb) bb[col] = xb;
        }
      }
    }
    if constexpr (STREAM) {
      __syncthreads();
      // this round's items tix in [tlo, thi): patch row mi, pair n/2
      const int tlo = base, thi = min(base + TPW, NT);
      for (int j = threadIdx.x; j < d.O1; j += wgs) {
        const int vcol = j + d.oc1;
        if (vcol >= 0) {
          int qh = (int)((float)vcol * inv_hop1);
          qh += ((qh + 1) * g.hop1 <= vcol) - (qh * g.hop1 > vcol);
          const int nhi = min(g.csf1 - 1, qh);
          const int lc0 = vcol - nhi * g.hop1;
          for (int mr = tlo / NP2; mr <= (thi - 1) / NP2; ++mr) {
            float sub = 0.f;
            for (int nn = nhi, lc = lc0; nn >= 0 && lc < d.ext1; --nn, lc += g.hop1) {
              const int tx = mr * NP2 + (nn >> 1);
              if (tx >= tlo && tx < thi) sub += rowbuf[(2 * (tx - tlo) + (nn & 1)) * d.ext1 + lc];
            }
            accl[j] += sub;
          }
        }
      }
      __syncthreads();  // the round's rows are read before the next round's transforms
    }
  }
  if constexpr (STREAM) {
    for (int j = threadIdx.x; j < d.O1; j += wgs) {
      const size_t o = ((size_t)fr * d.O0 + i) * d.O1 + j;
      float a = accl[j] * d.scale;
      if (d.sub) a -= d.sub[o];
      if (d.mul) a *= d.mul[o];
      if (d.axz) a += d.beta * d.axz[o];
      if (d.clip_lo) a = fmaxf(a, d.lo);
      d.out[o] = a;
    }
    return;
  }
  __syncthreads();
  // fixed-order gather over covering patches
  const int K = Mc * NC;
  for (int j = threadIdx.x; j < d.O1; j += wgs) {
    const int vcol = j + d.oc1;
    float acc = 0.f;
    if (NC) {
      // column vcol of every class row, in (patch row, class) order
      if (vcol >= 0 && vcol < RLc) {
        const float* cb = rowbuf + vcol;
        int k = 0;
        for (; k + 4 <= K; k += 4) {
          const float a0 = cb[k * RLc], a1 = cb[(k + 1) * RLc], a2 = cb[(k + 2) * RLc], a3 = cb[(k + 3) * RLc];
          acc += a0;
          acc += a1;
          acc += a2;
          acc += a3;
        }
        for (; k < K; ++k) acc += cb[k * RLc];
      }
    } else if (vcol >= 0) {
      // nhi = min(csf1-1, floor(vcol/hop1)) via a float reciprocal + exact fix-up
      int qh = (int)((float)vcol * inv_hop1);
      qh += ((qh + 1) * g.hop1 <= vcol) - (qh * g.hop1 > vcol);
      const int nhi = min(g.csf1 - 1, qh);
      const int lc0 = vcol - nhi * g.hop1;
      for (int mi = 0; mi < Mc; ++mi) {
        const float* rbm = rowbuf + mi * g.csf1 * d.ext1;
        // covering n are nhi, nhi-1, ... while the local column stays < ext1
        float sub = 0.f;
        for (int nn = nhi, lc = lc0; nn >= 0 && lc < d.ext1; --nn, lc += g.hop1) sub += rbm[nn * d.ext1 + lc];
        acc += sub;
      }
    }
    const size_t o = ((size_t)fr * d.O0 + i) * d.O1 + j;
    acc *= d.scale;
    if (d.sub) acc -= d.sub[o];
    if (d.mul) acc *= d.mul[o];
    if (d.axz) acc += d.beta * d.axz[o];
    if (d.clip_lo) acc = fmaxf(acc, d.lo);
    d.out[o] = acc;
  }
}

template <int TW, bool STREAM, bool WIN, int WGS>
__global__ __launch_bounds__(WGS, (TW >= 512 ? 3 : 1)) void k_rows_c2r_ola(Geo g, OlaDst d, const float2* __restrict__ tw) {
  extern __shared__ float2 dyn[];
  c2r_ola_body<TW, STREAM, WIN, WGS>(g, d, blockIdx.x, blockIdx.y, dyn, tw);
}

// Summed form of the overlap-add pass (class layout).  The inverse row
// transform is linear and the patches (m, n) of one patch column n land on
// the same output columns n*hop1 + c, so output row u is
//   sum_n IFFT_row( sum_m Y[m, n][u - m*hop0] )  at columns n*hop1 + c:
// one transform per (output row, pair of patch columns) instead of one per
// (covering patch row, pair), 2.4x fewer at 50 % overlap.  R = d.sumr output
// rows per workgroup share its TPW transforms (R * NP2 <= TPW).  A transform
// fetches its covering rows (the first three together, absent rows read the
// zero row) and adds them in ascending patch-row order; its two patches'
// rows go to their class rows (patch n to row n mod ncls of its output row),
// and each output column adds its ncls class values.
//
// WGT: separable windows nom(r, c) = a(r) b(c) over the normaliser wimg (the
// x-adjoint; olaGPU.py:305-309 applies ws = nom / wimg per patch): row m
// enters the sum weighted by a(r), the transform's column c by b(c), the
// output pixel by 1/wimg.  Where every covering window is at the floor (rows
// with zrow set, columns bcol) the reference's weights are 1/count: such a
// row takes unit weights (its normaliser is 1/count); such a column of the
// other rows is the unweighted sum of its covering patches' inverse
// transforms at that column, evaluated as one inverse DFT term per bin.
// Output rows i0 .. i0 + R - 1 of frame fr (dyn: ola_shared_bytes of LDS).
__host__ __device__ inline int ola_sum_wcol_offset_floats(const Geo& g, const OlaDst& d) {
  const int fl = (WG / (g.Tw / 16)) * (g.Tw + g.Tw / 16) * 2;
  const int rb = d.sumr * d.ncls * ola_class_len(g, d);
  return ((fl > rb ? fl : rb) + 3) & ~3;
}

template <int TW, bool WGT>
__device__ __forceinline__ void c2r_sum_body(Geo g, OlaDst d, int i0, int fr, float2* dyn,
                                             const float2* __restrict__ tw) {
  constexpr int TPT = TW / 16;
  constexpr int SL = FftLds<TW>::floats2;
  const int R = d.sumr;
  const int NP2 = (g.csf1 + 1) >> 1;
  const int NC = d.ncls;
  const int RLc = ola_class_len(g, d);
  float* rowbuf = (float*)dyn;  // class rows alias the FFT slices
  const int tr = threadIdx.x / TPT;
  const int t = threadIdx.x - tr * TPT;
  const int ri = tr / NP2;
  const int jp = tr - ri * NP2;
  const int i = i0 + ri;
  const bool act = ri < R && i < d.O0;
  const int u = i + d.oc0;
  // covering patch rows m: m*hop0 <= u < m*hop0 + ext0
  int mlo = 0, mhi = -1;
  if (act && u >= 0) {
    mhi = min(g.csf0 - 1, u / g.hop0);
    const int lo = u - d.ext0 + 1;
    mlo = lo <= 0 ? 0 : (lo + g.hop0 - 1) / g.hop0;
  }
  const int na = 2 * jp, nb = na + 1;
  const bool hasb = nb < g.csf1;
  const int kf = fr * g.Pf;
  float2* sl = dyn + tr * SL;
  // column weights staged in thread order (thread t weights columns
  // t + TPT*mm, kept at floats t*MW + mm): 16-byte LDS reads
  const int MW = ola_win_mw(g);
  float* wcs = (float*)dyn + ola_sum_wcol_offset_floats(g, d);
  if constexpr (WGT) {
    for (int e = threadIdx.x; e < TPT * MW; e += WG) {
      const int tt = e / MW;
      const int c = tt + TPT * (e - tt * MW);
      wcs[e] = c < g.sw1 ? d.wcol[c] : 0.f;
    }
  }
  // unit weights on a row whose covering windows are all at the floor
  const bool zr = WGT && act && d.zrow[i];
  // a wavefront whose transforms all lie past the workgroup's rows skips them
  const int trw = (int)(threadIdx.x >> 6) * (64 / TPT);
  const bool wact = trw < R * NP2 && i0 + trw / NP2 < d.O0;
  float2 v[16];
  float invA = 1.f;
  float s0 = 0.f;
  bool f0a = false;
  if (wact) {
    TwSeed<TW> S;  // twiddle loads in flight with the row fetches
    twiddles_seed<TW>(S, t, tw);
    // the first three covering patch rows (all of them at 50 % overlap) are
    // fetched together, any further ones one at a time; added in ascending
    // patch-row order, absent rows as zeros.  Returns the row's weight.
    int nok = 0;  // covering rows fetched
    auto fetch = [&](float4(&a)[4], float4(&b)[4], int m) {
      const int r = u - m * g.hop0;
      const bool ok = m <= mhi && r >= 0 && r < d.srow;
      const float2* ra = ok ? d.src + ((size_t)(kf + m * g.csf1 + na) * d.srow + r) * g.NBP : d.src;
      c2r_fetch<TW>(a, b, ra, ra + (size_t)d.srow * g.NBP, ok, ok && hasb, t);
      nok += ok;
      return ok ? ((WGT && !zr) ? d.wrow[r] : 1.f) : 0.f;
    };
    auto add = [&](float4(&A)[4], float4(&B)[4], const float4(&a)[4], const float4(&b)[4], float w) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if constexpr (WGT) {
          A[q].x = fmaf(w, a[q].x, A[q].x); A[q].y = fmaf(w, a[q].y, A[q].y);
          A[q].z = fmaf(w, a[q].z, A[q].z); A[q].w = fmaf(w, a[q].w, A[q].w);
          B[q].x = fmaf(w, b[q].x, B[q].x); B[q].y = fmaf(w, b[q].y, B[q].y);
          B[q].z = fmaf(w, b[q].z, B[q].z); B[q].w = fmaf(w, b[q].w, B[q].w);
        } else {
          A[q].x += a[q].x; A[q].y += a[q].y; A[q].z += a[q].z; A[q].w += a[q].w;
          B[q].x += b[q].x; B[q].y += b[q].y; B[q].z += b[q].z; B[q].w += b[q].w;
        }
      }
    };
    // the weighted form (the x-adjoint: two covering patch rows at 50 %
    // overlap) fetches two rows together, which keeps it within 128 VGPRs
    constexpr int NPEEL = WGT ? 2 : 3;
    // a patch of this pair whose local column 0 is a floor column of the
    // canvas (left edge): its unweighted inverse transforms at column 0,
    // X[0] + X[Tw/2] + 2 sum_{0<k<Tw/2} Re X[k], summed over the covering
    // rows from the fetched bins
    f0a = WGT && !zr && ((d.bcol[0] >= 0 && d.bc[0] == 0 && d.bn[0] == na) ||
                         (d.bcol[1] >= 0 && d.bc[1] == 0 && d.bn[1] == na));
    auto edge0 = [&](const float4(&a)[4]) {
#pragma unroll
      for (int q = 0; q < 4; ++q)
        s0 += (t == 0 && q == 0) ? (a[q].x + a[q].y) + 2.f * a[q].z : 2.f * (a[q].x + a[q].z);
    };
    float4 A[4], B[4], a1[4], b1[4], a2[4], b2[4];
    const float w0 = fetch(A, B, mlo);
    const float w1 = fetch(a1, b1, mlo + 1);
    float w2 = 0.f;
    if (NPEEL == 3 && d.mc >= 3) w2 = fetch(a2, b2, mlo + 2);
    if constexpr (WGT) {
      if (f0a) {
        edge0(A);
        edge0(a1);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        A[q].x *= w0; A[q].y *= w0; A[q].z *= w0; A[q].w *= w0;
        B[q].x *= w0; B[q].y *= w0; B[q].z *= w0; B[q].w *= w0;
      }
    }
    add(A, B, a1, b1, w1);
    if (NPEEL == 3 && d.mc >= 3) add(A, B, a2, b2, w2);
    float wsum = w0 + w1 + w2;
    for (int m0 = NPEEL; m0 < d.mc; ++m0) {
      const float wk = fetch(a1, b1, mlo + m0);
      if (WGT && f0a) edge0(a1);
      add(A, B, a1, b1, wk);
      wsum += wk;
    }
    if constexpr (WGT) {
      // 1/(sum of the row weights): the row part of 1/wimg (1/count on a
      // floor row, whose weights are 1)
      invA = wsum > 0.f ? 1.f / wsum : 0.f;
      if (f0a) {  // the reference weight there: 1/count, one covering patch column
#pragma unroll
        for (int off = 1; off < TPT; off <<= 1) s0 += __shfl_xor(s0, off, TPT);
        s0 = nok ? s0 / (float)nok : 0.f;
      }
    }
    Twiddles<TW> T;
    twiddles_from_seed<TW>(T, S);
    c2r_stage<TW>(v, A, B, sl, t);
    fft<TW, +1>(v, sl, t, T, tw);
  }
  __syncthreads();  // the slices become class rows; column weights staged
  {
    float4* z4 = (float4*)rowbuf;
    const int n4 = (R * NC * RLc) >> 2;
    for (int e = threadIdx.x; e < n4; e += WG) z4[e] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  __syncthreads();
  if (wact && act) {
    const int qa = na % NC;
    const int qb = qa + 1 == NC ? 0 : qa + 1;
    float* ba = rowbuf + (ri * NC + qa) * RLc + na * g.hop1;
    float* bb = rowbuf + (ri * NC + qb) * RLc + nb * g.hop1;
    const float* wct = wcs + t * MW;
    float4 qw = make_float4(1.f, 1.f, 1.f, 1.f);
#pragma unroll
    for (int mm = 0; mm < 16; ++mm) {
      const int col = t + TPT * mm;
      if (col < d.ext1) {
        float xa = v[mm].x, xb = v[mm].y;
        if constexpr (WGT) {
          if ((mm & 3) == 0 && !zr) qw = *(const float4*)(wct + mm);
          const float wc = ((mm & 3) == 0 ? qw.x : (mm & 3) == 1 ? qw.y : (mm & 3) == 2 ? qw.z : qw.w) * invA;
          xa *= wc;
          xb *= wc;
        }
        ba[col] = xa;
        if (hasb) bb[col] = xb;
      }
    }
    if (WGT && f0a && t == 0) wcs[TPT * MW + ri] = s0;
  }
  __syncthreads();
  // four output columns per thread and step: their class reads are issued
  // together
  for (int rr = 0; rr < R && i0 + rr < d.O0; ++rr) {
    const float* cr = rowbuf + rr * NC * RLc;
    const int irow = i0 + rr;
    const size_t orow = ((size_t)fr * d.O0 + irow) * d.O1;
    // floor columns of a row with real weights: the left one from its slot,
    // others evaluated apart below
    const bool fix = WGT && !d.zrow[irow];
    const float* cw = WGT ? d.colw + (fix ? 0 : d.O1) : nullptr;
    for (int j0 = threadIdx.x; j0 < d.O1; j0 += 4 * WG) {
      float acc[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int vcol = j0 + e * WG + d.oc1;
        acc[e] = 0.f;
        if (j0 + e * WG < d.O1 && vcol >= 0 && vcol < RLc)
          for (int q = 0; q < NC; ++q) acc[e] += cr[q * RLc + vcol];
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int j = j0 + e * WG;
        const int vcol = j + d.oc1;
        // (selects, not indexing: a runtime index would put d in scratch)
        const int eb = vcol == d.bcol[0] ? 0 : vcol == d.bcol[1] ? 1 : -1;
        const int bce = eb == 0 ? d.bc[0] : d.bc[1];
        if (j < d.O1 && !(fix && eb >= 0 && bce != 0)) {
          const size_t o = orow + j;
          float a;
          if (WGT && fix && eb >= 0)
            a = wcs[TPT * MW + rr] * d.scale;  // left floor column: mean of the covering patches
          else if constexpr (WGT)
            a = acc[e] * d.scale * cw[j];
          else
            a = acc[e] * d.scale;
          if (d.sub) a -= d.sub[o];
          if (d.mul) a *= d.mul[o];
          if (d.axz) a += d.beta * d.axz[o];
          if (d.clip_lo) a = fmaxf(a, d.lo);
          d.out[o] = a;
        }
      }
    }
  }
  if constexpr (WGT) {
    // floor columns of the rows with real weights: sum over the covering
    // patch rows m of patch column bn's inverse transform at column bc,
    // x(c) = X[0] + (-1)^c X[Tw/2] + 2 sum_{0<k<Tw/2} Re(X[k] e^{2 pi i k c / Tw}),
    // one wavefront per (row, column), bins over the lanes, a fixed-order
    // butterfly reduction
    const int lane = threadIdx.x & 63;
    for (int jb = (int)(threadIdx.x >> 6); jb < 2 * R; jb += WG / 64) {
      const int rr = jb >> 1, e = jb & 1;
      const int irow = i0 + rr;
      const int bcol = e ? d.bcol[1] : d.bcol[0], bn = e ? d.bn[1] : d.bn[0], bc = e ? d.bc[1] : d.bc[0];
      if (irow >= d.O0 || bcol < 0 || bc == 0 || d.zrow[irow]) continue;
      const int j = bcol - d.oc1;
      if (j < 0 || j >= d.O1) continue;
      const int uu = irow + d.oc0;
      const int mh = min(g.csf0 - 1, uu / g.hop0);
      const int l0 = uu - d.ext0 + 1;
      const int ml = l0 <= 0 ? 0 : (l0 + g.hop0 - 1) / g.hop0;
      const int c = bc;
      float sacc = 0.f;
      int cnt = 0;
      for (int m = ml; m <= mh; ++m) {
        const int r = uu - m * g.hop0;
        if (r < 0 || r >= d.srow) continue;
        ++cnt;
        const float2* row = d.src + ((size_t)(kf + m * g.csf1 + bn) * d.srow + r) * g.NBP;
        for (int k = lane; k < TW / 2; k += 64) {
          const float2 X = row[k];
          if (k == 0) {
            sacc += X.x + ((c & 1) ? -X.y : X.y);
          } else {
            const float2 w = tw[(k * c * (TW_LEN / TW)) & (TW_LEN - 1)];
            sacc += 2.f * fmaf(X.x, w.x, X.y * w.y);
          }
        }
      }
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) sacc += __shfl_xor(sacc, off, 64);
      if (lane == 0) {
        const size_t o = ((size_t)fr * d.O0 + irow) * d.O1 + j;
        float a = cnt ? sacc * d.scale / (float)cnt : 0.f;
        if (d.sub) a -= d.sub[o];
        if (d.mul) a *= d.mul[o];
        if (d.axz) a += d.beta * d.axz[o];
        if (d.clip_lo) a = fmaxf(a, d.lo);
        d.out[o] = a;
      }
    }
  }
}

// four workgroups per CU from 64 to 256 points (the LDS of the 256-point
// form allows four): at most 128 VGPRs
template <int TW, bool WGT>
__global__ __launch_bounds__(WG, (TW >= 512 ? 3 : TW >= 64 ? 4 : 1)) void k_rows_c2r_sum(Geo g, OlaDst d,
                                                                         const float2* __restrict__ tw) {
  extern __shared__ float2 dyn[];
  c2r_sum_body<TW, WGT>(g, d, (int)blockIdx.x * d.sumr, blockIdx.y, dyn, tw);
}

// float offset of the staged windows: past the FFT slices and the gather
// buffer (which alias each other in single-round launches), rounded up to
// 16 bytes for the windows' 16-byte loads
__host__ __device__ int ola_window_offset_floats(const Geo& g, const OlaDst& d) {
  const int TPT = g.Tw / 16;
  const int TPW = WG / TPT;  // launches with fewer threads are single-round: TPW unused
  const int nsl = d.alias ? d.mc * ((g.csf1 + 1) / 2) : TPW;  // FFT slices in use
  const int fl = nsl * (g.Tw + g.Tw / 16) * 2;
  const int rb = d.stream ? 2 * TPW * d.ext1 : d.ncls ? d.mc * d.ncls * ola_class_len(g, d) : d.mc * g.csf1 * d.ext1;
  const int off = (d.alias || d.stream) ? (fl > rb ? fl : rb) : fl + rb;
  return (off + 3) & ~3;
}

size_t ola_rowbuf_bytes(const Geo& g, const OlaDst& d) {
  const int TPW = WG / (g.Tw / 16);
  if (d.sumr) return (size_t)d.sumr * d.ncls * ola_class_len(g, d) * sizeof(float);
  if (!d.stream && d.ncls) return (size_t)d.mc * d.ncls * ola_class_len(g, d) * sizeof(float);
  return (size_t)(d.stream ? 2 * TPW : d.mc * g.csf1) * d.ext1 * sizeof(float);
}

size_t ola_shared_bytes(const Geo& g, const OlaDst& d) {
  if (d.sumr)  // all slices of a workgroup, aliased by the class rows; staged column
               // weights and one left-floor value per output row
    return (size_t)ola_sum_wcol_offset_floats(g, d) * sizeof(float) +
           (d.wcol ? ((size_t)ola_win_rw(g) + d.sumr) * sizeof(float) : 0);
  // staged window rows + slot table
  const size_t wl = d.win ? ((size_t)d.mc * d.nslot * ola_win_rw(g) + (size_t)d.mc * g.csf1) * sizeof(float) : 0;
  const size_t acc = d.stream ? (size_t)d.O1 * sizeof(float) : 0;
  return (size_t)ola_window_offset_floats(g, d) * sizeof(float) + wl + acc;
}

int ola_single_round(const Geo& g, const OlaDst& d) {
  const int TPW = WG / (g.Tw / 16);
  return d.mc * ((g.csf1 + 1) / 2) <= TPW;
}

int ola_stream_rounds(const Geo& g, const OlaDst& d) { return !ola_single_round(g, d) && d.O1 <= OLA_KJ * WG; }

// Single-round launches gather from class rows (ncls = ceil(ext1/hop1)
// classes, at most csf1; patches n and n + ncls of a class do not overlap):
// an output column then reads one column of each class row with fixed
// strides instead of searching the covering patches.  The buffer is larger
// than one row per patch when ncls > 2, so the layout is taken only if it
// keeps the resident wavefronts per CU for every workgroup size the launch
// may use.  Those are bounded by the LDS and by the kernels' registers
// (about 104-112 VGPRs at Tw <= 256, four wavefronts per SIMD; three at 512).
void ola_layout(const Geo& g, OlaDst& d) {
  d.alias = ola_single_round(g, d);
  d.stream = ola_stream_rounds(g, d);
  d.ncls = 0;
  d.sumr = 0;
  if (g.hop1 <= 0) return;
  if (!d.alias) {
    // the per-patch transforms of a row take several rounds; the summed form
    // needs one transform per pair of patch columns, however many patch
    // rows cover the row, so it runs in one round whenever those fit a
    // workgroup (and, with windows, they factor)
    const int NP2 = (g.csf1 + 1) / 2;
    const int TPW = WG / (g.Tw / 16);
    if ((d.win && !d.wrow) || NP2 > TPW) return;
    OlaDst f = d;
    f.ncls = std::min(g.csf1, (d.ext1 + g.hop1 - 1) / g.hop1);
    for (int R = TPW / NP2; R >= 1; --R) {
      f.sumr = R;
      if (ola_shared_bytes(g, f) <= (size_t)160 * 1024) {
        d.ncls = f.ncls;
        d.sumr = R;
        return;
      }
    }
    return;
  }
  OlaDst e = d;
  e.ncls = std::min(g.csf1, (d.ext1 + g.hop1 - 1) / g.hop1);
  const size_t b0 = ola_shared_bytes(g, d), b1 = ola_shared_bytes(g, e);
  const int cap = g.Tw >= 512 ? 12 : 16;  // wavefronts per CU the registers allow
  auto waves = [cap](size_t b, int wpg) {
    const int by_lds = b == 0 ? 64 : (int)((size_t)160 * 1024 / b);
    return std::min(cap / wpg, by_lds) * wpg;
  };
  for (int wgs = 128; wgs <= WG; wgs += 64)
    if ((wgs == WG || ola_fits(g, d, wgs)) && waves(b1, wgs / 64) < waves(b0, wgs / 64)) return;
  d.ncls = e.ncls;
  // without a window the covering patch rows can be summed before the
  // transform: as many output rows per workgroup as its transforms hold,
  // fewer if their class rows would cost resident wavefronts
  if (d.win && !d.wrow) return;
  const int NP2 = (g.csf1 + 1) / 2;
  for (int R = (WG / (g.Tw / 16)) / NP2; R >= 1; --R) {
    OlaDst f = d;
    f.sumr = R;
    if (waves(ola_shared_bytes(g, f), WG / 64) >= waves(b1, WG / 64)) {
      d.sumr = R;
      return;
    }
  }
}

// threads per workgroup: a single-round launch whose grid fills the GPU
// many times over takes 128, 192 or 256 threads, the first that covers its
// transforms (the x-adjoint at Tw 256: 10 transforms, 192 threads, five
// workgroups per CU instead of four).  Smaller grids keep WG: there the
// latency of one output row, not the number of resident rows, sets the time.
int ola_wg(const Geo& g, const OlaDst& d, long long rows) {
  if (!d.alias || rows < 8 * 1024) return WG;
  const int need = (g.Tw / 16) * d.mc * ((g.csf1 + 1) / 2);
  return need <= 128 ? 128 : need <= 192 ? 192 : WG;
}

// streamed launches always take WG threads (ola_wg)
template <int TW, int WGS>
static void c2r_ola_launch(const Geo& g, const OlaDst& d, const float2* tw, hipStream_t st, dim3 grid, size_t shm) {
  if constexpr (WGS == WG) {
    if (d.stream) {
      if (d.win)
        hipLaunchKernelGGL((k_rows_c2r_ola<TW, true, true, WG>), grid, dim3(WG), shm, st, g, d, tw);
      else
        hipLaunchKernelGGL((k_rows_c2r_ola<TW, true, false, WG>), grid, dim3(WG), shm, st, g, d, tw);
      return;
    }
  }
  if (d.win)
    hipLaunchKernelGGL((k_rows_c2r_ola<TW, false, true, WGS>), grid, dim3(WGS), shm, st, g, d, tw);
  else
    hipLaunchKernelGGL((k_rows_c2r_ola<TW, false, false, WGS>), grid, dim3(WGS), shm, st, g, d, tw);
}

// Output rows per workgroup of a summed-form launch of `rows` output rows
// (all frames): three at 32 frames of 1024 rows, fewer when the grid would
// not fill the GPU several times over (one 1024-row frame: one row per
// workgroup, 1024 workgroups).
void ola_fit_rows(OlaDst& d, long long rows) {
  if (d.sumr > 1) d.sumr = (int)std::min<long long>(d.sumr, std::max<long long>(1, rows / 2048));
}

int ola_fits(const Geo& g, const OlaDst& d, int wgs) {
  return d.alias && (g.Tw / 16) * d.mc * ((g.csf1 + 1) / 2) <= wgs;
}

template <int TW>
static hipError_t c2r_ola_t(const Geo& g, const OlaDst& d0, const float2* tw, hipStream_t st, int wgs) {
  const int frames = g.P / g.Pf;
  OlaDst d = d0;
  ola_fit_rows(d, (long long)d.O0 * frames);
  const size_t shm = ola_shared_bytes(g, d);
  ProfScope ps(d.win ? "rows_c2r_ola_win" : "rows_c2r_ola", st);
  if (d.sumr) {  // summed form: 256 threads, sumr output rows per workgroup
    const dim3 gs((unsigned)((d.O0 + d.sumr - 1) / d.sumr), (unsigned)frames);
    if (d.wrow)
      hipLaunchKernelGGL((k_rows_c2r_sum<TW, true>), gs, dim3(WG), shm, st, g, d, tw);
    else
      hipLaunchKernelGGL((k_rows_c2r_sum<TW, false>), gs, dim3(WG), shm, st, g, d, tw);
    return hipGetLastError();
  }
  const dim3 grid((unsigned)d.O0, (unsigned)frames);
  if (wgs == 0) wgs = ola_wg(g, d, (long long)d.O0 * frames);
  if (wgs != WG && !ola_fits(g, d, wgs)) return hipErrorInvalidValue;
  switch (wgs) {
    case 128: c2r_ola_launch<TW, 128>(g, d, tw, st, grid, shm); break;
    case 192: c2r_ola_launch<TW, 192>(g, d, tw, st, grid, shm); break;
    default: c2r_ola_launch<TW, WG>(g, d, tw, st, grid, shm); break;
  }
  return hipGetLastError();
}

hipError_t launch_rows_c2r_ola(const Geo& g, const OlaDst& d, const float2* tw, hipStream_t st, int wgs) {
  if (d.O0 <= 0 || d.O1 <= 0) return hipSuccess;
  if (g.NBP < g.Tw / 2 + 2) return hipErrorInvalidValue;  // two-bin loads need a padding bin per row
  if (wgs != 0 && wgs != 128 && wgs != 192 && wgs != WG) return hipErrorInvalidValue;
  switch (g.Tw) {
    case 16: return c2r_ola_t<16>(g, d, tw, st, wgs);
    case 32: return c2r_ola_t<32>(g, d, tw, st, wgs);
    case 64: return c2r_ola_t<64>(g, d, tw, st, wgs);
    case 128: return c2r_ola_t<128>(g, d, tw, st, wgs);
    case 256: return c2r_ola_t<256>(g, d, tw, st, wgs);
    case 512: return c2r_ola_t<512>(g, d, tw, st, wgs);
    default: return hipErrorInvalidValue;
  }
}

// ----------------------------------------------------------------------------
// rows_c2r_stack: (P, nrows, ncols) stack from the first rows/cols of each tile
// ----------------------------------------------------------------------------
template <int TW>
struct StackLds {
  static constexpr int v = (WG / (TW / 16)) * FftLds<TW>::floats2;
};

// One workgroup of the stack pass: row block by of patch p (fl:
// StackLds<TW>::v float2 of LDS).
template <int TW>
__device__ __forceinline__ void c2r_stack_body(Geo g, const float2* __restrict__ src, int srow,
                                               float* __restrict__ out, int nrows, int ncols, float scale,
                                               const float2* __restrict__ tw, const float* __restrict__ axz, float beta,
                                               int p, int by, float2* fl) {
  constexpr int TPT = TW / 16;
  constexpr int TPW = WG / TPT;
  constexpr int SL = FftLds<TW>::floats2;
  const int tr = threadIdx.x / TPT;
  const int t = threadIdx.x - tr * TPT;
  const int NQ = (nrows + 1) >> 1;
  const int q = by * TPW + tr;
  const bool act = q < NQ;
  // no workgroup barrier follows: a wavefront without row pairs is done
  if (by * TPW + (int)(threadIdx.x >> 6) * (64 / TPT) >= NQ) return;
  const int r0 = 2 * q, r1 = 2 * q + 1;
  const float2* ra = act ? src + ((size_t)p * srow + r0) * g.NBP : src;
  const bool hasb = act && r1 < nrows;
  const float2* rb = hasb ? ra + g.NBP : ra;
  Twiddles<TW> T;
  twiddles_init<TW>(T, t, tw);
  float2 v[16];
  float2* sl = fl + tr * SL;
  c2r_load<TW>(v, ra, rb, act, hasb, sl, t);
  fft<TW, +1>(v, sl, t, T, tw);
  if (act) {
#pragma unroll
    for (int mm = 0; mm < 16; ++mm) {
      const int col = t + TPT * mm;
      if (col < ncols) {
        const size_t o0 = ((size_t)p * nrows + r0) * ncols + col;
        const float v0 = v[mm].x * scale, v1 = v[mm].y * scale;
        out[o0] = axz ? v0 + beta * axz[o0] : v0;
        if (r1 < nrows) {
          const size_t o1 = o0 + ncols;
          out[o1] = axz ? v1 + beta * axz[o1] : v1;
        }
      }
    }
  }
}

template <int TW>
__global__ __launch_bounds__(WG) void k_rows_c2r_stack(Geo g, const float2* __restrict__ src, int srow,
                                                       float* __restrict__ out, int nrows, int ncols, float scale,
                                                       const float2* __restrict__ tw, const float* __restrict__ axz,
                                                       float beta) {
  __shared__ float2 fl[StackLds<TW>::v];
  c2r_stack_body<TW>(g, src, srow, out, nrows, ncols, scale, tw, axz, beta, blockIdx.x, blockIdx.y, fl);
}

template <int TW>
static hipError_t c2r_stack_t(const Geo& g, const float2* src, int srow, float* out, int nrows, int ncols,
                              float scale, const float2* tw, hipStream_t st, const float* axz, float beta) {
  constexpr int TPW = WG / (TW / 16);
  const int blocks = ((nrows + 1) / 2 + TPW - 1) / TPW;
  if (blocks <= 0 || g.P <= 0) return hipSuccess;
  ProfScope ps("rows_c2r_stack", st);
  hipLaunchKernelGGL(k_rows_c2r_stack<TW>, dim3((unsigned)g.P, (unsigned)blocks), dim3(WG), 0, st, g, src, srow,
                     out, nrows, ncols, scale, tw, axz, beta);
  return hipGetLastError();
}

hipError_t launch_rows_c2r_stack(const Geo& g, const float2* src, int srow, float* out, int nrows, int ncols,
                                 float scale, const float2* tw, hipStream_t st, const float* axz, float beta) {
  if (g.NBP < g.Tw / 2 + 2) return hipErrorInvalidValue;  // two-bin loads need a padding bin per row
  switch (g.Tw) {
    case 16: return c2r_stack_t<16>(g, src, srow, out, nrows, ncols, scale, tw, st, axz, beta);
    case 32: return c2r_stack_t<32>(g, src, srow, out, nrows, ncols, scale, tw, st, axz, beta);
    case 64: return c2r_stack_t<64>(g, src, srow, out, nrows, ncols, scale, tw, st, axz, beta);
    case 128: return c2r_stack_t<128>(g, src, srow, out, nrows, ncols, scale, tw, st, axz, beta);
    case 256: return c2r_stack_t<256>(g, src, srow, out, nrows, ncols, scale, tw, st, axz, beta);
    case 512: return c2r_stack_t<512>(g, src, srow, out, nrows, ncols, scale, tw, st, axz, beta);
    default: return hipErrorInvalidValue;
  }
}

// ----------------------------------------------------------------------------
// batches: one launch per pass over the groups of a table
// ----------------------------------------------------------------------------
// Each kernel finds its group from the launch's block ranges (a scalar scan
// of at most BATCH_MAX_GROUPS starts), takes that group's parameters from the
// table in device memory and runs the same workgroup body as the single-plan
// kernel with the block index local to the group.  Blocks past a group's
// count (its range is rounded up to a multiple of 8) exit at once.
__device__ __forceinline__ int batch_find(const BatchLaunch& L, int b) {
  int gi = 0;
  for (int j = 1; j < L.ng; ++j)
    if (b >= L.start[j]) gi = j;
  return gi;
}

template <int TW>
__global__ __launch_bounds__(WG) void k_batch_rows(const BatchGroup* __restrict__ G, BatchLaunch L, int res) {
  __shared__ float2 sm[RowsLds<TW>::v];
  const int gi = batch_find(L, (int)blockIdx.x);
  const int lb = (int)blockIdx.x - L.start[gi];
  const BatchGroup& B = G[L.gidx[gi]];
  const Geo& g = res == 1 ? B.gres : res == 2 ? B.gx : B.g;
  const RowJobs& jobs = res == 1 ? B.rows_res : res == 2 ? B.rows_x : B.rows_fwd;
  if (lb >= g.P * jobs.blk_end[jobs.n - 1]) return;
  int pu, ji, q0;
  rows_unit<TW>(g, jobs, lb, pu, ji, q0);
  const RowSrc s = jobs.s[ji];
  rows_r2c_body<TW>(g, s, jobs.out[ji], pu, q0, sm, B.tw);
}

template <int TH, int OP>
__global__ __launch_bounds__(WG, (TH >= 256 ? 4 : 1)) void k_batch_cols(const BatchGroup* __restrict__ G,
                                                                        BatchLaunch L) {
  constexpr int CS = WG / (TH / 16);
  __shared__ float2 sm[ColsLds<TH, OP>::v];
  const int gi = batch_find(L, (int)blockIdx.x);
  const int lb = (int)blockIdx.x - L.start[gi];
  const BatchGroup& B = G[L.gidx[gi]];
  const int strips = (B.g.NB - 1 + CS - 1) / CS;
  const int p = lb / strips;
  if (p >= B.g.P) return;
  if constexpr (OP == COL_FWD)
    cols_body<TH, COL_FWD>(B.g, B.cf_a, B.cf_b, B.cf_b, B.cf_o, B.cf_o, nullptr, lb - p * strips, p, sm, B.tw);
  else
    cols_body<TH, COL_ADJ2>(B.g, B.ca_a, B.ca_b, B.ca_c, B.ca_o1, B.ca_o2, nullptr, lb - p * strips, p, sm, B.tw);
}

// WGS < WG: every group of the launch holds its row in one round of WGS
// threads (batch_ola_threads), so only the single-round body is needed
template <int TW, bool WIN, int WGS>
__global__ __launch_bounds__(WGS, (TW >= 512 ? 3 : 1)) void k_batch_ola(const BatchGroup* __restrict__ G,
                                                                        BatchLaunch L) {
  extern __shared__ float2 dyn[];
  const int gi = batch_find(L, (int)blockIdx.x);
  const int lb = (int)blockIdx.x - L.start[gi];
  const BatchGroup& B = G[L.gidx[gi]];
  const OlaDst& d = WIN ? B.ola_adj : B.ola_fwd;
  if constexpr (WGS == WG) {
    if (d.sumr) {  // summed form: sumr output rows per workgroup
      const int nbr = (d.O0 + d.sumr - 1) / d.sumr;
      const int fs = lb / nbr;
      if (fs >= B.g.P / B.g.Pf) return;
      if (d.wrow)
        c2r_sum_body<TW, true>(B.g, d, (lb - fs * nbr) * d.sumr, fs, dyn, B.tw);
      else
        c2r_sum_body<TW, false>(B.g, d, (lb - fs * nbr) * d.sumr, fs, dyn, B.tw);
      return;
    }
  }
  const int fr = lb / d.O0;
  if (fr >= B.g.P / B.g.Pf) return;
  if constexpr (WGS == WG) {
    if (d.stream) {
      c2r_ola_body<TW, true, WIN, WG>(B.g, d, lb - fr * d.O0, fr, dyn, B.tw);
      return;
    }
  }
  c2r_ola_body<TW, false, WIN, WGS>(B.g, d, lb - fr * d.O0, fr, dyn, B.tw);
}

template <int TW>
__global__ __launch_bounds__(WG) void k_batch_stack(const BatchGroup* __restrict__ G, BatchLaunch L) {
  constexpr int TPW = WG / (TW / 16);
  __shared__ float2 fl[StackLds<TW>::v];
  const int gi = batch_find(L, (int)blockIdx.x);
  const int lb = (int)blockIdx.x - L.start[gi];
  const BatchGroup& B = G[L.gidx[gi]];
  const int by = lb / B.g.P;
  if (by * TPW >= (B.st_nrows + 1) / 2) return;
  c2r_stack_body<TW>(B.g, B.st_src, B.st_srow, B.st_out, B.st_nrows, B.st_ncols, B.st_scale, B.tw, nullptr, 0.f,
                     lb - by * B.g.P, by, fl);
}

int batch_tile(const BatchGroup& G, int pass) {
  return (pass == BP_COLS_FWD || pass == BP_COLS_ADJ) ? G.g.Th : G.g.Tw;
}

int batch_blocks(const BatchGroup& G, int pass) {
  const int TPW = WG / (G.g.Tw / 16);
  switch (pass) {
    case BP_ROWS_FWD: return G.g.P * G.rows_fwd.blk_end[G.rows_fwd.n - 1];
    case BP_ROWS_RES: return G.gres.P * G.rows_res.blk_end[G.rows_res.n - 1];
    case BP_ROWS_X: return G.rows_x.n ? G.gx.P * G.rows_x.blk_end[G.rows_x.n - 1] : 0;
    case BP_COLS_FWD:
    case BP_COLS_ADJ: {
      const int CS = WG / (G.g.Th / 16);
      return G.g.P * ((G.g.NB - 1 + CS - 1) / CS);
    }
    case BP_OLA_FWD:
      return (G.ola_fwd.sumr ? (G.ola_fwd.O0 + G.ola_fwd.sumr - 1) / G.ola_fwd.sumr : G.ola_fwd.O0) *
             (G.g.P / G.g.Pf);
    case BP_OLA_ADJ:
      return (G.ola_adj.sumr ? (G.ola_adj.O0 + G.ola_adj.sumr - 1) / G.ola_adj.sumr : G.ola_adj.O0) *
             (G.g.P / G.g.Pf);
    case BP_STACK: return G.g.P * (((G.st_nrows + 1) / 2 + TPW - 1) / TPW);
    default: return 0;
  }
}

size_t batch_shared_bytes(const BatchGroup& G, int pass) {
  if (pass == BP_OLA_FWD) return ola_shared_bytes(G.g, G.ola_fwd);
  if (pass == BP_OLA_ADJ) return ola_shared_bytes(G.g, G.ola_adj);
  return 0;
}

int batch_ola_threads(const BatchGroup* H, const BatchLaunch& L, int pass) {
  int need = 0;
  for (int i = 0; i < L.ng; ++i) {
    const BatchGroup& B = H[L.gidx[i]];
    const OlaDst& d = pass == BP_OLA_ADJ ? B.ola_adj : B.ola_fwd;
    if (!d.alias || d.sumr) return WG;
    need = std::max(need, (B.g.Tw / 16) * d.mc * ((B.g.csf1 + 1) / 2));
  }
  return need <= 128 ? 128 : need <= 192 ? 192 : WG;
}

template <int T, bool WIN>
static void batch_ola_launch(const BatchGroup* G, const BatchLaunch& L, int wgs, size_t shm, hipStream_t st,
                             dim3 grid) {
  switch (wgs) {
    case 128: hipLaunchKernelGGL((k_batch_ola<T, WIN, 128>), grid, dim3(128), shm, st, G, L); break;
    case 192: hipLaunchKernelGGL((k_batch_ola<T, WIN, 192>), grid, dim3(192), shm, st, G, L); break;
    default: hipLaunchKernelGGL((k_batch_ola<T, WIN, WG>), grid, dim3(WG), shm, st, G, L); break;
  }
}

template <int T>
static hipError_t batch_t(const BatchGroup* G, const BatchLaunch& L, int pass, int wgs, size_t shm,
                          hipStream_t st) {
  const dim3 grid((unsigned)L.start[L.ng]);
  if (grid.x == 0) return hipSuccess;
  static const char* names[BP_N] = {"batch_rows_fwd", "batch_rows_x", "batch_cols_fwd", "batch_ola_fwd", "batch_rows_res",
                                    "batch_cols_adj2", "batch_ola_win", "batch_stack"};
  ProfScope ps(names[pass], st);
  switch (pass) {
    case BP_ROWS_FWD: hipLaunchKernelGGL(k_batch_rows<T>, grid, dim3(WG), 0, st, G, L, 0); break;
    case BP_ROWS_RES: hipLaunchKernelGGL(k_batch_rows<T>, grid, dim3(WG), 0, st, G, L, 1); break;
    case BP_ROWS_X: hipLaunchKernelGGL(k_batch_rows<T>, grid, dim3(WG), 0, st, G, L, 2); break;
    case BP_COLS_FWD: hipLaunchKernelGGL((k_batch_cols<T, COL_FWD>), grid, dim3(WG), 0, st, G, L); break;
    case BP_COLS_ADJ: hipLaunchKernelGGL((k_batch_cols<T, COL_ADJ2>), grid, dim3(WG), 0, st, G, L); break;
    case BP_OLA_FWD: batch_ola_launch<T, false>(G, L, wgs, shm, st, grid); break;
    case BP_OLA_ADJ: batch_ola_launch<T, true>(G, L, wgs, shm, st, grid); break;
    case BP_STACK: hipLaunchKernelGGL(k_batch_stack<T>, grid, dim3(WG), 0, st, G, L); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_batch(const BatchGroup* dev_groups, const BatchLaunch& L, int pass, int tile, int wgs, size_t shm,
                        hipStream_t st) {
  switch (tile) {
    case 16: return batch_t<16>(dev_groups, L, pass, wgs, shm, st);
    case 32: return batch_t<32>(dev_groups, L, pass, wgs, shm, st);
    case 64: return batch_t<64>(dev_groups, L, pass, wgs, shm, st);
    case 128: return batch_t<128>(dev_groups, L, pass, wgs, shm, st);
    case 256: return batch_t<256>(dev_groups, L, pass, wgs, shm, st);
    case 512: return batch_t<512>(dev_groups, L, pass, wgs, shm, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace vmbd
