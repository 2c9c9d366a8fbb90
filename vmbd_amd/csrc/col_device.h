// Device helpers shared by the column-pass kernels (ola_kernels.hip,
// deconv_kernels.hip): workgroup size, the column loads of row-pass output or
// cached spectra, and the staged row-major store of a strip's columns.
#pragma once
#include <hip/hip_runtime.h>

#include "fft_core.h"
#include "ola_kernels.h"

namespace vmbd {

constexpr int WG = 256;
constexpr int NXCD = 8;  // XCDs of an MI355X; workgroups are dealt to them round-robin

template <int A, int B>
struct cmax {
  static constexpr int v = A > B ? A : B;
};

// 16-byte global load of 4 consecutive floats at a 4-byte-aligned address
// (gfx950 vector memory accepts unaligned addresses)
__device__ __forceinline__ void ld4(float (&d)[4], const float* p) { __builtin_memcpy(d, p, 16); }

// NZ < 16: only slots r < NZ can hold stored rows (rlo + nr <= NZ*TPT); the
// rest are zero without a load.  XS: the operand may be shared image rows
// with patch-row factors (in.rsc); only those call sites carry that path.
template <int TH, int NZ = 16, bool XS = false>
__device__ __forceinline__ void col_load(float2 (&v)[16], const ColIn& in, int p, int c, int NB, int t, bool active,
                                         int grp, int row0, const ColCtx& cx) {
  constexpr int TPT = TH / 16;
  if constexpr (NZ < 16) {
#pragma unroll
    for (int r = NZ; r < 16; ++r) v[r] = make_float2(0.f, 0.f);
  }
  if (in.spec) {
    const float2* col = in.p + ((size_t)p * NB + c) * TH;
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = active ? col[t + TPT * r] : make_float2(0.f, 0.f);
  } else if (XS && in.rsc) {
    // shared image rows, each scaled by its patch row's factor.  The factors
    // are loaded unconditionally (index clamped into the table, so no load
    // waits behind a branch) and applied once the column's loads are in.
    const float2* col = in.p + ((size_t)grp * NB + c) * in.pitch + row0;
    const float* rs = in.rsc + (size_t)cx.pm * in.nr;
    float sc[NZ];
#pragma unroll
    for (int r = 0; r < NZ; ++r) {
      const int rr = t + TPT * r - in.rlo;
      sc[r] = rs[min(max(rr, 0), in.nr - 1)];
    }
#pragma unroll
    for (int r = 0; r < NZ; ++r) {
      const int rr = t + TPT * r - in.rlo;
      v[r] = (active && rr >= 0 && rr < in.nr) ? col[rr] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int r = 0; r < NZ; ++r) {
      v[r].x *= sc[r];
      v[r].y *= sc[r];
    }
    // floor columns at the canvas edge: the pixel's own term, at its row's
    // weight (no patch-row factor), as bin c of a unit impulse at column ec:
    // e^{-2 pi i c ec / Tw}; the packed column holds bins 0 and Tw/2
    // (factors 1 and (-1)^ec)
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int ecol = e ? in.ecol[1] : in.ecol[0];
      const int en = e ? in.en[1] : in.en[0];
      const int ec = e ? in.ec[1] : in.ec[0];
      if (in.ex && ecol >= 0 && cx.pn == en) {
        // row q of the operand's term sits in lane q % 64 of table register
        // q / 64 (one gather per 64 rows) and reaches the lanes that hold it
        // in a transform by a cross-lane permute
        constexpr int K = (TPT * NZ + 63) / 64;
        const int lane = threadIdx.x & 63;
        float tab[K];
        const float* xf = in.ex + (size_t)cx.pfr * in.exf + ecol;
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const int q = 64 * k + lane - in.rlo;
          const int u = cx.pm * cx.hop0 + q;
          tab[k] = (q >= 0 && q < in.nr && u < in.exH) ? xf[(size_t)u * in.exld] * in.einv[u] : 0.f;
        }
        const float2 E = c == 0 ? make_float2(1.f, (ec & 1) ? -1.f : 1.f)
                                : cx.tw[((long long)c * ec * (TW_LEN / cx.Tw)) & (TW_LEN - 1)];
#pragma unroll
        for (int r = 0; r < NZ; ++r) {
          const float dl = __shfl(tab[(TPT * r) / 64], (TPT * r) % 64 + t, 64);
          if (active) {
            v[r].x += dl * E.x;
            v[r].y += dl * E.y;
          }
        }
      }
    }
  } else {
    const float2* col = in.shared ? in.p + ((size_t)grp * NB + c) * in.pitch + row0
                                  : in.p + ((size_t)p * NB + c) * in.pitch;
#pragma unroll
    for (int r = 0; r < NZ; ++r) {
      const int rr = t + TPT * r - in.rlo;
      v[r] = (active && rr >= 0 && rr < in.nr) ? col[rr] : make_float2(0.f, 0.f);
    }
  }
}

// stage rows [0,nrows) of the strip in LDS and store row-major [P][nrows][NBP]
// (WGC threads per workgroup).
// The staging area aliases the FFT slices: barriers on both sides, except
// after the kernel's last store (LAST), where no slice is used again.
// MM < 16: the caller guarantees nrows <= MM*TPT, so slots m >= MM are never
// stored (and the transform's outputs there are dead code).
template <int TH, int CS, bool LAST = false, int MM = 16, int WGC = WG>
__device__ __forceinline__ void col_store_rows(const float2 (&v)[16], float2* st, const ColOut& o, int p, int strip,
                                               int NB, int NBP, int cc, int t) {
  constexpr int TPT = TH / 16;
  __syncthreads();  // all transforms of the workgroup are done with the slices
#pragma unroll
  for (int m = 0; m < MM; ++m) {
    const int r = t + TPT * m;
    if (r < o.nrows) st[r * (CS + 1) + cc] = v[m];
  }
  __syncthreads();
  const int c0 = strip * CS;
  const int ci = threadIdx.x & (CS - 1);
  if (c0 + ci < NB - 1) {  // bin NB-1 travels packed in bin 0
    float2* dst = o.p + (size_t)p * o.nrows * NBP + c0 + ci;
    for (int r = threadIdx.x / CS; r < o.nrows; r += WGC / CS) dst[(size_t)r * NBP] = st[r * (CS + 1) + ci];
  }
  if (!LAST) __syncthreads();  // staging drained before the slices are reused
}

}  // namespace vmbd
