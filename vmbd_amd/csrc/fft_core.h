// Radix-16/8/4/2 Stockham FFT core for gfx950 (wave64).
//
// One complex transform of length N (16 <= N <= 1024, power of two) is
// carried by TPT = N/16 threads that each hold 16 points in registers.  All
// TPT threads of a transform sit inside one wavefront (TPT <= 64 and TPT
// divides 64), so the exchanges between passes go through the transform's own
// LDS slice with wavefront-level ordering only, no workgroup barrier.
//
// Register mappings (both validated against numpy.fft on the host):
//   input : thread t, slot r  holds  x[t + TPT*r]
//   output: thread t, slot m  holds  X[t + TPT*m]
// The identical mapping on both sides lets a forward transform, a pointwise
// product and an inverse transform run back to back in registers.
//
// Algorithm: self-sorting Stockham passes with radix 16 first, then the
// remaining factor (Govindaraju et al., "High performance discrete Fourier
// transforms on graphics processors", SC'08): pass with span Ns and radix R
// reads x[j + r*N/R], multiplies by w^(r*(j mod Ns)) with w = e^{S 2 pi i/(Ns R)},
// applies an R-point DFT and writes (j/Ns)*Ns*R + (j mod Ns) + r*Ns.
#pragma once
#include <hip/hip_runtime.h>

namespace vmbd {

// Twiddle table length.  tw[m] = exp(-2 pi i m / TW_LEN), rounded from FP64.
constexpr int TW_LEN = 1024;

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.y, b.x, -a.x * b.y));
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// S * i * a
template <int S>
__device__ __forceinline__ float2 mul_si(float2 a) { return make_float2(-(float)S * a.y, (float)S * a.x); }

// Wavefront-scope ordering point for LDS exchanges inside one wavefront.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------- small DFTs
// All produce outputs in natural order in place; S = -1 forward, +1 inverse.
template <int S>
__device__ __forceinline__ void dft2(float2& a, float2& b) {
  float2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <int S>
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  float2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
  float2 t2 = cadd(a1, a3), t3 = mul_si<S>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

#define VMBD_R2H 0.70710678118654752440f
#define VMBD_C8 0.92387953251128675613f  // cos(pi/8)
#define VMBD_S8 0.38268343236508977173f  // sin(pi/8)

template <int S>
__device__ __forceinline__ void dft8(float2* v) {
  // n = 2 n1 + n2, k = k1 + 4 k2
  dft4<S>(v[0], v[2], v[4], v[6]);
  dft4<S>(v[1], v[3], v[5], v[7]);
  // twiddle w8^(k1) on the odd branch; branch k1 sits in slot 1+2*k1
  v[3] = cmul(v[3], make_float2(VMBD_R2H, (float)S * VMBD_R2H));
  v[5] = mul_si<S>(v[5]);
  v[7] = cmul(v[7], make_float2(-VMBD_R2H, (float)S * VMBD_R2H));
  float2 o[8];
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    o[k1] = cadd(v[2 * k1], v[2 * k1 + 1]);
    o[k1 + 4] = csub(v[2 * k1], v[2 * k1 + 1]);
  }
#pragma unroll
  for (int i = 0; i < 8; ++i) v[i] = o[i];
}

template <int S>
__device__ __forceinline__ void dft16(float2* v) {
  // n = 4 n1 + n2, k = k1 + 4 k2.  Step 1: DFT4 over n1 for each n2.
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) dft4<S>(v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]);
  // now v[n2 + 4*k1] = Y[n2][k1]; twiddle by w16^(n2*k1)
  const float2 w1 = make_float2(VMBD_C8, (float)S * VMBD_S8);
  const float2 w2 = make_float2(VMBD_R2H, (float)S * VMBD_R2H);
  const float2 w3 = make_float2(VMBD_S8, (float)S * VMBD_C8);
  const float2 w6 = make_float2(-VMBD_R2H, (float)S * VMBD_R2H);
  const float2 w9 = make_float2(-VMBD_C8, -(float)S * VMBD_S8);
  v[1 + 4] = cmul(v[5], w1);
  v[1 + 8] = cmul(v[9], w2);
  v[1 + 12] = cmul(v[13], w3);
  v[2 + 4] = cmul(v[6], w2);
  v[2 + 8] = mul_si<S>(v[10]);
  v[2 + 12] = cmul(v[14], w6);
  v[3 + 4] = cmul(v[7], w3);
  v[3 + 8] = cmul(v[11], w6);
  v[3 + 12] = cmul(v[15], w9);
  // Step 3: DFT4 over n2 for each k1 -> X[k1 + 4 k2]
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) dft4<S>(v[4 * k1 + 0], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3]);
  // v[4*k1 + k2] = X[k1 + 4 k2]; transpose the 4x4 index grid to natural order
  float2 o[16];
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) o[k1 + 4 * k2] = v[4 * k1 + k2];
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = o[i];
}

template <int R, int S>
__device__ __forceinline__ void dftR(float2* v) {
  if constexpr (R == 16) dft16<S>(v);
  else if constexpr (R == 8) dft8<S>(v);
  else if constexpr (R == 4) dft4<S>(v[0], v[1], v[2], v[3]);
  else dft2<S>(v[0], v[1]);
}

// LDS slot of element i inside a transform's slice (one pad slot per 16).
__device__ __forceinline__ int lds_pad(int i) { return i + (i >> 4); }
template <int N>
struct FftLds {
  static constexpr int floats2 = N + N / 16;  // float2 slots per transform
};

// w^r for r = 0..R-1 where w = twiddle base.
template <int R>
__device__ __forceinline__ void twiddle_powers(float2 w1, float2 w4, float2* p) {
  p[0] = make_float2(1.f, 0.f);
  p[1] = w1;
  if constexpr (R > 2) {
    p[2] = cmul(w1, w1);
    p[3] = cmul(p[2], w1);
  }
  if constexpr (R > 4) {
    p[4] = w4;
    p[5] = cmul(w4, w1);
    p[6] = cmul(w4, p[2]);
    p[7] = cmul(w4, p[3]);
  }
  if constexpr (R > 8) {
    float2 w8 = cmul(w4, w4);
    p[8] = w8;
    p[9] = cmul(w8, w1);
    p[10] = cmul(w8, p[2]);
    p[11] = cmul(w8, p[3]);
    float2 w12 = cmul(w8, w4);
    p[12] = w12;
    p[13] = cmul(w12, w1);
    p[14] = cmul(w12, p[2]);
    p[15] = cmul(w12, p[3]);
  }
}

// Radix plan of a length-N transform after the first radix-16 pass.
template <int N>
struct Plan {
  static constexpr int R1 = (N >= 256) ? 16 : N / 16;  // second pass radix
  static constexpr int R2 = (N > 256) ? N / 256 : 1;   // third pass radix
};

// Per-thread forward twiddles of the second pass (span 16): the thread's
// 16/R1 butterflies j = t + b*TPT use w^(r*(j mod 16)), w = e^{-2 pi i/(16 R1)}.
// They depend only on t, so a kernel computes them once and reuses them for
// every transform it performs.  The (optional) third pass loads its twiddles.
template <int N>
struct Twiddles {
  static constexpr int R1 = Plan<N>::R1;
  static constexpr int NB1 = (R1 > 1) ? 16 / R1 : 1;
  static constexpr int CNT = (R1 > 1) ? NB1 * (R1 - 1) : 1;
  float2 w[CNT];
};

// LOAD=true reads every twiddle from the table (best when a thread performs
// several transforms); LOAD=false reads two and builds the powers by complex
// multiplication (fewer memory instructions for one transform per thread).
template <int N, bool LOAD = false>
__device__ __forceinline__ void twiddles_init(Twiddles<N>& T, int t, const float2* __restrict__ tw) {
  constexpr int TPT = N / 16;
  constexpr int R1 = Plan<N>::R1;
  if constexpr (R1 > 1) {
    constexpr int NB1 = 16 / R1;
    constexpr int STR = TW_LEN / (16 * R1);
#pragma unroll
    for (int b = 0; b < NB1; ++b) {
      const int k = (t + b * TPT) & 15;
      if constexpr (LOAD) {
#pragma unroll
        for (int r = 1; r < R1; ++r) T.w[b * (R1 - 1) + r - 1] = tw[(k * r * STR) & (TW_LEN - 1)];
      } else {
        float2 p[R1];
        twiddle_powers<R1>(tw[k * STR], tw[(4 * k * STR) & (TW_LEN - 1)], p);
#pragma unroll
        for (int r = 1; r < R1; ++r) T.w[b * (R1 - 1) + r - 1] = p[r];
      }
    }
  }
}

// The same twiddles in two steps: twiddles_seed issues the table loads (two
// per butterfly), twiddles_from_seed forms the powers, so the loads can be in
// flight with other work while only the seeds hold registers.
template <int N>
struct TwSeed {
  static constexpr int NB1 = Twiddles<N>::NB1;
  float2 w1[NB1], w4[NB1];
};
template <int N>
__device__ __forceinline__ void twiddles_seed(TwSeed<N>& S, int t, const float2* __restrict__ tw) {
  constexpr int TPT = N / 16;
  constexpr int R1 = Plan<N>::R1;
  if constexpr (R1 > 1) {
    constexpr int STR = TW_LEN / (16 * R1);
#pragma unroll
    for (int b = 0; b < TwSeed<N>::NB1; ++b) {
      const int k = (t + b * TPT) & 15;
      S.w1[b] = tw[k * STR];
      S.w4[b] = tw[(4 * k * STR) & (TW_LEN - 1)];
    }
  }
}
template <int N>
__device__ __forceinline__ void twiddles_from_seed(Twiddles<N>& T, const TwSeed<N>& S) {
  constexpr int R1 = Plan<N>::R1;
  if constexpr (R1 > 1) {
#pragma unroll
    for (int b = 0; b < TwSeed<N>::NB1; ++b) {
      float2 p[R1];
      twiddle_powers<R1>(S.w1[b], S.w4[b], p);
#pragma unroll
      for (int r = 1; r < R1; ++r) T.w[b * (R1 - 1) + r - 1] = p[r];
    }
  }
}

// One Stockham pass after the first (Ns > 1).  Reads 16 values of the slice,
// applies 16/R butterflies, writes 16 values (or keeps them in registers when
// LAST: slot m = b + r*(16/R) holds X[t + TPT*m]).  Twiddles come from `pre`
// (per-thread forward twiddles, conjugated for S=+1) or from the table.
template <int N, int R, int NS, int S, bool LAST, bool PRE>
__device__ __forceinline__ void stockham_pass(float2 (&v)[16], float2* sl, int t, const float2* pre,
                                              const float2* __restrict__ tw) {
  constexpr int TPT = N / 16;
  constexpr int NB = 16 / R;
  float2 in[16];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * TPT;
#pragma unroll
    for (int r = 0; r < R; ++r) in[b * R + r] = sl[lds_pad(j + r * (N / R))];
  }
  wave_sync();
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * TPT;
    const int k = j & (NS - 1);
    float2* q = in + b * R;
    if constexpr (PRE) {
#pragma unroll
      for (int r = 1; r < R; ++r) q[r] = (S < 0) ? cmul(q[r], pre[b * (R - 1) + r - 1]) : cmulc(q[r], pre[b * (R - 1) + r - 1]);
    } else {
      constexpr int STR = TW_LEN / (NS * R);
      float2 w1 = tw[k * STR];
      float2 w4 = tw[(4 * k * STR) & (TW_LEN - 1)];
      if (S > 0) {
        w1 = cconj(w1);
        w4 = cconj(w4);
      }
      float2 p[R];
      twiddle_powers<R>(w1, w4, p);
#pragma unroll
      for (int r = 1; r < R; ++r) q[r] = cmul(q[r], p[r]);
    }
    dftR<R, S>(q);
    if constexpr (LAST) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[b + r * NB] = q[r];
    } else {
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) sl[lds_pad(base + r * NS)] = q[r];
    }
  }
  if constexpr (!LAST) wave_sync();
}

// Full complex FFT of length N on the slice `sl` (FftLds<N>::floats2 float2).
// v: input mapping on entry, output mapping on exit.  S=-1 forward, +1 inverse
// (unscaled).  Must be called by all TPT threads of the transform.
// NZ < 16: the caller guarantees x[t + TPT*r] = 0 for r >= NZ; those slots are
// taken as constant zeros, so the first pass shrinks to the arithmetic the
// nonzero inputs need (outputs a caller never reads are removed the same way
// by the compiler once the call is inlined).
template <int N, int S, int NZ = 16>
__device__ __forceinline__ void fft(float2 (&v)[16], float2* sl, int t, const Twiddles<N>& T,
                                    const float2* __restrict__ tw) {
  static_assert(N >= 16 && N <= 1024 && (N & (N - 1)) == 0, "N must be a power of two in [16,1024]");
  static_assert(NZ >= 1 && NZ <= 16, "NZ in [1,16]");
  if constexpr (NZ < 16) {
#pragma unroll
    for (int r = NZ; r < 16; ++r) v[r] = make_float2(0.f, 0.f);
  }
  // pass 0: radix 16, Ns = 1, j = t, no twiddles; output index t*16 + r
  dft16<S>(v);
  if constexpr (N == 16) {
    return;
  } else {
#pragma unroll
    for (int r = 0; r < 16; ++r) sl[lds_pad(t * 16 + r)] = v[r];
    wave_sync();
    constexpr int R1 = Plan<N>::R1;
    constexpr int R2 = Plan<N>::R2;
    if constexpr (R2 == 1) {
      stockham_pass<N, R1, 16, S, true, true>(v, sl, t, T.w, tw);
    } else {
      stockham_pass<N, R1, 16, S, false, true>(v, sl, t, T.w, tw);
      stockham_pass<N, R2, 256, S, true, false>(v, sl, t, nullptr, tw);
    }
  }
}

}  // namespace vmbd
