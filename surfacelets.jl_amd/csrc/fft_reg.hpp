// Register-resident four-step FFT of one pencil held by R2 threads (gfx950).
//
// A length-N pencil (N = R1 R2, R2 | R1) lives in LDS at consecutive slots
// buf[slot(e)], e in [0, N).  Thread j in [0, R2) of the pencil:
//   phase 1: loads x[j + R2 r] (r < R1), runs an R1-point DFT in registers and
//            multiplies output s by w_N^(j s);
//   exchange: stores output s at slot s R2 + j;
//   phase 2: loads the R1 slots [j R1, (j+1) R1) -- R1/R2 rows of R2 -- and
//            runs R1/R2 R2-point DFTs: row s gives X[s + R1 q], q < R2.
// X[s + R1 q] = sum_j w_R2^(j q) w_N^(j s) sum_r x[j + R2 r] w_R1^(r s) is the
// four-step factorisation of the DFT; DIR = -1 forward, +1 inverse
// (unnormalised).  The caller places __syncthreads() between the phases.
#pragma once

#include <hip/hip_runtime.h>

#include "fft.hpp"

namespace surfhip {

template <int N> struct FourStep;
template <> struct FourStep<16> { static constexpr int R1 = 4, R2 = 4; };
template <> struct FourStep<32> { static constexpr int R1 = 8, R2 = 4; };
template <> struct FourStep<64> { static constexpr int R1 = 8, R2 = 8; };
template <> struct FourStep<128> { static constexpr int R1 = 16, R2 = 8; };
template <> struct FourStep<256> { static constexpr int R1 = 16, R2 = 16; };

template <int DIR> __device__ __forceinline__ void dft16(float2 *x) {
  // 16 = 4 x 4: DFT-4 over n1 for each n2, twiddle w16^(n2 k1), DFT-4 over n2.
  float2 a[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    float2 t[4] = {x[n2], x[4 + n2], x[8 + n2], x[12 + n2]};
    dft4<DIR>(t);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) a[n2][k1] = t[k1];
  }
  const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f, c2 = 0.70710678118654752440f;
  // w16^m, m = n2 k1 in {1, 2, 3, 2, 4, 6, 3, 6, 9}; forward exp(-2 pi i m / 16).
  // m = 4 is a quarter turn and m = 2, 6 have |c| = |s|: no multiplications
  // by 0 or 1 and one product per component.
  auto tw = [&](float2 v, int m) -> float2 {
    if (m == 4) return rot90<DIR>(v);
    if (m == 2)
      return DIR < 0 ? make_float2(c2 * (v.x + v.y), c2 * (v.y - v.x)) : make_float2(c2 * (v.x - v.y), c2 * (v.y + v.x));
    if (m == 6)
      return DIR < 0 ? make_float2(c2 * (v.y - v.x), -c2 * (v.x + v.y)) : make_float2(-c2 * (v.x + v.y), c2 * (v.x - v.y));
    float c, s;
    switch (m) {
      case 1: c = c1; s = s1; break;
      case 3: c = s1; s = c1; break;
      default: c = -c1; s = -s1; break;  // m = 9
    }
    // multiply by (c - i s) forward, (c + i s) inverse
    return DIR < 0 ? make_float2(v.x * c + v.y * s, v.y * c - v.x * s)
                   : make_float2(v.x * c - v.y * s, v.y * c + v.x * s);
  };
#pragma unroll
  for (int n2 = 1; n2 < 4; ++n2)
#pragma unroll
    for (int k1 = 1; k1 < 4; ++k1) a[n2][k1] = tw(a[n2][k1], n2 * k1);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    float2 t[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
    dft4<DIR>(t);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) x[k1 + 4 * k2] = t[k2];
  }
}

template <int R, int DIR> __device__ __forceinline__ void dftR(float2 *v) {
  if constexpr (R == 2) dft2<DIR>(v);
  else if constexpr (R == 4) dft4<DIR>(v);
  else if constexpr (R == 8) dft8<DIR>(v);
  else if constexpr (R == 16) dft16<DIR>(v);
}

// Twiddle w_N^m from a table of the N forward twiddles (tw[m] = exp(-2 pi i m / N)).
template <int DIR> __device__ __forceinline__ float2 tmul(float2 v, const float2 *tw, int m) {
  float2 w = tw[m];
  return DIR < 0 ? make_float2(v.x * w.x - v.y * w.y, v.x * w.y + v.y * w.x)
                 : make_float2(v.x * w.x + v.y * w.y, v.y * w.x - v.x * w.y);
}

template <int N, int DIR> struct RegFFT {
  static constexpr int R1 = FourStep<N>::R1, R2 = FourStep<N>::R2, ROWS = R1 / R2;
  // Phase 1 on values already in registers (v[r] = x[j + R2 r]).  twT is the
  // twiddle table in phase-1 order (fill_twiddles): thread j reads
  // twT[s R2 + j] = w_N^(j s), one base address and constant offsets.
  __device__ static __forceinline__ void phase1(float2 (&v)[R1], int j, const float2 *twT) {
    dftR<R1, DIR>(v);
    const float2 *tj = twT + j;
#pragma unroll
    for (int s = 1; s < R1; ++s) v[s] = tmul<DIR>(v[s], tj, s * R2);
  }
  // twT[s R2 + j] = tw[j s] from the N forward twiddles tw[m] =
  // exp(-2 pi i m / N); j s < N.  Threads t = first, first + step, ...
  __device__ static __forceinline__ void fill_twiddles(float2 *twT, const float2 *tw, int first, int step) {
    for (int m = first; m < N; m += step) twT[m] = tw[(m % R2) * (m / R2)];
  }
  // Phase 2 on v[u R2 + q'] = slot((j ROWS + u) R2 + q'); results in place:
  // v[u R2 + q] = X[(j ROWS + u) + R1 q].
  __device__ static __forceinline__ void phase2(float2 (&v)[R1]) {
#pragma unroll
    for (int u = 0; u < ROWS; ++u) dftR<R2, DIR>(v + u * R2);
  }
};

}  // namespace surfhip
