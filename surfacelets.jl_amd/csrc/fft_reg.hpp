// Register-resident four-step FFT of one pencil held by R2 threads (gfx950).
//
// A length-N pencil (N = R1 R2) lives in LDS at consecutive slots
// buf[slot(e)], e in [0, N).  Thread j in [0, R2) of the pencil:
//   phase 1: loads x[j + R2 r] (r < R1), runs an R1-point DFT in registers and
//            multiplies output s by w_N^(j s);
//   exchange: stores output s at slot s R2 + j;
//   phase 2: loads rows s = j ROWS + u (u < ROWS = ceil(R1 / R2), s < R1) of
//            R2 slots each and runs an R2-point DFT per row: row s gives
//            X[s + R1 q], q < R2.  When R2 does not divide R1 (N = 80, 96,
//            112: R1 = 10, 12, 14) the last threads hold fewer rows or none
//            (row_ok); their register rows are computed but carry nothing.
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
template <> struct FourStep<80> { static constexpr int R1 = 10, R2 = 8; };
template <> struct FourStep<96> { static constexpr int R1 = 12, R2 = 8; };
template <> struct FourStep<112> { static constexpr int R1 = 14, R2 = 8; };

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
  else dft_radix<R, DIR>(v);  // 10, 12, 14: two-level composites (fft.hpp)
}

// Twiddle w_N^m from a table of the N forward twiddles (tw[m] = exp(-2 pi i m / N)).
template <int DIR> __device__ __forceinline__ float2 tmul(float2 v, const float2 *tw, int m) {
  float2 w = tw[m];
  return DIR < 0 ? make_float2(v.x * w.x - v.y * w.y, v.x * w.y + v.y * w.x)
                 : make_float2(v.x * w.x + v.y * w.y, v.y * w.x - v.x * w.y);
}

template <int N, int DIR> struct RegFFT {
  static constexpr int R1 = FourStep<N>::R1, R2 = FourStep<N>::R2, ROWS = (R1 + R2 - 1) / R2;
  // Register array length of a pencil thread: R1 for phase 1, ROWS rows of R2
  // for phase 2.
  static constexpr int NV = ROWS * R2 > R1 ? ROWS * R2 : R1;
  // Whether row u of thread j exists (always when R2 divides R1).
  __device__ static __forceinline__ bool row_ok(int j, int u) { return R1 % R2 == 0 || j * ROWS + u < R1; }
  // Exchange slot of element e: one pad slot per ROWS R2 slots (the rows one
  // thread reads in phase 2; R1 slots when R2 divides R1), so that the
  // threads' rows start in different banks.  ROWS R2 is a power of two for
  // every length here, so e / (ROWS R2) is a shift.  A pencil's exchange spans
  // N + PAD slots.
  static constexpr int PADDIV = ROWS * R2;
  static constexpr int PAD = (N + PADDIV - 1) / PADDIV;
  __device__ static __forceinline__ int xpos(int e) { return e + e / PADDIV; }
  // Phase 1 on values already in registers (v[r] = x[j + R2 r]).  twT is the
  // twiddle table in phase-1 order (fill_twiddles): thread j reads
  // twT[s R2 + j] = w_N^(j s), one base address and constant offsets.
  __device__ static __forceinline__ void phase1(float2 (&v)[NV], int j, const float2 *twT) {
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
  // The exchange stores of phase 1: output s of thread j to slot s R2 + j.
  __device__ static __forceinline__ void store_cols(const float2 (&v)[NV], float2 *pb, int j) {
#pragma unroll
    for (int s = 0; s < R1; ++s) pb[xpos(s * R2 + j)] = v[s];
  }
  // The exchange loads of phase 2: v[u R2 + q] = slot (j ROWS + u) R2 + q for
  // the rows thread j holds (the R1 consecutive slots from j R1 when R2
  // divides R1).
  __device__ static __forceinline__ void load_rows(float2 (&v)[NV], const float2 *pb, int j) {
    // Thread j's rows are the ROWS R2 consecutive slots from j ROWS R2, one
    // pad unit: they sit at j (ROWS R2 + 1) + u R2 + q.
#pragma unroll
    for (int u = 0; u < ROWS; ++u)
      if (row_ok(j, u)) {
#pragma unroll
        for (int q = 0; q < R2; ++q) v[u * R2 + q] = pb[j * (PADDIV + 1) + u * R2 + q];
      }
  }
  // Phase 2 on v[u R2 + q'] = slot((j ROWS + u) R2 + q'); results in place:
  // v[u R2 + q] = X[(j ROWS + u) + R1 q].
  __device__ static __forceinline__ void phase2(float2 (&v)[NV]) {
#pragma unroll
    for (int u = 0; u < ROWS; ++u) dftR<R2, DIR>(v + u * R2);
  }
};

}  // namespace surfhip
