// Batched mixed-radix Stockham FFT on LDS-resident pencils (gfx950).
//
// A block holds G pencils of length n in LDS, element (t, g) at [t * G + g],
// so that consecutive lanes touch consecutive words.  Each pass of radix R
// reads a[j + r n/R], applies the twiddles w^(r (j mod Ns) n/(Ns R)), runs an
// R-point DFT in registers and writes b[(j/Ns) Ns R + (j mod Ns) + r Ns]
// (decimation in time, self-sorting: no bit reversal), ping-ponging two LDS
// buffers.  Radices 2, 3, 4, 5, 7, 8 are hand-written, and 6, 9, 10, 12, 14,
// 16 are two-level composites of them in registers; other primes up to
// FFT_MAX_RADIX use a direct DFT.  Twiddles w^m = exp(-2 pi i m / n) are made
// in FP64 on the host and rounded once to FP32.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#define FFT_MAX_PASSES 24
#define FFT_MAX_RADIX 32

namespace surfhip {

struct FftPlan1D {
  int32_t n;
  int32_t npass;
  int32_t tw;  // offset of this length's n twiddles in the pool
  int32_t radix[FFT_MAX_PASSES];
};

// Host: factor n (the hand-written radices, largest first, then the remaining
// primes) and append its twiddles to `pool`.  Throws for a prime factor >
// FFT_MAX_RADIX.
FftPlan1D make_fft_plan(int n, std::vector<float2> &pool);

// Division by d >= 1 as a multiply-high: q = floor(x / d) exactly for
// 0 <= x < 2^16 and d <= 2^16 (m = floor((2^32 - 1) / d) + 1 exceeds 2^32 / d
// by less than 1, so x m / 2^32 exceeds x / d by less than 2^-16 <= 1 / d).
// d = 1 (m = 2^32) is taken apart.
struct FastDiv {
  uint32_t m;
  int d;
};
__device__ __forceinline__ FastDiv fast_div(int d) { return {0xFFFFFFFFu / (uint32_t)d + 1u, d}; }
__device__ __forceinline__ int div_of(int x, FastDiv f) {
  return f.d == 1 ? x : (int)__umulhi((uint32_t)x, f.m);
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }
// a * (-i) for the forward transform, a * (+i) for the inverse.
template <int DIR> __device__ __forceinline__ float2 rot90(float2 a) {
  return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}
// Twiddle w^m of the forward (DIR = -1) or inverse (DIR = +1) transform.
template <int DIR> __device__ __forceinline__ float2 twid(const float2 *tw, int m) {
  float2 w = tw[m];
  return DIR < 0 ? w : make_float2(w.x, -w.y);
}

template <int DIR> __device__ __forceinline__ void dft2(float2 *v) {
  float2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <int DIR> __device__ __forceinline__ void dft4(float2 *v) {
  float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  float2 t2 = cadd(v[1], v[3]), t3 = rot90<DIR>(csub(v[1], v[3]));
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}

template <int DIR> __device__ __forceinline__ void dft8(float2 *v) {
  const float c = 0.70710678118654752440f;
  float2 e[4] = {v[0], v[2], v[4], v[6]};
  float2 o[4] = {v[1], v[3], v[5], v[7]};
  dft4<DIR>(e);
  dft4<DIR>(o);
  // w8^1, w8^2, w8^3 of the forward transform: (c,-c), (0,-1), (-c,-c).
  float2 o1 = DIR < 0 ? make_float2(c * (o[1].x + o[1].y), c * (o[1].y - o[1].x))
                      : make_float2(c * (o[1].x - o[1].y), c * (o[1].y + o[1].x));
  float2 o2 = rot90<DIR>(o[2]);
  float2 o3 = DIR < 0 ? make_float2(c * (o[3].y - o[3].x), -c * (o[3].x + o[3].y))
                      : make_float2(-c * (o[3].x + o[3].y), c * (o[3].x - o[3].y));
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

template <int DIR> __device__ __forceinline__ void dft3(float2 *v) {
  const float s3 = 0.86602540378443864676f;
  float2 s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
  float2 m = make_float2(v[0].x - 0.5f * s.x, v[0].y - 0.5f * s.y);
  // forward: X1 = m - i s3 d, X2 = m + i s3 d
  float2 r = rot90<DIR>(cscale(d, s3));
  v[0] = cadd(v[0], s);
  v[1] = cadd(m, r);
  v[2] = csub(m, r);
}

template <int DIR> __device__ __forceinline__ void dft5(float2 *v) {
  // cos/sin of 2 pi/5 and 4 pi/5.
  const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
  const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
  float2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
  float2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
  float2 m1 = make_float2(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
  float2 m2 = make_float2(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
  // forward: X1 = m1 - i (s1 b1 + s2 b2), X2 = m2 - i (s2 b1 - s1 b2)
  float2 n1 = rot90<DIR>(make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
  float2 n2 = rot90<DIR>(make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
  v[0] = cadd(v[0], cadd(a1, a2));
  v[1] = cadd(m1, n1);
  v[4] = csub(m1, n1);
  v[2] = cadd(m2, n2);
  v[3] = csub(m2, n2);
}

template <int DIR> __device__ __forceinline__ void dft7(float2 *v) {
  // cos/sin of 2 pi k / 7, k = 1, 2, 3.
  const float c1 = 0.62348980185873353053f, c2 = -0.22252093395631440429f, c3 = -0.90096886790241912624f;
  const float s1 = 0.78183148246802980871f, s2 = 0.97492791218182360702f, s3 = 0.43388373911755812048f;
  const float2 a1 = cadd(v[1], v[6]), b1 = csub(v[1], v[6]);
  const float2 a2 = cadd(v[2], v[5]), b2 = csub(v[2], v[5]);
  const float2 a3 = cadd(v[3], v[4]), b3 = csub(v[3], v[4]);
  const float2 m1 = make_float2(v[0].x + c1 * a1.x + c2 * a2.x + c3 * a3.x, v[0].y + c1 * a1.y + c2 * a2.y + c3 * a3.y);
  const float2 m2 = make_float2(v[0].x + c2 * a1.x + c3 * a2.x + c1 * a3.x, v[0].y + c2 * a1.y + c3 * a2.y + c1 * a3.y);
  const float2 m3 = make_float2(v[0].x + c3 * a1.x + c1 * a2.x + c2 * a3.x, v[0].y + c3 * a1.y + c1 * a2.y + c2 * a3.y);
  // forward: X_m = M_m - i N_m, X_{7-m} = M_m + i N_m
  const float2 n1 = rot90<DIR>(make_float2(s1 * b1.x + s2 * b2.x + s3 * b3.x, s1 * b1.y + s2 * b2.y + s3 * b3.y));
  const float2 n2 = rot90<DIR>(make_float2(s2 * b1.x - s3 * b2.x - s1 * b3.x, s2 * b1.y - s3 * b2.y - s1 * b3.y));
  const float2 n3 = rot90<DIR>(make_float2(s3 * b1.x - s1 * b2.x + s2 * b3.x, s3 * b1.y - s1 * b2.y + s2 * b3.y));
  v[0] = cadd(v[0], cadd(a1, cadd(a2, a3)));
  v[1] = cadd(m1, n1);
  v[6] = csub(m1, n1);
  v[2] = cadd(m2, n2);
  v[5] = csub(m2, n2);
  v[3] = cadd(m3, n3);
  v[4] = csub(m3, n3);
}

// Direct R-point DFT for other primes, twiddles from the length-n table.
// cos and sin of 2 pi m / R for the composite radices, FP64 values rounded
// once to FP32 (m = 0 .. R-1).
__host__ __device__ constexpr float unit_cos(int R, int m) {
  switch (R) {
    case 6: {
      constexpr float t[6] = {1.00000000000000000000f, 0.50000000000000011102f, -0.49999999999999977796f, -1.00000000000000000000f, -0.50000000000000044409f, 0.50000000000000011102f};
      return t[m];
    }
    case 9: {
      constexpr float t[9] = {1.00000000000000000000f, 0.76604444311897801345f, 0.17364817766693041445f, -0.49999999999999977796f, -0.93969262078590831688f, -0.93969262078590842791f, -0.50000000000000044409f, 0.17364817766692997036f, 0.76604444311897779141f};
      return t[m];
    }
    case 10: {
      constexpr float t[10] = {1.00000000000000000000f, 0.80901699437494745126f, 0.30901699437494745126f, -0.30901699437494734024f, -0.80901699437494734024f, -1.00000000000000000000f, -0.80901699437494756229f, -0.30901699437494756229f, 0.30901699437494722922f, 0.80901699437494734024f};
      return t[m];
    }
    case 12: {
      constexpr float t[12] = {1.00000000000000000000f, 0.86602540378443870761f, 0.50000000000000011102f, 0.00000000000000006123f, -0.49999999999999977796f, -0.86602540378443870761f, -1.00000000000000000000f, -0.86602540378443881863f, -0.50000000000000044409f, -0.00000000000000018370f, 0.50000000000000011102f, 0.86602540378443837454f};
      return t[m];
    }
    case 14: {
      constexpr float t[14] = {1.00000000000000000000f, 0.90096886790241914600f, 0.62348980185873359439f, 0.22252093395631444839f, -0.22252093395631433737f, -0.62348980185873348336f, -0.90096886790241903498f, -1.00000000000000000000f, -0.90096886790241914600f, -0.62348980185873370541f, -0.22252093395631458717f, 0.22252093395631333816f, 0.62348980185873337234f, 0.90096886790241936804f};
      return t[m];
    }
    case 16: {
      constexpr float t[16] = {1.00000000000000000000f, 0.92387953251128673848f, 0.70710678118654757274f, 0.38268343236508983729f, 0.00000000000000006123f, -0.38268343236508972627f, -0.70710678118654746172f, -0.92387953251128673848f, -1.00000000000000000000f, -0.92387953251128684951f, -0.70710678118654768376f, -0.38268343236509033689f, -0.00000000000000018370f, 0.38268343236509000382f, 0.70710678118654735069f, 0.92387953251128651644f};
      return t[m];
    }
    default: return 0.0f;
  }
}
__host__ __device__ constexpr float unit_sin(int R, int m) {
  switch (R) {
    case 6: {
      constexpr float t[6] = {0.00000000000000000000f, 0.86602540378443859659f, 0.86602540378443870761f, 0.00000000000000012246f, -0.86602540378443837454f, -0.86602540378443859659f};
      return t[m];
    }
    case 9: {
      constexpr float t[9] = {0.00000000000000000000f, 0.64278760968653925190f, 0.98480775301220802032f, 0.86602540378443870761f, 0.34202014332566887944f, -0.34202014332566865740f, -0.86602540378443837454f, -0.98480775301220813134f, -0.64278760968653958496f};
      return t[m];
    }
    case 10: {
      constexpr float t[10] = {0.00000000000000000000f, 0.58778525229247313710f, 0.95105651629515353118f, 0.95105651629515364220f, 0.58778525229247324813f, 0.00000000000000012246f, -0.58778525229247302608f, -0.95105651629515353118f, -0.95105651629515364220f, -0.58778525229247335915f};
      return t[m];
    }
    case 12: {
      constexpr float t[12] = {0.00000000000000000000f, 0.49999999999999994449f, 0.86602540378443859659f, 1.00000000000000000000f, 0.86602540378443870761f, 0.49999999999999994449f, 0.00000000000000012246f, -0.49999999999999972244f, -0.86602540378443837454f, -1.00000000000000000000f, -0.86602540378443859659f, -0.50000000000000044409f};
      return t[m];
    }
    case 14: {
      constexpr float t[14] = {0.00000000000000000000f, 0.43388373911755812040f, 0.78183148246802980363f, 0.97492791218182361934f, 0.97492791218182361934f, 0.78183148246802991466f, 0.43388373911755823142f, 0.00000000000000012246f, -0.43388373911755800938f, -0.78183148246802969261f, -0.97492791218182361934f, -0.97492791218182384139f, -0.78183148246802991466f, -0.43388373911755750978f};
      return t[m];
    }
    case 16: {
      constexpr float t[16] = {0.00000000000000000000f, 0.38268343236508978178f, 0.70710678118654746172f, 0.92387953251128673848f, 1.00000000000000000000f, 0.92387953251128673848f, 0.70710678118654757274f, 0.38268343236508989280f, 0.00000000000000012246f, -0.38268343236508967076f, -0.70710678118654746172f, -0.92387953251128651644f, -1.00000000000000000000f, -0.92387953251128662746f, -0.70710678118654768376f, -0.38268343236509039240f};
      return t[m];
    }
    default: return 0.0f;
  }
}

// a exp(DIR 2 pi i m / R), m a constant once unrolled: quarter and half turns
// exactly, the rest with the tabulated cos and sin.
template <int R, int DIR> __device__ __forceinline__ float2 root_mul(float2 a, int m) {
  m %= R;
  if (m == 0) return a;
  if (4 * m == R) return rot90<DIR>(a);
  if (2 * m == R) return make_float2(-a.x, -a.y);
  if (4 * m == 3 * R) return rot90<-DIR>(a);
  const float c = unit_cos(R, m), s = DIR < 0 ? -unit_sin(R, m) : unit_sin(R, m);
  return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
}

template <int N, int DIR> __device__ __forceinline__ void dft_radix(float2 *v);

// Composite radix R = P Q (Cooley-Tukey in registers): input r = q + Q p,
// output k = k1 + P k2.  P-point DFTs over p for each q, the internal twiddles
// w_R^(q k1), then Q-point DFTs over q for each k1.
template <int P, int Q, int DIR> __device__ __forceinline__ void dft_pq(float2 *v) {
  constexpr int R = P * Q;
  float2 y[Q][P];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
#pragma unroll
    for (int p = 0; p < P; ++p) y[q][p] = v[q + Q * p];
    dft_radix<P, DIR>(y[q]);
#pragma unroll
    for (int k1 = 1; k1 < P; ++k1) y[q][k1] = root_mul<R, DIR>(y[q][k1], q * k1);
  }
#pragma unroll
  for (int k1 = 0; k1 < P; ++k1) {
    float2 z[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) z[q] = y[q][k1];
    dft_radix<Q, DIR>(z);
#pragma unroll
    for (int k2 = 0; k2 < Q; ++k2) v[k1 + P * k2] = z[k2];
  }
}

// The hand-written butterflies: 2, 3, 4, 5, 7, 8 directly; 6, 9, 10, 12, 14
// and 16 as two-level composites of them.
template <int N, int DIR> __device__ __forceinline__ void dft_radix(float2 *v) {
  if constexpr (N == 2) dft2<DIR>(v);
  else if constexpr (N == 3) dft3<DIR>(v);
  else if constexpr (N == 4) dft4<DIR>(v);
  else if constexpr (N == 5) dft5<DIR>(v);
  else if constexpr (N == 7) dft7<DIR>(v);
  else if constexpr (N == 8) dft8<DIR>(v);
  else if constexpr (N == 6) dft_pq<2, 3, DIR>(v);
  else if constexpr (N == 9) dft_pq<3, 3, DIR>(v);
  else if constexpr (N == 10) dft_pq<2, 5, DIR>(v);
  else if constexpr (N == 12) dft_pq<4, 3, DIR>(v);
  else if constexpr (N == 14) dft_pq<2, 7, DIR>(v);
  else if constexpr (N == 16) dft_pq<4, 4, DIR>(v);
}

// Whether radix R has a hand-written butterfly.
__host__ __device__ constexpr bool fft_radix_fixed(int R) {
  return R == 2 || R == 3 || R == 4 || R == 5 || R == 6 || R == 7 || R == 8 || R == 9 || R == 10 || R == 12 ||
         R == 14 || R == 16;
}

template <int DIR> __device__ __noinline__ void dft_generic(float2 *v, int R, const float2 *tw, int n) {
  float2 out[FFT_MAX_RADIX];
  const int step = n / R;
  for (int s = 0; s < R; ++s) {
    float2 acc = v[0];
    int e = 0;
    for (int r = 1; r < R; ++r) {
      e += s;
      if (e >= R) e -= R;
      acc = cadd(acc, cmul(v[r], twid<DIR>(tw, e * step)));
    }
    out[s] = acc;
  }
  for (int s = 0; s < R; ++s) v[s] = out[s];
}

template <int R, int DIR> __device__ __forceinline__ void fft_pass_fixed(
    const float2 *__restrict__ a, float2 *__restrict__ b, int n, int G, FastDiv dG, int Ns, FastDiv dNs,
    const float2 *tw, int tid, int nthreads) {
  const int nR = n / R;
  const int tstep = n / (Ns * R);
  for (int w = tid; w < G * nR; w += nthreads) {
    const int j = div_of(w, dG), g = w - j * G;
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = a[(j + r * nR) * G + g];
    const int jn = div_of(j, dNs);
    const int k = j - jn * Ns;
    if (k) {
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twid<DIR>(tw, r * k * tstep));
    }
    dft_radix<R, DIR>(v);
    const int idxD = jn * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) b[(idxD + r * Ns) * G + g] = v[r];
  }
}

template <int DIR> __device__ __noinline__ void fft_pass_generic(
    const float2 *__restrict__ a, float2 *__restrict__ b, int n, int R, int G, int Ns,
    const float2 *tw, int tid, int nthreads) {
  const int nR = n / R;
  const int tstep = n / (Ns * R);
  for (int w = tid; w < G * nR; w += nthreads) {
    const int g = w % G, j = w / G;
    float2 v[FFT_MAX_RADIX];
    for (int r = 0; r < R; ++r) v[r] = a[(j + r * nR) * G + g];
    const int k = j % Ns;
    for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twid<DIR>(tw, (r * k * tstep) % n));
    dft_generic<DIR>(v, R, tw, n);
    const int idxD = (j / Ns) * Ns * R + k;
    for (int r = 0; r < R; ++r) b[(idxD + r * Ns) * G + g] = v[r];
  }
}

// Whether a plan has a radix without a hand-written butterfly (it then
// needs lds_fft's direct-DFT pass).
inline bool fft_plan_generic(const FftPlan1D &pl) {
  for (int p = 0; p < pl.npass; ++p)
    if (!fft_radix_fixed(pl.radix[p])) return true;
  return false;
}

// Transform the G pencils in `a` (LDS), using `b` as the other half of the
// ping-pong pair; returns the buffer holding the result.  `tw` holds the
// plan's n twiddles (LDS or global).  All threads of the block must call it;
// n G < 2^16.  GEN = false leaves out the direct-DFT pass (a call, whose
// register reservation would cap the kernel's occupancy): only for plans
// with fft_plan_generic(pl) false.
template <int DIR, bool GEN = true>
__device__ float2 *lds_fft(float2 *a, float2 *b, const FftPlan1D &pl, int G, const float2 *tw, int tid,
                           int nthreads) {
  int Ns = 1;
  const int n = pl.n;
  const FastDiv dG = fast_div(G);
  for (int p = 0; p < pl.npass; ++p) {
    const int R = pl.radix[p];
    const FastDiv dNs = fast_div(Ns);
    switch (R) {
      case 8: fft_pass_fixed<8, DIR>(a, b, n, G, dG, Ns, dNs, tw, tid, nthreads); break;
      case 4: fft_pass_fixed<4, DIR>(a, b, n, G, dG, Ns, dNs, tw, tid, nthreads); break;
      case 2: fft_pass_fixed<2, DIR>(a, b, n, G, dG, Ns, dNs, tw, tid, nthreads); break;
      case 3: fft_pass_fixed<3, DIR>(a, b, n, G, dG, Ns, dNs, tw, tid, nthreads); break;
      case 5: fft_pass_fixed<5, DIR>(a, b, n, G, dG, Ns, dNs, tw, tid, nthreads); break;
      case 7: fft_pass_fixed<7, DIR>(a, b, n, G, dG, Ns, dNs, tw, tid, nthreads); break;
      case 16: fft_pass_fixed<16, DIR>(a, b, n, G, dG, Ns, dNs, tw, tid, nthreads); break;
      case 14: fft_pass_fixed<14, DIR>(a, b, n, G, dG, Ns, dNs, tw, tid, nthreads); break;
      case 12: fft_pass_fixed<12, DIR>(a, b, n, G, dG, Ns, dNs, tw, tid, nthreads); break;
      case 10: fft_pass_fixed<10, DIR>(a, b, n, G, dG, Ns, dNs, tw, tid, nthreads); break;
      case 9: fft_pass_fixed<9, DIR>(a, b, n, G, dG, Ns, dNs, tw, tid, nthreads); break;
      case 6: fft_pass_fixed<6, DIR>(a, b, n, G, dG, Ns, dNs, tw, tid, nthreads); break;
      default:
        if (GEN) fft_pass_generic<DIR>(a, b, n, R, G, Ns, tw, tid, nthreads);
        break;
    }
    __syncthreads();
    float2 *t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

}  // namespace surfhip
