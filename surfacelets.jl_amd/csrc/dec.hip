// Decimated Surfacelet analyze/synthesize on gfx950.
//
// Analysis of one real channel (src/transforms/surfacelet.jl:369-381):
//   X = FFT_3(x) (complex, full grid);
//   per pyramid level j (finest first): split X into its highpass (in place)
//   and the resampled lowpass (fourier_pyramid.jl:108-118); then per
//   hourglass K: S_K = X[half_K] V_K / prod(subband size) and the IRC analysis
//   stages (irc.jl:127-134), one pass (k_irc_group), inverse FFTs along the two
//   other axes within each wedge block, and a c2r along K that writes each
//   block's real subband into the flat coefficient vector
//   (surfacelet_scale.jl:200-213);
//   coarse subband = real(IFFT_3(lowpass)).
// Synthesis mirrors it (surfacelet.jl:384-396, surfacelet_scale.jl:220-257).
// Every index map (pyramid boxes, block ranges, Hermitian mirror, subband
// offsets) is integer arithmetic fixed on the host and shipped as tables.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <memory>
#include <tuple>
#include <utility>
#include <vector>

#include "dec_tables.hpp"
#include "fft.hpp"
#include "fft_reg.hpp"
#include "hip_util.hpp"

namespace surfhip {

namespace {

constexpr int DT = 256;

// Exchange slot of element e: one pad slot per R1 (bank spread in phase 2).
template <int R1> __device__ __forceinline__ int xpos_c2c(int e) { return e + e / R1; }

__device__ __forceinline__ int64_t lin3(int i1, int i2, int i3, int m1, int m2) {
  return (int64_t)i1 + (int64_t)m1 * ((int64_t)i2 + (int64_t)m2 * i3);
}

// Complex FFT along AXIS (1, 2, 3) of an m1 x m2 x m3 column-major array, in
// segments of length `seg` (seg divides the axis; seg = the axis length for a
// whole transform).  G pencils per block: consecutive segments for AXIS 1
// (one contiguous run), consecutive i1 for AXIS 2 and 3.
template <int AXIS, int DIR, bool GEN>
__global__ __launch_bounds__(DT) void k_c2c(const float2 *in, float2 *out, int m1, int m2, int m3, int seg, int G,
                                           FftPlan1D pl, const float2 *__restrict__ twg) {
  extern __shared__ float2 smem[];
  float2 *tw = smem;
  float2 *a = tw + seg;
  float2 *b = a + (size_t)seg * G;
  const int tid = threadIdx.x;
  for (int k = tid; k < seg; k += DT) tw[k] = twg[pl.tw + k];
  int64_t base0 = 0, sa = 1, sg = 1;
  int valid = G;
  if (AXIS == 1) {
    const int64_t npen = (int64_t)(m1 / seg) * m2 * m3;
    const int64_t p0 = (int64_t)blockIdx.x * G;
    valid = (int)min((int64_t)G, npen - p0);
    base0 = p0 * seg;
    sa = 1;
    sg = seg;
  } else {
    const int ub = (m1 + G - 1) / G;
    const int u = blockIdx.x % ub;
    const int rest = blockIdx.x / ub;
    const int u0 = u * G;
    valid = min(G, m1 - u0);
    if (AXIS == 2) {
      const int ns = m2 / seg;
      const int s = rest % ns, i3 = rest / ns;
      base0 = lin3(u0, s * seg, i3, m1, m2);
      sa = m1;
    } else {
      const int i2 = rest % m2, s = rest / m2;
      base0 = lin3(u0, i2, s * seg, m1, m2);
      sa = (int64_t)m1 * m2;
    }
    sg = 1;
  }
  for (int idx = tid; idx < seg * G; idx += DT) {
    int t, g;
    if (AXIS == 1) {
      g = idx / seg;
      t = idx - g * seg;
    } else {
      t = idx / G;
      g = idx - t * G;
    }
    a[t * G + g] = g < valid ? in[base0 + g * sg + t * sa] : make_float2(0.0f, 0.0f);
  }
  __syncthreads();
  float2 *res = lds_fft<DIR, GEN>(a, b, pl, G, tw, tid, DT);
  for (int idx = tid; idx < seg * G; idx += DT) {
    int t, g;
    if (AXIS == 1) {
      g = idx / seg;
      t = idx - g * seg;
    } else {
      t = idx / G;
      g = idx - t * G;
    }
    if (g < valid) out[base0 + g * sg + t * sa] = res[t * G + g];
  }
}

// Power-of-two segments (16..128): the four-step register FFT of fft_reg.hpp
// with one padded LDS exchange, 256 threads = TP pencils of R2 threads.
// Pencil p (flattened: segment index, then the other axes) starts at
// pencil_base(p).  For AXIS 1 the R2 threads of a pencil are consecutive
// lanes (its segment is contiguous) and results leave through LDS in natural
// order; for AXIS 2 and 3 consecutive lanes hold consecutive pencils, which are
// consecutive in memory.
template <int SEG> constexpr int reg_tp() { return DT / FourStep<SEG>::R2; }
template <int SEG> constexpr int reg_ps() { return SEG + SEG / FourStep<SEG>::R1 + 1; }

template <int AXIS>
__device__ __forceinline__ int64_t pencil_base(int64_t p, int m1, int m2, int m3, int seg) {
  if (AXIS == 1) return p * seg;
  const int u = (int)(p % m1);
  const int64_t rest = p / m1;
  if (AXIS == 2) {
    const int ns = m2 / seg;
    return lin3(u, (int)(rest % ns) * seg, (int)(rest / ns), m1, m2);
  }
  return lin3(u, (int)(rest % m2), (int)(rest / m2) * seg, m1, m2);
}

// Optional real end of a volume transform: IO 1 reads the real input
// rin[i stride] scale (imaginary part 0) instead of A; IO 2 writes
// rout[i stride] = Re(result) scale instead of A (i = linear element index).
struct RealIO {
  const float *in;
  float *out;
  int64_t stride;
  float scale;
};

template <int AXIS, int DIR, int SEG, int IO = 0>
__global__ __launch_bounds__(DT) void k_c2c_reg(float2 *A, int m1, int m2, int m3, const float2 *__restrict__ twg,
                                               RealIO rio) {
  using F = RegFFT<SEG, DIR>;
  constexpr int R1 = F::R1, R2 = F::R2, ROWS = F::ROWS, TP = reg_tp<SEG>(), PS = reg_ps<SEG>();
  __shared__ float2 buf[TP * PS];
  __shared__ float2 tw[SEG];
  const int tid = threadIdx.x;
  F::fill_twiddles(tw, twg, tid, DT);
  const int pp = AXIS == 1 ? tid / R2 : tid % TP, j = AXIS == 1 ? tid % R2 : tid / TP;
  const int64_t npen = (int64_t)m1 * m2 * m3 / SEG;
  const int64_t p = (int64_t)blockIdx.x * TP + pp;
  const bool on = p < npen;
  const int64_t sa = AXIS == 1 ? 1 : (AXIS == 2 ? (int64_t)m1 : (int64_t)m1 * m2);
  const int64_t base = on ? pencil_base<AXIS>(p, m1, m2, m3, SEG) : 0;
  float2 v[R1];
#pragma unroll
  for (int r = 0; r < R1; ++r) {
    const int64_t i = base + (j + R2 * r) * sa;
    if (IO == 1) v[r] = make_float2(on ? rio.in[i * rio.stride] * rio.scale : 0.0f, 0.0f);
    else v[r] = on ? A[i] : make_float2(0.0f, 0.0f);
  }
  __syncthreads();  // twiddles
  F::phase1(v, j, tw);
  float2 *pb = buf + pp * PS;
#pragma unroll
  for (int s = 0; s < R1; ++s) pb[xpos_c2c<R1>(s * R2 + j)] = v[s];
  __syncthreads();
#pragma unroll
  for (int e = 0; e < R1; ++e) v[e] = pb[xpos_c2c<R1>(j * R1 + e)];
  F::phase2(v);
  if (AXIS == 1) {
    __syncthreads();
#pragma unroll
    for (int u = 0; u < ROWS; ++u)
#pragma unroll
      for (int q = 0; q < R2; ++q) pb[(j * ROWS + u) + R1 * q] = v[u * R2 + q];
    __syncthreads();
    // The block's TP segments are one contiguous run of TP SEG elements.
    const int64_t p0 = (int64_t)blockIdx.x * TP;
    const int nvalid = (int)min((int64_t)TP, npen - p0) * SEG;
    for (int e = tid; e < nvalid; e += DT) {
      const float2 z = buf[(e / SEG) * PS + e % SEG];
      if (IO == 2) rio.out[(p0 * SEG + e) * rio.stride] = z.x * rio.scale;
      else A[p0 * SEG + e] = z;
    }
  } else if (on) {
#pragma unroll
    for (int u = 0; u < ROWS; ++u)
#pragma unroll
      for (int q = 0; q < R2; ++q) {
        const int64_t i = base + (int64_t)((j * ROWS + u) + R1 * q) * sa;
        if (IO == 2) rio.out[i * rio.stride] = v[u * R2 + q].x * rio.scale;
        else A[i] = v[u * R2 + q];
      }
  }
}

// Both block FFTs of an hourglass K != 0 in one pass: a B1 x B2 tile of S
// (axis 0, which is contiguous, and axis m2, for one index along K) is loaded
// into LDS coalesced, transformed along axis 0 (one register FFT per tile
// row) and then along m2 (one per tile column), and stored coalesced.
// Block = (q1, q2, iK), q1 fastest.
template <int B1, int B2> constexpr int blk2d_threads() {
  constexpr int a = B2 * FourStep<B1>::R2, b = B1 * FourStep<B2>::R2;
  return ((a > b ? a : b) + 63) / 64 * 64;
}
template <int B1, int B2> constexpr int blk2d_xch() {
  constexpr int a = B2 * reg_ps<B1>(), b = B1 * reg_ps<B2>();
  return a > b ? a : b;
}

template <int DIR, int B1, int B2>
__device__ __forceinline__ void blk2d_body(float2 *S, int h0, int h1, int h2, int K, int m2,
                                           const float2 *__restrict__ tw1g, const float2 *__restrict__ tw2g) {
  using FA = RegFFT<B1, DIR>;
  using FB = RegFFT<B2, DIR>;
  constexpr int NTH = blk2d_threads<B1, B2>(), TS = B1 + 1;
  constexpr int RA1 = FA::R1, RA2 = FA::R2, RAR = FA::ROWS, PSA = reg_ps<B1>();
  constexpr int RB1 = FB::R1, RB2 = FB::R2, RBR = FB::ROWS, PSB = reg_ps<B2>();
  __shared__ float2 tile[B2 * TS];
  __shared__ float2 xch[blk2d_xch<B1, B2>()];
  __shared__ float2 twA[B1], twB[B2];
  const int tid = threadIdx.x;
  FA::fill_twiddles(twA, tw1g, tid, NTH);
  FB::fill_twiddles(twB, tw2g, tid, NTH);
  const int hd[3] = {h0, h1, h2};
  const int n1 = h0 / B1, n2 = hd[m2] / B2;
  const int q1 = (int)(blockIdx.x % n1);
  const int rest = (int)(blockIdx.x / n1);
  const int q2 = rest % n2, iK = rest / n2;
  const int64_t s2 = m2 == 1 ? (int64_t)h0 : (int64_t)h0 * h1;
  const int64_t sK = K == 1 ? (int64_t)h0 : (int64_t)h0 * h1;
  const int64_t base = (int64_t)q1 * B1 + (int64_t)q2 * B2 * s2 + (int64_t)iK * sK;
  for (int e = tid; e < B1 * B2; e += NTH) {
    const int a = e % B1, b = e / B1;
    tile[b * TS + a] = S[base + a + b * s2];
  }
  __syncthreads();  // tile and twiddles
  {
    // Along axis 0: pencil p = tile row p (B2 of them), RA2 consecutive lanes.
    const int p = tid / RA2, j = tid % RA2;
    const bool on = p < B2;
    float2 v[RA1];
    if (on) {
#pragma unroll
      for (int r = 0; r < RA1; ++r) v[r] = tile[p * TS + j + RA2 * r];
      FA::phase1(v, j, twA);
    }
    float2 *pb = xch + (on ? p : 0) * PSA;
    if (on) {
#pragma unroll
      for (int s = 0; s < RA1; ++s) pb[xpos_c2c<RA1>(s * RA2 + j)] = v[s];
    }
    __syncthreads();
    if (on) {
#pragma unroll
      for (int e = 0; e < RA1; ++e) v[e] = pb[xpos_c2c<RA1>(j * RA1 + e)];
      FA::phase2(v);
#pragma unroll
      for (int u = 0; u < RAR; ++u)
#pragma unroll
        for (int q = 0; q < RA2; ++q) tile[p * TS + (j * RAR + u) + RA1 * q] = v[u * RA2 + q];
    }
  }
  __syncthreads();
  {
    // Along m2: pencil p = tile column p (B1 of them).
    const int p = tid / RB2, j = tid % RB2;
    const bool on = p < B1;
    float2 v[RB1];
    if (on) {
#pragma unroll
      for (int r = 0; r < RB1; ++r) v[r] = tile[(j + RB2 * r) * TS + p];
      FB::phase1(v, j, twB);
    }
    float2 *pb = xch + (on ? p : 0) * PSB;
    if (on) {
#pragma unroll
      for (int s = 0; s < RB1; ++s) pb[xpos_c2c<RB1>(s * RB2 + j)] = v[s];
    }
    __syncthreads();
    if (on) {
#pragma unroll
      for (int e = 0; e < RB1; ++e) v[e] = pb[xpos_c2c<RB1>(j * RB1 + e)];
      FB::phase2(v);
#pragma unroll
      for (int u = 0; u < RBR; ++u)
#pragma unroll
        for (int q = 0; q < RB2; ++q) tile[((j * RBR + u) + RB1 * q) * TS + p] = v[u * RB2 + q];
    }
  }
  __syncthreads();
  for (int e = tid; e < B1 * B2; e += NTH) {
    const int a = e % B1, b = e / B1;
    S[base + a + b * s2] = tile[b * TS + a];
  }
}

template <int DIR, int B1, int B2>
__global__ __launch_bounds__((blk2d_threads<B1, B2>())) void k_blk2d(float2 *S, int h0, int h1, int h2, int K, int m2,
                                                                  const float2 *__restrict__ tw1g,
                                                                  const float2 *__restrict__ tw2g) {
  blk2d_body<DIR, B1, B2>(S, h0, h1, h2, K, m2, tw1g, tw2g);
}

// The block FFTs of hourglasses 2 and 3 (K = 1, 2) of one scale in one
// launch when both take the same tile: blockIdx.y selects the hourglass.
struct Blk2dPair {
  float2 *S[2];
  int h[2][3], K[2], m2[2];
  unsigned nb[2];
};
template <int DIR, int B1, int B2>
__global__ __launch_bounds__((blk2d_threads<B1, B2>())) void k_blk2d_pair(Blk2dPair P,
                                                                       const float2 *__restrict__ tw1g,
                                                                       const float2 *__restrict__ tw2g) {
  const int y = blockIdx.y;
  if (blockIdx.x >= P.nb[y]) return;  // the whole block: before any barrier
  blk2d_body<DIR, B1, B2>(P.S[y], P.h[y][0], P.h[y][1], P.h[y][2], P.K[y], P.m2[y], tw1g, tw2g);
}

// X = complex(x) * scale, x read with element stride `is`.
__global__ void k_to_complex(const float *__restrict__ x, int64_t is, float2 *__restrict__ X, int64_t n, float scale) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    X[i] = make_float2(x[i * is] * scale, 0.0f);
}

// out[i * os] = Re(X[i]) * scale.
__global__ void k_real_out(const float2 *__restrict__ X, float *__restrict__ out, int64_t os, float scale, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i * os] = X[i].x * scale;
}

struct PyrDev {
  int n1, n2, n3, l1, l2, l3;        // spectrum and lowpass sizes
  int len1, len2, len3;              // box extents (2P + 1)
  const int *idx, *idx_low;          // concatenated per axis
  const float *prof, *comp;          // profile and 1 - profile, per axis
  float D15;
};

__device__ __forceinline__ void pyr_factors(const PyrDev &P, int a, int b, int c, int64_t &xi, int64_t &li, float &L,
                                            float &s) {
  const int ob = P.len1, oc = P.len1 + P.len2;
  xi = lin3(P.idx[a], P.idx[ob + b], P.idx[oc + c], P.n1, P.n2);
  li = lin3(P.idx_low[a], P.idx_low[ob + b], P.idx_low[oc + c], P.l1, P.l2);
  const float p1 = P.prof[a], p2 = P.prof[ob + b], p3 = P.prof[oc + c];
  L = (p1 * p2) * p3;
  // 1 - L from the FP64 complements, then 1 - L^2 = (1 - L)(1 + L).
  const float om = P.comp[a] + p1 * P.comp[ob + b] + (p1 * p2) * P.comp[oc + c];
  const float r = om * (1.0f + L);
  s = sqrtf(r > 0.0f ? r : 0.0f);
}

// Box coordinates along one axis whose lowpass index is l: the box is
// 0..P then nl - P..nl - 1 of the lowpass axis (box coordinate a <= P, then
// a > P), so l has none, one, or (when 2P = nl, at l = P) two, ascending.
__device__ __forceinline__ int box_coords(int l, int nl, int P, int c[2]) {
  int k = 0;
  if (l <= P) c[k++] = l;
  if (l >= nl - P) c[k++] = l - nl + 2 * P + 1;
  return k;
}

// low[idx_low] = L X[idx] / D^1.5 and X[idx] *= sqrt(1 - L^2), as one pass
// over the lowpass array, which also writes its zeros outside the box.  Each
// box element of X belongs to exactly one lowpass element; where two box
// elements share a lowpass element (2P = nl), the later one in box order
// gives it its value, as the reference's array assignment does.
__global__ void k_pyr_split(float2 *__restrict__ X, float2 *__restrict__ low, PyrDev P) {
  const int64_t n = (int64_t)P.l1 * P.l2 * P.l3;
  const int P1 = (P.len1 - 1) / 2, P2 = (P.len2 - 1) / 2, P3 = (P.len3 - 1) / 2;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % P.l1);
    const int64_t r = e / P.l1;
    const int j = (int)(r % P.l2), k = (int)(r / P.l2);
    int ca[2], cb[2], cc[2];
    const int na = box_coords(i, P.l1, P1, ca), nb = box_coords(j, P.l2, P2, cb), nc = box_coords(k, P.l3, P3, cc);
    float2 out = make_float2(0.0f, 0.0f);
    for (int ia = 0; ia < na; ++ia)
      for (int ib = 0; ib < nb; ++ib)
        for (int ic = 0; ic < nc; ++ic) {
          int64_t xi, li;
          float L, s;
          pyr_factors(P, ca[ia], cb[ib], cc[ic], xi, li, L, s);
          const float2 v = X[xi];
          const float f = L / P.D15;
          out = make_float2(v.x * f, v.y * f);
          X[xi] = make_float2(v.x * s, v.y * s);
        }
    low[e] = out;
  }
}

// X[idx] = sqrt(1 - L^2) X[idx] + L fac low[idx_low]: fac = D^1.5 merges
// (synthesis); fac = 1 / D^1.5 is the adjoint of the split.
__global__ void k_pyr_merge(float2 *__restrict__ X, const float2 *__restrict__ low, PyrDev P, float fac) {
  const int64_t n = (int64_t)P.len1 * P.len2 * P.len3;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int a = e % P.len1;
    const int64_t r = e / P.len1;
    const int b = r % P.len2, c = r / P.len2;
    int64_t xi, li;
    float L, s;
    pyr_factors(P, a, b, c, xi, li, L, s);
    const float2 v = X[xi], w = low[li];
    const float f = L * fac;
    X[xi] = make_float2(s * v.x + f * w.x, s * v.y + f * w.y);
  }
}

// S = X[half] . V . scale over the half-spectrum hd of axis K.
__global__ void k_masked(const float2 *__restrict__ X, int n1, int n2, float2 *__restrict__ S, const float *__restrict__ V,
                         int h1, int h2, int h3, float scale) {
  const int64_t n = (int64_t)h1 * h2 * h3;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int i1 = e % h1;
    const int64_t r = e / h1;
    const int i2 = r % h2, i3 = r / h2;
    const float2 v = X[lin3(i1, i2, i3, n1, n2)];
    const float m = V[e] * scale;
    S[e] = make_float2(v.x * m, v.y * m);
  }
}

// One IRC stage on S (dims h, hourglass axis K, decimated axis m, block b):
// mode 0 analysis A <- F0(A) A + F0(B) B, B <- F1(A) A + F1(B) B;
// mode 1 synthesis A <- G0(A) A + G1(A) B, B <- G0(B) A + G1(B) B
// (irc.jl:105-146); mode 2 the adjoint of analysis, the conjugate transpose
// A <- conj F0(A) A + conj F1(A) B, B <- conj F0(B) A + conj F1(B) B.
__global__ void k_irc(float2 *__restrict__ S, int h1, int h2, int h3, int K, int m, int b,
                      const float2 *__restrict__ f0, const float2 *__restrict__ f1, int mode) {
  const int hd[3] = {h1, h2, h3};
  const int nm = hd[m], nKh = hd[K];
  const int64_t npairs = (int64_t)h1 * h2 * h3 / 2;
  const int64_t stride_m = m == 0 ? 1 : (m == 1 ? (int64_t)h1 : (int64_t)h1 * h2);
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < npairs; q += (int64_t)gridDim.x * blockDim.x) {
    // q enumerates (index along m restricted to A halves, other indices).
    int idx[3];
    int64_t r = q;
    for (int d = 0; d < 3; ++d) {
      const int ext = d == m ? nm / 2 : hd[d];
      idx[d] = r % ext;
      r /= ext;
    }
    const int pi = idx[m];
    const int cA = (pi / b) * 2 * b + (pi % b), cB = cA + b;
    idx[m] = cA;
    const int64_t eA = lin3(idx[0], idx[1], idx[2], h1, h2);
    const int64_t eB = eA + (int64_t)b * stride_m;
    const int kK = idx[K];
    const float2 A = S[eA], B = S[eB];
    const float2 f0A = f0[kK + (int64_t)nKh * cA], f0B = f0[kK + (int64_t)nKh * cB];
    const float2 f1A = f1[kK + (int64_t)nKh * cA], f1B = f1[kK + (int64_t)nKh * cB];
    float2 nA, nB;
    if (mode == 0) {
      nA = cadd(cmul(f0A, A), cmul(f0B, B));
      nB = cadd(cmul(f1A, A), cmul(f1B, B));
    } else if (mode == 1) {
      nA = cadd(cmul(f0A, A), cmul(f1A, B));
      nB = cadd(cmul(f0B, A), cmul(f1B, B));
    } else {
      nA = cadd(cmul(conjf2(f0A), A), cmul(conjf2(f1A), B));
      nB = cadd(cmul(conjf2(f0B), A), cmul(conjf2(f1B), B));
    }
    S[eA] = nA;
    S[eB] = nB;
  }
}

// Where the real subbands of hourglass K of one scale live: the element of
// line (o-indices) at position t along K.
struct BlockMap {
  int K;
  int h[3];          // half-spectrum dims (h[K] = nK/2+1)
  int nK;
  int bs[3];         // block size along the other axes (bs[K] unused)
  int cnt[3];        // block counts (cnt[K] = 1)
  int sb[3];         // subband shape: sb[K] = nK, bs elsewhere
  const int64_t *off;  // element offset of each block's subband
};

__device__ __forceinline__ int64_t block_elem(const BlockMap &M, int o[3], int t) {
  int q[3], l[3];
  for (int d = 0; d < 3; ++d) {
    if (d == M.K) {
      q[d] = 0;
      l[d] = t;
    } else {
      q[d] = o[d] / M.bs[d];
      l[d] = o[d] - q[d] * M.bs[d];
    }
  }
  const int i = q[0] + M.cnt[0] * (q[1] + M.cnt[1] * q[2]);
  return M.off[i] + lin3(l[0], l[1], l[2], M.sb[0], M.sb[1]);
}

// Lines along K: for K = 0 consecutive lines are consecutive (i2, i3) rows;
// otherwise consecutive i1 for fixed remaining index.
__device__ __forceinline__ void line_index(const BlockMap &M, int64_t line, int o[3]) {
  const int K = M.K;
  if (K == 0) {
    o[0] = 0;
    o[1] = line % M.h[1];
    o[2] = line / M.h[1];
  } else if (K == 1) {
    o[1] = 0;
    o[0] = line % M.h[0];
    o[2] = line / M.h[0];
  } else {
    o[2] = 0;
    o[0] = line % M.h[0];
    o[1] = line / M.h[0];
  }
}

__device__ __forceinline__ int64_t half_elem(const BlockMap &M, const int o[3], int t) {
  int i[3] = {o[0], o[1], o[2]};
  i[M.K] = t;
  return lin3(i[0], i[1], i[2], M.h[0], M.h[1]);
}

// Both element maps are linear in t along a line: the offsets of a block's
// G lines at t = 0 go to LDS once, and element t of line g is base + t stride.
constexpr int LINES_MAX = 16;
struct LineBases {
  int64_t s[LINES_MAX], c[LINES_MAX];  // half-spectrum and coefficient offsets at t = 0
};
__device__ __forceinline__ void line_bases(const BlockMap &M, int64_t l0, int valid, LineBases &B, int tid) {
  if (tid < valid) {
    int o[3];
    line_index(M, l0 + tid, o);
    B.s[tid] = half_elem(M, o, 0);
    B.c[tid] = block_elem(M, o, 0);
  }
}
__device__ __forceinline__ int64_t half_stride(const BlockMap &M) {
  return M.K == 0 ? 1 : (M.K == 1 ? (int64_t)M.h[0] : (int64_t)M.h[0] * M.h[1]);
}
__device__ __forceinline__ int64_t block_stride(const BlockMap &M) {
  return M.K == 0 ? 1 : (M.K == 1 ? (int64_t)M.sb[0] : (int64_t)M.sb[0] * M.sb[1]);
}

// f(inner, outer) for idx = tid, tid + DT, ... < n_inner n_outer, with
// inner = idx mod n_inner and outer = idx / n_inner, without a division per
// element: one division for the thread's start, then steps with a carry.
template <class Fn> __device__ __forceinline__ void walk2(int tid, int n_inner, int n_outer, Fn f) {
  int outer = tid / n_inner, inner = tid - outer * n_inner;
  const int dq = DT / n_inner, dr = DT - dq * n_inner;
  while (outer < n_outer) {
    f(inner, outer);
    inner += dr;
    outer += dq;
    if (inner >= n_inner) {
      inner -= n_inner;
      ++outer;
    }
  }
}
// Element (t, g) of a block's G lines of length n in the order that keeps
// global accesses coalesced: t fastest along axis 0 (K = 0, lines contiguous
// in t), g fastest otherwise (adjacent lines adjacent in memory).
template <class Fn> __device__ __forceinline__ void walk_lines(int K, int tid, int n, int G, Fn f) {
  if (K == 0)
    walk2(tid, n, G, [&](int t, int g) { f(t, g); });
  else
    walk2(tid, G, n, [&](int g, int t) { f(t, g); });
}

// Subband lines = c2r along K of S (unnormalised), scattered into the
// coefficients (element stride os).  G lines per block, two per complex FFT:
// lines 2f and 2f + 1 have Hermitian spectra A and B, so the inverse of
// A + iB is line 2f + i line 2f + 1.
template <bool GEN>
__device__ __forceinline__ void c2r_lines_body(const float2 *__restrict__ S, float *__restrict__ out, int64_t os,
                                               const BlockMap &M, int G, const FftPlan1D &pl,
                                               const float2 *__restrict__ twg) {
  extern __shared__ float2 smem[];
  const int nK = M.nK, hK = M.h[M.K], F = (G + 1) / 2;
  float2 *tw = smem;
  float2 *st = tw + nK;
  float2 *a = st + (size_t)hK * G;
  float2 *b = a + (size_t)nK * F;
  const int tid = threadIdx.x;
  for (int k = tid; k < nK; k += DT) tw[k] = twg[pl.tw + k];
  const int64_t nlines = (int64_t)M.h[0] * M.h[1] * M.h[2] / hK;
  const int64_t l0 = (int64_t)blockIdx.x * G;
  const int valid = (int)min((int64_t)G, nlines - l0);
  __shared__ LineBases LB;
  line_bases(M, l0, valid, LB, tid);
  const int64_t sS = half_stride(M), sC = block_stride(M);
  __syncthreads();
  walk_lines(M.K, tid, hK, G, [&](int t, int g) {
    st[t * G + g] = g < valid ? S[LB.s[g] + t * sS] : make_float2(0.0f, 0.0f);
  });
  __syncthreads();
  const bool even = (nK & 1) == 0;
  // Bin t of line g through the Hermitian extension (bins 0 and nK/2 real).
  auto herm = [&](int t, int g) {
    if (g >= G) return make_float2(0.0f, 0.0f);
    float2 z;
    if (t < hK) {
      z = st[t * G + g];
      if (t == 0 || (even && t == nK / 2)) z.y = 0.0f;
    } else {
      z = conjf2(st[(nK - t) * G + g]);
    }
    return z;
  };
  walk2(tid, F, nK, [&](int f, int t) {
    const float2 A = herm(t, 2 * f), B = herm(t, 2 * f + 1);
    a[t * F + f] = make_float2(A.x - B.y, A.y + B.x);
  });
  __syncthreads();
  float2 *res = lds_fft<1, GEN>(a, b, pl, F, tw, tid, DT);
  walk_lines(M.K, tid, nK, G, [&](int t, int g) {
    if (g >= valid) return;
    const float2 z = res[t * F + (g >> 1)];
    out[(LB.c[g] + t * sC) * os] = (g & 1) ? z.y : z.x;
  });
}

template <bool GEN>
__global__ __launch_bounds__(DT) void k_c2r_lines(const float2 *__restrict__ S, float *__restrict__ out, int64_t os,
                                                 BlockMap M, int G, FftPlan1D pl, const float2 *__restrict__ twg) {
  c2r_lines_body<GEN>(S, out, os, M, G, pl, twg);
}

// The line transforms of the three hourglasses of one scale in one launch
// (blockIdx.y = K) when their lines have one length: S[K] is hourglass K's
// half spectrum, M[K] its block map, nb[K] its blocks of G lines.
struct Lines3 {
  float2 *S[3];
  BlockMap M[3];
  unsigned nb[3];
};
template <bool GEN>
__global__ __launch_bounds__(DT) void k_c2r_lines3(Lines3 L, float *__restrict__ out, int64_t os, int G, FftPlan1D pl,
                                                  const float2 *__restrict__ twg) {
  const int K = blockIdx.y;
  if (blockIdx.x >= L.nb[K]) return;  // the whole block: before any barrier
  c2r_lines_body<GEN>(L.S[K], out, os, L.M[K], G, pl, twg);
}

// S lines = r2c along K of the subbands' lines (the first nK/2+1 bins).
// With `dbl` the interior bins 1 .. nK/2 - 1 are doubled: the adjoint of the
// unnormalised c2r along K.  Lines 2f and 2f + 1 share one complex FFT
// Z = FFT(line 2f + i line 2f + 1): A[t] = (Z[t] + conj Z[-t]) / 2 and
// B[t] = (Z[t] - conj Z[-t]) / 2i.
template <bool GEN>
__device__ __forceinline__ void r2c_lines_body(const float *__restrict__ in, int64_t is, float2 *__restrict__ S,
                                               const BlockMap &M, int G, const FftPlan1D &pl,
                                               const float2 *__restrict__ twg, int dbl) {
  extern __shared__ float2 smem[];
  const int nK = M.nK, hK = M.h[M.K], F = (G + 1) / 2;
  float2 *tw = smem;
  float2 *a = tw + nK;
  float2 *b = a + (size_t)nK * F;
  const int tid = threadIdx.x;
  for (int k = tid; k < nK; k += DT) tw[k] = twg[pl.tw + k];
  const int64_t nlines = (int64_t)M.h[0] * M.h[1] * M.h[2] / hK;
  const int64_t l0 = (int64_t)blockIdx.x * G;
  const int valid = (int)min((int64_t)G, nlines - l0);
  __shared__ LineBases LB;
  line_bases(M, l0, valid, LB, tid);
  const int64_t sS = half_stride(M), sC = block_stride(M);
  __syncthreads();
  float *af = reinterpret_cast<float *>(a);
  // Component g & 1 of FFT g / 2 (slot g = G of an odd G is zero).
  walk_lines(M.K, tid, nK, 2 * F, [&](int t, int g) {
    af[2 * (t * F + (g >> 1)) + (g & 1)] = g < valid ? in[(LB.c[g] + t * sC) * is] : 0.0f;
  });
  __syncthreads();
  float2 *res = lds_fft<-1, GEN>(a, b, pl, F, tw, tid, DT);
  walk_lines(M.K, tid, hK, G, [&](int t, int g) {
    if (g >= valid) return;
    const int f = g >> 1;
    const float2 Z = res[t * F + f], Zm = res[(t == 0 ? 0 : nK - t) * F + f];
    float2 v = (g & 1) ? make_float2(0.5f * (Z.y + Zm.y), 0.5f * (Zm.x - Z.x))
                       : make_float2(0.5f * (Z.x + Zm.x), 0.5f * (Z.y - Zm.y));
    if (dbl && t >= 1 && t < nK - t) v = cscale(v, 2.0f);
    S[LB.s[g] + t * sS] = v;
  });
}

template <bool GEN>
__global__ __launch_bounds__(DT) void k_r2c_lines(const float *__restrict__ in, int64_t is, float2 *__restrict__ S,
                                                 BlockMap M, int G, FftPlan1D pl, const float2 *__restrict__ twg,
                                                 int dbl) {
  r2c_lines_body<GEN>(in, is, S, M, G, pl, twg, dbl);
}

template <bool GEN>
__global__ __launch_bounds__(DT) void k_r2c_lines3(Lines3 L, const float *__restrict__ in, int64_t is, int G,
                                                  FftPlan1D pl, const float2 *__restrict__ twg, int dbl) {
  const int K = blockIdx.y;
  if (blockIdx.x >= L.nb[K]) return;  // the whole block: before any barrier
  r2c_lines_body<GEN>(in, is, L.S[K], L.M[K], G, pl, twg, dbl);
}

// All IRC stages of hourglass K at one scale in one pass.  The stages along
// axis m (level le = 1..L, block size b = h_m >> le) pair elements at
// distance b, so the 2^L elements i_m = r + a (h_m >> L), a < 2^L, of one
// residue r are closed under them: stage le acts on bit L - le of a.  A
// thread holds such a group for both transverse axes m1 < m2 (group index
// g = a + 2^L1 b; stage le of m2 acts on bit L1 + L2 - le) and one index
// along K, and applies the stages in registers with the arithmetic of
// k_irc: mode 0 analysis in stage order, on X[half_K] V_K scale loaded here
// (k_masked); modes 1 (synthesis) and 2 (adjoint) in reverse stage order, in
// place on S.
struct IrcGroup {
  const float2 *f0[8], *f1[8];  // per stage in stage order: m1 levels 1..L1, then m2 levels 1..L2
  int sc[8], sk[8];             // value (k, c) of a plane at c sc + k sk
};

template <int L1, int L2>
__device__ __forceinline__ void irc_group_body(const float2 *__restrict__ X, int n1, int n2,
                                               const float *__restrict__ V, float scale, float2 *__restrict__ S,
                                               int h1, int h2, int h3, int K, int m1, int m2, const IrcGroup &F,
                                               int mode) {
  constexpr int T = L1 + L2, NG = 1 << T;
  const int hd[3] = {h1, h2, h3};
  const int B1 = hd[m1] >> L1, B2 = hd[m2] >> L2, nKh = hd[K];
  const int64_t ngroups = (int64_t)hd[K] * B1 * B2;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < ngroups; q += (int64_t)gridDim.x * blockDim.x) {
    // The axis that is contiguous in memory varies fastest: K when K = 0,
    // else m1 (= 0).
    int iK, r1, r2;
    if (K == 0) {
      iK = (int)(q % nKh);
      const int64_t t = q / nKh;
      r1 = (int)(t % B1);
      r2 = (int)(t / B1);
    } else {
      r1 = (int)(q % B1);
      const int64_t t = q / B1;
      iK = (int)(t % nKh);
      r2 = (int)(t / nKh);
    }
    auto coords = [&](int g, int idx[3]) {
      idx[K] = iK;
      idx[m1] = r1 + (g & ((1 << L1) - 1)) * B1;
      idx[m2] = r2 + (g >> L1) * B2;
    };
    float2 v[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      int idx[3];
      coords(g, idx);
      const int64_t e = lin3(idx[0], idx[1], idx[2], h1, h2);
      if (mode == 0) {
        const float2 x = X[lin3(idx[0], idx[1], idx[2], n1, n2)];
        const float m = V[e] * scale;
        v[g] = make_float2(x.x * m, x.y * m);
      } else {
        v[g] = S[e];
      }
    }
    // Stage s (stage order) acts on bit sb(s) of g along axis m(s); the
    // coordinate of element g along that axis is c(g).
    auto stage = [&](int s) {
      const bool on1 = s < L1;
      const int bit = on1 ? L1 - 1 - s : T - 1 - (s - L1);
      const float2 *f0 = F.f0[s], *f1 = F.f1[s];
      const int sc = F.sc[s], sk = F.sk[s];
      auto at = [&](const float2 *f, int c) { return f[(int64_t)c * sc + (int64_t)iK * sk]; };
#pragma unroll
      for (int gA = 0; gA < NG; ++gA) {
        if (gA & (1 << bit)) continue;
        const int gB = gA | (1 << bit);
        const int cA = on1 ? r1 + (gA & ((1 << L1) - 1)) * B1 : r2 + (gA >> L1) * B2;
        const int cB = on1 ? r1 + (gB & ((1 << L1) - 1)) * B1 : r2 + (gB >> L1) * B2;
        const float2 A = v[gA], B = v[gB];
        const float2 f0A = at(f0, cA), f0B = at(f0, cB);
        const float2 f1A = at(f1, cA), f1B = at(f1, cB);
        if (mode == 0) {
          v[gA] = cadd(cmul(f0A, A), cmul(f0B, B));
          v[gB] = cadd(cmul(f1A, A), cmul(f1B, B));
        } else if (mode == 1) {
          v[gA] = cadd(cmul(f0A, A), cmul(f1A, B));
          v[gB] = cadd(cmul(f0B, A), cmul(f1B, B));
        } else {
          v[gA] = cadd(cmul(conjf2(f0A), A), cmul(conjf2(f1A), B));
          v[gB] = cadd(cmul(conjf2(f0B), A), cmul(conjf2(f1B), B));
        }
      }
    };
    if (mode == 0) {
#pragma unroll
      for (int s = 0; s < T; ++s) stage(s);
    } else {
#pragma unroll
      for (int s = T - 1; s >= 0; --s) stage(s);
    }
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      int idx[3];
      coords(g, idx);
      S[lin3(idx[0], idx[1], idx[2], h1, h2)] = v[g];
    }
  }
}

template <int L1, int L2>
__global__ __launch_bounds__(DT) void k_irc_group(const float2 *__restrict__ X, int n1, int n2,
                                                  const float *__restrict__ V, float scale, float2 *__restrict__ S,
                                                  int h1, int h2, int h3, int K, int m1, int m2, IrcGroup F,
                                                  int mode) {
  irc_group_body<L1, L2>(X, n1, n2, V, scale, S, h1, h2, h3, K, m1, m2, F, mode);
}

// The grouped IRC passes of the three hourglasses of one scale in one launch
// (blockIdx.y = K) when all three have the same (L1, L2).
struct IrcGroup3 {
  const float *V[3];
  float scale[3];
  float2 *S[3];
  int h[3][3], m1[3], m2[3];
  IrcGroup F[3];
};
template <int L1, int L2>
__global__ __launch_bounds__(DT) void k_irc_group3(const float2 *__restrict__ X, int n1, int n2, IrcGroup3 P, int mode) {
  const int K = blockIdx.y;
  irc_group_body<L1, L2>(X, n1, n2, P.V[K], P.scale[K], P.S[K], P.h[K][0], P.h[K][1], P.h[K][2], K, P.m1[K],
                         P.m2[K], P.F[K], mode);
}

// X = sum over K of c_K M_K S_K on the half of axis K, plus (mirror) c_K M_K
// conj(S_K) on its strictly interior mirror, index i -> (n - i) mod n: the
// sum of one scale (surfacelet_scale.jl:246-257) as one write of X: for each
// element the contributions are added in K order, and at most one of the two
// (direct, mirrored) per K reaches it.
struct HermSum {
  const float2 *S[3];
  const float *M[3];
  float c[3];
};

// Source index along one axis of the Hermitian mirror X[-w] += c M conj(S[w])
// (src/directional/surfacelet_scale.jl:246-257, whose mirror of 1-based i is
// mod(1 - i, n) + 1): on axis K an element j >= n/2 + 1 reads n - j, and -1
// marks j on the direct half; on the other axes every j reads (n - j) mod n.
__host__ __device__ __forceinline__ int herm_mirror_src(int j, int n, bool onK) {
  if (onK) return j < n / 2 + 1 ? -1 : n - j;
  return (n - j) % n;
}

// With `merge`, the pyramid merge of this level follows in the same pass:
// an element inside the box becomes sqrt(1 - L^2) X + L fac low[idx_low]
// (k_pyr_merge's arithmetic, src/pyramid/fourier_pyramid.jl:122-127); the
// box is 0..P and n - P..n - 1 along each axis, so its coordinate is found
// in O(1) and each box element reads one lowpass element.
__global__ void k_hermitian_sum(float2 *__restrict__ X, int n1, int n2, int n3, HermSum H, int mirror,
                                const float2 *__restrict__ low, PyrDev P, float fac, int merge) {
  const int n[3] = {n1, n2, n3};
  const int64_t tot = (int64_t)n1 * n2 * n3;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    int j[3];
    j[0] = (int)(e % n1);
    const int64_t r = e / n1;
    j[1] = (int)(r % n2);
    j[2] = (int)(r / n2);
    float2 acc = make_float2(0.0f, 0.0f);
#pragma unroll
    for (int K = 0; K < 3; ++K) {
      const int hK = n[K] / 2 + 1;
      int h[3] = {n1, n2, n3};
      h[K] = hK;
      if (j[K] < hK) {
        const int64_t s = lin3(j[0], j[1], j[2], h[0], h[1]);
        const float w = H.c[K] * H.M[K][s];
        const float2 v = H.S[K][s];
        acc = make_float2(acc.x + w * v.x, acc.y + w * v.y);
      } else if (mirror) {
        int i[3];
        for (int d = 0; d < 3; ++d) i[d] = herm_mirror_src(j[d], n[d], d == K);
        const int64_t s = lin3(i[0], i[1], i[2], h[0], h[1]);
        const float w = H.c[K] * H.M[K][s];
        const float2 v = H.S[K][s];
        acc = make_float2(acc.x + w * v.x, acc.y - w * v.y);
      }
    }
    if (merge) {
      const int len[3] = {P.len1, P.len2, P.len3};
      int bc[3];
      bool in = true;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int Pd = (len[d] - 1) / 2;
        bc[d] = j[d] <= Pd ? j[d] : (j[d] >= n[d] - Pd ? j[d] - n[d] + 2 * Pd + 1 : -1);
        in = in && bc[d] >= 0;
      }
      if (in) {
        int64_t xi, li;
        float L, sq;
        pyr_factors(P, bc[0], bc[1], bc[2], xi, li, L, sq);
        const float2 w = low[li];
        const float f = L * fac;
        acc = make_float2(sq * acc.x + f * w.x, sq * acc.y + f * w.y);
      }
    }
    X[e] = acc;
  }
}

// Integer maps for inspection (surfhip_s_index_maps), from the same device
// functions the transform kernels use: kind 2, the box coordinates
// box_coords gives each lowpass bin (two per bin, -1 when absent); kind 3,
// herm_mirror_src along one axis.
__global__ void k_index_map(int kind, int n, int nl, int P, int onK, int64_t *__restrict__ out) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    if (kind == 2) {
      int c[2];
      const int k = box_coords(e, nl, P, c);
      out[2 * (int64_t)e] = k > 0 ? c[0] : -1;
      out[2 * (int64_t)e + 1] = k > 1 ? c[1] : -1;
    } else {
      out[e] = herm_mirror_src(e, n, onK != 0);
    }
  }
}

// The whole 3-D FFT of a 16 x 16 x 16 volume in one block of 256 threads
// (the coarse lowpass of the pyramids): the volume in LDS (row pitch 17, so
// no pass meets bank conflicts), one 16-point pencil per thread per axis.
// io 1: real input rio.in[e stride] scale; io 2: the real part, times
// rio.scale, to rio.out[e stride]; io 0: complex A in and out.
template <int DIR>
__global__ __launch_bounds__(256) void k_fft3_16(float2 *__restrict__ A, RealIO rio, int io) {
  constexpr int P1 = 17, P2 = 17 * 17;
  __shared__ float2 V[16 * P2];
  const int t = threadIdx.x;
  for (int e = t; e < 4096; e += 256) {
    const int i0 = e & 15, i1 = (e >> 4) & 15, i2 = e >> 8;
    V[i0 + P1 * i1 + P2 * i2] =
        io == 1 ? make_float2(rio.in[(int64_t)e * rio.stride] * rio.scale, 0.0f) : A[e];
  }
  __syncthreads();
  const int a = t & 15, b = t >> 4;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const int base = ax == 0 ? P1 * a + P2 * b : (ax == 1 ? a + P2 * b : a + P1 * b);
    const int step = ax == 0 ? 1 : (ax == 1 ? P1 : P2);
    float2 v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = V[base + k * step];
    dft16<DIR>(v);
#pragma unroll
    for (int k = 0; k < 16; ++k) V[base + k * step] = v[k];
    __syncthreads();
  }
  for (int e = t; e < 4096; e += 256) {
    const int i0 = e & 15, i1 = (e >> 4) & 15, i2 = e >> 8;
    const float2 z = V[i0 + P1 * i1 + P2 * i2];
    if (io == 2)
      rio.out[(int64_t)e * rio.stride] = z.x * rio.scale;
    else
      A[e] = z;
  }
}

int grid_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + DT - 1) / DT, 4096)); }

}  // namespace

}  // namespace surfhip

using namespace surfhip;

struct DecLevelDev {
  PyrDev pd;
  DevBuf<int> idx, idx_low;
  DevBuf<float> prof, comp;
};

struct DecScaleDev {
  std::array<DevBuf<int64_t>, 3> boff;  // per K: subband offset of each block
  std::array<int, 3> first;            // 0-based subband index of each K's first block
};

struct surfhip_s_plan {
  DecTables T;
  int dtype = SURFHIP_F32, nch = 1;
  hipStream_t stream = nullptr;
  // The three hourglasses of a scale are independent until their spectra are
  // summed: their passes run on one stream each, forked from and joined back
  // to `stream` with events.  `ls` is the stream the launch helpers use.
  hipStream_t ls = nullptr;
  hipStream_t kst[3] = {nullptr, nullptr, nullptr};
  hipEvent_t kfork = nullptr, kjoin[3] = {nullptr, nullptr, nullptr};

  // Launch recording.  A pair issues about a hundred launches, which costs
  // the host about as long as the GPU takes to run them.  So a call that
  // repeats the previous call's data pointers is recorded once as a hipGraph
  // (kernel nodes added explicitly, no stream capture) and replayed with one
  // hipGraphLaunch.  While `rec` is set, go() adds a node after the previous
  // nodes of lane `lane` (0 = `stream`, K + 1 = hourglass K's stream), and
  // fork_k/join_k add the edges their events stand for.
  struct Rec {
    hipGraph_t g = nullptr;
    std::array<std::vector<hipGraphNode_t>, 4> deps;
    std::vector<std::shared_ptr<void>> keep;  // kernel arguments of the nodes
  };
  Rec *rec = nullptr;
  int lane = 0;
  void use_lane(int L) {
    lane = L;
    ls = L == 0 ? stream : kst[L - 1];
  }
  template <class Tup, size_t... I> static void arg_ptrs(Tup &t, void **kp, std::index_sequence<I...>) {
    ((kp[I] = (void *)&std::get<I>(t)), ...);
  }
  template <typename... P, typename... A>
  void go(void (*k)(P...), dim3 grid, dim3 block, size_t lds, A... a) {
    static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
    auto t = std::make_shared<std::tuple<P...>>(static_cast<P>(a)...);
    void *kp[sizeof...(P) > 0 ? sizeof...(P) : 1];
    arg_ptrs(*t, kp, std::index_sequence_for<P...>{});
    if (!rec) {
      SH_HIP(hipLaunchKernel((const void *)k, grid, block, kp, lds, ls));
      return;
    }
    hipKernelNodeParams p{};
    p.func = (void *)k;
    p.gridDim = grid;
    p.blockDim = block;
    p.sharedMemBytes = (unsigned)lds;
    p.kernelParams = kp;
    p.extra = nullptr;
    hipGraphNode_t n;
    std::vector<hipGraphNode_t> &d = rec->deps[lane];
    SH_HIP(hipGraphAddKernelNode(&n, rec->g, d.empty() ? nullptr : d.data(), d.size(), &p));
    d.assign(1, n);
    rec->keep.push_back(t);
  }
  // One recorded call per direction (0 analyze, 1 synthesize, 2 adjoint).
  // `done` is recorded after every replay: an executable graph is destroyed
  // only once its last replay has finished, never while it may still run.
  struct Cached {
    const void *in = nullptr, *out = nullptr;
    int seen = 0;
    Rec rec;
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;
    bool pending = false;
  };
  Cached cache[3];
  static void drop(Cached &c) {
    if (c.pending && c.done) (void)hipEventSynchronize(c.done);
    c.pending = false;
    if (c.exec) (void)hipGraphExecDestroy(c.exec);
    if (c.rec.g) (void)hipGraphDestroy(c.rec.g);
    c.exec = nullptr;
    c.rec = Rec{};
  }
  // A recorded call replaced by a call with other pointers is kept, not
  // destroyed, until the plan is destroyed (at most kRetired of them; beyond
  // that the oldest is destroyed once its last replay has completed).  A
  // test that destroyed replaced graphs while the plan lived ended twice in
  // a host segmentation fault in a runtime thread a few calls later, with no
  // cause found in this code.
  static constexpr size_t kRetired = 8;
  std::vector<Cached> retired;
  void retire(Cached &c) {
    if (!c.exec && !c.rec.g) return;
    retired.push_back(std::move(c));
    c = Cached{};
    if (retired.size() > kRetired) {
      drop(retired.front());
      if (retired.front().done) (void)hipEventDestroy(retired.front().done);
      retired.erase(retired.begin());
    }
  }
  // body() issues the call's launches.  Direct launches unless the call
  // repeats the previous one's pointers; a stream being captured by the
  // caller always gets direct launches.
  template <class F> void run(int kind, const void *in, const void *out, F body) {
    Cached &c = cache[kind];
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &st) != hipSuccess) {
      (void)hipGetLastError();
      st = hipStreamCaptureStatusActive;
    }
    if (st != hipStreamCaptureStatusNone) {
      body();
      return;
    }
    if (c.in != in || c.out != out) {
      retire(c);
      c.in = in;
      c.out = out;
      c.seen = 0;
    }
    if (!c.exec && ++c.seen >= 2) {
      SH_HIP(hipGraphCreate(&c.rec.g, 0));
      rec = &c.rec;
      try {
        body();
      } catch (...) {
        rec = nullptr;
        drop(c);
        throw;
      }
      rec = nullptr;
      SH_HIP(hipGraphInstantiate(&c.exec, c.rec.g, nullptr, nullptr, 0));
    }
    if (c.exec) {
      if (!c.done) SH_HIP(hipEventCreateWithFlags(&c.done, hipEventDisableTiming));
      SH_HIP(hipGraphLaunch(c.exec, stream));
      SH_HIP(hipEventRecord(c.done, stream));
      c.pending = true;
      return;
    }
    body();
  }

  ~surfhip_s_plan() {
    for (Cached &c : cache) {
      drop(c);
      if (c.done) (void)hipEventDestroy(c.done);
    }
    for (Cached &c : retired) {
      drop(c);
      if (c.done) (void)hipEventDestroy(c.done);
    }
    for (int K = 0; K < 3; ++K) {
      if (kst[K]) (void)hipStreamDestroy(kst[K]);
      if (kjoin[K]) (void)hipEventDestroy(kjoin[K]);
    }
    if (kfork) (void)hipEventDestroy(kfork);
  }
  void make_lanes() {
    for (int K = 0; K < 3; ++K) {
      SH_HIP(hipStreamCreateWithFlags(&kst[K], hipStreamNonBlocking));
      SH_HIP(hipEventCreateWithFlags(&kjoin[K], hipEventDisableTiming));
    }
    SH_HIP(hipEventCreateWithFlags(&kfork, hipEventDisableTiming));
  }
  // Later work of the hourglass streams follows everything already on `stream`.
  void fork_k() {
    if (rec) {
      // A lane keeps the order of its own earlier nodes, as a stream does.
      for (int L = 1; L <= 3; ++L) {
        std::vector<hipGraphNode_t> &d = rec->deps[L];
        d.insert(d.end(), rec->deps[0].begin(), rec->deps[0].end());
        std::sort(d.begin(), d.end());
        d.erase(std::unique(d.begin(), d.end()), d.end());
      }
      return;
    }
    SH_HIP(hipEventRecord(kfork, stream));
    for (int K = 0; K < 3; ++K) SH_HIP(hipStreamWaitEvent(kst[K], kfork, 0));
  }
  // Later work of `stream` follows everything already on the hourglass streams.
  void join_k() {
    if (rec) {
      std::vector<hipGraphNode_t> &d0 = rec->deps[0];
      for (int L = 1; L <= 3; ++L) {
        d0.insert(d0.end(), rec->deps[L].begin(), rec->deps[L].end());
        rec->deps[L].clear();
      }
      std::sort(d0.begin(), d0.end());
      d0.erase(std::unique(d0.begin(), d0.end()), d0.end());
      return;
    }
    for (int K = 0; K < 3; ++K) {
      SH_HIP(hipEventRecord(kjoin[K], kst[K]));
      SH_HIP(hipStreamWaitEvent(stream, kjoin[K], 0));
    }
  }
  // Later work of `stream` follows everything already on lane L (1..3;
  // nothing for lane 0, `stream` itself).
  void join_lane(int L) {
    if (L == 0) return;
    if (rec) {
      std::vector<hipGraphNode_t> &d0 = rec->deps[0];
      d0.insert(d0.end(), rec->deps[L].begin(), rec->deps[L].end());
      std::sort(d0.begin(), d0.end());
      d0.erase(std::unique(d0.begin(), d0.end()), d0.end());
      return;
    }
    SH_HIP(hipEventRecord(kjoin[L - 1], kst[L - 1]));
    SH_HIP(hipStreamWaitEvent(stream, kjoin[L - 1], 0));
  }
  // The lane of scale j's merged hourglass passes: one lane per scale (of
  // three).  Any branch makes a recorded graph's replay cost the host more
  // (at 128^3 per pair: about 0.08 ms with no branch, 0.25 ms with one,
  // 0.3 ms with one per scale), but the overlap pays: the GPU then needs
  // about 0.40, 0.35 and 0.31 ms.
  static int scale_lane(int j) { return 1 + j % 3; }
  std::vector<std::unique_ptr<DecLevelDev>> lev;
  std::vector<std::unique_ptr<DecScaleDev>> sc;
  DevBuf<float> masks;
  DevBuf<float2> filt, tw;
  DevBuf<float2> filt_t;  // transposed planes of the stages along axis 0 of hourglasses 2 and 3
  std::vector<DevBuf<float2>> spectra;  // [j] the spectrum entering level j; [J] the coarse lowpass
  // Half spectra of the three hourglasses, one set per scale, so that the
  // scales' hourglass passes may overlap.
  std::vector<std::array<DevBuf<float2>, 3>> Sbuf;
  std::vector<std::pair<int, FftPlan1D>> plans;

  const FftPlan1D &plan_for(int n) const {
    for (auto &p : plans)
      if (p.first == n) return p.second;
    fail(SURFHIP_ERR_DEVICE, fmt("no FFT plan for length %d", n));
  }

  static bool reg_len(int n) { return n == 16 || n == 32 || n == 64 || n == 128; }
  // io: 0 complex A in and out; 1 (axis 0, forward) real input rio; 2 (axis 2,
  // inverse) real output rio.  io != 0 needs a register-FFT length.
  void launch_c2c(int axis, int dir, float2 *A, Dims3 m, int seg, int io = 0, RealIO rio = RealIO{}) {
    const FftPlan1D &pl = plan_for(seg);
    if (io == 1 && !(axis == 0 && dir < 0 && reg_len(seg))) fail(SURFHIP_ERR_DEVICE, "internal: real-input FFT pass");
    if (io == 2 && !(axis == 2 && dir > 0 && reg_len(seg))) fail(SURFHIP_ERR_DEVICE, "internal: real-output FFT pass");
    if (reg_len(seg)) {
      const int64_t npen = (int64_t)m[0] * m[1] * m[2] / seg;
      const float2 *twp = tw.p + pl.tw;
      if (io) {
#define SH_C2CIO(AX, D, SG, IO)                                                                                 \
  go((k_c2c_reg<AX, D, SG, IO>), dim3((unsigned)((npen + reg_tp<SG>() - 1) / reg_tp<SG>())), dim3(DT), 0, A, m[0], \
     m[1], m[2], twp, rio)
#define SH_C2CIO_SEG(AX, D, IO)                 \
  switch (seg) {                                \
    case 16: SH_C2CIO(AX, D, 16, IO); break;    \
    case 32: SH_C2CIO(AX, D, 32, IO); break;    \
    case 64: SH_C2CIO(AX, D, 64, IO); break;    \
    default: SH_C2CIO(AX, D, 128, IO); break;   \
  }
        if (io == 1) { SH_C2CIO_SEG(1, -1, 1) } else { SH_C2CIO_SEG(3, 1, 2) }
#undef SH_C2CIO_SEG
#undef SH_C2CIO
        return;
      }
#define SH_C2CR(AX, D, SG)                                                                                       \
  go((k_c2c_reg<AX, D, SG>), dim3((unsigned)((npen + reg_tp<SG>() - 1) / reg_tp<SG>())), dim3(DT), 0, A, m[0], m[1], m[2], twp, \
     RealIO{})
#define SH_C2CR_SEG(AX, D)                   \
  switch (seg) {                             \
    case 16: SH_C2CR(AX, D, 16); break;      \
    case 32: SH_C2CR(AX, D, 32); break;      \
    case 64: SH_C2CR(AX, D, 64); break;      \
    default: SH_C2CR(AX, D, 128); break;     \
  }
      if (axis == 0) { if (dir < 0) { SH_C2CR_SEG(1, -1) } else { SH_C2CR_SEG(1, 1) } }
      else if (axis == 1) { if (dir < 0) { SH_C2CR_SEG(2, -1) } else { SH_C2CR_SEG(2, 1) } }
      else { if (dir < 0) { SH_C2CR_SEG(3, -1) } else { SH_C2CR_SEG(3, 1) } }
#undef SH_C2CR_SEG
#undef SH_C2CR
      return;
    }
    int G = std::max(1, std::min(16, 4096 / seg));
    const size_t lds = sizeof(float2) * ((size_t)seg + 2 * (size_t)seg * G);
    if (lds > 65536) fail(SURFHIP_ERR_ARG, fmt("axis length %d is too long for the decimated path", seg));
    int64_t blocks;
    if (axis == 0) {
      const int64_t npen = (int64_t)(m[0] / seg) * m[1] * m[2];
      blocks = (npen + G - 1) / G;
    } else {
      const int64_t ub = (m[0] + G - 1) / G;
      blocks = axis == 1 ? ub * (m[1] / seg) * m[2] : ub * m[1] * (m[2] / seg);
    }
#define SH_C2C(AX, D)                                                                                       \
  go(fft_plan_generic(pl) ? &k_c2c<AX, D, true> : &k_c2c<AX, D, false>, dim3((unsigned)blocks), dim3(DT), lds, A, \
     A, m[0], m[1], m[2], seg, G, pl, tw.p)
    if (axis == 0) { if (dir < 0) SH_C2C(1, -1); else SH_C2C(1, 1); }
    else if (axis == 1) { if (dir < 0) SH_C2C(2, -1); else SH_C2C(2, 1); }
    else { if (dir < 0) SH_C2C(3, -1); else SH_C2C(3, 1); }
#undef SH_C2C
  }
  void fft3(int dir, float2 *A, Dims3 m) {
    for (int ax = 0; ax < 3; ++ax) launch_c2c(ax, dir, A, m, m[ax]);
  }
  // The block FFTs of hourglass K on S (dims hd, block sizes bsize) along
  // the two axes other than K: one tile pass when K != 0 and the tile fits,
  // else one segmented pass per axis.
  // Tile sizes (B1 along axis 0, B2 along the other block axis m2) of the
  // one-pass 2-D block FFT of hourglass K, or false when K takes one
  // segmented pass per axis (K = 0: axis 0 is its own axis; tiles that are
  // not register-FFT lengths or exceed 4096 elements).
  static bool blk_tile(const Dims3 &bsize, int K, int &B1, int &B2, int &m2) {
    m2 = K == 1 ? 2 : 1;
    B1 = bsize[0];
    B2 = bsize[m2];
    return K != 0 && reg_len(B1) && reg_len(B2) && B1 * B2 <= 4096;
  }
// Dispatch a tile launch LAUNCH(D, A, B) on the runtime tile B1 x B2.
#define SH_BLK_B(LAUNCH, D, A)          \
  switch (B2) {                         \
    case 16: LAUNCH(D, A, 16); break;   \
    case 32: LAUNCH(D, A, 32); break;   \
    case 64: LAUNCH(D, A, 64); break;   \
    default: LAUNCH(D, A, 128); break;  \
  }
#define SH_BLK_A(LAUNCH, D)                                    \
  switch (B1) {                                                \
    case 16: SH_BLK_B(LAUNCH, D, 16) break;                    \
    case 32: SH_BLK_B(LAUNCH, D, 32) break;                    \
    case 64:                                                   \
      switch (B2) {                                            \
        case 16: LAUNCH(D, 64, 16); break;                     \
        case 32: LAUNCH(D, 64, 32); break;                     \
        default: LAUNCH(D, 64, 64); break;                     \
      }                                                        \
      break;                                                   \
    default:                                                   \
      if (B2 == 16) LAUNCH(D, 128, 16); else LAUNCH(D, 128, 32); \
      break;                                                   \
  }
  // The block FFTs of hourglass K on S (dims hd, block sizes bsize) along
  // the two axes other than K: one tile pass when K != 0 and the tile fits,
  // else one segmented pass per axis.
  void block_ffts(int dir, float2 *S, Dims3 hd, Dims3 bsize, int K) {
    int B1, B2, m2;
    if (blk_tile(bsize, K, B1, B2, m2)) {
      const int64_t nb = (int64_t)(hd[0] / B1) * (hd[m2] / B2) * hd[K];
      const float2 *t1 = tw.p + plan_for(B1).tw, *t2 = tw.p + plan_for(B2).tw;
#define SH_BLK(D, A, B) go((k_blk2d<D, A, B>), dim3((unsigned)nb), dim3(blk2d_threads<A, B>()), 0, S, hd[0], hd[1], hd[2], K, m2, t1, t2)
      if (dir < 0) { SH_BLK_A(SH_BLK, -1) } else { SH_BLK_A(SH_BLK, 1) }
#undef SH_BLK
      return;
    }
    for (int d = 0; d < 3; ++d)
      if (d != K) launch_c2c(d, dir, S, hd, bsize[d]);
  }

  // ---- One launch per stage for the three hourglasses of a scale.
  // A scale qualifies when its spectrum is a cube (the three hourglasses
  // then have lines of one length and equally many IRC groups), every
  // hourglass runs its IRC stages as one grouped pass of the same (L1, L2),
  // and hourglasses 2 and 3 take the same 2-D block tile.  Hourglass 1's
  // block FFTs stay two segmented passes.
  bool merged_scale(const DecScaleT &s) const {
    if (!(s.n[0] == s.n[1] && s.n[1] == s.n[2])) return false;
    int L[3][2];
    for (int K = 0; K < 3; ++K) {
      if (!grouped(s, K)) return false;
      int m1, m2;
      transverse(K, m1, m2);
      L[K][0] = s.lv[K][m1];
      L[K][1] = s.lv[K][m2];
      if ((int)s.stages[K].size() != L[K][0] + L[K][1]) return false;
    }
    if (L[1][0] != L[0][0] || L[1][1] != L[0][1] || L[2][0] != L[0][0] || L[2][1] != L[0][1]) return false;
    int B[3][3];
    for (int K = 1; K < 3; ++K) {
      Dims3 c, bs;
      scale_blocks(s, K, c, bs);
      if (!blk_tile(bs, K, B[K][0], B[K][1], B[K][2])) return false;
    }
    return B[1][0] == B[2][0] && B[1][1] == B[2][1];
  }
  void block_ffts3(int dir, int j) {
    const DecScaleT &s = T.scales[j];
    Dims3 c, bs;
    scale_blocks(s, 0, c, bs);
    block_ffts(dir, Sbuf[j][0].p, half_dims(s.n, 0), bs, 0);
    Blk2dPair P{};
    int B1 = 0, B2 = 0;
    for (int y = 0; y < 2; ++y) {
      const int K = y + 1;
      int m2;
      scale_blocks(s, K, c, bs);
      blk_tile(bs, K, B1, B2, m2);
      const Dims3 hd = half_dims(s.n, K);
      P.S[y] = Sbuf[j][K].p;
      for (int d = 0; d < 3; ++d) P.h[y][d] = hd[d];
      P.K[y] = K;
      P.m2[y] = m2;
      P.nb[y] = (unsigned)((int64_t)(hd[0] / B1) * (hd[m2] / B2) * hd[K]);
    }
    const unsigned nb = std::max(P.nb[0], P.nb[1]);
    const float2 *t1 = tw.p + plan_for(B1).tw, *t2 = tw.p + plan_for(B2).tw;
#define SH_BLKP(D, A, B) go((k_blk2d_pair<D, A, B>), dim3(nb, 2), dim3(blk2d_threads<A, B>()), 0, P, t1, t2)
    if (dir < 0) { SH_BLK_A(SH_BLKP, -1) } else { SH_BLK_A(SH_BLKP, 1) }
#undef SH_BLKP
  }
#undef SH_BLK_A
#undef SH_BLK_B
  // A = FFT_3(rin[i is] scale), the conversion folded into the first pass
  // when every axis has a register-FFT length.
  void fft3_from_real(float2 *A, Dims3 m, const float *rin, int64_t is, float scale) {
    const int64_t n = (int64_t)m[0] * m[1] * m[2];
    if (m[0] == 16 && m[1] == 16 && m[2] == 16) {
      go(k_fft3_16<-1>, dim3(1), dim3(256), 0, A, RealIO{rin, nullptr, is, scale}, 1);
      return;
    }
    if (!(reg_len(m[0]) && reg_len(m[1]) && reg_len(m[2]))) {
      go(k_to_complex, dim3(grid_for(n)), dim3(DT), 0, rin, is, A, n, scale);
      fft3(-1, A, m);
      return;
    }
    launch_c2c(0, -1, A, m, m[0], 1, RealIO{rin, nullptr, is, scale});
    launch_c2c(1, -1, A, m, m[1]);
    launch_c2c(2, -1, A, m, m[2]);
  }
  // rout[i os] = Re(IFFT_3(A))[i] scale (unnormalised inverse), the
  // conversion folded into the last pass when it can be.
  void ifft3_to_real(float2 *A, Dims3 m, float *rout, int64_t os, float scale) {
    const int64_t n = (int64_t)m[0] * m[1] * m[2];
    if (m[0] == 16 && m[1] == 16 && m[2] == 16) {
      go(k_fft3_16<1>, dim3(1), dim3(256), 0, A, RealIO{nullptr, rout, os, scale}, 2);
      return;
    }
    if (!(reg_len(m[0]) && reg_len(m[1]) && reg_len(m[2]))) {
      fft3(1, A, m);
      go(k_real_out, dim3(grid_for(n)), dim3(DT), 0, A, rout, os, scale, n);
      return;
    }
    launch_c2c(0, 1, A, m, m[0]);
    launch_c2c(1, 1, A, m, m[1]);
    launch_c2c(2, 1, A, m, m[2], 2, RealIO{nullptr, rout, os, scale});
  }
  static void transverse(int K, int &m1, int &m2) {
    m1 = K == 0 ? 1 : 0;
    m2 = K == 2 ? 1 : 2;
  }
  // Hourglasses whose IRC groups hold at most 16 elements run their stages
  // (and, in analysis, the ritf mask) as one k_irc_group pass.
  static bool grouped(const DecScaleT &s, int K) {
    int m1, m2;
    transverse(K, m1, m2);
    return s.lv[K][m1] + s.lv[K][m2] <= 4;
  }
  // The filter planes of hourglass K's stages for a grouped pass (mode 1
  // reads the synthesis pair, modes 0 and 2 the analysis pair), with their
  // strides: the transposed copy where the device keeps one.
  IrcGroup irc_tables(const DecScaleT &s, int K, int mode) const {
    int m1, m2;
    transverse(K, m1, m2);
    const Dims3 hd = half_dims(s.n, K);
    IrcGroup F{};
    const auto &stg = s.stages[K];
    if ((int)stg.size() != s.lv[K][m1] + s.lv[K][m2]) fail(SURFHIP_ERR_DEVICE, "internal: IRC stage count");
    for (int q = 0; q < (int)stg.size(); ++q) {
      const int a = mode == 1 ? 2 : 0;
      const bool tr = stg[q].offT[a] >= 0;
      const float2 *base = tr ? filt_t.p : filt.p;
      const int *off = tr ? stg[q].offT : stg[q].off;
      F.f0[q] = base + off[a];
      F.f1[q] = base + off[a + 1];
      F.sc[q] = tr ? 1 : hd[K];
      F.sk[q] = tr ? hd[stg[q].m] : 1;
    }
    return F;
  }
  // mode 0: S = IRC_analysis(X[half_K] V scale); 1: S = IRC_synthesis(S);
  // 2: S = IRC_analysis^H(S).
  void irc_group(const DecScaleT &s, int K, int mode, const float2 *X, const float *V, float scale, float2 *S) {
    int m1, m2;
    transverse(K, m1, m2);
    const int L1 = s.lv[K][m1], L2 = s.lv[K][m2];
    const Dims3 hd = half_dims(s.n, K);
    const IrcGroup F = irc_tables(s, K, mode);
    const int64_t ng = (int64_t)hd[K] * (hd[m1] >> L1) * (hd[m2] >> L2);
    const dim3 grid(grid_for(ng)), block(DT);
#define SH_IRCG(A, B)                                                                                                \
  go((k_irc_group<A, B>), grid, block, 0, X, s.n[0], s.n[1], V, scale, S, hd[0], hd[1], hd[2], K, \
                     m1, m2, F, mode)
    switch (L1 * 5 + L2) {
      case 0: SH_IRCG(0, 0); break;
      case 1: SH_IRCG(0, 1); break;
      case 2: SH_IRCG(0, 2); break;
      case 3: SH_IRCG(0, 3); break;
      case 4: SH_IRCG(0, 4); break;
      case 5: SH_IRCG(1, 0); break;
      case 6: SH_IRCG(1, 1); break;
      case 7: SH_IRCG(1, 2); break;
      case 8: SH_IRCG(1, 3); break;
      case 10: SH_IRCG(2, 0); break;
      case 11: SH_IRCG(2, 1); break;
      case 12: SH_IRCG(2, 2); break;
      case 15: SH_IRCG(3, 0); break;
      case 16: SH_IRCG(3, 1); break;
      case 20: SH_IRCG(4, 0); break;
      default: fail(SURFHIP_ERR_DEVICE, "internal: IRC group size");
    }
#undef SH_IRCG
  }
  // X of scale j = the hourglass spectra summed back (with their Hermitian
  // mirrors when `mirror`), each scaled by c[K] and its ritf mask.
  // With merge_fac > 0 the pyramid merge of level j (low = the coarser
  // level's spectrum, factor merge_fac) is folded into the same pass.
  void hermitian_sum(int j, float2 *X, const float c[3], int mirror, float merge_fac = 0.0f) {
    const DecScaleT &s = T.scales[j];
    HermSum H{};
    for (int K = 0; K < 3; ++K) {
      H.S[K] = Sbuf[j][K].p;
      H.M[K] = masks.p + s.mask_off[K];
      H.c[K] = c[K];
    }
    const int64_t NX = (int64_t)s.n[0] * s.n[1] * s.n[2];
    go(k_hermitian_sum, dim3(grid_for(NX)), dim3(DT), 0, X, s.n[0], s.n[1], s.n[2], H, mirror,
       (const float2 *)spectra[j + 1].p, lev[j]->pd, merge_fac, merge_fac > 0.0f ? 1 : 0);
  }
  BlockMap block_map(int j, int K) const {
    const DecScaleT &s = T.scales[j];
    Dims3 counts, bsize;
    scale_blocks(s, K, counts, bsize);
    Dims3 hd = half_dims(s.n, K);
    Dims3 sb = subband_size(s, K);
    BlockMap M;
    M.K = K;
    M.nK = s.n[K];
    for (int d = 0; d < 3; ++d) {
      M.h[d] = hd[d];
      M.bs[d] = bsize[d];
      M.cnt[d] = d == K ? 1 : counts[d];
      M.sb[d] = sb[d];
    }
    M.off = sc[j]->boff[K].p;
    return M;
  }
  void lines(bool c2r, int j, int K, float2 *S, float *coef, int64_t stride, int dbl = 0) {
    BlockMap M = block_map(j, K);
    const int nK = M.nK, hK = M.h[K];
    const FftPlan1D &pl = plan_for(nK);
    const int G = std::max(1, std::min(LINES_MAX, 2048 / nK));
    const int F = (G + 1) / 2;
    const size_t lds = sizeof(float2) * ((size_t)nK + (size_t)hK * G + 2 * (size_t)nK * F);
    if (lds > 65536) fail(SURFHIP_ERR_ARG, fmt("axis length %d is too long for the decimated path", nK));
    const int64_t nlines = (int64_t)M.h[0] * M.h[1] * M.h[2] / hK;
    const unsigned blocks = (unsigned)((nlines + G - 1) / G);
    // Without the direct-DFT pass where every radix is hand-written.
    const bool gen = fft_plan_generic(pl);
    if (c2r)
      go(gen ? k_c2r_lines<true> : k_c2r_lines<false>, dim3(blocks), dim3(DT), lds, S, coef, stride, M, G, pl, tw.p);
    else
      go(gen ? k_r2c_lines<true> : k_r2c_lines<false>, dim3(blocks), dim3(DT), lds, (const float *)coef, stride, S, M,
         G, pl, tw.p, dbl);
  }

  void irc_group3(int j, int mode, const float2 *X, const float *scale3) {
    const DecScaleT &s = T.scales[j];
    IrcGroup3 P{};
    int L1 = 0, L2 = 0;
    int64_t ng = 0;
    for (int K = 0; K < 3; ++K) {
      int m1, m2;
      transverse(K, m1, m2);
      L1 = s.lv[K][m1];
      L2 = s.lv[K][m2];
      const Dims3 hd = half_dims(s.n, K);
      P.V[K] = masks.p + s.mask_off[K];
      P.scale[K] = scale3 ? scale3[K] : 1.0f;
      P.S[K] = Sbuf[j][K].p;
      for (int d = 0; d < 3; ++d) P.h[K][d] = hd[d];
      P.m1[K] = m1;
      P.m2[K] = m2;
      P.F[K] = irc_tables(s, K, mode);
      ng = std::max(ng, (int64_t)hd[K] * (hd[m1] >> L1) * (hd[m2] >> L2));
    }
    const dim3 grid(grid_for(ng), 3), block(DT);
#define SH_IRCG3(A, B) go((k_irc_group3<A, B>), grid, block, 0, X, s.n[0], s.n[1], P, mode)
    switch (L1 * 5 + L2) {
      case 0: SH_IRCG3(0, 0); break;
      case 1: SH_IRCG3(0, 1); break;
      case 2: SH_IRCG3(0, 2); break;
      case 3: SH_IRCG3(0, 3); break;
      case 4: SH_IRCG3(0, 4); break;
      case 5: SH_IRCG3(1, 0); break;
      case 6: SH_IRCG3(1, 1); break;
      case 7: SH_IRCG3(1, 2); break;
      case 8: SH_IRCG3(1, 3); break;
      case 10: SH_IRCG3(2, 0); break;
      case 11: SH_IRCG3(2, 1); break;
      case 12: SH_IRCG3(2, 2); break;
      case 15: SH_IRCG3(3, 0); break;
      case 16: SH_IRCG3(3, 1); break;
      case 20: SH_IRCG3(4, 0); break;
      default: fail(SURFHIP_ERR_DEVICE, "internal: IRC group size");
    }
#undef SH_IRCG3
  }
  void lines3(bool c2r, int j, float *coef, int64_t stride, int dbl = 0) {
    Lines3 L{};
    int G = 1;
    size_t lds = 0;
    unsigned nbmax = 0;
    int nK = 0;
    for (int K = 0; K < 3; ++K) {
      BlockMap M = block_map(j, K);
      nK = M.nK;
      const int hK = M.h[K];
      G = std::max(1, std::min(LINES_MAX, 2048 / nK));
      const int F = (G + 1) / 2;
      lds = sizeof(float2) * ((size_t)nK + (size_t)hK * G + 2 * (size_t)nK * F);
      if (lds > 65536) fail(SURFHIP_ERR_ARG, fmt("axis length %d is too long for the decimated path", nK));
      const int64_t nlines = (int64_t)M.h[0] * M.h[1] * M.h[2] / hK;
      L.S[K] = Sbuf[j][K].p;
      L.M[K] = M;
      L.nb[K] = (unsigned)((nlines + G - 1) / G);
      nbmax = std::max(nbmax, L.nb[K]);
    }
    const FftPlan1D &pl = plan_for(nK);
    const bool gen = fft_plan_generic(pl);
    if (c2r)
      go(gen ? k_c2r_lines3<true> : k_c2r_lines3<false>, dim3(nbmax, 3), dim3(DT), lds, L, coef, stride, G, pl, tw.p);
    else
      go(gen ? k_r2c_lines3<true> : k_r2c_lines3<false>, dim3(nbmax, 3), dim3(DT), lds, L, (const float *)coef, stride,
         G, pl, tw.p, dbl);
  }

  void analyze_channel(const float *x, float *coef) {
    use_lane(0);
    const int64_t os = nch;
    Dims3 n = T.dims;
    float2 *X = spectra[0].p;
    fft3_from_real(X, n, x, (int64_t)nch, 1.0f);
    for (int j = 0; j < T.J; ++j) {
      const PyrLevel &pl = T.pyr[j];
      float2 *low = spectra[j + 1].p;
      const PyrDev &pd = lev[j]->pd;
      go(k_pyr_split, dim3(grid_for((int64_t)pl.nlow[0] * pl.nlow[1] * pl.nlow[2])), dim3(DT), 0,
                         spectra[j].p, low, pd);
      const DecScaleT &s = T.scales[j];
      if (merged_scale(s)) {
        // The scale's hourglass passes after the split, the finest scale on
        // its own lane, so that the next split (of the lowpass) and the
        // coarser scales do not wait for it.
        if (scale_lane(j) != 0) fork_k();
        use_lane(scale_lane(j));
        float sc3[3];
        for (int K = 0; K < 3; ++K) {
          const Dims3 sb = subband_size(s, K);
          sc3[K] = (float)(1.0 / ((double)sb[0] * sb[1] * sb[2]));
        }
        irc_group3(j, 0, spectra[j].p, sc3);
        block_ffts3(1, j);
        lines3(true, j, coef, os);
        use_lane(0);
        continue;
      }
      fork_k();
      for (int K = 0; K < 3; ++K) {
        use_lane(K + 1);
        Dims3 hd = half_dims(s.n, K);
        Dims3 sb = subband_size(s, K);
        const float scale = (float)(1.0 / ((double)sb[0] * sb[1] * sb[2]));
        float2 *S = Sbuf[j][K].p;
        const int64_t HN = (int64_t)hd[0] * hd[1] * hd[2];
        if (grouped(s, K)) {
          irc_group(s, K, 0, spectra[j].p, masks.p + s.mask_off[K], scale, S);
        } else {
          go(k_masked, dim3(grid_for(HN)), dim3(DT), 0, spectra[j].p, s.n[0], s.n[1], S,
                             masks.p + s.mask_off[K], hd[0], hd[1], hd[2], scale);
          for (const IRCStageT &st : s.stages[K]) {
            go(k_irc, dim3(grid_for(HN / 2)), dim3(DT), 0, S, hd[0], hd[1], hd[2], K, st.m, st.b,
                               filt.p + st.off[0], filt.p + st.off[1], 0);
          }
        }
        Dims3 counts, bsize;
        scale_blocks(s, K, counts, bsize);
        block_ffts(1, S, hd, bsize, K);
        lines(true, j, K, S, coef, os);
      }
      use_lane(0);
    }
    const Dims3 nl = T.pyr.back().nlow;
    const int64_t NL = (int64_t)nl[0] * nl[1] * nl[2];
    float2 *C = spectra[T.J].p;
    ifft3_to_real(C, nl, coef + T.offsets[0] * nch, os, (float)(1.0 / (double)NL));
    join_k();
  }

  void synthesize_channel(const float *coef, float *x) {
    use_lane(0);
    const int64_t is = nch;
    const Dims3 nl = T.pyr.back().nlow;
    float2 *C = spectra[T.J].p;
    // The merged scales' hourglass passes read only the coefficients: all of
    // them start at once, each scale on its own lane.
    fork_k();
    for (int j = T.J - 1; j >= 0; --j) {
      if (!merged_scale(T.scales[j])) continue;
      use_lane(scale_lane(j));
      lines3(false, j, const_cast<float *>(coef), is);
      block_ffts3(-1, j);
      irc_group3(j, 1, nullptr, nullptr);
    }
    use_lane(0);
    fft3_from_real(C, nl, coef + T.offsets[0] * nch, is, 1.0f);
    for (int j = T.J - 1; j >= 0; --j) {
      const DecScaleT &s = T.scales[j];
      float2 *X = spectra[j].p;
      if (merged_scale(s)) {
        join_lane(scale_lane(j));
      } else {
      fork_k();
      for (int K = 0; K < 3; ++K) {
        use_lane(K + 1);
        Dims3 hd = half_dims(s.n, K);
        float2 *S = Sbuf[j][K].p;
        lines(false, j, K, S, const_cast<float *>(coef), is);
        Dims3 counts, bsize;
        scale_blocks(s, K, counts, bsize);
        block_ffts(-1, S, hd, bsize, K);
        const int64_t HN = (int64_t)hd[0] * hd[1] * hd[2];
        if (grouped(s, K)) {
          if (!s.stages[K].empty()) irc_group(s, K, 1, nullptr, nullptr, 1.0f, S);
        } else {
          for (int q = (int)s.stages[K].size() - 1; q >= 0; --q) {
            const IRCStageT &st = s.stages[K][q];
            go(k_irc, dim3(grid_for(HN / 2)), dim3(DT), 0, S, hd[0], hd[1], hd[2], K, st.m,
                               st.b, filt.p + st.off[2], filt.p + st.off[3], 1);
          }
        }
      }
      use_lane(0);
      join_k();
      }
      {
        const float c[3] = {(float)(1.0 / 3.0), (float)(1.0 / 3.0), (float)(1.0 / 3.0)};
        hermitian_sum(j, X, c, 1, lev[j]->pd.D15);
      }
    }
    const Dims3 n = T.dims;
    const int64_t N = (int64_t)n[0] * n[1] * n[2];
    ifft3_to_real(spectra[0].p, n, x, (int64_t)nch, (float)(1.0 / (double)N));
  }

  // x = A^T c, the adjoint of the analysis (not its inverse: the transform
  // is a frame but not a tight one, src/transforms/surfacelet.jl:24-26).
  void adjoint_channel(const float *coef, float *x) {
    use_lane(0);
    const int64_t is = nch;
    const Dims3 nl = T.pyr.back().nlow;
    const int64_t NL = (int64_t)nl[0] * nl[1] * nl[2];
    float2 *C = spectra[T.J].p;
    fork_k();
    for (int j = T.J - 1; j >= 0; --j) {
      if (!merged_scale(T.scales[j])) continue;
      use_lane(scale_lane(j));
      lines3(false, j, const_cast<float *>(coef), is, 1);
      block_ffts3(-1, j);
      irc_group3(j, 2, nullptr, nullptr);
    }
    use_lane(0);
    fft3_from_real(C, nl, coef + T.offsets[0] * nch, is, (float)(1.0 / (double)NL));
    for (int j = T.J - 1; j >= 0; --j) {
      const DecScaleT &s = T.scales[j];
      float2 *X = spectra[j].p;
      if (merged_scale(s)) {
        join_lane(scale_lane(j));
      } else {
      fork_k();
      for (int K = 0; K < 3; ++K) {
        use_lane(K + 1);
        Dims3 hd = half_dims(s.n, K);
        float2 *S = Sbuf[j][K].p;
        lines(false, j, K, S, const_cast<float *>(coef), is, 1);
        Dims3 counts, bsize;
        scale_blocks(s, K, counts, bsize);
        block_ffts(-1, S, hd, bsize, K);
        const int64_t HN = (int64_t)hd[0] * hd[1] * hd[2];
        if (grouped(s, K)) {
          if (!s.stages[K].empty()) irc_group(s, K, 2, nullptr, nullptr, 1.0f, S);
        } else {
          for (int q = (int)s.stages[K].size() - 1; q >= 0; --q) {
            const IRCStageT &st = s.stages[K][q];
            go(k_irc, dim3(grid_for(HN / 2)), dim3(DT), 0, S, hd[0], hd[1], hd[2], K, st.m,
                               st.b, filt.p + st.off[0], filt.p + st.off[1], 2);
          }
        }
      }
      use_lane(0);
      join_k();
      }
      {
        float c[3];
        for (int K = 0; K < 3; ++K) {
          const Dims3 sb = subband_size(s, K);
          c[K] = (float)(1.0 / ((double)sb[0] * sb[1] * sb[2]));
        }
        hermitian_sum(j, X, c, 0, 1.0f / lev[j]->pd.D15);
      }
    }
    const Dims3 n = T.dims;
    ifft3_to_real(spectra[0].p, n, x, (int64_t)nch, 1.0f);
  }
};

namespace {

surfhip_s_params checked(const surfhip_s_params *p) {
  if (!p) fail(SURFHIP_ERR_ARG, "null argument");
  dec_check_params(*p);
  return *p;
}

}  // namespace

extern "C" {

int surfhip_s_check_size(const surfhip_s_params *p, int32_t n1, int32_t n2, int32_t n3) {
  return guarded([&] { dec_check_size(checked(p), Dims3{n1, n2, n3}); });
}

int surfhip_s_counts(const surfhip_s_params *p, int32_t n1, int32_t n2, int32_t n3, int32_t *nsub, int64_t *ncoef) {
  return guarded([&] {
    auto sizes = dec_layout_sizes(checked(p), Dims3{n1, n2, n3});
    int64_t tot = 0;
    for (auto &s : sizes) tot += (int64_t)s[0] * s[1] * s[2];
    if (nsub) *nsub = (int32_t)sizes.size();
    if (ncoef) *ncoef = tot;
  });
}

int surfhip_s_layout(const surfhip_s_params *p, int32_t n1, int32_t n2, int32_t n3, int64_t *sizes3, int64_t *offsets,
                     int32_t cap) {
  return guarded([&] {
    auto sizes = dec_layout_sizes(checked(p), Dims3{n1, n2, n3});
    if ((int)sizes.size() > cap) fail(SURFHIP_ERR_DIM, fmt("layout has %zu subbands, buffer holds %d", sizes.size(), cap));
    int64_t off = 0;
    for (size_t i = 0; i < sizes.size(); ++i) {
      for (int d = 0; d < 3; ++d)
        if (sizes3) sizes3[3 * i + d] = sizes[i][d];
      if (offsets) offsets[i] = off;
      off += (int64_t)sizes[i][0] * sizes[i][1] * sizes[i][2];
    }
  });
}

int surfhip_s_plan_create(surfhip_s_plan **out, const surfhip_s_params *p, int32_t n1, int32_t n2, int32_t n3,
                          int32_t dtype, void *stream) {
  return guarded([&] {
    if (!out) fail(SURFHIP_ERR_ARG, "null argument");
    *out = nullptr;
    if (dtype != SURFHIP_F32 && dtype != SURFHIP_C64) fail(SURFHIP_ERR_ARG, fmt("unknown dtype %d", dtype));
    auto P = std::make_unique<surfhip_s_plan>();
    dec_build_tables(checked(p), Dims3{n1, n2, n3}, P->T);
    P->dtype = dtype;
    P->nch = dtype == SURFHIP_C64 ? 2 : 1;
    P->stream = (hipStream_t)stream;
    P->ls = P->stream;
    P->make_lanes();
    const DecTables &T = P->T;
    // FFT plans of every length used: level sizes, block sizes, coarse size.
    std::vector<int> lens;
    auto add = [&](int n) {
      if (std::find(lens.begin(), lens.end(), n) == lens.end()) lens.push_back(n);
    };
    std::vector<float2> pool;
    for (int d = 0; d < 3; ++d) add(T.pyr.back().nlow[d]);
    for (const DecScaleT &s : T.scales)
      for (int K = 0; K < 3; ++K) {
        Dims3 c, b;
        scale_blocks(s, K, c, b);
        add(s.n[K]);
        for (int d = 0; d < 3; ++d) {
          add(s.n[d]);
          if (d != K) add(b[d]);
        }
      }
    for (int n : lens) P->plans.push_back({n, make_fft_plan(n, pool)});
    P->tw.upload(pool);
    P->masks.upload(T.masks);
    if (T.filt.empty()) P->T.filt.push_back(make_float2(0, 0));
    P->filt.upload(P->T.filt);
    {
      // In k_irc_group the lanes of hourglass K != 0 run along axis 0, so a
      // stage along axis 0 reads its planes across c: keep those transposed.
      std::vector<float2> ft;
      for (DecScaleT &s : P->T.scales)
        for (int K = 1; K < 3; ++K) {
          const Dims3 hd = half_dims(s.n, K);
          for (IRCStageT &st : s.stages[K]) {
            if (st.m != 0) continue;
            const int nKh = hd[K], nm = hd[0];
            for (int f = 0; f < 4; ++f) {
              st.offT[f] = (int)ft.size();
              for (int k = 0; k < nKh; ++k)
                for (int c = 0; c < nm; ++c) ft.push_back(P->T.filt[(size_t)st.off[f] + k + (size_t)nKh * c]);
            }
          }
        }
      if (ft.empty()) ft.push_back(make_float2(0, 0));
      P->filt_t.upload(ft);
    }
    // Pyramid tables.
    for (const PyrLevel &pl : T.pyr) {
      auto L = std::make_unique<DecLevelDev>();
      std::vector<int> ix, il;
      std::vector<float> pr, cm;
      for (int a = 0; a < 3; ++a) {
        ix.insert(ix.end(), pl.idx[a].begin(), pl.idx[a].end());
        il.insert(il.end(), pl.idx_low[a].begin(), pl.idx_low[a].end());
        for (double v : pl.prof[a]) {
          pr.push_back((float)v);
          cm.push_back((float)(1.0 - v));
        }
      }
      L->idx.upload(ix);
      L->idx_low.upload(il);
      L->prof.upload(pr);
      L->comp.upload(cm);
      PyrDev &pd = L->pd;
      pd.n1 = pl.n[0];
      pd.n2 = pl.n[1];
      pd.n3 = pl.n[2];
      pd.l1 = pl.nlow[0];
      pd.l2 = pl.nlow[1];
      pd.l3 = pl.nlow[2];
      pd.len1 = (int)pl.idx[0].size();
      pd.len2 = (int)pl.idx[1].size();
      pd.len3 = (int)pl.idx[2].size();
      pd.idx = L->idx.p;
      pd.idx_low = L->idx_low.p;
      pd.prof = L->prof.p;
      pd.comp = L->comp.p;
      pd.D15 = (float)pl.D15;
      P->lev.push_back(std::move(L));
    }
    // Block offsets per scale and hourglass.
    for (size_t j = 0; j < T.scales.size(); ++j) {
      const DecScaleT &s = T.scales[j];
      auto D = std::make_unique<DecScaleDev>();
      int sub = s.first_subband;
      for (int K = 0; K < 3; ++K) {
        Dims3 c, b;
        scale_blocks(s, K, c, b);
        const int nb = c[0] * c[1] * c[2];
        std::vector<int64_t> off(nb);
        for (int i = 0; i < nb; ++i) off[i] = T.offsets[sub + i];
        D->first[K] = sub;
        D->boff[K].upload(off);
        sub += nb;
      }
      P->sc.push_back(std::move(D));
    }
    P->Sbuf.resize(T.J);
    for (int j = 0; j < T.J; ++j)
      for (int K = 0; K < 3; ++K) {
        const Dims3 hd = half_dims(T.scales[j].n, K);
        P->Sbuf[j][K].alloc((size_t)hd[0] * hd[1] * hd[2]);
      }
    P->spectra.resize(T.J + 1);
    for (int j = 0; j <= T.J; ++j) {
      Dims3 n = j < T.J ? T.pyr[j].n : T.pyr.back().nlow;
      P->spectra[j].alloc((size_t)n[0] * n[1] * n[2]);
    }
    SH_HIP(hipDeviceSynchronize());
    *out = P.release();
  });
}

int surfhip_s_plan_destroy(surfhip_s_plan *plan) {
  return guarded([&] {
    if (!plan) return;
    (void)hipStreamSynchronize(plan->stream);
    delete plan;
  });
}

int surfhip_s_set_stream(surfhip_s_plan *plan, void *stream) {
  return guarded([&] {
    if (!plan) fail(SURFHIP_ERR_ARG, "null plan");
    plan->stream = (hipStream_t)stream;
  });
}

}  // extern "C"

namespace {
// Extents a call reads or writes through caller pointers: the volume and the
// flat coefficient vector (complex data: two floats per element).
void check_dec_buffers(const surfhip_s_plan *plan, const void *x, const void *coeffs) {
  const size_t esz = (size_t)plan->nch * sizeof(float);
  const Dims3 &d = plan->T.dims;
  check_device_buffer(x, (size_t)d[0] * d[1] * d[2] * esz, "volume");
  check_device_buffer(coeffs, (size_t)plan->T.ncoef * esz, "coefficient vector");
}
}  // namespace

extern "C" {

int surfhip_s_index_maps(surfhip_s_plan *plan, int32_t kind, int32_t level, int32_t hourglass, int32_t axis,
                         int64_t *out, int64_t cap, int64_t *count) {
  return guarded([&] {
    if (!plan || !count) fail(SURFHIP_ERR_ARG, "null argument");
    if (axis < 0 || axis > 2) fail(SURFHIP_ERR_ARG, fmt("axis %d outside 0:2", axis));
    const int nlev = (int)plan->lev.size();
    if (level < 0 || level >= nlev) fail(SURFHIP_ERR_ARG, fmt("level %d outside 0:%d", level, nlev - 1));
    if (kind < 0 || kind > 4) fail(SURFHIP_ERR_ARG, fmt("unknown map kind %d", kind));
    if ((kind == 3 || kind == 4) && (hourglass < 0 || hourglass > 2))
      fail(SURFHIP_ERR_ARG, fmt("hourglass %d outside 0:2", hourglass));
    const PyrDev &pd = plan->lev[level]->pd;
    const int len[3] = {pd.len1, pd.len2, pd.len3};
    const int nl[3] = {pd.l1, pd.l2, pd.l3};
    const int off = axis == 0 ? 0 : (axis == 1 ? len[0] : len[0] + len[1]);
    std::vector<int64_t> host;
    if (kind == 0 || kind == 1) {
      std::vector<int> v(len[axis]);
      const int *src = (kind == 0 ? plan->lev[level]->idx.p : plan->lev[level]->idx_low.p) + off;
      SH_HIP(hipMemcpy(v.data(), src, sizeof(int) * v.size(), hipMemcpyDeviceToHost));
      host.assign(v.begin(), v.end());
    } else if (kind == 4) {
      const DevBuf<int64_t> &b = plan->sc[level]->boff[hourglass];
      host.resize(b.n);
      if (b.n) SH_HIP(hipMemcpy(host.data(), b.p, sizeof(int64_t) * b.n, hipMemcpyDeviceToHost));
    } else {
      const Dims3 &n = plan->T.scales[level].n;
      const int m = kind == 2 ? nl[axis] : n[axis];
      const int64_t total = kind == 2 ? 2 * (int64_t)m : m;
      DevBuf<int64_t> d;
      d.alloc((size_t)total);
      hipLaunchKernelGGL(k_index_map, dim3((unsigned)grid_for(m)), dim3(DT), 0, plan->stream, kind, m, nl[axis],
                         (len[axis] - 1) / 2, axis == hourglass ? 1 : 0, d.p);
      SH_LAUNCHED();
      host.resize((size_t)total);
      SH_HIP(hipMemcpyAsync(host.data(), d.p, sizeof(int64_t) * total, hipMemcpyDeviceToHost, plan->stream));
      SH_HIP(hipStreamSynchronize(plan->stream));
    }
    *count = (int64_t)host.size();
    if (out) std::copy(host.begin(), host.begin() + (ptrdiff_t)std::min<int64_t>(cap, *count), out);
  });
}

int surfhip_s_analyze(surfhip_s_plan *plan, const void *x, void *coeffs) {
  return guarded([&] {
    if (!plan || !x || !coeffs) fail(SURFHIP_ERR_ARG, "null argument");
    check_dec_buffers(plan, x, coeffs);
    plan->run(0, x, coeffs, [&] {
      for (int ch = 0; ch < plan->nch; ++ch) plan->analyze_channel((const float *)x + ch, (float *)coeffs + ch);
    });
  });
}

int surfhip_s_adjoint(surfhip_s_plan *plan, const void *coeffs, void *x) {
  return guarded([&] {
    if (!plan || !x || !coeffs) fail(SURFHIP_ERR_ARG, "null argument");
    check_dec_buffers(plan, x, coeffs);
    plan->run(2, coeffs, x, [&] {
      for (int ch = 0; ch < plan->nch; ++ch) plan->adjoint_channel((const float *)coeffs + ch, (float *)x + ch);
    });
  });
}

int surfhip_s_synthesize(surfhip_s_plan *plan, const void *coeffs, void *x) {
  return guarded([&] {
    if (!plan || !x || !coeffs) fail(SURFHIP_ERR_ARG, "null argument");
    check_dec_buffers(plan, x, coeffs);
    plan->run(1, coeffs, x, [&] {
      for (int ch = 0; ch < plan->nch; ++ch) plan->synthesize_channel((const float *)coeffs + ch, (float *)x + ch);
    });
  });
}

}  // extern "C"
