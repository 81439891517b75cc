// NSSurfacelet analyze/synthesize on gfx950.
//
// Every subband is irfft(X . M_k) with X = rfft(x) and M_k the real mask of
// subband k (src/transforms/ns_surfacelet.jl:8-32); synthesis is
// irfft(sum_k rfft(c_k) . M_k).  This file holds the general-size path and
// the plan; grids with every length in {32, 64, 80, 96, 112, 128} run the
// factored kernels of ns_fused.hip.
// A 3-D real FFT here is a paired-row r2c/c2r along axis 1 (the contiguous
// axis of the volume) and one LDS-resident 2-D transform of each (k2, k3)
// plane (k_plane), or where a plane exceeds LDS one pass per axis over
// strips of pencils (k_strip, else k_axis).  The mask is evaluated on the
// fly from FP32 tables (built in FP64) inside the plane or axis-2 pass:
// before the inverse transform in analysis, after the forward one in
// synthesis, whose masked spectra are then summed in subband order (k_sum).
//
// Spectra use the rfft grid in plane-major order: S[(k1 n2 + k2) n3 + k3],
// h1 = n1/2 + 1, the order of the group-factor tables.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "fft.hpp"
#include "hip_util.hpp"
#include "ns_fused.hpp"
#include "ns_plan_api.hpp"
#include "ns_scratch.hpp"
#include "ns_tables.hpp"

namespace surfhip {

constexpr int NT = 256;

// ------------------------------------------------------------------ kernels

// Row rp of a block's rows in plane-major order (rp = k2 n3 + k3) is the
// real row (k2, k3), stored at row k3 n2 + k2 of x.  The G row offsets of a
// block are made once, into LDS.
__device__ __forceinline__ void real_rows(int64_t *roff, int r0, int G, int n2, int n3, int64_t n1, int tid) {
  for (int g = tid; g < G; g += NT) {
    const int rp = r0 + g;
    const int k2 = rp / n3;
    roff[g] = ((int64_t)(rp - k2 * n3) * n2 + k2) * n1;
  }
}

// Real rows of length n1 -> half-spectrum bins out[k R + rp], R = n2 n3.
// A block takes G rows consecutive in rp, so the stores of each bin k are
// runs of G words.  Two rows share one complex FFT: z = a + i b, then
// A[k] = (Z[k] + conj Z[-k]) / 2, B[k] = (Z[k] - conj Z[-k]) / 2i.
template <bool GEN>
__global__ __launch_bounds__(NT) void k_rows_r2c(const float *__restrict__ x, int64_t xs, int64_t x_bs,
                                                 float2 *__restrict__ out, int64_t out_bs, int n1, int h1,
                                                 int n2, int n3, int G, FftPlan1D pl,
                                                 const float2 *__restrict__ twg, int vec, int wrap_at,
                                                 int64_t wrap_off) {
  // Batch entry blockIdx.y: its own input and output; entries from wrap_at on
  // read x + wrap_off + (b - wrap_at) x_bs instead.
  const int eb = blockIdx.y;
  x += eb < wrap_at ? eb * x_bs : wrap_off + (eb - wrap_at) * x_bs;
  out += eb * out_bs;
  extern __shared__ float2 smem[];
  const int nrows = n2 * n3;
  const int P = G / 2;
  float2 *tw = smem;
  float2 *a = tw + n1;
  float2 *b = a + (size_t)n1 * P;
  int64_t *roff = reinterpret_cast<int64_t *>(b + (size_t)n1 * P);
  const int tid = threadIdx.x;
  for (int k = tid; k < n1; k += NT) tw[k] = twg[pl.tw + k];
  const int r0 = blockIdx.x * G;
  real_rows(roff, r0, G, n2, n3, n1, tid);
  __syncthreads();
  const FastDiv dn1 = fast_div(n1), dG = fast_div(G);
  if (vec) {
    // Unit-stride rows, n1 a multiple of 4, 16-byte aligned: 16-byte loads.
    const int n4 = n1 >> 2;
    const FastDiv dn4 = fast_div(n4);
    for (int idx = tid; idx < G * n4; idx += NT) {
      const int g = div_of(idx, dn4), t = 4 * (idx - g * n4);
      const int rp = r0 + g;
      const float4 v = rp < nrows ? *reinterpret_cast<const float4 *>(x + roff[g] + t) : make_float4(0.f, 0.f, 0.f, 0.f);
      float *ap = reinterpret_cast<float *>(&a[t * P + (g >> 1)]) + (g & 1);
      ap[0] = v.x;
      ap[2 * P] = v.y;
      ap[4 * P] = v.z;
      ap[6 * P] = v.w;
    }
  } else {
    for (int idx = tid; idx < G * n1; idx += NT) {
      const int g = div_of(idx, dn1), t = idx - g * n1;
      const int rp = r0 + g;
      float v = rp < nrows ? x[(roff[g] + t) * xs] : 0.0f;
      float *ap = reinterpret_cast<float *>(&a[t * P + (g >> 1)]);
      ap[g & 1] = v;
    }
  }
  __syncthreads();
  float2 *res = lds_fft<-1, GEN>(a, b, pl, P, tw, tid, NT);
  for (int idx = tid; idx < G * h1; idx += NT) {
    const int k = div_of(idx, dG), g = idx - k * G;
    const int rp = r0 + g;
    if (rp >= nrows) continue;
    const int p = g >> 1;
    const float2 Z = res[k * P + p];
    const float2 Zc = conjf2(res[((n1 - k) % n1) * P + p]);
    float2 v;
    if ((g & 1) == 0) {
      v = cscale(cadd(Z, Zc), 0.5f);
    } else {
      const float2 d = csub(Z, Zc);
      v = make_float2(0.5f * d.y, -0.5f * d.x);
    }
    out[(int64_t)k * nrows + rp] = v;
  }
}

// Half-spectrum bins in[k R + rp] -> real rows (row (k2, k3) at
// out[((k3 n2 + k2) n1 + t) os]) * scale.  Pairs of rows form z = a + i b
// through their Hermitian extensions; the imaginary parts of the DC and
// (even n1) Nyquist bins are ignored, as a c2r transform does.
template <bool GEN>
__global__ __launch_bounds__(NT) void k_rows_c2r(const float2 *__restrict__ in, int64_t in_bs,
                                                 float *__restrict__ out, int64_t os, int64_t out_bs, float scale,
                                                 int n1, int h1, int n2, int n3, int G, FftPlan1D pl,
                                                 const float2 *__restrict__ twg, int vec, int wrap_at,
                                                 int64_t wrap_off) {
  // Entries from wrap_at on write out + wrap_off + (b - wrap_at) out_bs.
  const int eb = blockIdx.y;
  in += eb * in_bs;
  out += eb < wrap_at ? eb * out_bs : wrap_off + (eb - wrap_at) * out_bs;
  extern __shared__ float2 smem[];
  const int nrows = n2 * n3;
  const int P = G / 2;
  float2 *tw = smem;
  float2 *s = tw + n1;
  float2 *a = s + (size_t)G * h1;
  float2 *b = a + (size_t)n1 * P;
  int64_t *roff = reinterpret_cast<int64_t *>(b + (size_t)n1 * P);
  const int tid = threadIdx.x;
  for (int k = tid; k < n1; k += NT) tw[k] = twg[pl.tw + k];
  const int r0 = blockIdx.x * G;
  real_rows(roff, r0, G, n2, n3, n1, tid);
  const FastDiv dn1 = fast_div(n1), dG = fast_div(G), dP = fast_div(P);
  if (vec & 1) {
    // Even row count and 16-byte aligned bins: two rows' bins per load.
    for (int idx = tid; idx < P * h1; idx += NT) {
      const int k = div_of(idx, dP), g = 2 * (idx - k * P);
      const int rp = r0 + g;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (rp < nrows) v = *reinterpret_cast<const float4 *>(in + (int64_t)k * nrows + rp);
      s[g * h1 + k] = make_float2(v.x, v.y);
      s[(g + 1) * h1 + k] = make_float2(v.z, v.w);
    }
  } else {
    for (int idx = tid; idx < G * h1; idx += NT) {
      const int k = div_of(idx, dG), g = idx - k * G;
      const int rp = r0 + g;
      s[g * h1 + k] = rp < nrows ? in[(int64_t)k * nrows + rp] : make_float2(0.0f, 0.0f);
    }
  }
  __syncthreads();
  const bool even = (n1 & 1) == 0;
  for (int idx = tid; idx < P * n1; idx += NT) {
    const int t = div_of(idx, dP), p = idx - t * P;
    float2 A, B;
    if (t < h1) {
      A = s[(2 * p) * h1 + t];
      B = s[(2 * p + 1) * h1 + t];
      if (t == 0 || (even && t == n1 / 2)) {
        A.y = 0.0f;
        B.y = 0.0f;
      }
    } else {
      A = conjf2(s[(2 * p) * h1 + (n1 - t)]);
      B = conjf2(s[(2 * p + 1) * h1 + (n1 - t)]);
    }
    a[t * P + p] = make_float2(A.x - B.y, A.y + B.x);
  }
  __syncthreads();
  float2 *res = lds_fft<1, GEN>(a, b, pl, P, tw, tid, NT);
  if (vec & 2) {
    // Unit-stride rows, n1 a multiple of 4, 16-byte aligned: 16-byte stores.
    const int n4 = n1 >> 2;
    const FastDiv dn4 = fast_div(n4);
    for (int idx = tid; idx < G * n4; idx += NT) {
      const int g = div_of(idx, dn4), t = 4 * (idx - g * n4);
      const int rp = r0 + g;
      if (rp >= nrows) continue;
      const float *rv = reinterpret_cast<const float *>(&res[t * P + (g >> 1)]) + (g & 1);
      *reinterpret_cast<float4 *>(out + roff[g] + t) =
          make_float4(rv[0] * scale, rv[2 * P] * scale, rv[4 * P] * scale, rv[6 * P] * scale);
    }
  } else {
    for (int idx = tid; idx < G * n1; idx += NT) {
      const int g = div_of(idx, dn1), t = idx - g * n1;
      const int rp = r0 + g;
      if (rp >= nrows) continue;
      const float2 v = res[t * P + (g >> 1)];
      out[(roff[g] + t) * os] = ((g & 1) ? v.y : v.x) * scale;
    }
  }
}

// The mask of one subband over the (k2, k3) plane of one k1: the group
// table row k1 and each wedge plane as a base and (k2, k3) strides, so an
// element costs three loads.  Same values, and the same product
// G (w_a w_b), as ns_mask.
struct PlaneMask {
  const float *g, *wa, *wb;  // wa, wb null: factor 1
  int a2, a3, b2, b3;
  bool unit;
};
__device__ __forceinline__ const float *plane_wedge(const NSMaskView &v, int32_t off, int32_t pair, int k1, int &s2,
                                                    int &s3) {
  if (off < 0) return nullptr;
  const float *base = v.planes + off;
  if (pair == 0) {  // (k1, k2)
    s2 = 1;
    s3 = 0;
    return base + (size_t)k1 * v.n2;
  }
  if (pair == 1) {  // (k1, k3)
    s2 = 0;
    s3 = 1;
    return base + (size_t)k1 * v.n3;
  }
  s2 = v.n3;  // (k2, k3)
  s3 = 1;
  return base;
}
__device__ __forceinline__ PlaneMask plane_mask(const NSMaskView &v, const NSItem &it, int k1, bool analysis) {
  PlaneMask m{};
  m.g = ns_group_table(v, ns_group_index(v, it), analysis) + (size_t)k1 * v.n2 * v.n3;
  m.unit = it.lowpass || ns_unit_wedge(v, analysis);
  if (!m.unit) {
    m.wa = plane_wedge(v, it.pa_off, it.pa_pair, k1, m.a2, m.a3);
    m.wb = plane_wedge(v, it.pb_off, it.pb_pair, k1, m.b2, m.b3);
  }
  return m;
}
// Mask at plane element i = k2 n3 + k3.
__device__ __forceinline__ float plane_mask_at(const PlaneMask &m, int i, int k2, int k3) {
  const float g = m.g[i];
  if (m.unit) return g * 1.0f;
  const float wa = m.wa ? m.wa[k2 * m.a2 + k3 * m.a3] : 1.0f;
  const float wb = m.wb ? m.wb[k2 * m.b2 + k3 * m.b3] : 1.0f;
  return g * (wa * wb);
}

constexpr int AXIS_MAXE = 16;  // k_axis elements per thread: n G <= 4096

// Complex FFT along axis AXIS (2 or 3) of a plane-major spectrum, G pencils
// per block: for axis 3 G consecutive (k1, k2) rows (contiguous pencils),
// for axis 2 G consecutive k3 of one k1 (pencils of stride n3).  MODE 0:
// out = FFT(in); MODE 1: out = FFT(in . mask) (analysis); MODE 2:
// out = mask . FFT(in) (a masked synthesis term).  Batch entry
// b = blockIdx.y reads in + b in_bs, writes out + b out_bs and uses
// items[b].  This is the path for planes that do not fit one block's LDS.
template <int AXIS, int DIR, int MODE, bool GEN>
__global__ __launch_bounds__(NT) void k_axis(const float2 *in, int64_t in_bs, float2 *out, int64_t out_bs,
                                             int h1, int n2, int n3, int G, FftPlan1D pl,
                                             const float2 *__restrict__ twg, NSMaskView mv,
                                             const NSItem *__restrict__ items) {
  in += blockIdx.y * in_bs;
  out += blockIdx.y * out_bs;
  extern __shared__ float2 smem[];
  const int n = AXIS == 2 ? n2 : n3;
  float2 *tw = smem;
  float2 *a = tw + n;
  float2 *b = a + (size_t)n * G;
  const int tid = threadIdx.x;
  for (int k = tid; k < n; k += NT) tw[k] = twg[pl.tw + k];
  const NSItem it = items[blockIdx.y];
  // Pencil g of the block: first element at start + g gs, elements t es apart.
  // An axis-2 block lies in one k1 plane: its mask is a PlaneMask there.
  int k1, q0 = 0, u0 = 0, npen;
  int64_t start, gs, es;
  if (AXIS == 3) {
    q0 = blockIdx.x * G;  // rows q = k1 n2 + k2
    npen = h1 * n2 - q0;
    k1 = 0;
    start = (int64_t)q0 * n3;
    gs = n3;
    es = 1;
  } else {
    const int ublocks = (n3 + G - 1) / G;
    k1 = blockIdx.x / ublocks;
    u0 = (blockIdx.x - k1 * ublocks) * G;
    npen = n3 - u0;
    start = (int64_t)k1 * n2 * n3 + u0;
    gs = 1;
    es = n3;
  }
  PlaneMask pm{};
  if (AXIS == 2 && MODE != 0) pm = plane_mask(mv, it, k1, MODE == 1);
  // Axis 3 maps consecutive threads along the pencil, axis 2 across pencils:
  // both read and write runs of consecutive words.
  auto coords = [&](int idx, int &t, int &g) {
    if (AXIS == 3) {
      g = idx / n;
      t = idx - g * n;
    } else {
      t = idx / G;
      g = idx - t * G;
    }
  };
  auto bin = [&](int t, int g, int &b1, int &b2, int &b3) {
    if (AXIS == 3) {
      const int q = q0 + g;
      b1 = q / n2;
      b2 = q - b1 * n2;
      b3 = t;
    } else {
      b1 = k1;
      b2 = t;
      b3 = u0 + g;
    }
  };
  for (int idx = tid; idx < n * G; idx += NT) {
    int t, g;
    coords(idx, t, g);
    float2 val = make_float2(0.0f, 0.0f);
    if (g < npen) {
      val = in[start + g * gs + t * es];
      if (MODE == 1) {
        int b1, b2, b3;
        bin(t, g, b1, b2, b3);
        val = cscale(val, AXIS == 2 ? plane_mask_at(pm, b2 * n3 + b3, b2, b3) : ns_mask(mv, it, b1, b2, b3, true));
      }
    }
    a[t * G + g] = val;
  }
  __syncthreads();
  float2 *res = lds_fft<DIR, GEN>(a, b, pl, G, tw, tid, NT);
  for (int idx = tid; idx < n * G; idx += NT) {
    int t, g;
    coords(idx, t, g);
    if (g >= npen) continue;
    float2 r = res[t * G + g];
    if (MODE == 2) {
      int b1, b2, b3;
      bin(t, g, b1, b2, b3);
      r = cscale(r, AXIS == 2 ? plane_mask_at(pm, b2 * n3 + b3, b2, b3) : ns_mask(mv, it, b1, b2, b3, false));
    }
    out[start + g * gs + t * es] = r;
  }
}

// One radix-R Stockham pass over all npen pencils of one axis of an
// LDS-resident plane, in place: pencil p's element t at P[p ps + t es].
// Every thread reads the inputs of its (at most PLANE_V / R) butterflies
// into registers, the block synchronises, and the outputs overwrite the
// plane; the same index and twiddle rules as fft_pass_fixed.  The pass is a
// call of its own (inlined together the six radices hold more registers
// than a 1024-thread block has) and takes the plane and twiddles as offsets
// into the block's dynamic LDS, so that its accesses stay LDS accesses.
constexpr int PT = 1024;                 // threads of a plane block
constexpr int PLANE_V = 16;              // values a thread holds across one pass
constexpr size_t PLANE_LDS = 160 * 1024;  // LDS of one CU on gfx950
// TWO: instantiated per register budget (a callee compiles to the tightest
// budget among its callers).
template <int R, int DIR, bool TWO>
__device__ __noinline__ void plane_pass(int poff, int twoff, int n, int npen, int es, int ps, int Ns, int tid) {
  extern __shared__ float2 smem[];
  float2 *P = smem + poff;
  const float2 *tw = smem + twoff;
  constexpr int Q = PLANE_V / R;
  const int nR = n / R;
  const int tstep = n / (Ns * R);
  const int nbf = npen * nR;
  const FastDiv dpen = fast_div(npen), dns = fast_div(Ns);
  float2 v[Q][R];
  int at[Q];  // p ps + j es of each butterfly
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int w = tid + q * PT;
    if (w < nbf) {
      const int j = div_of(w, dpen), p = w - j * npen;
      at[q] = p * ps + j * es;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = P[at[q] + r * nR * es];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int w = tid + q * PT;
    if (w < nbf) {
      const int j = div_of(w, dpen);
      const int jn = div_of(j, dns);
      const int k = j - jn * Ns;
      if (k) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[q][r] = cmul(v[q][r], twid<DIR>(tw, r * k * tstep));
      }
      dft_radix<R, DIR>(v[q]);
      // output index jn Ns R + k = j + jn Ns (R - 1)
      const int o = at[q] + jn * Ns * (R - 1) * es;
#pragma unroll
      for (int r = 0; r < R; ++r) P[o + r * Ns * es] = v[q][r];
    }
  }
  __syncthreads();
}

// Whether radix R fits the plane pass for npen pencils of length n.
inline bool plane_radix_fits(int R, int64_t elems) {
  return fft_radix_fixed(R) && (elems / R + PT - 1) / PT <= PLANE_V / R;
}

template <int DIR, bool TWO>
__device__ void plane_fft(int poff, int twoff, const FftPlan1D &pl, int npen, int es, int ps, int tid) {
  int Ns = 1;
  for (int p = 0; p < pl.npass; ++p) {
    const int R = pl.radix[p];
    switch (R) {
      case 8: plane_pass<8, DIR, TWO>(poff, twoff, pl.n, npen, es, ps, Ns, tid); break;
      case 4: plane_pass<4, DIR, TWO>(poff, twoff, pl.n, npen, es, ps, Ns, tid); break;
      case 2: plane_pass<2, DIR, TWO>(poff, twoff, pl.n, npen, es, ps, Ns, tid); break;
      case 3: plane_pass<3, DIR, TWO>(poff, twoff, pl.n, npen, es, ps, Ns, tid); break;
      case 5: plane_pass<5, DIR, TWO>(poff, twoff, pl.n, npen, es, ps, Ns, tid); break;
      case 7: plane_pass<7, DIR, TWO>(poff, twoff, pl.n, npen, es, ps, Ns, tid); break;
      case 16: plane_pass<16, DIR, TWO>(poff, twoff, pl.n, npen, es, ps, Ns, tid); break;
      case 14: plane_pass<14, DIR, TWO>(poff, twoff, pl.n, npen, es, ps, Ns, tid); break;
      case 12: plane_pass<12, DIR, TWO>(poff, twoff, pl.n, npen, es, ps, Ns, tid); break;
      case 10: plane_pass<10, DIR, TWO>(poff, twoff, pl.n, npen, es, ps, Ns, tid); break;
      case 9: plane_pass<9, DIR, TWO>(poff, twoff, pl.n, npen, es, ps, Ns, tid); break;
      default: plane_pass<6, DIR, TWO>(poff, twoff, pl.n, npen, es, ps, Ns, tid); break;
    }
    Ns *= R;
  }
}

// 2-D FFT over the (k2, k3) plane of one k1 (blockIdx.x) of a plane-major
// spectrum, the plane held in LDS with row pitch n3p (odd, so that strided
// pencils fall in distinct banks).  Modes as k_axis: 0 plain, 1 mask before
// the (inverse) transform, 2 mask after the (forward) transform; batch
// entry blockIdx.y.  Replaces one k_axis pass per axis.  TWO: built for two
// blocks per CU (64 VGPRs), for planes whose LDS lets two blocks share a CU.
template <int DIR, int MODE, bool TWO>
__global__ __launch_bounds__(PT, TWO ? 8 : 4) void k_plane(const float2 *in, int64_t in_bs, float2 *out, int64_t out_bs,
                                              int n2, int n3, int n3p, FftPlan1D p2, FftPlan1D p3,
                                              const float2 *__restrict__ twg, NSMaskView mv,
                                              const NSItem *__restrict__ items) {
  in += blockIdx.y * in_bs;
  out += blockIdx.y * out_bs;
  extern __shared__ float2 smem[];
  float2 *t2 = smem;
  float2 *t3 = t2 + n2;
  float2 *P = t3 + n3;
  const int tid = threadIdx.x;
  for (int k = tid; k < n2; k += PT) t2[k] = twg[p2.tw + k];
  for (int k = tid; k < n3; k += PT) t3[k] = twg[p3.tw + k];
  const int k1 = blockIdx.x;
  const int nel = n2 * n3;
  const int64_t base = (int64_t)k1 * nel;
  const PlaneMask pm = plane_mask(mv, items[blockIdx.y], k1, MODE == 1);
  const FastDiv d3 = fast_div(n3);
  for (int i = tid; i < nel; i += PT) {
    const int k2 = div_of(i, d3), k3 = i - k2 * n3;
    float2 v = in[base + i];
    if (MODE == 1) v = cscale(v, plane_mask_at(pm, i, k2, k3));
    P[k2 * n3p + k3] = v;
  }
  __syncthreads();
  plane_fft<DIR, TWO>(n2 + n3, n2, p3, n2, 1, n3p, tid);  // along k3: pencils are rows
  plane_fft<DIR, TWO>(n2 + n3, 0, p2, n3, n3p, 1, tid);   // along k2: pencils are columns
  for (int i = tid; i < nel; i += PT) {
    const int k2 = div_of(i, d3), k3 = i - k2 * n3;
    float2 v = P[k2 * n3p + k3];
    if (MODE == 2) v = cscale(v, plane_mask_at(pm, i, k2, k3));
    out[base + i] = v;
  }
}

// FFT along axis AXIS (2 or 3) of a plane-major spectrum for a strip of
// G pencils held in LDS, with the plane kernel's in-place passes: for axis 3,
// G consecutive (k1, k2) rows (one contiguous run, LDS [g][t]); for axis 2,
// G consecutive k3 of one k1 (runs of G words, LDS [t][g]).  The LDS pitch
// is odd.  Modes as k_axis; an axis-2 strip lies in one k1 plane, so its
// mask is a PlaneMask there (axis-3 strips are MODE 0 only).  For planes
// too large for k_plane; k_axis stays for lengths the strip cannot take.
template <int AXIS, int DIR, int MODE, bool TWO>
__global__ __launch_bounds__(PT, TWO ? 8 : 4) void k_strip(const float2 *in, int64_t in_bs, float2 *out,
                                                           int64_t out_bs, int h1, int n2, int n3, int G, int pitch,
                                                           FftPlan1D pl, const float2 *__restrict__ twg,
                                                           NSMaskView mv, const NSItem *__restrict__ items) {
  static_assert(AXIS == 2 || MODE == 0, "axis-3 strips are unmasked");
  in += blockIdx.y * in_bs;
  out += blockIdx.y * out_bs;
  extern __shared__ float2 smem[];
  const int n = AXIS == 2 ? n2 : n3;
  float2 *S = smem + n;
  const int tid = threadIdx.x;
  for (int k = tid; k < n; k += PT) smem[k] = twg[pl.tw + k];
  int64_t start;
  int npen, k1 = 0, u0 = 0;
  if (AXIS == 3) {
    const int q0 = blockIdx.x * G;
    npen = min(G, h1 * n2 - q0);
    start = (int64_t)q0 * n3;
  } else {
    const int ub = (n3 + G - 1) / G;
    k1 = blockIdx.x / ub;
    u0 = (blockIdx.x - k1 * ub) * G;
    npen = min(G, n3 - u0);
    start = (int64_t)k1 * n2 * n3 + u0;
  }
  PlaneMask pm{};
  if (AXIS == 2 && MODE != 0) pm = plane_mask(mv, items[blockIdx.y], k1, MODE == 1);
  // Element e of the strip: axis 3 e = g n3 + t at start + e; axis 2
  // e = t npen + g at start + t n3 + g.
  const int inner = AXIS == 3 ? n3 : npen;
  const int nel = n * npen;
  const FastDiv dd = fast_div(inner);
  for (int e = tid; e < nel; e += PT) {
    const int hi = div_of(e, dd), lo = e - hi * inner;
    float2 v = in[AXIS == 3 ? start + e : start + (int64_t)hi * n3 + lo];
    if (AXIS == 2 && MODE == 1) v = cscale(v, plane_mask_at(pm, hi * n3 + u0 + lo, hi, u0 + lo));
    S[hi * pitch + lo] = v;
  }
  __syncthreads();
  if (AXIS == 3)
    plane_fft<DIR, TWO>(n, 0, pl, npen, 1, pitch, tid);
  else
    plane_fft<DIR, TWO>(n, 0, pl, npen, pitch, 1, tid);
  for (int e = tid; e < nel; e += PT) {
    const int hi = div_of(e, dd), lo = e - hi * inner;
    float2 v = S[hi * pitch + lo];
    if (AXIS == 2 && MODE == 2) v = cscale(v, plane_mask_at(pm, hi * n3 + u0 + lo, hi, u0 + lo));
    out[AXIS == 3 ? start + e : start + (int64_t)hi * n3 + lo] = v;
  }
}

// acc += W_0 + W_1 + ... + W_{nb-1}, added in that order (W_b at W + b bs),
// so a batch sums exactly as nb single-entry calls do.  from_zero: the sum
// starts from 0 instead of acc, as after zeroing acc.
__global__ __launch_bounds__(NT) void k_sum(float2 *__restrict__ acc, const float2 *__restrict__ W, int64_t bs,
                                            int nb, int64_t n, int from_zero) {
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
    float2 s = from_zero ? make_float2(0.0f, 0.0f) : acc[i];
    for (int b = 0; b < nb; ++b) {
      const float2 w = W[b * bs + i];
      s = make_float2(s.x + w.x, s.y + w.y);
    }
    acc[i] = s;
  }
}

// The group-factor tables on the device, one voxel per thread, from the
// small FP64 inputs the host prepared (the same ns_group_voxel the host build
// runs, so the two agree bit for bit).
__global__ __launch_bounds__(NT) void k_group_tables(NSGroupInputs in, float *__restrict__ ga,
                                                     float *__restrict__ gs, int64_t V) {
  for (int64_t q = (int64_t)blockIdx.x * NT + threadIdx.x; q < V; q += (int64_t)gridDim.x * NT) {
    const int k3 = (int)(q % in.n3);
    const int64_t r = q / in.n3;
    ns_group_voxel(in, (int)(r / in.n2), (int)(r % in.n2), k3, ga, gs, (size_t)V);
  }
}

__global__ void k_zero(float2 *__restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = make_float2(0.0f, 0.0f);
}

}  // namespace surfhip

// ---------------------------------------------------------------------- plan

using namespace surfhip;

// A plan runs one of two kernel sets on the same tables: the factored path
// (ns_fused.hip) on the grids it covers (every length in {32, 64, 80, 96,
// 112, 128}), or the general path above on any other size.  Both evaluate
// the same masks.
struct surfhip_ns_plan {
  NSTables T;
  int dtype = SURFHIP_F32;
  int nch = 1;
  hipStream_t stream = nullptr;
  // The two orientations of the factored path are independent until the
  // synthesis sums meet: AX2 work runs on `aux`, forked from and joined back
  // to `stream` with events (not while per-launch timing is on).
  hipStream_t aux = nullptr;
  hipEvent_t efork = nullptr, ejoin = nullptr;
  int64_t N = 0, H = 0;
  int n1 = 0, n2 = 0, n3 = 0, h1 = 0;
  FftPlan1D f1{}, f2{}, f3{};
  int g_rows = 16, g_axis = 16;
  int n3p = 1;        // plane row pitch: n3 rounded up to odd
  bool plane = false;  // axes 2 and 3 by the plane kernel
  bool strip[2] = {false, false};  // axis 2 / axis 3 by the strip kernel (when not plane)
  int nwork = 1;                    // subbands per general-path launch (work holds nwork spectra)
  float tuned_ms = 0.0f;            // timed sequence of the chosen configuration (0: not tuned)
  DevBuf<float> gana, gsyn, planes;
  DevBuf<float2> tw, X[2], acc[2], work;
  DevBuf<NSItem> ditems;
  DevBuf<NSItem> dvisit;  // items in visiting order: 1 .. S-1, then the lowpass 0
  NSMaskView dview{};
  bool streaming_fwd = false;

  // ---- factored path
  bool fused = false;
  FusedOps fops;
  int nslabs = 0;
  std::vector<GroupDev> groups;     // 3 J groups, then the lowpass
  std::vector<int> slab_group;      // group of each slab
  std::vector<uint32_t> live;       // nslabs x h1 tile words
  std::vector<int4> units;          // analysis pencil units, group-major, AX3 groups first
  std::vector<int> ubeg, uend;      // units of each group
  int ax_ubeg[2] = {0, 0}, ax_uend[2] = {0, 0};
  std::vector<PlaneJob> jobs;       // [analysis AX3, AX2][synthesis AX3, AX2][whole grid][S fresh][S continuing]
  int ja_ax[2] = {0, 0}, jna_ax[2] = {0, 0}, js_ax[2] = {0, 0}, jns_ax[2] = {0, 0};
  int j_full = 0, j_fresh = 0, j_cont = 0;
  // Synthesis slab order {slab, group, item of q, 0}: AX2 groups, then AX3
  // groups and the lowpass, q ascending within a group.
  std::vector<int4> slist;
  int sl_beg[2] = {0, 0}, sl_n[2] = {0, 0};
  DevBuf<GroupDev> dgroups;
  DevBuf<uint32_t> dlive;
  DevBuf<int4> dunits;
  DevBuf<PlaneJob> djobs;
  DevBuf<int4> dslist;
  DevBuf<float2> U[2], R[2], TV, accf;
  int cached_group = -1;            // streaming analysis: the group whose slabs U holds
  std::vector<char> touched;        // streaming synthesis: slabs holding a partial sum

  // ---- optional per-class event timing
  bool timing = false;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> tev;
  int64_t tcount[SURFHIP_NS_TIMING_CLASSES] = {0};
  double tms[SURFHIP_NS_TIMING_CLASSES] = {0};

  // Run `launch` bracketed by events of class `cls` when timing is on.
  template <class F> void timed(int cls, F &&launch) {
    if (!timing) {
      launch();
      return;
    }
    hipEvent_t a, b;
    SH_HIP(hipEventCreate(&a));
    SH_HIP(hipEventCreate(&b));
    SH_HIP(hipEventRecord(a, stream));
    launch();
    SH_HIP(hipEventRecord(b, stream));
    tev.push_back({cls, {a, b}});
  }
  void collect_timing() {
    if (tev.empty()) return;
    SH_HIP(hipStreamSynchronize(stream));
    for (auto &e : tev) {
      float ms = 0.0f;
      SH_HIP(hipEventElapsedTime(&ms, e.second.first, e.second.second));
      tcount[e.first] += 1;
      tms[e.first] += ms;
      (void)hipEventDestroy(e.second.first);
      (void)hipEventDestroy(e.second.second);
    }
    tev.clear();
  }
  ~surfhip_ns_plan() {
    for (auto &e : tev) {
      (void)hipEventDestroy(e.second.first);
      (void)hipEventDestroy(e.second.second);
    }
    if (aux) (void)hipStreamDestroy(aux);
    if (efork) (void)hipEventDestroy(efork);
    if (ejoin) (void)hipEventDestroy(ejoin);
  }

  FusedCtx ctx() const {
    return FusedCtx{stream, n1, n2, n3, h1, H, dview, ditems.p, dgroups.p, dlive.p,
                    tw.p + f1.tw, tw.p + f2.tw, tw.p + f3.tw};
  }
  int group_of(int i) const { return T.items[i].lowpass ? (int)groups.size() - 1 : T.items[i].group; }

  // ---- factored-path set-up: groups, slabs, live tiles, units and jobs.
  void build_fused() {
    const int G = (int)T.group_first.size();
    const int TW = fops.tile;
    groups.clear();
    slab_group.clear();
    for (int g = 0; g <= G; ++g) {
      GroupDev gd{};
      if (g < G) {
        gd.first = T.group_first[g];
        gd.nA = T.group_na[g];
        gd.nB = T.group_nb[g];
        gd.orient = T.items[gd.first].hourglass == 3 ? AX2 : AX3;
        gd.lowpass = 0;
      } else {
        gd = GroupDev{0, 1, 1, AX3, 0, 1};
      }
      gd.slab0 = (int)slab_group.size();
      for (int q = 0; q < gd.nB; ++q) slab_group.push_back(g);
      groups.push_back(gd);
    }
    nslabs = (int)slab_group.size();
    // Live tiles: tile t of row k1 of slab (g, q) holds a nonzero iff some bin
    // of it has a nonzero group factor and b-plane, for the analysis or the
    // synthesis masks.  Exact zeros of the tables decide it, so skipping dead
    // tiles changes no result.
    // The group tables live on the device; the liveness scan reads a host copy.
    const size_t VT = (size_t)H * T.ngroups;
    std::vector<float> hga(VT), hgs(T.tight ? 0 : VT);
    SH_HIP(hipMemcpy(hga.data(), gana.p, VT * sizeof(float), hipMemcpyDeviceToHost));
    if (!T.tight) SH_HIP(hipMemcpy(hgs.data(), gsyn.p, VT * sizeof(float), hipMemcpyDeviceToHost));
    NSMaskView hv = T.view();
    hv.gana = hga.data();
    hv.gsyn = T.tight ? hga.data() : hgs.data();
    const bool noprune = std::getenv("SURFHIP_NS_NO_PRUNE") != nullptr;
    live.assign((size_t)nslabs * h1, 0u);
    const size_t V = (size_t)H;
    for (int g = 0; g <= G; ++g) {
      const GroupDev &gd = groups[g];
      const bool ax3 = gd.orient == AX3;
      const int nt = ax3 ? n2 : n3, np = ax3 ? n3 : n2, ntiles = nt / TW;
      const uint32_t all = ntiles >= 32 ? 0xffffffffu : ((1u << ntiles) - 1u);
      const float *ga = hv.gana + (size_t)g * V;
      const float *gs = hv.gsyn + (size_t)g * V;
      for (int q = 0; q < gd.nB; ++q) {
        const NSItem &it = T.items[gd.first + q];
        uint32_t *lw = live.data() + (size_t)(gd.slab0 + q) * h1;
        for (int k1 = 0; k1 < h1; ++k1) {
          if (noprune) {
            lw[k1] = all;
            continue;
          }
          uint32_t w = 0;
          for (int kt = 0; kt < nt; ++kt) {
            const int t = kt / TW;
            if ((w >> t) & 1u) continue;
            for (int kp = 0; kp < np; ++kp) {
              const int k2 = ax3 ? kt : kp, k3 = ax3 ? kp : kt;
              const size_t q3 = ((size_t)k1 * n2 + k2) * n3 + k3;
              const float wb = it.lowpass ? 1.0f : ns_plane(hv, it.pb_off, it.pb_pair, k1, k2, k3);
              const bool la = ga[q3] != 0.0f && wb != 0.0f;
              const bool ls = gs[q3] != 0.0f && (!T.tight || wb != 0.0f);
              if (la || ls) {
                w |= 1u << t;
                break;
              }
            }
          }
          lw[k1] = w;
        }
      }
    }
    // Analysis units, AX3 groups first.
    units.clear();
    ubeg.assign(G + 1, 0);
    uend.assign(G + 1, 0);
    for (int o = 0; o < 2; ++o) {
      ax_ubeg[o] = (int)units.size();
      for (int g = 0; g <= G; ++g) {
        const GroupDev &gd = groups[g];
        if (gd.orient != o) continue;
        const int ntiles = (o == AX3 ? n2 : n3) / TW;
        ubeg[g] = (int)units.size();
        for (int k1 = 0; k1 < h1; ++k1)
          for (int t = 0; t < ntiles; ++t) {
            bool any = false;
            for (int q = 0; q < gd.nB; ++q) any |= ((live[(size_t)(gd.slab0 + q) * h1 + k1] >> t) & 1u) != 0;
            if (any) units.push_back(make_int4(g, k1, t * TW, 0));
          }
        uend[g] = (int)units.size();
      }
      ax_uend[o] = (int)units.size();
    }
    // Plane job of slab (g, q), a-indices [p0, p1): plane a of item (p, q)
    // sits at pa0 + (p - p0) pa_step in the pool (the planes of one wedge
    // level are stored back to back).
    auto mkjob = [&](int g, int q, int p0, int p1, int init, int out0, int step) {
      const GroupDev &gd = groups[g];
      PlaneJob jb{};
      jb.slab = gd.slab0 + q;
      jb.group = g;
      jb.q = q;
      jb.p0 = p0;
      jb.p1 = p1;
      jb.init = init;
      jb.out0 = out0;
      jb.out_step = step;
      const int first = gd.first;
      jb.pa0 = gd.lowpass ? -1 : T.items[first + p0 * gd.nB + q].pa_off;
      jb.pa_step = (gd.lowpass || gd.nA < 2) ? 0 : T.items[first + gd.nB + q].pa_off - T.items[first + q].pa_off;
      for (int p = p0; p < p1 && !gd.lowpass; ++p)
        if (T.items[first + p * gd.nB + q].pa_off != (jb.pa0 < 0 ? -1 : jb.pa0 + (p - p0) * jb.pa_step))
          fail(SURFHIP_ERR_ARG, "internal: wedge planes of a level are not evenly spaced");
      return jb;
    };
    // Plane jobs.  Analysis: one job per slab, its nA subbands one after the
    // other in each block, the longest jobs first: blocks of unequal length
    // drift apart in time, so that some load while others compute.
    // Synthesis: one job per slab, summing its nA subbands in order.
    jobs.clear();
    for (int o = 0; o < 2; ++o) {
      ja_ax[o] = (int)jobs.size();
      for (int g = 0; g <= G; ++g) {
        const GroupDev &gd = groups[g];
        if (gd.orient != o) continue;
        for (int q = 0; q < gd.nB; ++q) jobs.push_back(mkjob(g, q, 0, gd.nA, 1, gd.first + q, gd.nB));
      }
      std::stable_sort(jobs.begin() + ja_ax[o], jobs.end(),
                       [](const PlaneJob &a, const PlaneJob &b) { return a.p1 - a.p0 > b.p1 - b.p0; });
      jna_ax[o] = (int)jobs.size() - ja_ax[o];
    }
    for (int o = 0; o < 2; ++o) {
      js_ax[o] = (int)jobs.size();
      for (int g = 0; g <= G; ++g) {
        const GroupDev &gd = groups[g];
        if (gd.orient != o) continue;
        for (int q = 0; q < gd.nB; ++q)
          jobs.push_back(mkjob(g, q, 0, gd.nA, 1, gd.first + q, gd.nB));
      }
      jns_ax[o] = (int)jobs.size() - js_ax[o];
    }
    j_full = (int)jobs.size();
    {
      PlaneJob full{};
      full.group = -1;
      full.p1 = 1;
      full.init = 1;
      full.pa0 = -1;
      jobs.push_back(full);
    }
    for (int fresh = 1; fresh >= 0; --fresh) {
      (fresh ? j_fresh : j_cont) = (int)jobs.size();
      for (int s = 0; s < T.S; ++s) {
        const int g = group_of(s);
        const GroupDev &gd = groups[g];
        const int w = s - gd.first, p = w / gd.nB, q = w % gd.nB;
        jobs.push_back(mkjob(g, q, p, p + 1, fresh, 0, 0));
      }
    }
    // Synthesis group order.
    slist.clear();
    for (int o : {AX2, AX3}) {
      sl_beg[o] = (int)slist.size();
      for (int g = 0; g <= G; ++g)
        if (groups[g].orient == o)
          for (int q = 0; q < groups[g].nB; ++q) slist.push_back(make_int4(groups[g].slab0 + q, g, groups[g].first + q, 0));
      sl_n[o] = (int)slist.size() - sl_beg[o];
    }
    dgroups.upload(groups);
    dlive.upload(live);
    dunits.upload(units);
    djobs.upload(jobs);
    dslist.upload(slist);
    for (int ch = 0; ch < nch; ++ch) {
      U[ch].alloc((size_t)nslabs * H);
      R[ch].alloc((size_t)nslabs * H);
    }
    TV.alloc(H);
    accf.alloc(H);
    touched.assign(nslabs, 0);
  }

  // Rows kernels: twiddles, (c2r: the G half-spectrum rows,) the ping-pong
  // pair and the G row offsets.
  size_t lds_rows_r2c() const { return sizeof(float2) * ((size_t)n1 + 2 * (size_t)n1 * (g_rows / 2) + g_rows); }
  size_t lds_rows_c2r() const {
    return sizeof(float2) * ((size_t)n1 + (size_t)g_rows * h1 + 2 * (size_t)n1 * (g_rows / 2) + g_rows);
  }
  size_t lds_axis(int n) const { return sizeof(float2) * ((size_t)n + 2 * (size_t)n * g_axis); }

  int rows() const { return n2 * n3; }
  // The plane kernel's LDS: both axes' twiddles and one (k2, k3) plane of
  // row pitch n3p.
  size_t lds_plane() const { return sizeof(float2) * ((size_t)n2 + n3 + (size_t)n2 * n3p); }
  bool plane_two() const { return 2 * lds_plane() <= PLANE_LDS; }
  // Whether the plane kernel takes axes 2 and 3: the plane fits one block's
  // LDS and every radix of both plans has a hand-written pass that holds its
  // butterflies in registers.
  bool plane_fits() const {
    if (lds_plane() > PLANE_LDS) return false;
    for (const FftPlan1D *pl : {&f2, &f3})
      for (int k = 0; k < pl->npass; ++k)
        if (!plane_radix_fits(pl->radix[k], (int64_t)n2 * n3)) return false;
    return true;
  }

  // General-path launchers; nb > 1 runs a batch of nb subbands (grid.y),
  // entry b at in + b in_bs / out + b out_bs with item item0 + b.
  // Entries from wrap_at on (if any) take their input at x + wrap_off.
  void r2c(const float *x, int64_t xs, float2 *out, int nb = 1, int64_t x_bs = 0, int64_t out_bs = 0,
           int wrap_at = -1, int64_t wrap_off = 0) {
    if (wrap_at < 0) wrap_at = nb;
    int grid = (rows() + g_rows - 1) / g_rows;
    const int vec = xs == 1 && n1 % 4 == 0 && x_bs % 4 == 0 && wrap_off % 4 == 0 &&
                    (reinterpret_cast<uintptr_t>(x) & 15) == 0;
    timed(0, [&] {
      hipLaunchKernelGGL(fft_plan_generic(f1) ? k_rows_r2c<true> : k_rows_r2c<false>, dim3(grid, nb), dim3(NT),
                         lds_rows_r2c(), stream, x, xs, x_bs, out, out_bs, n1, h1, n2, n3, g_rows, f1, tw.p, vec,
                         wrap_at, wrap_off);
    });
    SH_LAUNCHED();
  }
  // Entries from wrap_at on (if any) write at out + wrap_off.
  void c2r(const float2 *in, float *out, int64_t os, float scale, int nb = 1, int64_t in_bs = 0,
           int64_t out_bs = 0, int wrap_at = -1, int64_t wrap_off = 0) {
    if (wrap_at < 0) wrap_at = nb;
    int grid = (rows() + g_rows - 1) / g_rows;
    const int vin = rows() % 2 == 0 && in_bs % 2 == 0 && (reinterpret_cast<uintptr_t>(in) & 15) == 0;
    const int vout = os == 1 && n1 % 4 == 0 && out_bs % 4 == 0 && wrap_off % 4 == 0 &&
                     (reinterpret_cast<uintptr_t>(out) & 15) == 0;
    timed(1, [&] {
      hipLaunchKernelGGL(fft_plan_generic(f1) ? k_rows_c2r<true> : k_rows_c2r<false>, dim3(grid, nb), dim3(NT),
                         lds_rows_c2r(), stream, in, in_bs, out, os, out_bs, scale, n1, h1, n2, n3, g_rows, f1, tw.p,
                         vin | (vout << 1), wrap_at, wrap_off);
    });
    SH_LAUNCHED();
  }
  // Strip kernel of axis a (2 or 3): pencils per strip, LDS pitch and bytes.
  int strip_g(int a) const { return std::max(1, std::min(32, 8192 / (a == 2 ? n2 : n3))); }
  int strip_pitch(int a) const { return a == 2 ? (strip_g(2) | 1) : (n3 | 1); }
  size_t lds_strip(int a) const {
    const int n = a == 2 ? n2 : n3;
    return sizeof(float2) * ((size_t)n + (size_t)(a == 2 ? n2 : strip_g(3)) * strip_pitch(a));
  }
  // Whether the strip kernel takes axis a: every radix hand-written and each
  // pass's butterflies held in registers.
  bool strip_fits(int a) const {
    const FftPlan1D &pl = a == 2 ? f2 : f3;
    if (lds_strip(a) > PLANE_LDS) return false;
    for (int k = 0; k < pl.npass; ++k)
      if (!plane_radix_fits(pl.radix[k], (int64_t)pl.n * strip_g(a))) return false;
    return true;
  }
  template <int AXIS, int DIR, int MODE>
  void axis(const float2 *in, float2 *out, int item0 = 0, int nb = 1, int64_t in_bs = 0, int64_t out_bs = 0,
            const NSItem *items = nullptr) {
    if (!items) items = ditems.p + item0;
    if (strip[AXIS - 2]) {
      const int G = strip_g(AXIS);
      const int grid = AXIS == 3 ? (h1 * n2 + G - 1) / G : h1 * ((n3 + G - 1) / G);
      const size_t lds = lds_strip(AXIS);
      auto *kernel = 2 * lds <= PLANE_LDS ? &k_strip<AXIS, DIR, MODE, true> : &k_strip<AXIS, DIR, MODE, false>;
      timed(DIR < 0 ? 2 : 3, [&] {
        hipLaunchKernelGGL(kernel, dim3(grid, nb), dim3(PT), lds, stream, in, in_bs, out, out_bs, h1, n2, n3, G,
                           strip_pitch(AXIS), AXIS == 2 ? f2 : f3, tw.p, dview, items);
      });
      SH_LAUNCHED();
      return;
    }
    const int G = g_axis;
    const int grid = AXIS == 3 ? (h1 * n2 + G - 1) / G : h1 * ((n3 + G - 1) / G);
    const FftPlan1D &pl = AXIS == 2 ? f2 : f3;
    if ((int64_t)pl.n * G > (int64_t)AXIS_MAXE * NT)
      fail(SURFHIP_ERR_ARG, fmt("axis block of %d x %d elements exceeds %d", pl.n, G, AXIS_MAXE * NT));
    auto *kernel = fft_plan_generic(pl) ? &k_axis<AXIS, DIR, MODE, true> : &k_axis<AXIS, DIR, MODE, false>;
    timed(DIR < 0 ? 2 : 3, [&] {
      hipLaunchKernelGGL(kernel, dim3(grid, nb), dim3(NT), lds_axis(pl.n), stream, in, in_bs, out, out_bs, h1, n2,
                         n3, G, pl, tw.p, dview, items);
    });
    SH_LAUNCHED();
  }
  // FFT over axes 2 and 3 (DIR), modes as k_axis: the plane kernel where the
  // plane fits, else one k_axis pass per axis.
  template <int DIR, int MODE>
  void axes23(const float2 *in, float2 *out, int item0 = 0, int nb = 1, int64_t in_bs = 0, int64_t out_bs = 0,
              const NSItem *items = nullptr) {
    if (!items) items = ditems.p + item0;
    if (plane) {
      auto *kernel = plane_two() ? &k_plane<DIR, MODE, true> : &k_plane<DIR, MODE, false>;
      timed(DIR < 0 ? 2 : 3, [&] {
        hipLaunchKernelGGL(kernel, dim3(h1, nb), dim3(PT), lds_plane(), stream, in, in_bs, out, out_bs, n2, n3, n3p,
                           f2, f3, tw.p, dview, items);
      });
      SH_LAUNCHED();
      return;
    }
    // Both masks on the axis-2 pass, whose blocks each lie in one k1 plane:
    // analysis transforms axis 2 first, synthesis last.
    if (MODE == 1) {
      axis<2, DIR, 1>(in, out, item0, nb, in_bs, out_bs, items);
      axis<3, DIR, 0>(out, out, item0, nb, out_bs, out_bs, items);
    } else {
      axis<3, DIR, 0>(in, out, item0, nb, in_bs, out_bs, items);
      axis<2, DIR, MODE == 2 ? 2 : 0>(out, out, item0, nb, out_bs, out_bs, items);
    }
  }
  // acc += work_0 + ... + work_{nb-1} in order (from_zero: acc = 0 first).
  void sum_into(float2 *a, const float2 *w, int nb, int from_zero = 0) {
    const int64_t blocks = std::min<int64_t>((H + NT - 1) / NT, 4096);
    timed(4, [&] {
      hipLaunchKernelGGL(k_sum, dim3((unsigned)blocks), dim3(NT), 0, stream, a, w, H, nb, H, from_zero);
    });
    SH_LAUNCHED();
  }
  void zero(float2 *p, int64_t n) {
    hipLaunchKernelGGL(k_zero, dim3(1024), dim3(NT), 0, stream, p, n);
    SH_LAUNCHED();
  }

  // Autotuner of the general path: block shapes (g_rows rows per axis-1
  // block, g_axis pencils per axis-2/3 block) and the radix order of each
  // axis's Stockham plan, chosen by timing this plan's own launch sequence
  // for one batch of nwork subbands (forward spectrum, a batched masked
  // analysis, a batched masked accumulation: every kernel of a step, at the
  // batch size a step launches) on the plan's stream.  Coordinate descent:
  // each knob in turn over its candidates, the others fixed.
  float time_config(const float *x, float *band) {
    hipEvent_t a, b;
    SH_HIP(hipEventCreate(&a));
    SH_HIP(hipEventCreate(&b));
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      SH_HIP(hipEventRecord(a, stream));
      forward_spectrum(x);
      const int nw = std::min(nwork, T.S - 1);
      subbands_generic(1, nw, band);
      accumulate_generic(1, nw, band);
      SH_HIP(hipEventRecord(b, stream));
      SH_HIP(hipEventSynchronize(b));
      float ms = 0.0f;
      SH_HIP(hipEventElapsedTime(&ms, a, b));
      best = std::min(best, ms);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return best;
  }
  static void reverse_radices(FftPlan1D &pl) { std::reverse(pl.radix, pl.radix + pl.npass); }
  void tune() {
    if (fused) return;
    DevBuf<float> x, band;
    x.alloc((size_t)N * nch);
    band.alloc((size_t)N * nch * std::max(1, std::min(nwork, T.S - 1)));
    SH_HIP(hipMemsetAsync(x.p, 0, sizeof(float) * (size_t)N * nch, stream));
    float best = time_config(x.p, band.p);
    auto try_knob = [&](int *knob, std::initializer_list<int> cands, auto fits) {
      const int keep = *knob;
      int pick = keep;
      for (int v : cands) {
        if (v == keep) continue;
        *knob = v;
        if (!fits()) continue;
        const float ms = time_config(x.p, band.p);
        if (ms < best) {
          best = ms;
          pick = v;
        }
      }
      *knob = pick;
    };
    if (!plane)
      try_knob(&g_axis, {1, 2, 4, 8, 16, 32},
               [&] { return lds_axis(n2) <= 65536 && lds_axis(n3) <= 65536; });
    try_knob(&g_rows, {2, 4, 6, 8, 14, 16, 30, 32},
             [&] { return lds_rows_r2c() <= 65536 && lds_rows_c2r() <= 65536; });
    for (FftPlan1D *pl : {&f1, &f2, &f3}) {
      if (pl->npass < 2) continue;
      reverse_radices(*pl);
      const float ms = time_config(x.p, band.p);
      if (ms < best)
        best = ms;
      else
        reverse_radices(*pl);
    }
    tuned_ms = best;
  }

  // X[ch] = rfft(channel ch of x)
  void forward_spectrum(const float *x) {
    cached_group = -1;
    for (int ch = 0; ch < nch; ++ch) {
      if (fused) {
        const FusedCtx c = ctx();
        timed(2, [&] { fops.q_syn[AX3](c, AX3, x + ch, 0, nch, TV.p, djobs.p + j_full, 1); });
        timed(4, [&] { fops.p_fwd(c, TV.p, X[ch].p); });
        continue;
      }
      r2c(x + ch, nch, X[ch].p);
      axes23<-1, 0>(X[ch].p, X[ch].p);
    }
  }
  // Generic path: subbands item0 .. item0+nw-1 into out (subband stride
  // N nch), up to nwork of them per launch.
  void subbands_generic(int item0, int nw, float *out) {
    const float scale = 1.0f / (float)N;
    for (int ch = 0; ch < nch; ++ch)
      for (int w0 = 0; w0 < nw; w0 += nwork) {
        const int nb = std::min(nwork, nw - w0);
        axes23<1, 1>(X[ch].p, work.p, item0 + w0, nb, 0, H);
        c2r(work.p, out + (int64_t)w0 * N * nch + ch, nch, scale, nb, H, N * nch);
      }
  }
  // Generic path: acc += sum over subbands item0 .. item0+nw-1 (in order) of
  // rfft(band) . M_i, bands at in (subband stride N nch): the masked spectra
  // of a batch go to work, then one pass adds them to acc in order.
  void accumulate_generic(int item0, int nw, const float *in) {
    for (int ch = 0; ch < nch; ++ch)
      for (int w0 = 0; w0 < nw; w0 += nwork) {
        const int nb = std::min(nwork, nw - w0);
        r2c(in + (int64_t)w0 * N * nch + ch, nch, work.p, nb, N * nch, H);
        axes23<-1, 2>(work.p, work.p, item0 + w0, nb, H, H);
        sum_into(acc[ch].p, work.p, nb);
      }
  }
  // The fork and join events only order two streams of this device, so they
  // skip the system-scope fence (cache writeback) that an event otherwise
  // performs when it is recorded: with it, each fork or join left the GPU
  // idle for 7-18 us between the kernels it orders.
  void make_aux() {
    SH_HIP(hipStreamCreateWithFlags(&aux, hipStreamNonBlocking));
    SH_HIP(hipEventCreateWithFlags(&efork, hipEventDisableTiming | hipEventDisableSystemFence));
    SH_HIP(hipEventCreateWithFlags(&ejoin, hipEventDisableTiming | hipEventDisableSystemFence));
  }
  void fork_aux() {
    SH_HIP(hipEventRecord(efork, stream));
    SH_HIP(hipStreamWaitEvent(aux, efork, 0));
  }
  void join_aux() {
    SH_HIP(hipEventRecord(ejoin, aux));
    SH_HIP(hipStreamWaitEvent(stream, ejoin, 0));
  }
  // Under stream capture the orientations stay on `stream`: a capture that
  // forked to `aux` with these events crashed inside the runtime at capture
  // end, so the captured sequence is the single-stream one.
  bool concurrent() const {
    if (!fused || !aux || timing) return false;
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &st) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    return st == hipStreamCaptureStatusNone;
  }

  void analyze(const float *x, float *coeffs) {
    forward_spectrum(x);
    if (!fused) {
      // Batches of nwork subbands in visiting order (position p holds subband
      // p + 1, the last position the lowpass, subband 0, which sits N nch
      // before subband 1: the rows kernels' wrap offset).
      const float scale = 1.0f / (float)N;
      for (int ch = 0; ch < nch; ++ch)
        for (int p0 = 0; p0 < T.S; p0 += nwork) {
          const int nb = std::min(nwork, T.S - p0);
          axes23<1, 1>(X[ch].p, work.p, 0, nb, 0, H, dvisit.p + p0);
          c2r(work.p, coeffs + (int64_t)(p0 + 1) * N * nch + ch, nch, scale, nb, H, N * nch, T.S - 1 - p0,
              -(int64_t)(p0 + 1) * N * nch);
        }
      return;
    }
    const FusedCtx c = ctx();
    if (concurrent()) {
      FusedCtx ca = c;
      ca.stream = aux;
      for (int ch = 0; ch < nch; ++ch) {
        fork_aux();
        fops.p_ana[AX2](ca, X[ch].p, U[ch].p, dunits.p + ax_ubeg[AX2], ax_uend[AX2] - ax_ubeg[AX2], AX2);
        fops.q_ana[AX2](ca, AX2, U[ch].p, djobs.p + ja_ax[AX2], jna_ax[AX2], coeffs + ch, N * nch, nch,
                        1.0f / (float)N);
        fops.p_ana[AX3](c, X[ch].p, U[ch].p, dunits.p + ax_ubeg[AX3], ax_uend[AX3] - ax_ubeg[AX3], AX3);
        fops.q_ana[AX3](c, AX3, U[ch].p, djobs.p + ja_ax[AX3], jna_ax[AX3], coeffs + ch, N * nch, nch,
                        1.0f / (float)N);
        join_aux();
      }
      cached_group = -1;
      return;
    }
    for (int ch = 0; ch < nch; ++ch) {
      for (int o = 0; o < 2; ++o)
        timed(0, [&] {
          fops.p_ana[o](c, X[ch].p, U[ch].p, dunits.p + ax_ubeg[o], ax_uend[o] - ax_ubeg[o], o);
        });
      for (int o = 0; o < 2; ++o)
        timed(1, [&] {
          fops.q_ana[o](c, o, U[ch].p, djobs.p + ja_ax[o], jna_ax[o], coeffs + ch, N * nch, nch, 1.0f / (float)N);
        });
    }
    cached_group = -1;
  }
  // Working memory of analyze() carved from one region: pieces of `need`
  // bytes each, 256-byte aligned.
  static size_t aligned(size_t b) { return (b + 255) & ~(size_t)255; }
  size_t working_bytes(int work_spectra) const {
    const size_t spec = aligned(sizeof(float2) * (size_t)H);
    if (fused) return spec * (1 + nch) + aligned(sizeof(float2) * (size_t)nslabs * H);
    return spec * nch + aligned(sizeof(float2) * (size_t)H * work_spectra);
  }
  // analyze() with TV, X, U (factored) or X and work (general) pointing into
  // [ws, ws + bytes); the plan's own buffers are put back afterwards.  The
  // general path's batch shrinks to the work spectra that fit, which changes
  // no coefficient (each subband's arithmetic is independent of its batch).
  // Channels share the slabs: each channel's analysis joins both
  // orientations before the next starts.
  void analyze_in(char *ws, size_t bytes, const float *x, float *coeffs) {
    struct Saved {
      surfhip_ns_plan *p;
      float2 *tv, *x0, *x1, *u0, *u1, *work;
      int nwork;
      ~Saved() {
        p->TV.p = tv;
        p->X[0].p = x0;
        p->X[1].p = x1;
        p->U[0].p = u0;
        p->U[1].p = u1;
        p->work.p = work;
        p->nwork = nwork;
      }
    } saved{this, TV.p, X[0].p, X[1].p, U[0].p, U[1].p, work.p, nwork};
    const size_t spec = aligned(sizeof(float2) * (size_t)H);
    char *at = ws;
    auto take = [&](size_t b) {
      char *q = at;
      at += aligned(b);
      if ((size_t)(at - ws) > bytes) fail(SURFHIP_ERR_ARG, "working region too small");
      return reinterpret_cast<float2 *>(q);
    };
    if (fused) {
      TV.p = take(spec);
      for (int ch = 0; ch < nch; ++ch) X[ch].p = take(spec);
      U[0].p = U[1].p = take(sizeof(float2) * (size_t)nslabs * H);
    } else {
      for (int ch = 0; ch < nch; ++ch) X[ch].p = take(spec);
      const size_t left = bytes - (size_t)(at - ws);
      const int nb = (int)std::min<size_t>((size_t)nwork, left / (sizeof(float2) * (size_t)H));
      if (nb < 1) fail(SURFHIP_ERR_ARG, "working region too small");
      work.p = take(sizeof(float2) * (size_t)H * nb);
      nwork = nb;
    }
    analyze(x, coeffs);
  }

  // Streaming analysis: subband i (0-based) into band.
  void stream_subband(int i, float *band) {
    if (!fused) {
      subbands_generic(i, 1, band);
      return;
    }
    const FusedCtx c = ctx();
    const int g = group_of(i);
    const int o = groups[g].orient;
    if (cached_group != g) {
      for (int ch = 0; ch < nch; ++ch)
        timed(0, [&] { fops.p_ana[o](c, X[ch].p, U[ch].p, dunits.p + ubeg[g], uend[g] - ubeg[g], o); });
      cached_group = g;
    }
    for (int ch = 0; ch < nch; ++ch)
      timed(1, [&] { fops.q_ana[o](c, o, U[ch].p, djobs.p + j_fresh + i, 1, band + ch, 0, nch, 1.0f / (float)N); });
  }
  // Close the synthesis sum of channel ch from its slabs and write the volume.
  void finish_channel(int ch, float *x) {
    const FusedCtx c = ctx();
    const bool has2 = sl_n[AX2] > 0;
    if (has2)
      timed(3, [&] {
        fops.p_syn[AX2](c, AX2, R[ch].p, dslist.p + sl_beg[AX2], sl_n[AX2], accf.p, 0, nullptr);
      });
    timed(3, [&] {
      fops.p_syn[AX3](c, AX3, R[ch].p, dslist.p + sl_beg[AX3], sl_n[AX3], accf.p, has2 ? 1 : 0, TV.p);
    });
    timed(1, [&] { fops.q_ana[AX3](c, AX3, TV.p, djobs.p + j_full, 1, x + ch, 0, nch, 1.0f / (float)N); });
  }
  void synthesize(const float *coeffs, float *x) {
    if (!fused) {
      // Batches in visiting order, summed in that order: the order of
      // streaming accumulation over subbands 1 .. S-1 and then the lowpass.
      // The first batch's sum starts from zero, as after inverse_begin.
      for (int ch = 0; ch < nch; ++ch)
        for (int p0 = 0; p0 < T.S; p0 += nwork) {
          const int nb = std::min(nwork, T.S - p0);
          r2c(coeffs + (int64_t)(p0 + 1) * N * nch + ch, nch, work.p, nb, N * nch, H, T.S - 1 - p0,
              -(int64_t)(p0 + 1) * N * nch);
          axes23<-1, 2>(work.p, work.p, 0, nb, H, H, dvisit.p + p0);
          sum_into(acc[ch].p, work.p, nb, p0 == 0);
        }
      finish(x);
      return;
    }
    const FusedCtx c = ctx();
    if (concurrent()) {
      FusedCtx ca = c;
      ca.stream = aux;
      const bool has2 = sl_n[AX2] > 0;
      for (int ch = 0; ch < nch; ++ch) {
        fork_aux();
        fops.q_syn[AX2](ca, AX2, coeffs + ch, N * nch, nch, R[ch].p, djobs.p + js_ax[AX2], jns_ax[AX2]);
        if (has2) fops.p_syn[AX2](ca, AX2, R[ch].p, dslist.p + sl_beg[AX2], sl_n[AX2], accf.p, 0, nullptr);
        fops.q_syn[AX3](c, AX3, coeffs + ch, N * nch, nch, R[ch].p, djobs.p + js_ax[AX3], jns_ax[AX3]);
        join_aux();
        fops.p_syn[AX3](c, AX3, R[ch].p, dslist.p + sl_beg[AX3], sl_n[AX3], accf.p, has2 ? 1 : 0, TV.p);
        fops.q_ana[AX3](c, AX3, TV.p, djobs.p + j_full, 1, x + ch, 0, nch, 1.0f / (float)N);
      }
      return;
    }
    for (int ch = 0; ch < nch; ++ch) {
      for (int o = 0; o < 2; ++o)
        timed(2, [&] { fops.q_syn[o](c, o, coeffs + ch, N * nch, nch, R[ch].p, djobs.p + js_ax[o], jns_ax[o]); });
      finish_channel(ch, x);
    }
  }
  void inverse_begin() {
    if (fused) {
      std::fill(touched.begin(), touched.end(), 0);
      return;
    }
    for (int ch = 0; ch < nch; ++ch) zero(acc[ch].p, H);
  }
  // acc += rfft(band) . M_i
  void accumulate(int i, const float *band) {
    if (fused) {
      const FusedCtx c = ctx();
      const int g = group_of(i);
      const int o = groups[g].orient;
      const int u = groups[g].slab0 + (i - groups[g].first) % groups[g].nB;
      const int jb = (touched[u] ? j_cont : j_fresh) + i;
      touched[u] = 1;
      for (int ch = 0; ch < nch; ++ch)
        timed(2, [&] { fops.q_syn[o](c, o, band + ch, 0, nch, R[ch].p, djobs.p + jb, 1); });
      return;
    }
    accumulate_generic(i, 1, band);
  }
  void finish(float *x) {
    if (fused) {
      // Slabs that received no subband hold no partial sum: clear them.
      for (int u = 0; u < nslabs; ++u)
        if (!touched[u])
          for (int ch = 0; ch < nch; ++ch) timed(4, [&] { zero(R[ch].p + (int64_t)u * H, H); });
      for (int ch = 0; ch < nch; ++ch) finish_channel(ch, x);
      return;
    }
    for (int ch = 0; ch < nch; ++ch) {
      axes23<1, 0>(acc[ch].p, work.p);
      c2r(work.p, x + ch, nch, 1.0f / (float)N);
    }
  }
};

namespace {

void check_lds(size_t bytes, const char *what) {
  // 160 KiB of LDS per CU on gfx950; keep each block within 64 KiB so two or
  // more blocks fit per CU.
  if (bytes > 65536)
    fail(SURFHIP_ERR_ARG, fmt("%s needs %zu bytes of LDS per block; the axis is too long for this path", what, bytes));
}

int pick_group(int n) {
  // G pencils of length n per block, G a multiple of 2, G*n about 2048.
  int g = 2048 / std::max(n, 1);
  g = std::max(2, std::min(32, g));
  return g & ~1;
}

}  // namespace

// Run `body` with the mask view swapped for an adjoint (restored after).
template <class F> void with_masks(surfhip_ns_plan *plan, int syn_uses_ana, int ana_uses_syn, F &&body) {
  const NSMaskView keep = plan->dview;
  plan->dview.syn_uses_ana = syn_uses_ana;
  plan->dview.ana_uses_syn = ana_uses_syn;
  try {
    body();
  } catch (...) {
    plan->dview = keep;
    throw;
  }
  plan->dview = keep;
}

extern "C" {

int surfhip_ns_nsubbands(const surfhip_ns_params *p, int32_t *out) {
  return guarded([&] {
    if (!p || !out) fail(SURFHIP_ERR_ARG, "null argument");
    *out = ns_check_params(*p);
  });
}

}  // extern "C"

namespace {

void create_plan(surfhip_ns_plan **out, const surfhip_ns_params *p, int32_t n1, int32_t n2, int32_t n3,
                 int32_t dtype, void *stream, bool pyramid) {
    if (!out || !p) fail(SURFHIP_ERR_ARG, "null argument");
    *out = nullptr;
    if (dtype != SURFHIP_F32 && dtype != SURFHIP_C64) fail(SURFHIP_ERR_ARG, fmt("unknown dtype %d", dtype));
    auto P = std::make_unique<surfhip_ns_plan>();
    ns_build_tables(*p, n1, n2, n3, P->T, pyramid, /*device_groups=*/true);
    P->dtype = dtype;
    P->nch = dtype == SURFHIP_C64 ? 2 : 1;
    P->stream = (hipStream_t)stream;
    P->n1 = n1;
    P->n2 = n2;
    P->n3 = n3;
    P->h1 = n1 / 2 + 1;
    P->N = (int64_t)n1 * n2 * n3;
    P->H = (int64_t)P->h1 * n2 * n3;
    std::vector<float2> pool;
    P->f1 = make_fft_plan(n1, pool);
    P->f2 = make_fft_plan(n2, pool);
    P->f3 = make_fft_plan(n3, pool);
    P->fused = std::getenv("SURFHIP_NS_GENERIC") == nullptr && fused_ops_for(n1, n2, n3, P->fops);
    if (!P->fused) {
      // Rows per axis-1 block: the most (up to 30) whose LDS fits 64 KiB, with
      // an odd number of row pairs, so that the 16-byte row loads' and
      // stores' transposing LDS accesses spread over 8 banks instead of one.
      for (int g : {30, 14, 6, 2}) {
        P->g_rows = g;
        if (std::max(P->lds_rows_r2c(), P->lds_rows_c2r()) <= 65536) break;
      }
      P->g_axis = std::min(pick_group(std::max(n2, n3)), 16);
      check_lds(P->lds_rows_r2c(), "the axis-1 transform");
      check_lds(P->lds_rows_c2r(), "the axis-1 transform");
      check_lds(P->lds_axis(n2), "the axis-2 transform");
      check_lds(P->lds_axis(n3), "the axis-3 transform");
      // Axes 2 and 3 by the plane kernel where it fits (SURFHIP_NS_NO_PLANE
      // selects the per-axis passes, for comparison).
      P->n3p = n3 | 1;
      P->plane = std::getenv("SURFHIP_NS_NO_PLANE") == nullptr && P->plane_fits();
      if (!P->plane && std::getenv("SURFHIP_NS_NO_STRIP") == nullptr)
        for (int a = 2; a <= 3; ++a) {
          P->strip[a - 2] = P->strip_fits(a);
          if (!P->strip[a - 2]) continue;
          const int lds = (int)P->lds_strip(a);
          const void *k2[] = {(const void *)&k_strip<2, -1, 0, true>,  (const void *)&k_strip<2, 1, 0, true>,
                              (const void *)&k_strip<2, 1, 1, true>,   (const void *)&k_strip<2, -1, 2, true>,
                              (const void *)&k_strip<2, -1, 0, false>, (const void *)&k_strip<2, 1, 0, false>,
                              (const void *)&k_strip<2, 1, 1, false>,  (const void *)&k_strip<2, -1, 2, false>};
          const void *k3[] = {(const void *)&k_strip<3, -1, 0, true>, (const void *)&k_strip<3, 1, 0, true>,
                              (const void *)&k_strip<3, -1, 0, false>, (const void *)&k_strip<3, 1, 0, false>};
          if (a == 2)
            for (const void *k : k2) SH_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
          else
            for (const void *k : k3) SH_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
        }
      if (P->plane) {
        const int lds = (int)P->lds_plane();
        for (const void *k :
             {(const void *)&k_plane<-1, 0, false>, (const void *)&k_plane<1, 0, false>,
              (const void *)&k_plane<1, 1, false>, (const void *)&k_plane<-1, 2, false>,
              (const void *)&k_plane<-1, 0, true>, (const void *)&k_plane<1, 0, true>,
              (const void *)&k_plane<1, 1, true>, (const void *)&k_plane<-1, 2, true>})
          SH_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
      }
    }
    P->tw.upload(pool);
    {
      // Group-factor tables built on the device in FP64 (rounded once).
      const size_t V = (size_t)P->h1 * n2 * n3;
      P->gana.alloc((size_t)P->T.ngroups * V);
      if (!P->T.tight) P->gsyn.alloc((size_t)P->T.ngroups * V);
      DevBuf<double> t2, b3, hg2, hg2c, acc, lp;
      t2.upload(P->T.g_t2);
      b3.upload(P->T.g_b3);
      hg2.upload(P->T.g_hg2);
      hg2c.upload(P->T.g_hg2c);
      acc.upload(P->T.g_acc);
      lp.upload(P->T.g_lp);
      NSGroupInputs in = P->T.group_inputs();
      in.t2 = t2.p;
      in.b3 = b3.p;
      in.hg2 = hg2.p;
      in.hg2c = hg2c.p;
      in.acc = acc.p;
      in.lp = lp.p;
      const int64_t blocks = std::min<int64_t>(((int64_t)V + NT - 1) / NT, 65536);
      hipLaunchKernelGGL(k_group_tables, dim3((unsigned)blocks), dim3(NT), 0, P->stream, in, P->gana.p,
                         P->T.tight ? nullptr : P->gsyn.p, (int64_t)V);
      SH_LAUNCHED();
      SH_HIP(hipStreamSynchronize(P->stream));
    }
    if (P->T.planes.empty()) P->T.planes.push_back(0.0f);
    P->planes.upload(P->T.planes);
    P->ditems.upload(P->T.items);
    {
      std::vector<NSItem> visit(P->T.items.begin() + 1, P->T.items.end());
      visit.push_back(P->T.items[0]);
      P->dvisit.upload(visit);
    }
    P->dview = NSMaskView{P->gana.p, P->T.tight ? P->gana.p : P->gsyn.p, P->planes.p, P->h1, n2, n3,
                          P->T.ngroups, P->T.tight, 0, 0};
    for (int ch = 0; ch < P->nch; ++ch) P->X[ch].alloc(P->H);
    if (P->fused) {
      P->build_fused();
      P->make_aux();
    } else {
      for (int ch = 0; ch < P->nch; ++ch) P->acc[ch].alloc(P->H);
      // Batched general-path launches: up to nwork subbands per launch, their
      // work spectra within 1 GiB (at least 8 of them).  Large batches make
      // launches of many blocks, so the last wave of blocks costs little.
      const int64_t fit = ((int64_t)1 << 30) / ((int64_t)sizeof(float2) * P->H);
      P->nwork = (int)std::max<int64_t>(1, std::min<int64_t>(P->T.S, std::max<int64_t>(8, fit)));
      // SURFHIP_NS_NWORK caps the batch (at least 1): small volumes then take
      // the several-batch schedule that large ones (256^3: 15 per batch) use.
      if (const char *cap = std::getenv("SURFHIP_NS_NWORK"))
        P->nwork = std::max(1, std::min(P->nwork, std::atoi(cap)));
      P->work.alloc((size_t)P->H * P->nwork);
    }
    // The group-table inputs are no longer needed on the host.
    std::vector<double>().swap(P->T.g_t2);
    std::vector<double>().swap(P->T.g_b3);
    SH_HIP(hipDeviceSynchronize());
    *out = P.release();
}

// One directional bank (src/directional/ndfb.jl:91-127) as the NS parameters
// of a single scale without the pyramid.
surfhip_ns_params ndfb_as_ns(const surfhip_ndfb_params *p) {
  if (p->levels < 0) fail(SURFHIP_ERR_ARG, "levels must be ≥ 0");
  if (!(p->transition >= 0)) fail(SURFHIP_ERR_ARG, "transition must be ≥ 0");
  surfhip_ns_params q{};
  q.scales = 1;
  q.direction_levels[0] = p->levels;
  q.bernstein_order = p->bernstein_order;
  q.halfband_order = p->halfband_order;
  q.transition = p->transition;
  q.tight = p->tight;
  return q;
}

void mask_to_host(const NSTables &T, int32_t i, float *out) {
  if (i < 1 || i > T.S) fail(SURFHIP_ERR_ARG, fmt("subband index %d outside 1:%d", i, T.S));
  const NSMaskView v = T.view();
  const NSItem &it = T.items[i - 1];
  for (int k3 = 0; k3 < T.n3; ++k3)
    for (int k2 = 0; k2 < T.n2; ++k2)
      for (int k1 = 0; k1 < T.h1; ++k1)
        out[(size_t)k1 + (size_t)T.h1 * ((size_t)k2 + (size_t)T.n2 * k3)] = ns_mask(v, it, k1, k2, k3, true);
}

// Extents the plan reads or writes through caller pointers: one volume, or
// the S subbands (complex data: two floats per element).
void check_volume_buffer(const surfhip_ns_plan *plan, const void *p, const char *what) {
  check_device_buffer(p, (size_t)plan->N * plan->nch * sizeof(float), what);
}
void check_coefficient_buffer(const surfhip_ns_plan *plan, const void *p) {
  check_device_buffer(p, (size_t)plan->N * plan->T.S * plan->nch * sizeof(float), "coefficient array");
}

}  // namespace

namespace surfhip {

size_t ns_analyze_working_bytes(const surfhip_ns_plan *plan) { return plan->working_bytes(1); }

size_t ns_volume_bytes(const surfhip_ns_plan *plan) { return (size_t)plan->N * plan->nch * sizeof(float); }
size_t ns_coefficient_bytes(const surfhip_ns_plan *plan) { return ns_volume_bytes(plan) * plan->T.S; }
void ns_analyze_raw(surfhip_ns_plan *plan, const void *x, void *coeffs) {
  plan->analyze((const float *)x, (float *)coeffs);
}
void ns_synthesize_raw(surfhip_ns_plan *plan, const void *coeffs, void *x) {
  plan->synthesize((const float *)coeffs, (float *)x);
}

int ns_plan_device(const surfhip_ns_plan *plan) {
  int dev = 0;
  if (plan->stream) {
    SH_HIP(hipStreamGetDevice(plan->stream, &dev));
  } else {
    SH_HIP(hipGetDevice(&dev));
  }
  return dev;
}

void ns_analyze_in(surfhip_ns_plan *plan, void *ws, size_t bytes, const void *x, void *coeffs) {
  if (!plan || !x || !coeffs || !ws) fail(SURFHIP_ERR_ARG, "null argument");
  check_volume_buffer(plan, x, "volume");
  check_coefficient_buffer(plan, coeffs);
  plan->analyze_in(static_cast<char *>(ws), bytes, (const float *)x, (float *)coeffs);
}

}  // namespace surfhip

extern "C" {

int surfhip_ns_plan_create(surfhip_ns_plan **out, const surfhip_ns_params *p, int32_t n1, int32_t n2, int32_t n3,
                           int32_t dtype, void *stream) {
  return guarded([&] { create_plan(out, p, n1, n2, n3, dtype, stream, true); });
}

int surfhip_ndfb_plan_create(surfhip_ns_plan **out, const surfhip_ndfb_params *p, int32_t n1, int32_t n2,
                             int32_t n3, int32_t dtype, void *stream) {
  return guarded([&] {
    if (!p) fail(SURFHIP_ERR_ARG, "null argument");
    const surfhip_ns_params q = ndfb_as_ns(p);
    create_plan(out, &q, n1, n2, n3, dtype, stream, false);
  });
}

int surfhip_ndfb_mask_host(const surfhip_ndfb_params *p, int32_t n1, int32_t n2, int32_t n3, int32_t idx,
                           float *out) {
  return guarded([&] {
    if (!p || !out) fail(SURFHIP_ERR_ARG, "null argument");
    const surfhip_ns_params q = ndfb_as_ns(p);
    NSTables T;
    ns_build_tables(q, n1, n2, n3, T, false);
    if (idx < 1 || idx > T.S - 1) fail(SURFHIP_ERR_ARG, fmt("subband index %d outside 1:%d", idx, T.S - 1));
    mask_to_host(T, idx + 1, out);
  });
}

int surfhip_ns_group_table(surfhip_ns_plan *plan, int32_t g, int32_t synthesis, float *out) {
  return guarded([&] {
    if (!plan || !out) fail(SURFHIP_ERR_ARG, "null argument");
    if (g < 0 || g >= plan->T.ngroups) fail(SURFHIP_ERR_ARG, fmt("group %d outside 0:%d", g, plan->T.ngroups - 1));
    const float *src = (synthesis && !plan->T.tight ? plan->gsyn.p : plan->gana.p) + (size_t)g * plan->H;
    SH_HIP(hipMemcpy(out, src, sizeof(float) * (size_t)plan->H, hipMemcpyDeviceToHost));
  });
}

int surfhip_ns_group_table_host(const surfhip_ns_params *p, int32_t n1, int32_t n2, int32_t n3, int32_t g,
                                int32_t synthesis, float *out) {
  return guarded([&] {
    if (!p || !out) fail(SURFHIP_ERR_ARG, "null argument");
    NSTables T;
    ns_build_tables(*p, n1, n2, n3, T);
    if (g < 0 || g >= T.ngroups) fail(SURFHIP_ERR_ARG, fmt("group %d outside 0:%d", g, T.ngroups - 1));
    const std::vector<float> &src = synthesis && !T.tight ? T.gsyn : T.gana;
    const size_t V = (size_t)T.h1 * n2 * n3;
    std::copy(src.begin() + (ptrdiff_t)(g * V), src.begin() + (ptrdiff_t)((g + 1) * V), out);
  });
}

int surfhip_ns_plan_destroy(surfhip_ns_plan *plan) {
  return guarded([&] {
    if (!plan) return;
    (void)hipStreamSynchronize(plan->stream);
    delete plan;
  });
}

int surfhip_ns_plan_tune(surfhip_ns_plan *plan, int32_t *config) {
  return guarded([&] {
    if (!plan) fail(SURFHIP_ERR_ARG, "null plan");
    plan->tune();
    if (config) {
      config[0] = plan->fused ? 1 : 0;
      config[1] = plan->g_rows;
      config[2] = plan->g_axis;
      const FftPlan1D *pl[3] = {&plan->f1, &plan->f2, &plan->f3};
      for (int a = 0; a < 3; ++a) {
        int32_t code = 0;
        for (int k = 0; k < pl[a]->npass; ++k) code = code * 100 + pl[a]->radix[k];
        config[3 + a] = code;
      }
    }
  });
}

int surfhip_ns_set_stream(surfhip_ns_plan *plan, void *stream) {
  return guarded([&] {
    if (!plan) fail(SURFHIP_ERR_ARG, "null plan");
    plan->stream = (hipStream_t)stream;
  });
}

int surfhip_ns_analyze(surfhip_ns_plan *plan, const void *x, void *coeffs) {
  return guarded([&] {
    if (!plan || !x || !coeffs) fail(SURFHIP_ERR_ARG, "null argument");
    check_volume_buffer(plan, x, "volume");
    check_coefficient_buffer(plan, coeffs);
    plan->analyze((const float *)x, (float *)coeffs);
  });
}

int surfhip_ns_synthesize(surfhip_ns_plan *plan, const void *coeffs, void *x) {
  return guarded([&] {
    if (!plan || !x || !coeffs) fail(SURFHIP_ERR_ARG, "null argument");
    check_volume_buffer(plan, x, "volume");
    check_coefficient_buffer(plan, coeffs);
    plan->synthesize((const float *)coeffs, (float *)x);
  });
}

int surfhip_ns_adjoint(surfhip_ns_plan *plan, const void *coeffs, void *x) {
  return guarded([&] {
    if (!plan || !x || !coeffs) fail(SURFHIP_ERR_ARG, "null argument");
    check_volume_buffer(plan, x, "volume");
    check_coefficient_buffer(plan, coeffs);
    with_masks(plan, 1, 0, [&] { plan->synthesize((const float *)coeffs, (float *)x); });
  });
}

int surfhip_ns_synthesize_adjoint(surfhip_ns_plan *plan, const void *x, void *coeffs) {
  return guarded([&] {
    if (!plan || !x || !coeffs) fail(SURFHIP_ERR_ARG, "null argument");
    check_volume_buffer(plan, x, "volume");
    check_coefficient_buffer(plan, coeffs);
    with_masks(plan, 0, 1, [&] { plan->analyze((const float *)x, (float *)coeffs); });
  });
}

int surfhip_ns_stream_begin(surfhip_ns_plan *plan, const void *x) {
  return guarded([&] {
    if (!plan || !x) fail(SURFHIP_ERR_ARG, "null argument");
    check_volume_buffer(plan, x, "volume");
    plan->forward_spectrum((const float *)x);
    plan->streaming_fwd = true;
  });
}

int surfhip_ns_stream_subband(surfhip_ns_plan *plan, int32_t i, void *band) {
  return guarded([&] {
    if (!plan || !band) fail(SURFHIP_ERR_ARG, "null argument");
    if (!plan->streaming_fwd) fail(SURFHIP_ERR_ARG, "surfhip_ns_stream_begin was not called");
    if (i < 1 || i > plan->T.S) fail(SURFHIP_ERR_ARG, fmt("subband index %d outside 1:%d", i, plan->T.S));
    check_volume_buffer(plan, band, "subband buffer");
    plan->stream_subband(i - 1, (float *)band);
  });
}

int surfhip_ns_stream_inverse_begin(surfhip_ns_plan *plan) {
  return guarded([&] {
    if (!plan) fail(SURFHIP_ERR_ARG, "null plan");
    plan->inverse_begin();
  });
}

int surfhip_ns_stream_accumulate(surfhip_ns_plan *plan, int32_t i, const void *band) {
  return guarded([&] {
    if (!plan || !band) fail(SURFHIP_ERR_ARG, "null argument");
    if (i < 1 || i > plan->T.S) fail(SURFHIP_ERR_ARG, fmt("subband index %d outside 1:%d", i, plan->T.S));
    check_volume_buffer(plan, band, "subband buffer");
    plan->accumulate(i - 1, (const float *)band);
  });
}

int surfhip_ns_stream_finish(surfhip_ns_plan *plan, void *x) {
  return guarded([&] {
    if (!plan || !x) fail(SURFHIP_ERR_ARG, "null argument");
    check_volume_buffer(plan, x, "volume");
    plan->finish((float *)x);
  });
}

int surfhip_ns_set_timing(surfhip_ns_plan *plan, int32_t enable) {
  return guarded([&] {
    if (!plan) fail(SURFHIP_ERR_ARG, "null plan");
    plan->collect_timing();
    for (int c = 0; c < SURFHIP_NS_TIMING_CLASSES; ++c) {
      plan->tcount[c] = 0;
      plan->tms[c] = 0.0;
    }
    plan->timing = enable != 0;
  });
}

int surfhip_ns_timing(surfhip_ns_plan *plan, int64_t *launches, double *ms, char *names, size_t n) {
  return guarded([&] {
    if (!plan) fail(SURFHIP_ERR_ARG, "null plan");
    plan->collect_timing();
    for (int c = 0; c < SURFHIP_NS_TIMING_CLASSES; ++c) {
      if (launches) launches[c] = plan->tcount[c];
      if (ms) ms[c] = plan->tms[c];
    }
    const char *s = plan->fused ? "pencil_analysis,plane_inverse,plane_forward,pencil_synthesis,other"
                               : "rows_r2c,rows_c2r,axes23_forward,axes23_inverse,other";
    if (names && n) {
      size_t m = std::min(n - 1, std::strlen(s));
      std::memcpy(names, s, m);
      names[m] = '\0';
    }
  });
}

int surfhip_ns_mask_host(const surfhip_ns_params *p, int32_t n1, int32_t n2, int32_t n3, int32_t i, float *out) {
  return guarded([&] {
    if (!p || !out) fail(SURFHIP_ERR_ARG, "null argument");
    NSTables T;
    ns_build_tables(*p, n1, n2, n3, T);
    mask_to_host(T, i, out);
  });
}

}  // extern "C"
