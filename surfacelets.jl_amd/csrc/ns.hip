// NSSurfacelet analyze/synthesize on gfx950.
//
// Every subband is irfft(X . M_k) with X = rfft(x) and M_k the real mask of
// subband k (src/transforms/ns_surfacelet.jl:8-32); synthesis is
// irfft(sum_k rfft(c_k) . M_k).  The 3-D real FFTs run as three LDS passes:
// a paired-row r2c/c2r along axis 1 (the contiguous axis) and complex passes
// along axes 2 and 3, each block holding G pencils that are contiguous along
// k1 so that every global access is coalesced.  The mask is evaluated on the
// fly from FP32 tables (built in FP64 on the host) inside the first inverse
// pass of analysis and the last forward pass of synthesis, so no masked
// spectrum is ever stored.
//
// Spectra use the rfft grid in column-major order: S[k1 + h1 (k2 + n2 k3)],
// h1 = n1/2 + 1.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "fft.hpp"
#include "hip_util.hpp"
#include "ns_fast.hpp"
#include "ns_tables.hpp"

namespace surfhip {

constexpr int NT = 256;

// ------------------------------------------------------------------ kernels

// Real rows of length n1 (row r at x[(r n1 + t) xs]) -> half-spectrum rows
// out[r h1 + k].  Two rows share one complex FFT: z = a + i b, then
// A[k] = (Z[k] + conj Z[-k]) / 2, B[k] = (Z[k] - conj Z[-k]) / 2i.
__global__ __launch_bounds__(NT) void k_rows_r2c(const float *__restrict__ x, int64_t xs,
                                                 float2 *__restrict__ out, int n1, int h1, int nrows,
                                                 int G, FftPlan1D pl, const float2 *__restrict__ twg) {
  extern __shared__ float2 smem[];
  const int P = G / 2;
  float2 *tw = smem;
  float2 *a = tw + n1;
  float2 *b = a + (size_t)n1 * P;
  const int tid = threadIdx.x;
  for (int k = tid; k < n1; k += NT) tw[k] = twg[pl.tw + k];
  const int r0 = blockIdx.x * G;
  for (int idx = tid; idx < G * n1; idx += NT) {
    const int g = idx / n1, t = idx - g * n1;
    const int r = r0 + g;
    float v = r < nrows ? x[((int64_t)r * n1 + t) * xs] : 0.0f;
    float *ap = reinterpret_cast<float *>(&a[t * P + (g >> 1)]);
    ap[g & 1] = v;
  }
  __syncthreads();
  float2 *res = lds_fft<-1>(a, b, pl, P, tw, tid, NT);
  for (int idx = tid; idx < G * h1; idx += NT) {
    const int g = idx / h1, k = idx - g * h1;
    const int r = r0 + g;
    if (r >= nrows) continue;
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
    out[(int64_t)r * h1 + k] = v;
  }
}

// Half-spectrum rows in[r h1 + k] -> real rows out[(r n1 + t) os] * scale.
// Pairs of rows form z = a + i b through their Hermitian extensions; the
// imaginary parts of the DC and (even n1) Nyquist bins are ignored, as a c2r
// transform does.
__global__ __launch_bounds__(NT) void k_rows_c2r(const float2 *__restrict__ in, float *__restrict__ out,
                                                 int64_t os, float scale, int n1, int h1, int nrows, int G,
                                                 FftPlan1D pl, const float2 *__restrict__ twg) {
  extern __shared__ float2 smem[];
  const int P = G / 2;
  float2 *tw = smem;
  float2 *s = tw + n1;
  float2 *a = s + (size_t)G * h1;
  float2 *b = a + (size_t)n1 * P;
  const int tid = threadIdx.x;
  for (int k = tid; k < n1; k += NT) tw[k] = twg[pl.tw + k];
  const int r0 = blockIdx.x * G;
  for (int idx = tid; idx < G * h1; idx += NT) {
    const int g = idx / h1, k = idx - g * h1;
    const int r = r0 + g;
    s[idx] = r < nrows ? in[(int64_t)r * h1 + k] : make_float2(0.0f, 0.0f);
  }
  __syncthreads();
  const bool even = (n1 & 1) == 0;
  for (int idx = tid; idx < P * n1; idx += NT) {
    const int t = idx / P, p = idx - t * P;
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
  float2 *res = lds_fft<1>(a, b, pl, P, tw, tid, NT);
  for (int idx = tid; idx < G * n1; idx += NT) {
    const int g = idx / n1, t = idx - g * n1;
    const int r = r0 + g;
    if (r >= nrows) continue;
    const float2 v = res[t * P + (g >> 1)];
    out[((int64_t)r * n1 + t) * os] = ((g & 1) ? v.y : v.x) * scale;
  }
}

// Complex FFT along axis AXIS (2 or 3) of an rfft-grid array, G pencils per
// block that are consecutive in k1.  MODE 0: out = FFT(in); MODE 1:
// out = FFT(in . mask) (analysis); MODE 2: out += mask . FFT(in) (synthesis).
template <int AXIS, int DIR, int MODE>
__global__ __launch_bounds__(NT) void k_axis(const float2 *in, float2 *out, int h1,
                                             int n2, int n3, int G, FftPlan1D pl,
                                             const float2 *__restrict__ twg, NSMaskView mv, NSItem it) {
  extern __shared__ float2 smem[];
  const int n = AXIS == 2 ? n2 : n3;
  float2 *tw = smem;
  float2 *a = tw + n;
  float2 *b = a + (size_t)n * G;
  const int tid = threadIdx.x;
  for (int k = tid; k < n; k += NT) tw[k] = twg[pl.tw + k];
  const int ublocks = (h1 + G - 1) / G;
  const int ub = blockIdx.x % ublocks;
  const int v = blockIdx.x / ublocks;
  const int u0 = ub * G;
  const int64_t sa = AXIS == 2 ? (int64_t)h1 : (int64_t)h1 * n2;
  const int64_t base = AXIS == 2 ? (int64_t)v * h1 * n2 : (int64_t)v * h1;
  for (int idx = tid; idx < n * G; idx += NT) {
    const int t = idx / G, g = idx - t * G;
    const int k1 = u0 + g;
    float2 val = make_float2(0.0f, 0.0f);
    if (k1 < h1) {
      val = in[base + k1 + t * sa];
      if (MODE == 1) {
        const int k2 = AXIS == 2 ? t : v, k3 = AXIS == 2 ? v : t;
        val = cscale(val, ns_mask(mv, it, k1, k2, k3, true));
      }
    }
    a[idx] = val;
  }
  __syncthreads();
  float2 *res = lds_fft<DIR>(a, b, pl, G, tw, tid, NT);
  for (int idx = tid; idx < n * G; idx += NT) {
    const int t = idx / G, g = idx - t * G;
    const int k1 = u0 + g;
    if (k1 >= h1) continue;
    const int64_t o = base + k1 + t * sa;
    float2 r = res[idx];
    if (MODE == 2) {
      const int k2 = AXIS == 2 ? t : v, k3 = AXIS == 2 ? v : t;
      const float m = ns_mask(mv, it, k1, k2, k3, false);
      float2 acc = out[o];
      out[o] = make_float2(acc.x + m * r.x, acc.y + m * r.y);
    } else {
      out[o] = r;
    }
  }
}

__global__ void k_zero(float2 *__restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = make_float2(0.0f, 0.0f);
}

}  // namespace surfhip

// ---------------------------------------------------------------------- plan

using namespace surfhip;

// A plan runs one of two kernel sets on the same tables: the fast fused path
// (ns_fast.hip) on power-of-two grids it covers, or the general three-pass
// path above on any other size.  Both evaluate the same mask expression.
struct surfhip_ns_plan {
  NSTables T;
  int dtype = SURFHIP_F32;
  int nch = 1;
  hipStream_t stream = nullptr;
  int64_t N = 0, H = 0;
  int n1 = 0, n2 = 0, n3 = 0, h1 = 0;
  FftPlan1D f1{}, f2{}, f3{};
  int g_rows = 16, g_axis = 16;
  DevBuf<float> hg1, hg1c, hg2, hg2c, a, l, c, planes;
  DevBuf<float2> tw, X[2], acc[2], work;
  DevBuf<NSItem> ditems;
  NSMaskView dview{};
  bool streaming_fwd = false;
  // fast path
  bool fast = false;
  FastOps ops;
  int maxw = 1;                  // largest group
  DevBuf<float2> VT, Vfin, accg[2];
  // optional per-class event timing
  bool timing = false;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> tev;
  int64_t tcount[SURFHIP_NS_TIMING_CLASSES] = {0};
  double tms[SURFHIP_NS_TIMING_CLASSES] = {0};
  int open_group = -2;           // streaming synthesis: group being accumulated (-2: none)
  int open_item = 0;             // an item of that group
  bool open_fresh = true;        // no term accumulated yet in the open group

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
  }

  FastCtx ctx() const {
    return FastCtx{stream, n1, n2, n3, h1, H, dview, ditems.p, tw.p + f1.tw, tw.p + f2.tw, tw.p + f3.tw};
  }
  int group_of(int i) const { return T.items[i].lowpass ? -1 : T.items[i].group; }
  int group_size(int i) const { return T.items[i].lowpass ? 1 : T.group_count[T.items[i].group]; }
  // Frequency support of each group (and the lowpass): see Support.
  std::vector<Support> gsup;
  Support lowsup{}, fullsup{};
  Support support_of(int i) const { return T.items[i].lowpass ? lowsup : gsup[T.items[i].group]; }
  Support level_support(int lev) const {
    const int L = h1 + n2 + n3;
    const float *a = T.a.data() + (size_t)lev * L;
    int K1 = 0, K2 = 0;
    for (int k = 0; k < h1; ++k)
      if (a[k] != 0.0f) K1 = k;
    for (int k = 0; k < n2; ++k)
      if (a[h1 + k] != 0.0f) K2 = std::max(K2, std::min(k, n2 - k));
    return Support{K1, K2};
  }
  void compute_supports() {
    fullsup = Support{h1 - 1, n2 / 2};
    gsup.clear();
    for (size_t g = 0; g < T.group_first.size(); ++g) gsup.push_back(level_support(T.items[T.group_first[g]].level));
    lowsup = level_support(T.J);
  }

  size_t lds_rows_r2c() const { return sizeof(float2) * ((size_t)n1 + 2 * (size_t)n1 * (g_rows / 2)); }
  size_t lds_rows_c2r() const {
    return sizeof(float2) * ((size_t)n1 + (size_t)g_rows * h1 + 2 * (size_t)n1 * (g_rows / 2));
  }
  size_t lds_axis(int n) const { return sizeof(float2) * ((size_t)n + 2 * (size_t)n * g_axis); }

  int rows() const { return n2 * n3; }

  void r2c(const float *x, int64_t xs, float2 *out) {
    int grid = (rows() + g_rows - 1) / g_rows;
    hipLaunchKernelGGL(k_rows_r2c, dim3(grid), dim3(NT), lds_rows_r2c(), stream, x, xs, out, n1, h1, rows(),
                       g_rows, f1, tw.p);
    SH_LAUNCHED();
  }
  void c2r(const float2 *in, float *out, int64_t os, float scale) {
    int grid = (rows() + g_rows - 1) / g_rows;
    hipLaunchKernelGGL(k_rows_c2r, dim3(grid), dim3(NT), lds_rows_c2r(), stream, in, out, os, scale, n1, h1,
                       rows(), g_rows, f1, tw.p);
    SH_LAUNCHED();
  }
  template <int AXIS, int DIR, int MODE> void axis(const float2 *in, float2 *out, const NSItem &it) {
    const int ublocks = (h1 + g_axis - 1) / g_axis;
    const int grid = ublocks * (AXIS == 2 ? n3 : n2);
    const FftPlan1D &pl = AXIS == 2 ? f2 : f3;
    hipLaunchKernelGGL((k_axis<AXIS, DIR, MODE>), dim3(grid), dim3(NT), lds_axis(pl.n), stream, in, out, h1, n2,
                       n3, g_axis, pl, tw.p, dview, it);
    SH_LAUNCHED();
  }
  void zero(float2 *p) {
    hipLaunchKernelGGL(k_zero, dim3(1024), dim3(NT), 0, stream, p, H);
    SH_LAUNCHED();
  }

  // X[ch] = rfft(channel ch of x)
  void forward_spectrum(const float *x) {
    static const NSItem none{};
    for (int ch = 0; ch < nch; ++ch) {
      if (fast) {
        const FastCtx c = ctx();
        timed(4, [&] { ops.b_fwd(c, x + ch, 0, nch, 1, VT.p, fullsup); });
        timed(4, [&] { ops.a_fwd(c, VT.p, X[ch].p); });
        continue;
      }
      r2c(x + ch, nch, X[ch].p);
      axis<2, -1, 0>(X[ch].p, X[ch].p, none);
      axis<3, -1, 0>(X[ch].p, X[ch].p, none);
    }
  }
  // Subbands item0 .. item0+nw-1 (one group) into out (subband stride N nch).
  void subbands(int item0, int nw, float *out) {
    static const NSItem none{};
    const float scale = 1.0f / (float)N;
    for (int ch = 0; ch < nch; ++ch) {
      if (fast) {
        const FastCtx c = ctx();
        const Support sp = support_of(item0);
        timed(0, [&] { ops.a_ana(c, X[ch].p, VT.p, item0, nw, group_size(item0), sp); });
        timed(1, [&] { ops.b_inv(c, VT.p, nw, out + ch, (size_t)N * nch, (size_t)nch, scale, sp); });
        continue;
      }
      for (int w = 0; w < nw; ++w) {
        axis<3, 1, 1>(X[ch].p, work.p, T.items[item0 + w]);
        axis<2, 1, 0>(work.p, work.p, none);
        c2r(work.p, out + (int64_t)w * N * nch + ch, nch, scale);
      }
    }
  }
  void analyze(const float *x, float *coeffs) {
    forward_spectrum(x);
    for (size_t g = 0; g < T.group_first.size(); ++g) {
      const int i0 = T.group_first[g], nw = T.group_count[g];
      subbands(i0, nw, coeffs + (int64_t)i0 * N * nch);
    }
    subbands(0, 1, coeffs);
  }
  void synthesize(const float *coeffs, float *x) {
    if (!fast) {
      inverse_begin();
      for (int i = 1; i < T.S; ++i) accumulate(i, coeffs + (int64_t)i * N * nch);
      accumulate(0, coeffs);
      finish(x);
      return;
    }
    const FastCtx c = ctx();
    for (int ch = 0; ch < nch; ++ch) timed(4, [&] { zero(acc[ch].p); });
    const int G = (int)T.group_first.size();
    for (int g = 0; g <= G; ++g) {
      const int i0 = g < G ? T.group_first[g] : 0, nw = g < G ? T.group_count[g] : 1;
      const bool last = g == G;
      for (int ch = 0; ch < nch; ++ch) {
        const Support sp = support_of(i0);
        timed(2, [&] {
          ops.b_fwd(c, coeffs + (int64_t)i0 * N * nch + ch, (size_t)N * nch, (size_t)nch, nw, VT.p, sp);
        });
        timed(3, [&] { ops.a_syn(c, VT.p, nw, acc[ch].p, nullptr, nullptr, Vfin.p, i0, 1, last ? 1 : 0, nw, sp); });
        if (last) timed(4, [&] { ops.b_inv(c, Vfin.p, 1, x + ch, 0, (size_t)nch, 1.0f / (float)N, fullsup); });
      }
    }
    open_group = -2;
  }
  void inverse_begin() {
    for (int ch = 0; ch < nch; ++ch) zero(acc[ch].p);
    open_group = -2;
    open_fresh = true;
  }
  void close_open() {
    if (open_group == -2) return;
    if (!open_fresh) {
      const FastCtx c = ctx();
      for (int ch = 0; ch < nch; ++ch) ops.a_close(c, acc[ch].p, accg[ch].p, open_item, support_of(open_item));
    }
    open_group = -2;
    open_fresh = true;
  }
  // acc += rfft(band) . M_i
  void accumulate(int i, const float *band) {
    static const NSItem none{};
    if (fast) {
      const int g = group_of(i);
      if (g != open_group) {
        close_open();
        open_group = g;
        open_item = i;
        open_fresh = true;
      }
      const FastCtx c = ctx();
      for (int ch = 0; ch < nch; ++ch) {
        const Support sp = support_of(i);
        ops.b_fwd(c, band + ch, 0, (size_t)nch, 1, VT.p, sp);
        ops.a_syn(c, VT.p, 1, acc[ch].p, open_fresh ? nullptr : accg[ch].p, accg[ch].p, nullptr, i, 0, 0,
                  group_size(i), sp);
      }
      open_fresh = false;
      return;
    }
    const NSItem &it = T.items[i];
    for (int ch = 0; ch < nch; ++ch) {
      r2c(band + ch, nch, work.p);
      axis<2, -1, 0>(work.p, work.p, none);
      axis<3, -1, 2>(work.p, acc[ch].p, it);
    }
  }
  void finish(float *x) {
    static const NSItem none{};
    if (fast) {
      close_open();
      const FastCtx c = ctx();
      for (int ch = 0; ch < nch; ++ch) {
        ops.a_inv(c, acc[ch].p, Vfin.p);
        ops.b_inv(c, Vfin.p, 1, x + ch, 0, (size_t)nch, 1.0f / (float)N, fullsup);
      }
      return;
    }
    for (int ch = 0; ch < nch; ++ch) {
      axis<3, 1, 0>(acc[ch].p, work.p, none);
      axis<2, 1, 0>(work.p, work.p, none);
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

int surfhip_ns_plan_create(surfhip_ns_plan **out, const surfhip_ns_params *p, int32_t n1, int32_t n2, int32_t n3,
                           int32_t dtype, void *stream) {
  return guarded([&] {
    if (!out || !p) fail(SURFHIP_ERR_ARG, "null argument");
    *out = nullptr;
    if (dtype != SURFHIP_F32 && dtype != SURFHIP_C64) fail(SURFHIP_ERR_ARG, fmt("unknown dtype %d", dtype));
    auto P = std::make_unique<surfhip_ns_plan>();
    ns_build_tables(*p, n1, n2, n3, P->T);
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
    P->fast = std::getenv("SURFHIP_NS_GENERIC") == nullptr && fast_ops_for(n1, n2, n3, P->ops);
    if (!P->fast) {
      P->g_rows = std::max(2, std::min(16, pick_group(n1)));
      P->g_axis = std::min(pick_group(std::max(n2, n3)), 16);
      check_lds(P->lds_rows_r2c(), "the axis-1 transform");
      check_lds(P->lds_rows_c2r(), "the axis-1 transform");
      check_lds(P->lds_axis(n2), "the axis-2 transform");
      check_lds(P->lds_axis(n3), "the axis-3 transform");
    }
    P->tw.upload(pool);
    P->hg1.upload(P->T.hg1);
    P->hg1c.upload(P->T.hg1c);
    P->hg2.upload(P->T.hg2);
    P->hg2c.upload(P->T.hg2c);
    P->a.upload(P->T.a);
    P->l.upload(P->T.l);
    P->c.upload(P->T.c);
    if (P->T.planes.empty()) P->T.planes.push_back(0.0f);
    P->planes.upload(P->T.planes);
    P->ditems.upload(P->T.items);
    P->dview = NSMaskView{P->hg1.p, P->hg1c.p, P->hg2.p, P->hg2c.p, P->a.p, P->l.p, P->c.p, P->planes.p,
                          P->h1,    n2,         n3,        P->T.J,    P->T.tight, 0, 0};
    for (int ch = 0; ch < P->nch; ++ch) {
      P->X[ch].alloc(P->H);
      P->acc[ch].alloc(P->H);
    }
    if (P->fast) {
      P->compute_supports();
      if (std::getenv("SURFHIP_NS_NO_PRUNE")) {
        for (auto &g : P->gsup) g = P->fullsup;
        P->lowsup = P->fullsup;
      }
      for (int c : P->T.group_count) P->maxw = std::max(P->maxw, c);
      P->VT.alloc((size_t)P->maxw * P->H);
      P->Vfin.alloc(P->H);
      for (int ch = 0; ch < P->nch; ++ch) P->accg[ch].alloc(P->H);
    } else {
      P->work.alloc(P->H);
    }
    // The host copies of the big tables are no longer needed.
    std::vector<float>().swap(P->T.hg1);
    std::vector<float>().swap(P->T.hg1c);
    SH_HIP(hipDeviceSynchronize());
    *out = P.release();
  });
}

int surfhip_ns_plan_destroy(surfhip_ns_plan *plan) {
  return guarded([&] {
    if (!plan) return;
    (void)hipStreamSynchronize(plan->stream);
    delete plan;
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
    plan->analyze((const float *)x, (float *)coeffs);
  });
}

int surfhip_ns_synthesize(surfhip_ns_plan *plan, const void *coeffs, void *x) {
  return guarded([&] {
    if (!plan || !x || !coeffs) fail(SURFHIP_ERR_ARG, "null argument");
    plan->synthesize((const float *)coeffs, (float *)x);
  });
}

int surfhip_ns_adjoint(surfhip_ns_plan *plan, const void *coeffs, void *x) {
  return guarded([&] {
    if (!plan || !x || !coeffs) fail(SURFHIP_ERR_ARG, "null argument");
    with_masks(plan, 1, 0, [&] { plan->synthesize((const float *)coeffs, (float *)x); });
  });
}

int surfhip_ns_synthesize_adjoint(surfhip_ns_plan *plan, const void *x, void *coeffs) {
  return guarded([&] {
    if (!plan || !x || !coeffs) fail(SURFHIP_ERR_ARG, "null argument");
    with_masks(plan, 0, 1, [&] { plan->analyze((const float *)x, (float *)coeffs); });
  });
}

int surfhip_ns_stream_begin(surfhip_ns_plan *plan, const void *x) {
  return guarded([&] {
    if (!plan || !x) fail(SURFHIP_ERR_ARG, "null argument");
    plan->forward_spectrum((const float *)x);
    plan->streaming_fwd = true;
  });
}

int surfhip_ns_stream_subband(surfhip_ns_plan *plan, int32_t i, void *band) {
  return guarded([&] {
    if (!plan || !band) fail(SURFHIP_ERR_ARG, "null argument");
    if (!plan->streaming_fwd) fail(SURFHIP_ERR_ARG, "surfhip_ns_stream_begin was not called");
    if (i < 1 || i > plan->T.S) fail(SURFHIP_ERR_ARG, fmt("subband index %d outside 1:%d", i, plan->T.S));
    plan->subbands(i - 1, 1, (float *)band);
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
    plan->accumulate(i - 1, (const float *)band);
  });
}

int surfhip_ns_stream_finish(surfhip_ns_plan *plan, void *x) {
  return guarded([&] {
    if (!plan || !x) fail(SURFHIP_ERR_ARG, "null argument");
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
    const char *s = plan->fast ? "pencil_analysis,plane_analysis,plane_synthesis,pencil_synthesis,other"
                               : "unclassified,unclassified,unclassified,unclassified,unclassified";
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
    if (i < 1 || i > T.S) fail(SURFHIP_ERR_ARG, fmt("subband index %d outside 1:%d", i, T.S));
    const NSMaskView v = T.view();
    const NSItem &it = T.items[i - 1];
    for (int k3 = 0; k3 < n3; ++k3)
      for (int k2 = 0; k2 < n2; ++k2)
        for (int k1 = 0; k1 < T.h1; ++k1)
          out[(size_t)k1 + (size_t)T.h1 * ((size_t)k2 + (size_t)n2 * k3)] = ns_mask(v, it, k1, k2, k3, true);
  });
}

}  // extern "C"
