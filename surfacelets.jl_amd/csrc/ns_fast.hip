// Fast NSSurfacelet path for power-of-two grids (n1, n2 in {32, 64, 128},
// n3 in {32, 64, 128, 256}).
//
// A 3-D real FFT is split into two fused kernels with compile-time lengths:
//   pencil pass (axis 3):  spectra X / acc are pencil-major, [k1][k2][k3], so a
//     block of TP pencils (k1 fixed, TP consecutive k2) reads and writes one
//     contiguous run; the mask is evaluated in the same pass.
//   plane pass (axes 2 and 1): intermediates V / T are plane-major,
//     [i3][k1][k2], so a block reads one contiguous (k1, k2) plane, transforms
//     it along k2 and then along k1 (a real transform, two columns per complex
//     FFT) and writes one contiguous (i1, i2) output plane.
// Every FFT is a four-step FFT in registers (fft_reg.hpp) with one LDS
// exchange.  A group of subbands (one scale and hourglass, its wedges) runs in
// one launch per pass; subbands never share a kernel's arithmetic, so the
// streaming entry points (one subband per launch) give bit-identical results.
//
// Synthesis accumulates per group as the reference's engine does
// (src/transforms/ns_surfacelet.jl:23-32): acc_g = sum_w F_w . wedge_w, then
// acc += acc_g . G_g, with explicit FMAs in a fixed order.
#include <hip/hip_runtime.h>

#include "fft_reg.hpp"
#include "hip_util.hpp"
#include "ns_fast.hpp"
#include "ns_tables.hpp"

namespace surfhip {

constexpr int TP = 16;  // pencils per pencil-pass block

template <int N> constexpr int pencil_threads() { return TP * FourStep<N>::R2; }

template <int N1, int N2> constexpr int plane_threads() {
  constexpr int H1 = N1 / 2 + 1;
  constexpr int a = H1 * FourStep<N2>::R2, b = (N2 / 2) * FourStep<N1>::R2;
  constexpr int m = a > b ? a : b;
  return (m + 63) / 64 * 64;
}

// ----------------------------------------------------------- pencil helpers

// In-place FFT of the TP pencils held in buf (pencil p at buf + p * PS),
// R2 threads per pencil.  All threads of the block call it.
template <int N, int DIR, int PS>
__device__ __forceinline__ void pencil_fft(float2 *buf, const float2 *tw, int tid) {
  using F = RegFFT<N, DIR>;
  constexpr int R1 = F::R1, R2 = F::R2, ROWS = F::ROWS;
  const int p = tid / R2, j = tid % R2;
  float2 *pb = buf + p * PS;
  float2 v[R1];
#pragma unroll
  for (int r = 0; r < R1; ++r) v[r] = pb[j + R2 * r];
  F::phase1(v, j, tw);
  __syncthreads();
#pragma unroll
  for (int s = 0; s < R1; ++s) pb[s * R2 + j] = v[s];
  __syncthreads();
#pragma unroll
  for (int e = 0; e < R1; ++e) v[e] = pb[j * R1 + e];
  F::phase2(v);
  __syncthreads();
#pragma unroll
  for (int u = 0; u < ROWS; ++u)
#pragma unroll
    for (int q = 0; q < R2; ++q) pb[(j * ROWS + u) + R1 * q] = v[u * R2 + q];
  __syncthreads();
}

__device__ __forceinline__ float2 fma2(float2 a, float s, float2 c) {
  return make_float2(__fmaf_rn(a.x, s, c.x), __fmaf_rn(a.y, s, c.y));
}

// ------------------------------------------------------------ pencil passes

// X[k1][k2][:] = FFT_3(T[:][k1][k2]) (forward; T plane-major).
template <int N3>
__global__ __launch_bounds__(pencil_threads<N3>()) void kA_fwd(const float2 *__restrict__ T,
                                                               float2 *__restrict__ X, int h1, int n2,
                                                               const float2 *__restrict__ tw3) {
  constexpr int NTH = pencil_threads<N3>(), PS = N3 + 1;
  __shared__ float2 buf[TP * PS];
  __shared__ float2 tw[N3];
  const int tid = threadIdx.x;
  const int k2_0 = blockIdx.x * TP, k1 = blockIdx.y;
  for (int m = tid; m < N3; m += NTH) tw[m] = tw3[m];
  for (int e = tid; e < TP * N3; e += NTH) {
    const int i3 = e / TP, p = e % TP;
    buf[p * PS + i3] = T[((size_t)i3 * h1 + k1) * n2 + k2_0 + p];
  }
  __syncthreads();
  pencil_fft<N3, -1, PS>(buf, tw, tid);
  float2 *dst = X + ((size_t)k1 * n2 + k2_0) * N3;
  for (int e = tid; e < TP * N3; e += NTH) dst[e] = buf[(e / N3) * PS + e % N3];
}

// Analysis: V_w[i3][k1][k2] = IFFT_3(X . M_w), w = item0 + blockIdx.z.
template <int N3>
__global__ __launch_bounds__(pencil_threads<N3>()) void kA_ana(const float2 *__restrict__ X,
                                                               float2 *__restrict__ V, int h1, int n2,
                                                               NSMaskView mv, const NSItem *__restrict__ items,
                                                               int item0, size_t vstride,
                                                               const float2 *__restrict__ tw3) {
  constexpr int NTH = pencil_threads<N3>(), PS = N3 + 1;
  __shared__ float2 buf[TP * PS];
  __shared__ float2 tw[N3];
  const int tid = threadIdx.x;
  const int k2_0 = blockIdx.x * TP, k1 = blockIdx.y, w = blockIdx.z;
  const NSItem it = items[item0 + w];
  for (int m = tid; m < N3; m += NTH) tw[m] = tw3[m];
  const float2 *src = X + ((size_t)k1 * n2 + k2_0) * N3;
  for (int e = tid; e < TP * N3; e += NTH) {
    const int p = e / N3, k3 = e % N3;
    const float m = ns_mask(mv, it, k1, k2_0 + p, k3, true);
    const float2 v = src[e];
    buf[p * PS + k3] = make_float2(v.x * m, v.y * m);
  }
  __syncthreads();
  pencil_fft<N3, 1, PS>(buf, tw, tid);
  float2 *dst = V + (size_t)w * vstride;
  for (int e = tid; e < TP * N3; e += NTH) {
    const int i3 = e / TP, p = e % TP;
    dst[((size_t)i3 * h1 + k1) * n2 + k2_0 + p] = buf[p * PS + i3];
  }
}

// Synthesis pencil pass over the items [item0, item0 + nw) of one group:
//   acc_g = accg_in (or 0) + sum_w FFT_3(T_w) . wedge_w
// then either accg_out = acc_g (an open group, streaming) or
//   acc = acc + acc_g . G  (closing the group), and with `finalize` the
//   closed pencil is transformed back along axis 3 into V (plane-major).
template <int N3>
__global__ __launch_bounds__(pencil_threads<N3>()) void kA_syn(
    const float2 *__restrict__ T, size_t tstride, int nw, float2 *__restrict__ acc,
    const float2 *__restrict__ accg_in, float2 *__restrict__ accg_out, float2 *__restrict__ V, int h1, int n2,
    NSMaskView mv, const NSItem *__restrict__ items, int item0, int close, int finalize,
    const float2 *__restrict__ tw3) {
  constexpr int NTH = pencil_threads<N3>(), PS = N3 + 1, PER = TP * N3 / NTH;
  __shared__ float2 buf[TP * PS];
  __shared__ float2 tw[N3];
  const int tid = threadIdx.x;
  const int k2_0 = blockIdx.x * TP, k1 = blockIdx.y;
  for (int m = tid; m < N3; m += NTH) tw[m] = tw3[m];
  const size_t base = ((size_t)k1 * n2 + k2_0) * N3;
  float2 g[PER];
#pragma unroll
  for (int m = 0; m < PER; ++m) g[m] = accg_in ? accg_in[base + tid + NTH * m] : make_float2(0.0f, 0.0f);
  for (int w = 0; w < nw; ++w) {
    const NSItem it = items[item0 + w];
    const float2 *src = T + (size_t)w * tstride;
    for (int e = tid; e < TP * N3; e += NTH) {
      const int i3 = e / TP, p = e % TP;
      buf[p * PS + i3] = src[((size_t)i3 * h1 + k1) * n2 + k2_0 + p];
    }
    __syncthreads();
    pencil_fft<N3, -1, PS>(buf, tw, tid);
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      const int e = tid + NTH * m, p = e / N3, k3 = e % N3;
      const float wd = ns_wedge(mv, it, k1, k2_0 + p, k3, false);
      g[m] = fma2(buf[p * PS + k3], wd, g[m]);
    }
    __syncthreads();
  }
  if (!close) {
#pragma unroll
    for (int m = 0; m < PER; ++m) accg_out[base + tid + NTH * m] = g[m];
    return;
  }
  const NSItem it0 = items[item0];
#pragma unroll
  for (int m = 0; m < PER; ++m) {
    const int e = tid + NTH * m, p = e / N3, k3 = e % N3;
    const float G = ns_group(mv, it0, k1, k2_0 + p, k3, false);
    const float2 a = fma2(g[m], G, acc[base + e]);
    if (finalize) buf[p * PS + k3] = a;
    else acc[base + e] = a;
  }
  if (!finalize) return;
  __syncthreads();
  pencil_fft<N3, 1, PS>(buf, tw, tid);
  for (int e = tid; e < TP * N3; e += NTH) {
    const int i3 = e / TP, p = e % TP;
    V[((size_t)i3 * h1 + k1) * n2 + k2_0 + p] = buf[p * PS + i3];
  }
}

// V[i3][k1][k2] = IFFT_3(acc) (the inverse of the accumulated spectrum).
template <int N3>
__global__ __launch_bounds__(pencil_threads<N3>()) void kA_inv(const float2 *__restrict__ acc,
                                                               float2 *__restrict__ V, int h1, int n2,
                                                               const float2 *__restrict__ tw3) {
  constexpr int NTH = pencil_threads<N3>(), PS = N3 + 1;
  __shared__ float2 buf[TP * PS];
  __shared__ float2 tw[N3];
  const int tid = threadIdx.x;
  const int k2_0 = blockIdx.x * TP, k1 = blockIdx.y;
  for (int m = tid; m < N3; m += NTH) tw[m] = tw3[m];
  const float2 *src = acc + ((size_t)k1 * n2 + k2_0) * N3;
  for (int e = tid; e < TP * N3; e += NTH) buf[(e / N3) * PS + e % N3] = src[e];
  __syncthreads();
  pencil_fft<N3, 1, PS>(buf, tw, tid);
  for (int e = tid; e < TP * N3; e += NTH) {
    const int i3 = e / TP, p = e % TP;
    V[((size_t)i3 * h1 + k1) * n2 + k2_0 + p] = buf[p * PS + i3];
  }
}

// acc += accg . G (closing an open streaming group).
template <int N3>
__global__ __launch_bounds__(256) void kA_close(float2 *__restrict__ acc, const float2 *__restrict__ accg, int h1,
                                                int n2, NSMaskView mv, const NSItem *__restrict__ items, int item0) {
  const NSItem it0 = items[item0];
  const size_t total = (size_t)h1 * n2 * N3;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int k3 = e % N3;
    const size_t r = e / N3;
    const int k2 = r % n2, k1 = r / n2;
    const float G = ns_group(mv, it0, k1, k2, k3, false);
    acc[e] = fma2(accg[e], G, acc[e]);
  }
}

// ------------------------------------------------------------- plane passes

// Pair slot e of column pair p in a plane stored [k1][k2] (row stride PS):
// the pair's N1-point sequence occupies column 2p rows 0..N1/2 and column
// 2p+1 rows 1..N1/2-1.
template <int N1, int PS> __device__ __forceinline__ int pair_slot(int p, int e) {
  return e <= N1 / 2 ? e * PS + 2 * p : (e - N1 / 2) * PS + 2 * p + 1;
}

// In-place FFT along k2 of rows 0..H1-1 of the plane (R2 threads per row).
template <int N2, int DIR, int H1, int PS, int NTH>
__device__ __forceinline__ void plane_rows_fft(float2 *L, const float2 *tw, int tid) {
  using F = RegFFT<N2, DIR>;
  constexpr int R1 = F::R1, R2 = F::R2, ROWS = F::ROWS;
  const int row = tid / R2, j = tid % R2;
  const bool on = row < H1;
  float2 *rb = L + (on ? row : 0) * PS;
  float2 v[R1];
  if (on) {
#pragma unroll
    for (int r = 0; r < R1; ++r) v[r] = rb[j + R2 * r];
    F::phase1(v, j, tw);
  }
  __syncthreads();
  if (on) {
#pragma unroll
    for (int s = 0; s < R1; ++s) rb[s * R2 + j] = v[s];
  }
  __syncthreads();
  if (on) {
#pragma unroll
    for (int e = 0; e < R1; ++e) v[e] = rb[j * R1 + e];
    F::phase2(v);
  }
  __syncthreads();
  if (on) {
#pragma unroll
    for (int u = 0; u < ROWS; ++u)
#pragma unroll
      for (int q = 0; q < R2; ++q) rb[(j * ROWS + u) + R1 * q] = v[u * R2 + q];
  }
  __syncthreads();
}

// Analysis plane pass: out_w plane i3 = c2r_1(IFFT_2(V_w[i3])) * scale.
// out element (i1, i2, i3) at out[(i1 + N1 (i2 + N2 i3)) * os] + w * owstride.
template <int N1, int N2>
__global__ __launch_bounds__((plane_threads<N1, N2>())) void kB_inv(const float2 *__restrict__ V, size_t vstride,
                                                                  float *__restrict__ out, size_t owstride,
                                                                  size_t os, float scale,
                                                                  const float2 *__restrict__ tw1g,
                                                                  const float2 *__restrict__ tw2g) {
  constexpr int H1 = N1 / 2 + 1, PS = N2 + 1, NTH = plane_threads<N1, N2>();
  __shared__ float2 L[H1 * PS];
  __shared__ float2 tw1[N1], tw2[N2];
  const int tid = threadIdx.x;
  const int i3 = blockIdx.x, w = blockIdx.y;
  for (int m = tid; m < N1; m += NTH) tw1[m] = tw1g[m];
  for (int m = tid; m < N2; m += NTH) tw2[m] = tw2g[m];
  const float2 *src = V + (size_t)w * vstride + (size_t)i3 * H1 * N2;
  for (int e = tid; e < H1 * N2; e += NTH) L[(e / N2) * PS + e % N2] = src[e];
  __syncthreads();
  plane_rows_fft<N2, 1, H1, PS, NTH>(L, tw2, tid);
  // c2r along k1: pair p = columns (2p, 2p+1), z = a + i b.
  using F = RegFFT<N1, 1>;
  constexpr int R1 = F::R1, R2 = F::R2, ROWS = F::ROWS;
  const int p = tid / R2, j = tid % R2;
  const bool on = p < N2 / 2;
  float2 v[R1];
  if (on) {
#pragma unroll
    for (int r = 0; r < R1; ++r) {
      const int t = j + R2 * r;
      float2 A, B;
      if (t <= N1 / 2) {
        A = L[t * PS + 2 * p];
        B = L[t * PS + 2 * p + 1];
        if (t == 0 || t == N1 / 2) {
          A.y = 0.0f;
          B.y = 0.0f;
        }
      } else {
        A = conjf2(L[(N1 - t) * PS + 2 * p]);
        B = conjf2(L[(N1 - t) * PS + 2 * p + 1]);
      }
      v[r] = make_float2(A.x - B.y, A.y + B.x);
    }
    F::phase1(v, j, tw1);
  }
  __syncthreads();
  if (on) {
#pragma unroll
    for (int s = 0; s < R1; ++s) L[pair_slot<N1, PS>(p, s * R2 + j)] = v[s];
  }
  __syncthreads();
  if (on) {
#pragma unroll
    for (int e = 0; e < R1; ++e) v[e] = L[pair_slot<N1, PS>(p, j * R1 + e)];
    F::phase2(v);
  }
  __syncthreads();
  float *O = reinterpret_cast<float *>(L);  // [i2][i1], N1 * N2 floats
  if (on) {
#pragma unroll
    for (int u = 0; u < ROWS; ++u)
#pragma unroll
      for (int q = 0; q < R2; ++q) {
        const int i1 = (j * ROWS + u) + R1 * q;
        const float2 z = v[u * R2 + q];
        O[(2 * p) * N1 + i1] = z.x * scale;
        O[(2 * p + 1) * N1 + i1] = z.y * scale;
      }
  }
  __syncthreads();
  float *dst = out + (size_t)w * owstride + (size_t)i3 * N1 * N2 * os;
  for (int e = tid; e < N1 * N2; e += NTH) dst[(size_t)e * os] = O[e];
}

// Synthesis plane pass: T_w[i3] = FFT_2(r2c_1(in_w plane i3)), T plane-major.
template <int N1, int N2>
__global__ __launch_bounds__((plane_threads<N1, N2>())) void kB_fwd(const float *__restrict__ in, size_t iwstride,
                                                                  size_t is, float2 *__restrict__ T, size_t tstride,
                                                                  const float2 *__restrict__ tw1g,
                                                                  const float2 *__restrict__ tw2g) {
  constexpr int H1 = N1 / 2 + 1, PS = N2 + 1, NTH = plane_threads<N1, N2>();
  __shared__ float2 L[H1 * PS];
  __shared__ float2 tw1[N1], tw2[N2];
  const int tid = threadIdx.x;
  const int i3 = blockIdx.x, w = blockIdx.y;
  for (int m = tid; m < N1; m += NTH) tw1[m] = tw1g[m];
  for (int m = tid; m < N2; m += NTH) tw2[m] = tw2g[m];
  const float *src = in + (size_t)w * iwstride + (size_t)i3 * N1 * N2 * is;
  for (int e = tid; e < N1 * N2; e += NTH) {
    const int i2 = e / N1, i1 = e % N1;
    float *f = reinterpret_cast<float *>(&L[pair_slot<N1, PS>(i2 >> 1, i1)]);
    f[i2 & 1] = src[(size_t)e * is];
  }
  __syncthreads();
  using F = RegFFT<N1, -1>;
  constexpr int R1 = F::R1, R2 = F::R2, ROWS = F::ROWS;
  const int p = tid / R2, j = tid % R2;
  const bool on = p < N2 / 2;
  float2 v[R1];
  if (on) {
#pragma unroll
    for (int r = 0; r < R1; ++r) v[r] = L[pair_slot<N1, PS>(p, j + R2 * r)];
    F::phase1(v, j, tw1);
  }
  __syncthreads();
  if (on) {
#pragma unroll
    for (int s = 0; s < R1; ++s) L[pair_slot<N1, PS>(p, s * R2 + j)] = v[s];
  }
  __syncthreads();
  if (on) {
#pragma unroll
    for (int e = 0; e < R1; ++e) v[e] = L[pair_slot<N1, PS>(p, j * R1 + e)];
    F::phase2(v);
  }
  __syncthreads();
  if (on) {
#pragma unroll
    for (int u = 0; u < ROWS; ++u)
#pragma unroll
      for (int q = 0; q < R2; ++q) L[pair_slot<N1, PS>(p, (j * ROWS + u) + R1 * q)] = v[u * R2 + q];
  }
  __syncthreads();
  // Unpack: A[k] = (Z[k] + conj Z[-k]) / 2, B[k] = (Z[k] - conj Z[-k]) / 2i.
  constexpr int KPER = (H1 + R2 - 1) / R2;
  float2 A[KPER], B[KPER];
  if (on) {
#pragma unroll
    for (int m = 0; m < KPER; ++m) {
      const int k = j + R2 * m;
      if (k < H1) {
        const float2 Z = L[pair_slot<N1, PS>(p, k)];
        const float2 Zc = conjf2(L[pair_slot<N1, PS>(p, (N1 - k) % N1)]);
        A[m] = cscale(cadd(Z, Zc), 0.5f);
        const float2 d = csub(Z, Zc);
        B[m] = make_float2(0.5f * d.y, -0.5f * d.x);
      }
    }
  }
  __syncthreads();
  if (on) {
#pragma unroll
    for (int m = 0; m < KPER; ++m) {
      const int k = j + R2 * m;
      if (k < H1) {
        L[k * PS + 2 * p] = A[m];
        L[k * PS + 2 * p + 1] = B[m];
      }
    }
  }
  __syncthreads();
  plane_rows_fft<N2, -1, H1, PS, NTH>(L, tw2, tid);
  float2 *dst = T + (size_t)w * tstride + (size_t)i3 * H1 * N2;
  for (int e = tid; e < H1 * N2; e += NTH) dst[e] = L[(e / N2) * PS + e % N2];
}

// ------------------------------------------------------------------ host side

namespace {

template <int N3> struct PencilLaunch {
  static void fwd(const FastCtx &c, const float2 *T, float2 *X) {
    dim3 grid(c.n2 / TP, c.h1);
    hipLaunchKernelGGL(kA_fwd<N3>, grid, dim3(pencil_threads<N3>()), 0, c.stream, T, X, c.h1, c.n2, c.tw3);
    SH_LAUNCHED();
  }
  static void ana(const FastCtx &c, const float2 *X, float2 *V, int item0, int nw) {
    dim3 grid(c.n2 / TP, c.h1, nw);
    hipLaunchKernelGGL(kA_ana<N3>, grid, dim3(pencil_threads<N3>()), 0, c.stream, X, V, c.h1, c.n2, c.mv, c.items,
                       item0, (size_t)c.H, c.tw3);
    SH_LAUNCHED();
  }
  static void syn(const FastCtx &c, const float2 *T, int nw, float2 *acc, const float2 *gin, float2 *gout, float2 *V,
                  int item0, int close, int finalize) {
    dim3 grid(c.n2 / TP, c.h1);
    hipLaunchKernelGGL(kA_syn<N3>, grid, dim3(pencil_threads<N3>()), 0, c.stream, T, (size_t)c.H, nw, acc, gin, gout,
                       V, c.h1, c.n2, c.mv, c.items, item0, close, finalize, c.tw3);
    SH_LAUNCHED();
  }
  static void inv(const FastCtx &c, const float2 *acc, float2 *V) {
    dim3 grid(c.n2 / TP, c.h1);
    hipLaunchKernelGGL(kA_inv<N3>, grid, dim3(pencil_threads<N3>()), 0, c.stream, acc, V, c.h1, c.n2, c.tw3);
    SH_LAUNCHED();
  }
  static void close(const FastCtx &c, float2 *acc, const float2 *accg, int item0) {
    hipLaunchKernelGGL(kA_close<N3>, dim3(2048), dim3(256), 0, c.stream, acc, accg, c.h1, c.n2, c.mv, c.items, item0);
    SH_LAUNCHED();
  }
};

template <int N1, int N2> struct PlaneLaunch {
  static void inv(const FastCtx &c, const float2 *V, int nw, float *out, size_t owstride, size_t os, float scale) {
    dim3 grid(c.n3, nw);
    hipLaunchKernelGGL((kB_inv<N1, N2>), grid, dim3(plane_threads<N1, N2>()), 0, c.stream, V, (size_t)c.H, out,
                       owstride, os, scale, c.tw1, c.tw2);
    SH_LAUNCHED();
  }
  static void fwd(const FastCtx &c, const float *in, size_t iwstride, size_t is, int nw, float2 *T) {
    dim3 grid(c.n3, nw);
    hipLaunchKernelGGL((kB_fwd<N1, N2>), grid, dim3(plane_threads<N1, N2>()), 0, c.stream, in, iwstride, is, T,
                       (size_t)c.H, c.tw1, c.tw2);
    SH_LAUNCHED();
  }
};

template <int N3> void pencil_dispatch_set(FastOps &o) {
  o.a_fwd = &PencilLaunch<N3>::fwd;
  o.a_ana = &PencilLaunch<N3>::ana;
  o.a_syn = &PencilLaunch<N3>::syn;
  o.a_inv = &PencilLaunch<N3>::inv;
  o.a_close = &PencilLaunch<N3>::close;
}

template <int N1, int N2> void plane_dispatch_set(FastOps &o) {
  o.b_inv = &PlaneLaunch<N1, N2>::inv;
  o.b_fwd = &PlaneLaunch<N1, N2>::fwd;
}

}  // namespace

bool fast_ops_for(int n1, int n2, int n3, FastOps &o) {
  o = FastOps{};
  switch (n3) {
    case 32: pencil_dispatch_set<32>(o); break;
    case 64: pencil_dispatch_set<64>(o); break;
    case 128: pencil_dispatch_set<128>(o); break;
    case 256: pencil_dispatch_set<256>(o); break;
    default: return false;
  }
#define SH_PLANE(a, b) \
  if (n1 == a && n2 == b) { plane_dispatch_set<a, b>(o); return true; }
  SH_PLANE(32, 32) SH_PLANE(32, 64) SH_PLANE(32, 128)
  SH_PLANE(64, 32) SH_PLANE(64, 64) SH_PLANE(64, 128)
  SH_PLANE(128, 32) SH_PLANE(128, 64) SH_PLANE(128, 128)
#undef SH_PLANE
  return false;
}

}  // namespace surfhip
