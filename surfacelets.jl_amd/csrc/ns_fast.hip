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

#include <cstdint>

#include "fft_reg.hpp"
#include "hip_util.hpp"
#include "ns_fast.hpp"
#include "ns_tables.hpp"

namespace surfhip {

// Pencils per block (per wedge): 16 (8 for 256-point pencils, to stay within
// 1024 threads and 160 KiB of LDS with eight wedges in flight).
template <int N> constexpr int tp_of() { return N >= 256 ? 8 : 16; }
template <int N> constexpr int sub_threads() { return tp_of<N>() * FourStep<N>::R2; }
// Slots of one pencil in LDS: N data slots, the padded exchange needs
// N + N/R1, and an odd stride keeps transposed accesses conflict-free.
template <int N> constexpr int pencil_stride() { return N + N / FourStep<N>::R1 + 1; }

// LDS of a plane block (float2 slots): the plane rows, the pair-contiguous
// scratch of the axis-1 real transform, and the real output plane.
template <int N1, int N2> constexpr int plane_slots() {
  constexpr int a = (N1 / 2 + 1) * pencil_stride<N2>(), b = (N2 / 2) * pencil_stride<N1>(), c = N1 * N2 / 2;
  constexpr int m = a > b ? a : b;
  return m > c ? m : c;
}

template <int N1, int N2> constexpr int plane_threads() {
  constexpr int H1 = N1 / 2 + 1;
  constexpr int a = H1 * FourStep<N2>::R2, b = (N2 / 2) * FourStep<N1>::R2;
  constexpr int m = a > b ? a : b;
  return (m + 63) / 64 * 64;
}

// Exchange position of slot e: one pad slot per R1, so that the R1
// consecutive slots each thread reads in phase 2 start in different banks.
template <int R1> __device__ __forceinline__ int xpos(int e) { return e + e / R1; }

// In-place FFT of one pencil at pb (natural order in slots 0..N-1), by the
// R2 threads j of the pencil; all threads of the block call it (it holds
// __syncthreads), `active` masks the work of threads without a pencil.
template <int N, int DIR>
__device__ __forceinline__ void pencil_fft(float2 *pb, const float2 *tw, int j, bool active) {
  using F = RegFFT<N, DIR>;
  constexpr int R1 = F::R1, R2 = F::R2, ROWS = F::ROWS;
  float2 v[R1];
  if (active) {
#pragma unroll
    for (int r = 0; r < R1; ++r) v[r] = pb[j + R2 * r];
    F::phase1(v, j, tw);
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int s = 0; s < R1; ++s) pb[xpos<R1>(s * R2 + j)] = v[s];
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int e = 0; e < R1; ++e) v[e] = pb[xpos<R1>(j * R1 + e)];
    F::phase2(v);
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int u = 0; u < ROWS; ++u)
#pragma unroll
      for (int q = 0; q < R2; ++q) pb[(j * ROWS + u) + R1 * q] = v[u * R2 + q];
  }
  __syncthreads();
}

__device__ __forceinline__ float2 fma2(float2 a, float s, float2 c) {
  return make_float2(__fmaf_rn(a.x, s, c.x), __fmaf_rn(a.y, s, c.y));
}

// Pencil tiles (TP consecutive k2) of a support: the first `lo` tiles from
// k2 = 0 and the last `hi` tiles before k2 = n2 (the negative frequencies).
struct Tiles {
  int lo, hi, all;
};
__host__ __device__ inline Tiles support_tiles(Support sp, int n2, int TP) {
  const int all = n2 / TP;
  int lo = (sp.K2 + 1 + TP - 1) / TP, hi = (sp.K2 + TP - 1) / TP;
  if (lo + hi >= all) {
    lo = all;
    hi = 0;
  }
  return Tiles{lo, hi, all};
}
__device__ __forceinline__ int tile_start(Tiles t, int b, int n2, int TP) {
  return b < t.lo ? b * TP : n2 - (t.lo + t.hi - b) * TP;
}
// Whether column k2 lies in the support's tiles.
__device__ __forceinline__ bool tile_col(Tiles t, int k2, int n2, int TP) {
  const int b = k2 / TP;
  return b < t.lo || b >= t.all - t.hi;
}

// ------------------------------------------------------------ pencil passes

// X[k1][k2][:] = FFT_3(T[:][k1][k2]) (forward; T plane-major).
template <int N3>
__global__ __launch_bounds__(sub_threads<N3>()) void kA_fwd(const float2 *__restrict__ T, float2 *__restrict__ X,
                                                            int h1, int n2, const float2 *__restrict__ tw3) {
  constexpr int TP = tp_of<N3>(), NTH = sub_threads<N3>(), PS = pencil_stride<N3>(), R2 = FourStep<N3>::R2;
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
  pencil_fft<N3, -1>(buf + (tid / R2) * PS, tw, tid % R2, true);
  float2 *dst = X + ((size_t)k1 * n2 + k2_0) * N3;
  for (int e = tid; e < TP * N3; e += NTH) dst[e] = buf[(e / N3) * PS + e % N3];
}

// Analysis over the items [item0, item0 + nw) of one group, NW of them at a
// time in parallel: Y = X . G once per element, then per wedge
// V_w[i3][k1][k2] = IFFT_3(Y . W_w)  (src/transforms/ns_surfacelet.jl:23-32).
template <int N3, int NW>
__global__ __launch_bounds__(NW * sub_threads<N3>()) void kA_ana(const float2 *__restrict__ X,
                                                                 float2 *__restrict__ V, int h1, int n2,
                                                                 NSMaskView mv, const NSItem *__restrict__ items,
                                                                 int item0, int nw, size_t vstride,
                                                                 const float2 *__restrict__ tw3, Support sp) {
  constexpr int TP = tp_of<N3>(), SUB = sub_threads<N3>(), NTH = NW * SUB, PS = pencil_stride<N3>();
  constexpr int R2 = FourStep<N3>::R2, TILE = TP * PS;
  __shared__ float2 wt[NW * TILE];
  __shared__ float2 tw[N3];
  __shared__ NSItem its[NW];
  const int tid = threadIdx.x;
  const int k2_0 = tile_start(support_tiles(sp, n2, TP), blockIdx.x, n2, TP), k1 = blockIdx.y;
  const int sg = tid / SUB, lt = tid % SUB;
  for (int m = tid; m < N3; m += NTH) tw[m] = tw3[m];
  const NSItem it0 = items[item0];
  const float2 *src = X + ((size_t)k1 * n2 + k2_0) * N3;
  for (int c0 = 0; c0 < nw; c0 += NW) {
    const int cw = min(NW, nw - c0);
    if (tid < cw) its[tid] = items[item0 + c0 + tid];
    __syncthreads();
    for (int e = tid; e < TP * N3; e += NTH) {
      const int p = e / N3, k3 = e % N3;
      const float G = ns_group(mv, it0, k1, k2_0 + p, k3, true);
      const float2 x = src[e];
      const float2 y = make_float2(x.x * G, x.y * G);
      for (int w = 0; w < cw; ++w) {
        const float W = ns_wedge(mv, its[w], k1, k2_0 + p, k3, true);
        wt[w * TILE + p * PS + k3] = make_float2(y.x * W, y.y * W);
      }
    }
    __syncthreads();
    pencil_fft<N3, 1>(wt + sg * TILE + (lt / R2) * PS, tw, lt % R2, sg < cw);
    for (int q = tid; q < cw * TP * N3; q += NTH) {
      const int w = q / (TP * N3), e = q % (TP * N3);
      const int i3 = e / TP, p = e % TP;
      V[(size_t)(c0 + w) * vstride + ((size_t)i3 * h1 + k1) * n2 + k2_0 + p] = wt[w * TILE + p * PS + i3];
    }
    __syncthreads();
  }
}

// Synthesis over the items [item0, item0 + nw) of one group, NW at a time:
//   acc_g = accg_in (or 0); for w in order: acc_g = fma(FFT_3(T_w), W_w, acc_g)
// then accg_out = acc_g (group left open: streaming) or acc = fma(acc_g, G, acc)
// (group closed); with `finalize` the closed pencils go back along axis 3
// into V (plane-major).  The fixed FMA order makes one-item launches
// (streaming) bit-identical to a whole group.
template <int N3, int NW>
__global__ __launch_bounds__(NW * sub_threads<N3>()) void kA_syn(
    const float2 *__restrict__ T, size_t tstride, int nw, float2 *__restrict__ acc,
    const float2 *__restrict__ accg_in, float2 *__restrict__ accg_out, float2 *__restrict__ V, int h1, int n2,
    NSMaskView mv, const NSItem *__restrict__ items, int item0, int close, int finalize,
    const float2 *__restrict__ tw3, Support sp) {
  constexpr int TP = tp_of<N3>(), SUB = sub_threads<N3>(), NTH = NW * SUB, PS = pencil_stride<N3>();
  constexpr int R2 = FourStep<N3>::R2, TILE = TP * PS, PER = (TP * N3 + NTH - 1) / NTH;
  __shared__ float2 wt[NW * TILE];
  __shared__ float2 tw[N3];
  __shared__ NSItem its[NW];
  const int tid = threadIdx.x;
  // A finalizing launch covers the whole grid (it transforms all of acc);
  // blocks outside the group's support then add nothing.
  const Tiles tl = support_tiles(sp, n2, TP);
  const int k2_0 = finalize ? blockIdx.x * TP : tile_start(tl, blockIdx.x, n2, TP), k1 = blockIdx.y;
  const bool inside = k1 <= sp.K1 && tile_col(tl, k2_0, n2, TP);
  const int sg = tid / SUB, lt = tid % SUB;
  for (int m = tid; m < N3; m += NTH) tw[m] = tw3[m];
  const size_t base = ((size_t)k1 * n2 + k2_0) * N3;
  float2 g[PER];
#pragma unroll
  for (int m = 0; m < PER; ++m) {
    const int e = tid + NTH * m;
    g[m] = (accg_in && e < TP * N3) ? accg_in[base + e] : make_float2(0.0f, 0.0f);
  }
  for (int c0 = 0; inside && c0 < nw; c0 += NW) {
    const int cw = min(NW, nw - c0);
    if (tid < cw) its[tid] = items[item0 + c0 + tid];
    for (int q = tid; q < cw * TP * N3; q += NTH) {
      const int w = q / (TP * N3), e = q % (TP * N3);
      const int i3 = e / TP, p = e % TP;
      wt[w * TILE + p * PS + i3] = T[(size_t)(c0 + w) * tstride + ((size_t)i3 * h1 + k1) * n2 + k2_0 + p];
    }
    __syncthreads();
    pencil_fft<N3, -1>(wt + sg * TILE + (lt / R2) * PS, tw, lt % R2, sg < cw);
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      const int e = tid + NTH * m;
      if (e < TP * N3) {
        const int p = e / N3, k3 = e % N3;
        for (int w = 0; w < cw; ++w) {
          const float W = ns_wedge(mv, its[w], k1, k2_0 + p, k3, false);
          g[m] = fma2(wt[w * TILE + p * PS + k3], W, g[m]);
        }
      }
    }
    __syncthreads();
  }
  if (!close) {
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      const int e = tid + NTH * m;
      if (e < TP * N3) accg_out[base + e] = g[m];
    }
    return;
  }
  const NSItem it0 = items[item0];
#pragma unroll
  for (int m = 0; m < PER; ++m) {
    const int e = tid + NTH * m;
    if (e < TP * N3) {
      const int p = e / N3, k3 = e % N3;
      const float G = ns_group(mv, it0, k1, k2_0 + p, k3, false);
      const float2 a = fma2(g[m], G, acc[base + e]);
      if (finalize) wt[p * PS + k3] = a;
      else acc[base + e] = a;
    }
  }
  if (!finalize) return;
  __syncthreads();
  pencil_fft<N3, 1>(wt + (lt / R2) * PS, tw, lt % R2, sg == 0);
  for (int e = tid; e < TP * N3; e += NTH) {
    const int i3 = e / TP, p = e % TP;
    V[((size_t)i3 * h1 + k1) * n2 + k2_0 + p] = wt[p * PS + i3];
  }
}

// V[i3][k1][k2] = IFFT_3(acc) (the inverse of the accumulated spectrum).
template <int N3>
__global__ __launch_bounds__(sub_threads<N3>()) void kA_inv(const float2 *__restrict__ acc, float2 *__restrict__ V,
                                                            int h1, int n2, const float2 *__restrict__ tw3) {
  constexpr int TP = tp_of<N3>(), NTH = sub_threads<N3>(), PS = pencil_stride<N3>(), R2 = FourStep<N3>::R2;
  __shared__ float2 buf[TP * PS];
  __shared__ float2 tw[N3];
  const int tid = threadIdx.x;
  const int k2_0 = blockIdx.x * TP, k1 = blockIdx.y;
  for (int m = tid; m < N3; m += NTH) tw[m] = tw3[m];
  const float2 *src = acc + ((size_t)k1 * n2 + k2_0) * N3;
  for (int e = tid; e < TP * N3; e += NTH) buf[(e / N3) * PS + e % N3] = src[e];
  __syncthreads();
  pencil_fft<N3, 1>(buf + (tid / R2) * PS, tw, tid % R2, true);
  for (int e = tid; e < TP * N3; e += NTH) {
    const int i3 = e / TP, p = e % TP;
    V[((size_t)i3 * h1 + k1) * n2 + k2_0 + p] = buf[p * PS + i3];
  }
}

// acc += accg . G (closing an open streaming group).
template <int N3>
__global__ __launch_bounds__(256) void kA_close(float2 *__restrict__ acc, const float2 *__restrict__ accg, int h1,
                                                int n2, NSMaskView mv, const NSItem *__restrict__ items, int item0,
                                                Support sp) {
  const NSItem it0 = items[item0];
  const Tiles tl = support_tiles(sp, n2, tp_of<N3>());
  const size_t total = (size_t)h1 * n2 * N3;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int k3 = e % N3;
    const size_t r = e / N3;
    const int k2 = r % n2, k1 = r / n2;
    if (k1 > sp.K1 || !tile_col(tl, k2, n2, tp_of<N3>())) continue;
    const float G = ns_group(mv, it0, k1, k2, k3, false);
    acc[e] = fma2(accg[e], G, acc[e]);
  }
}

// ------------------------------------------------------------- plane passes

// In-place FFT along k2 of rows 0..H1-1 of the plane (R2 threads per row).
template <int N2, int DIR, int PS>
__device__ __forceinline__ void plane_rows_fft(float2 *L, const float2 *tw, int tid, int nrows) {
  constexpr int R2 = FourStep<N2>::R2;
  const int row = tid / R2;
  pencil_fft<N2, DIR>(L + (row < nrows ? row : 0) * PS, tw, tid % R2, row < nrows);
}

// Analysis plane pass: out_w plane i3 = c2r_1(IFFT_2(V_w[i3])) * scale.
// out element (i1, i2, i3) at out[(i1 + N1 (i2 + N2 i3)) * os] + w * owstride.
template <int N1, int N2>
__global__ __launch_bounds__((plane_threads<N1, N2>())) void kB_inv(const float2 *__restrict__ V, size_t vstride,
                                                                    float *__restrict__ out, size_t owstride,
                                                                    size_t os, float scale,
                                                                    const float2 *__restrict__ tw1g,
                                                                    const float2 *__restrict__ tw2g, Support sp,
                                                                    int TPA) {
  constexpr int H1 = N1 / 2 + 1, PS = pencil_stride<N2>(), PP = pencil_stride<N1>(), NTH = plane_threads<N1, N2>();
  __shared__ __align__(16) float2 L[plane_slots<N1, N2>()];
  __shared__ float2 tw1[N1], tw2[N2];
  const int tid = threadIdx.x;
  const int i3 = blockIdx.x, w = blockIdx.y;
  for (int m = tid; m < N1; m += NTH) tw1[m] = tw1g[m];
  for (int m = tid; m < N2; m += NTH) tw2[m] = tw2g[m];
  // Rows k1 <= K1 and the support's columns were written by the pencil pass;
  // the rest of the plane is zero.
  const int K1 = sp.K1;
  const Tiles tl = support_tiles(sp, N2, TPA);
  const float2 *src = V + (size_t)w * vstride + (size_t)i3 * H1 * N2;
  const float4 *src4 = reinterpret_cast<const float4 *>(src);  // H1 * N2 is even: 16-byte aligned planes
  for (int e2 = tid; e2 < (K1 + 1) * N2 / 2; e2 += NTH) {
    const int e = 2 * e2, k2 = e % N2;
    float2 a = make_float2(0.0f, 0.0f), b = a;
    if (tile_col(tl, k2, N2, TPA)) {
      const float4 v = src4[e2];
      a = make_float2(v.x, v.y);
      b = make_float2(v.z, v.w);
    }
    float2 *row = L + (e / N2) * PS;
    row[k2] = a;
    row[k2 + 1] = b;
  }
  __syncthreads();
  plane_rows_fft<N2, 1, PS>(L, tw2, tid, K1 + 1);
  // c2r along k1: pair p = columns (2p, 2p+1), z = a + i b through the
  // Hermitian extension; the pair's sequence then lives pair-contiguous.
  using F = RegFFT<N1, 1>;
  constexpr int R1 = F::R1, R2 = F::R2, ROWS = F::ROWS;
  const int p = tid / R2, j = tid % R2;
  const bool on = p < N2 / 2;
  float2 v[R1];
  if (on) {
#pragma unroll
    for (int r = 0; r < R1; ++r) {
      const int t = j + R2 * r;
      const int tt = t <= N1 / 2 ? t : N1 - t;
      float2 A = make_float2(0.0f, 0.0f), B = A;
      if (tt <= K1) {
        A = L[tt * PS + 2 * p];
        B = L[tt * PS + 2 * p + 1];
      }
      if (t == 0 || t == N1 / 2) {
        A.y = 0.0f;
        B.y = 0.0f;
      } else if (t > N1 / 2) {
        A = conjf2(A);
        B = conjf2(B);
      }
      v[r] = make_float2(A.x - B.y, A.y + B.x);
    }
    F::phase1(v, j, tw1);
  }
  __syncthreads();
  float2 *pb = L + p * PP;
  if (on) {
#pragma unroll
    for (int s = 0; s < R1; ++s) pb[xpos<R1>(s * R2 + j)] = v[s];
  }
  __syncthreads();
  if (on) {
#pragma unroll
    for (int e = 0; e < R1; ++e) v[e] = pb[xpos<R1>(j * R1 + e)];
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
  if (os == 1 && (reinterpret_cast<uintptr_t>(dst) & 15) == 0) {
    float4 *d4 = reinterpret_cast<float4 *>(dst);
    const float4 *o4 = reinterpret_cast<const float4 *>(O);
    for (int e = tid; e < N1 * N2 / 4; e += NTH) d4[e] = o4[e];
  } else {
    for (int e = tid; e < N1 * N2; e += NTH) dst[(size_t)e * os] = O[e];
  }
}

// Synthesis plane pass: T_w[i3] = FFT_2(r2c_1(in_w plane i3)), T plane-major.
template <int N1, int N2>
__global__ __launch_bounds__((plane_threads<N1, N2>())) void kB_fwd(const float *__restrict__ in, size_t iwstride,
                                                                    size_t is, float2 *__restrict__ T,
                                                                    size_t tstride,
                                                                    const float2 *__restrict__ tw1g,
                                                                    const float2 *__restrict__ tw2g, Support sp,
                                                                    int TPA) {
  constexpr int H1 = N1 / 2 + 1, PS = pencil_stride<N2>(), PP = pencil_stride<N1>(), NTH = plane_threads<N1, N2>();
  __shared__ __align__(16) float2 L[plane_slots<N1, N2>()];
  __shared__ float2 tw1[N1], tw2[N2];
  const int tid = threadIdx.x;
  const int i3 = blockIdx.x, w = blockIdx.y;
  for (int m = tid; m < N1; m += NTH) tw1[m] = tw1g[m];
  for (int m = tid; m < N2; m += NTH) tw2[m] = tw2g[m];
  // rows (i2, i2 + 1) of the input plane -> pair i2 / 2, pair-contiguous.
  const float *src = in + (size_t)w * iwstride + (size_t)i3 * N1 * N2 * is;
  if (is == 1 && (reinterpret_cast<uintptr_t>(src) & 15) == 0) {
    const float4 *s4 = reinterpret_cast<const float4 *>(src);
    for (int e4 = tid; e4 < N1 * N2 / 4; e4 += NTH) {
      const int e = 4 * e4, i2 = e / N1, i1 = e % N1;
      const float4 v = s4[e4];
      float *f = reinterpret_cast<float *>(&L[(i2 >> 1) * PP + i1]) + (i2 & 1);
      f[0] = v.x;
      f[2] = v.y;
      f[4] = v.z;
      f[6] = v.w;
    }
  } else {
    for (int e = tid; e < N1 * N2; e += NTH) {
      const int i2 = e / N1, i1 = e % N1;
      float *f = reinterpret_cast<float *>(&L[(i2 >> 1) * PP + i1]);
      f[i2 & 1] = src[(size_t)e * is];
    }
  }
  __syncthreads();
  using F = RegFFT<N1, -1>;
  constexpr int R2 = F::R2;
  const int p = tid / R2, j = tid % R2;
  const bool on = p < N2 / 2;
  pencil_fft<N1, -1>(L + (on ? p : 0) * PP, tw1, j, on);
  // Unpack rows k <= K1 (the rest is outside the support):
  // A[k] = (Z[k] + conj Z[-k]) / 2, B[k] = (Z[k] - conj Z[-k]) / 2i.
  const int K1 = sp.K1;
  constexpr int KPER = (H1 + R2 - 1) / R2;
  float2 A[KPER], B[KPER];
  if (on) {
    const float2 *pb = L + p * PP;
#pragma unroll
    for (int m = 0; m < KPER; ++m) {
      const int k = j + R2 * m;
      if (k <= K1) {
        const float2 Z = pb[k];
        const float2 Zc = conjf2(pb[(N1 - k) % N1]);
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
      if (k <= K1) {
        L[k * PS + 2 * p] = A[m];
        L[k * PS + 2 * p + 1] = B[m];
      }
    }
  }
  __syncthreads();
  plane_rows_fft<N2, -1, PS>(L, tw2, tid, K1 + 1);
  const Tiles tl = support_tiles(sp, N2, TPA);
  float2 *dst = T + (size_t)w * tstride + (size_t)i3 * H1 * N2;
  float4 *dst4 = reinterpret_cast<float4 *>(dst);
  for (int e2 = tid; e2 < (K1 + 1) * N2 / 2; e2 += NTH) {
    const int e = 2 * e2, k2 = e % N2;
    if (tile_col(tl, k2, N2, TPA)) {
      const float2 *row = L + (e / N2) * PS;
      const float2 a = row[k2], b = row[k2 + 1];
      dst4[e2] = make_float4(a.x, a.y, b.x, b.y);
    }
  }
}

// ------------------------------------------------------------------ host side

namespace {

int nw_class(int nw) { return nw <= 1 ? 1 : (nw <= 2 ? 2 : (nw <= 4 ? 4 : 8)); }

template <int N3> struct PencilLaunch {
  static dim3 grid(const FastCtx &c) { return dim3(c.n2 / tp_of<N3>(), c.h1); }
  static dim3 grid(const FastCtx &c, Support sp) {
    const Tiles t = support_tiles(sp, c.n2, tp_of<N3>());
    return dim3(t.lo + t.hi, sp.K1 + 1);
  }
  static void fwd(const FastCtx &c, const float2 *T, float2 *X) {
    hipLaunchKernelGGL(kA_fwd<N3>, grid(c), dim3(sub_threads<N3>()), 0, c.stream, T, X, c.h1, c.n2, c.tw3);
    SH_LAUNCHED();
  }
  template <int NW> static void ana_nw(const FastCtx &c, const float2 *X, float2 *V, int item0, int nw, Support sp) {
    hipLaunchKernelGGL((kA_ana<N3, NW>), grid(c, sp), dim3(NW * sub_threads<N3>()), 0, c.stream, X, V, c.h1, c.n2,
                       c.mv, c.items, item0, nw, (size_t)c.H, c.tw3, sp);
    SH_LAUNCHED();
  }
  static void ana(const FastCtx &c, const float2 *X, float2 *V, int item0, int nw, int cls, Support sp) {
    switch (nw_class(cls)) {
      case 1: ana_nw<1>(c, X, V, item0, nw, sp); break;
      case 2: ana_nw<2>(c, X, V, item0, nw, sp); break;
      case 4: ana_nw<4>(c, X, V, item0, nw, sp); break;
      default: ana_nw<8>(c, X, V, item0, nw, sp); break;
    }
  }
  template <int NW>
  static void syn_nw(const FastCtx &c, const float2 *T, int nw, float2 *acc, const float2 *gin, float2 *gout,
                     float2 *V, int item0, int close, int finalize, Support sp) {
    hipLaunchKernelGGL((kA_syn<N3, NW>), finalize ? grid(c) : grid(c, sp), dim3(NW * sub_threads<N3>()), 0, c.stream,
                       T, (size_t)c.H, nw, acc, gin, gout, V, c.h1, c.n2, c.mv, c.items, item0, close, finalize,
                       c.tw3, sp);
    SH_LAUNCHED();
  }
  static void syn(const FastCtx &c, const float2 *T, int nw, float2 *acc, const float2 *gin, float2 *gout, float2 *V,
                  int item0, int close, int finalize, int cls, Support sp) {
    switch (nw_class(cls)) {
      case 1: syn_nw<1>(c, T, nw, acc, gin, gout, V, item0, close, finalize, sp); break;
      case 2: syn_nw<2>(c, T, nw, acc, gin, gout, V, item0, close, finalize, sp); break;
      case 4: syn_nw<4>(c, T, nw, acc, gin, gout, V, item0, close, finalize, sp); break;
      default: syn_nw<8>(c, T, nw, acc, gin, gout, V, item0, close, finalize, sp); break;
    }
  }
  static void inv(const FastCtx &c, const float2 *acc, float2 *V) {
    hipLaunchKernelGGL(kA_inv<N3>, grid(c), dim3(sub_threads<N3>()), 0, c.stream, acc, V, c.h1, c.n2, c.tw3);
    SH_LAUNCHED();
  }
  static void close(const FastCtx &c, float2 *acc, const float2 *accg, int item0, Support sp) {
    hipLaunchKernelGGL(kA_close<N3>, dim3(2048), dim3(256), 0, c.stream, acc, accg, c.h1, c.n2, c.mv, c.items, item0,
                       sp);
    SH_LAUNCHED();
  }
};

// Pencil tile width of the pencil pass that pairs with a plane pass (the
// support's columns come in whole tiles of it).
int tpa_of(int n3) { return n3 >= 256 ? 8 : 16; }

template <int N1, int N2> struct PlaneLaunch {
  static void inv(const FastCtx &c, const float2 *V, int nw, float *out, size_t owstride, size_t os, float scale,
                  Support sp) {
    dim3 grid(c.n3, nw);
    hipLaunchKernelGGL((kB_inv<N1, N2>), grid, dim3(plane_threads<N1, N2>()), 0, c.stream, V, (size_t)c.H, out,
                       owstride, os, scale, c.tw1, c.tw2, sp, tpa_of(c.n3));
    SH_LAUNCHED();
  }
  static void fwd(const FastCtx &c, const float *in, size_t iwstride, size_t is, int nw, float2 *T, Support sp) {
    dim3 grid(c.n3, nw);
    hipLaunchKernelGGL((kB_fwd<N1, N2>), grid, dim3(plane_threads<N1, N2>()), 0, c.stream, in, iwstride, is, T,
                       (size_t)c.H, c.tw1, c.tw2, sp, tpa_of(c.n3));
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
