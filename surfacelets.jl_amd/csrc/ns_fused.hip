// Factored NSSurfacelet kernels for grids with n1, n2, n3 in {32, 64, 80, 96,
// 112, 128} (the factorisation is described in ns_fused.hpp).
//
// Pencil passes (one 1-D transform along the group's pencil axis per block
// of TP pencils, a "tile"): the spectra X / acc are pencil-major
// [k1][k2][k3]; the slabs are plane-major, so a tile reads and writes rows
// of TP consecutive elements.  Plane passes (two 1-D transforms per block of
// one plane): a slab plane holds rows k1 <= n1/2 of the tile axis; the real
// transform along axis 1 runs on column pairs (two real columns per complex
// FFT), and rows k1 = 0 and k1 = n1/2, whose transforms enter the real
// transform through their real parts only, share one complex FFT, so a block
// needs n1/2 row slots.  Every FFT is the four-step register FFT of
// fft_reg.hpp with one padded LDS exchange.
//
// Addresses are formed as block-uniform base pointers plus 32-bit offsets.
// Every kernel asks for at least 4 waves per SIMD (at most 128 VGPRs): two
// 8-wave plane blocks or four 4-wave pencil blocks per CU, which is what
// their LDS allows; plane blocks of other sizes ask for the fewest waves that
// keep their block count (plane_waves).
//
// Bitwise reproducibility: the streaming entry points run these same
// kernels with smaller job or unit lists; the arithmetic of a subband never
// depends on which other subbands share the launch, and every sum has a
// fixed order with explicit FMAs.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fft_reg.hpp"
#include "hip_util.hpp"
#include "ns_fused.hpp"
#include "ns_tables.hpp"

namespace surfhip {

// Pencils per tile and threads per pencil block (R2 threads per pencil).
template <int N> constexpr int tp_of() { return 16; }
template <int N> constexpr int pencil_threads() { return tp_of<N>() * FourStep<N>::R2; }
// LDS slots of one pencil: the padded exchange needs N + PAD, and the
// stride spreads the eight pencils that 16 lanes touch at once over the
// banks: for the lengths with 8 lanes per pencil and two rows per thread
// (80, 96, 112, 128) a stride of 6 or 10 mod 16 (138 at 128) makes the
// loads and both sides of the exchange conflict-free; other lengths take
// the next odd stride.
template <int N> constexpr int pencil_stride() {
  constexpr int need = N + RegFFT<N, 1>::PAD;
  if constexpr (N == 128 || FourStep<N>::R1 % FourStep<N>::R2 != 0) {
    // 10 or 6 mod 16 (the lengths with 8 lanes per pencil and two rows per
    // thread): loads and both sides of the exchange conflict-free
    int p = need;
    while (p % 16 != 6 && p % 16 != 10) ++p;
    return p;
  } else {
    return need + 1;
  }
}

// Plane blocks: n1/2 row slots of the tile axis (slot 0 carries rows 0 and
// n1/2), the pair-contiguous columns of the axis-1 transform, the real output.
// A row or column pair takes N + PAD slots (the padded exchange), rounded up
// to 8 mod 16 for 8 lanes per line (136 at 128), so the two rows that 16
// lanes of an 8-byte LDS access touch fall in opposite halves of the 32
// banks.  The transposed
// layouts written between the two passes are XOR-swizzled (plane_col_swz,
// ana_row_swz) instead of padded, so every wide access of the 128 x 128
// plane kernels is free of bank conflicts.
template <int N> constexpr int plane_stride() {
  constexpr int need = N + RegFFT<N, 1>::PAD;
  if constexpr (FourStep<N>::R2 == 8) {
    int p = need;  // 8 mod 16 (see above)
    while (p % 16 != 8) ++p;
    return p;
  } else {
    return need;
  }
}
template <int N1, int NC> constexpr int plane_slots() {
  constexpr int a = (N1 / 2) * plane_stride<NC>(), b = (NC / 2) * plane_stride<N1>(), c = N1 * NC / 2;
  constexpr int m = a > b ? a : b;
  return m > c ? m : c;
}
template <int N1, int NC> constexpr int plane_threads() {
  constexpr int a = (N1 / 2) * FourStep<NC>::R2, b = (NC / 2) * FourStep<N1>::R2;
  constexpr int m = a > b ? a : b;
  return (m + 63) / 64 * 64;
}
// Waves per SIMD a plane kernel of NTH threads and LDS bytes of shared memory
// asks for: the fewest that still let a CU hold as many blocks as 4 waves per
// SIMD would (LDS permitting).  A 6-wave block fits twice either way, so it
// gets 3 waves per SIMD and 168 registers instead of 128 (no spills); 8-wave
// blocks keep 4.
template <int NTH, int LDS> constexpr int plane_waves() {
  constexpr int W = NTH / 64, lds_blocks = 160 * 1024 / LDS;
  constexpr int target = 16 / W < lds_blocks ? 16 / W : lds_blocks;
  for (int m = 1; m < 4; ++m)
    if ((4 * m / W < lds_blocks ? 4 * m / W : lds_blocks) >= target) return m;
  return 4;
}

// Synchronisation of the lanes of one wave around LDS traffic that only they
// share (the exchange of a pencil whose threads are consecutive lanes): a
// wave executes its LDS instructions in order, so ordering the compiler's
// view of memory is enough; no block barrier is needed.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A copy of v the compiler cannot see through: index math derived from it
// inside a loop is recomputed per iteration instead of being hoisted and kept
// live (in registers) across the loop.
__device__ __forceinline__ int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

// The thread index rebuilt where it is used, from the wave's first index
// (w64, wave-uniform: an SGPR) and the lane id (v_mbcnt in a volatile asm,
// so it is recomputed rather than kept live): no VGPR holds the index across
// a loop.
__device__ __forceinline__ int thread_index(int w64) {
  int l;
  asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(l));
  return w64 + l;
}
__device__ __forceinline__ int wave_base() { return __builtin_amdgcn_readfirstlane((int)threadIdx.x & ~63); }

// base[byte offset]: a block-uniform base plus a 32-bit byte offset, the
// form the scalar-base global loads and stores take.
template <class T> __device__ __forceinline__ const T &at_b(const T *base, unsigned bytes) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + bytes);
}
template <class T> __device__ __forceinline__ T &at_b(T *base, unsigned bytes) {
  return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + bytes);
}

__device__ __forceinline__ float2 fma2(float2 a, float s, float2 c) {
  return make_float2(__fmaf_rn(a.x, s, c.x), __fmaf_rn(a.y, s, c.y));
}

// Mirror X[(N - k) mod N] of element (u, q) of a pencil thread after
// RegFFT::phase2 (v[u R2 + q] = X[(j ROWS + u) + R1 q]).  The R2 threads of a
// pencil are consecutive lanes, and the mirror of element (u, q) of thread j
// is element (u, R2 - 1 - q) of thread (R1 - j ROWS - u) / ROWS (element 0 of
// thread 0 mirrors its own element (0, -q mod R2)), so one shuffle per
// component fetches it.  Every lane of the pencil must call it.
template <int N>
__device__ __forceinline__ float2 fft_mirror(const float2 (&v)[RegFFT<N, 1>::NV], int j, int u, int q) {
  constexpr int R1 = FourStep<N>::R1, R2 = FourStep<N>::R2, ROWS = RegFFT<N, 1>::ROWS;
  static_assert(R1 % R2 == 0 || (ROWS == 2 && R1 % 2 == 0), "the mirror row keeps the row index u");
  const int s = j * ROWS + u;
  const int src = s == 0 ? 0 : (R1 - s) / ROWS;
  const float2 x = v[u * R2 + (R2 - 1 - q)];
  float2 y = make_float2(__shfl(x.x, src, R2), __shfl(x.y, src, R2));
  if (u == 0 && j == 0) y = v[(R2 - q) % R2];
  return y;
}

// In-place FFT of one pencil at pb (natural order in slots 0..N-1) by its R2
// threads j; every thread of the block calls it (it holds barriers).
template <int N, int DIR>
__device__ __forceinline__ void pencil_fft(float2 *pb, const float2 *tw, int j, bool active) {
  using F = RegFFT<N, DIR>;
  constexpr int R1 = F::R1, R2 = F::R2, ROWS = F::ROWS;
  float2 v[F::NV];
  if (active) {
#pragma unroll
    for (int r = 0; r < R1; ++r) v[r] = pb[j + R2 * r];
    F::phase1(v, j, tw);
  }
  __syncthreads();
  if (active) F::store_cols(v, pb, j);
  __syncthreads();
  if (active) {
    F::load_rows(v, pb, j);
    F::phase2(v);
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int u = 0; u < ROWS; ++u)
      if (F::row_ok(j, u)) {
#pragma unroll
        for (int q = 0; q < R2; ++q) pb[(j * ROWS + u) + R1 * q] = v[u * R2 + q];
      }
  }
  __syncthreads();
}

// One wedge plane as seen from row k1: value at (k2, k3) = p[k2 s2 + k3 s3],
// or 1 everywhere (unit) for an absent plane, the lowpass or a unit wedge.
struct PlaneRef {
  const float *p;
  int s2, s3;
  bool unit;
  __device__ __forceinline__ float at(int k2, int k3) const {
    const float v = p[(unsigned)(k2 * s2 + k3 * s3)];
    return unit ? 1.0f : v;
  }
};
__device__ __forceinline__ PlaneRef plane_ref(const NSMaskView &v, int32_t off, int32_t pair, int k1, bool unit) {
  PlaneRef r{v.planes, 0, 0, true};
  if (off < 0 || unit) return r;
  r.unit = false;
  if (pair == 0) {
    r.p = v.planes + off + (size_t)k1 * v.n2;
    r.s2 = 1;
  } else if (pair == 1) {
    r.p = v.planes + off + (size_t)k1 * v.n3;
    r.s3 = 1;
  } else {
    r.p = v.planes + off;
    r.s2 = v.n3;
    r.s3 = 1;
  }
  return r;
}
__device__ __forceinline__ PlaneRef wedge_b_ref(const NSMaskView &v, const NSItem &it, int k1, bool analysis) {
  return plane_ref(v, it.pb_off, it.pb_pair, k1, it.lowpass || ns_unit_wedge(v, analysis));
}

// ------------------------------------------------------------ pencil passes
//
// A pencil block has two halves of TP R2 threads; each half transforms one
// slab's tile at a time (consecutive b-indices q, q + 1), and the tile's
// elements are shared out over all threads of the block (TileElem).

template <int NP> constexpr int half_threads() { return tp_of<NP>() * FourStep<NP>::R2; }
template <int NP> constexpr int pblock_threads() { return 2 * half_threads<NP>(); }

// Element m of a thread's share of a tile (TP pencils of length NP), chosen so
// that a warp touches contiguous memory of the pencil-major spectrum: along
// the pencil for AX3 (pencils are contiguous), across the tile for AX2.
// `off` is the element's offset inside the k1 slab of an [k1][k2][k3] array.
template <int NP, int OR, int NTH = pblock_threads<NP>()> struct TileElem {
  int pp, kp, k2, k3;
  unsigned off;
  __device__ __forceinline__ TileElem(int tid, int m, int c0, int n3) {
    constexpr int TP = tp_of<NP>();
    const int e = tid + NTH * m;
    if (OR == AX3) {
      pp = e / NP;
      kp = e % NP;
      k2 = c0 + pp;
      k3 = kp;
    } else {
      pp = e % TP;
      kp = e / TP;
      k2 = kp;
      k3 = c0 + pp;
    }
    off = (unsigned)(k2 * n3 + k3);
  }
};

// Inverse (DIR = +1) or forward FFT of the TP pencils in buf (natural order)
// by one half-block: thread (pq, j); the result stays in registers in phase-2
// order, v[uu R2 + qq] = element (j ROWS + uu) + R1 qq of pencil pq.
// WAVE: the R2 threads of each pencil are lanes of one wave, so the
// exchange needs only wave-level fences.
template <int NP, int DIR, bool WAVE = false>
__device__ __forceinline__ void tile_fft(float2 *pb, const float2 *tw, int j, bool active,
                                         float2 (&v)[RegFFT<NP, 1>::NV]) {
  using F = RegFFT<NP, DIR>;
  constexpr int R1 = F::R1, R2 = F::R2;
  if (active) {
#pragma unroll
    for (int r = 0; r < R1; ++r) v[r] = pb[j + R2 * r];
    F::phase1(v, j, tw);
  }
  if (WAVE)
    wave_sync();
  else
    __syncthreads();
  if (active) F::store_cols(v, pb, j);
  if (WAVE)
    wave_sync();
  else
    __syncthreads();
  if (active) {
    F::load_rows(v, pb, j);
    F::phase2(v);
  }
}

// Analysis pencil pass.  Unit {g, k1, c0}: Y = X G_g over the tile, then for
// each live b-index q: slab_q[i][k1][c0 + pp] = IFFT_pencil(Y w_b,q)[i].
template <int NP, int OR>
__global__ __launch_bounds__(pblock_threads<NP>(), 4) void kP_ana(const float2 *__restrict__ X,
                                                                  float2 *__restrict__ slabs, int64_t H, int n2,
                                                                  int n3, NSMaskView mv,
                                                                  const NSItem *__restrict__ items,
                                                                  const GroupDev *__restrict__ groups,
                                                                  const uint32_t *__restrict__ live,
                                                                  const int4 *__restrict__ units,
                                                                  const float2 *__restrict__ twg) {
  constexpr int TP = tp_of<NP>(), NTH = pblock_threads<NP>(), HT = half_threads<NP>(), PS = pencil_stride<NP>();
  using F = RegFFT<NP, 1>;
  constexpr int R1 = F::R1, R2 = F::R2, ROWS = F::ROWS, E = TP * NP / NTH;
  static_assert(E * NTH == TP * NP, "a tile's elements share out evenly");
  __shared__ float2 buf[2][TP * PS];
  __shared__ float2 tw[NP];
  const int tid = threadIdx.x;
  const int4 un = units[blockIdx.x];
  const int g = un.x, k1 = un.y, c0 = un.z, tile = c0 / TP;
  const GroupDev gd = groups[g];
  const int h1 = mv.h1, nt = OR == AX3 ? n2 : n3;
  RegFFT<NP, 1>::fill_twiddles(tw, twg, tid, NTH);
  const size_t kb = (size_t)k1 * n2 * n3;
  const float2 *Xk = X + kb;
  const float *Gk = ns_group_table(mv, g, true) + kb;
  float2 y[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const TileElem<NP, OR> t(tid, m, c0, n3);
    y[m] = cscale(Xk[t.off], Gk[t.off]);
  }
  // The R2 lanes of a pencil lie in one wave (see kP_syn).
  constexpr int PL = 64 / R2;
  static_assert(HT % 64 == 0 && TP % PL == 0, "pencils of a half fill whole waves");
  const int half = __builtin_amdgcn_readfirstlane(tid / HT), lt = tid % HT;
  const int j = (lt / PL) % R2, pq = lt % PL + PL * (lt / 64);
  const int rs = h1 * nt;
  for (int q0 = 0; q0 < gd.nB; q0 += 2) {
    bool lv[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int q = q0 + h;
      lv[h] = q < gd.nB && ((live[(size_t)(gd.slab0 + q) * h1 + k1] >> tile) & 1u);
      if (!lv[h]) continue;
      const PlaneRef wb = wedge_b_ref(mv, items[gd.first + q], k1, true);
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const TileElem<NP, OR> t(tid, m, c0, n3);
        buf[h][t.pp * PS + t.kp] = cscale(y[m], wb.at(t.k2, t.k3));
      }
    }
    __syncthreads();
    const bool act = lv[half];
    float2 v[F::NV];
    tile_fft<NP, 1, true>(buf[half] + pq * PS, tw, j, act, v);
    if (act) {
      float2 *dst = slabs + (size_t)(gd.slab0 + q0 + half) * H + (size_t)k1 * nt + c0;
#pragma unroll
      for (int uu = 0; uu < ROWS; ++uu)
        if (F::row_ok(j, uu)) {
#pragma unroll
          for (int qq = 0; qq < R2; ++qq)
            dst[(unsigned)(((j * ROWS + uu) + R1 * qq) * rs + pq)] = v[uu * R2 + qq];
        }
    }
    __syncthreads();
  }
}

// Synthesis pencil pass over one tile (k1, c0) of orientation OR:
//   a = acc (AX3 with has_acc) or 0;
//   for each slab s of slist, in order, whose tile is live:
//     a = fma(FFT_pencil(slab_s), G_g(s) w_b(s), a)
// AX2 writes a into acc; AX3 writes IFFT_3(a) into V ([i3][k1][k2]).  The
// NQ quarters of the block transform NQ consecutive slabs of the list at a
// time; the sum still runs over them in list order.
template <int NP> constexpr int syn_quarters() { return 4; }
template <int NP> constexpr int syn_threads() { return syn_quarters<NP>() * half_threads<NP>(); }

template <int NP, int OR>
__global__ __launch_bounds__(syn_threads<NP>(), 4) void kP_syn(const float2 *__restrict__ slabs, int64_t H,
                                                               int n2, int n3, NSMaskView mv,
                                                               const NSItem *__restrict__ items,
                                                               const uint32_t *__restrict__ live,
                                                               const int4 *__restrict__ slist, int ns,
                                                               float2 *__restrict__ acc, int has_acc,
                                                               float2 *__restrict__ V,
                                                               const float2 *__restrict__ twg) {
  constexpr int TP = tp_of<NP>(), NQ = syn_quarters<NP>(), NTH = syn_threads<NP>(), QT = half_threads<NP>();
  constexpr int PS = pencil_stride<NP>();
  using F = RegFFT<NP, -1>;
  constexpr int R1 = F::R1, R2 = F::R2, ROWS = F::ROWS, E = (TP * NP + NTH - 1) / NTH;
  // Whether element m of this thread's share lies in the tile (the last
  // element of some threads does not when NTH does not divide TP NP).
  auto in_tile = [&](int m) { return E * NTH == TP * NP || m < E - 1 || (int)threadIdx.x + NTH * m < TP * NP; };
  using TE = TileElem<NP, OR, NTH>;
  __shared__ float2 buf[NQ][TP * PS];
  __shared__ float2 tw[NP];
  const int tid = threadIdx.x;
  const int h1 = mv.h1, nt = OR == AX3 ? n2 : n3, ntiles = nt / TP;
  const int k1 = blockIdx.x / ntiles, tile = blockIdx.x % ntiles, c0 = tile * TP;
  RegFFT<NP, 1>::fill_twiddles(tw, twg, tid, NTH);
  const size_t kb = (size_t)k1 * n2 * n3;
  float2 a[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const TE t(tid, m, c0, n3);
    a[m] = (OR == AX3 && has_acc && in_tile(m)) ? acc[kb + t.off] : make_float2(0.0f, 0.0f);
  }
  __syncthreads();
  // Thread lt of a quarter: pencil pq = pq_lo + PL pq_hi, lane j; the R2
  // lanes of a pencil lie in one wave (PL lanes apart), so its exchange needs
  // only a wave-level fence, and PL consecutive lanes read PL consecutive
  // pencils.
  constexpr int PL = 64 / R2;
  static_assert(QT % 64 == 0 && TP % PL == 0, "pencils of a quarter fill whole waves");
  // A quarter is whole waves, so `quarter` is wave-uniform: the slab list and
  // live bits it indexes are read with scalar loads.
  const int quarter = __builtin_amdgcn_readfirstlane(tid / QT), lt = tid % QT;
  const int j = (lt / PL) % R2, pq = lt % PL + PL * (lt / 64);
  const int rs = h1 * nt;
  // G of the group last used (slabs of one group are adjacent in the list and
  // share it).
  int gcur = -1;
  float gv[E];
  // The first step (NQ consecutive slabs of the list) at or after s with a
  // live slab in this tile, and its live bits; ns when there is none.
  auto seek = [&](int s, bool(&lv)[NQ]) {
    for (; s < ns; s += NQ) {
      bool any = false;
#pragma unroll
      for (int h = 0; h < NQ; ++h) {
        lv[h] = s + h < ns && ((live[(size_t)slist[s + h].x * h1 + k1] >> tile) & 1u);
        any = any || lv[h];
      }
      if (any) return s;
    }
    return ns;
  };
  // The pencils of this quarter's slab of step s.
  auto load_step = [&](int s, float2(&v)[F::NV]) {
    const float2 *src = slabs + (size_t)slist[s + quarter].x * H + (size_t)k1 * nt + c0;
#pragma unroll
    for (int r = 0; r < R1; ++r) v[r] = src[(unsigned)((j + R2 * r) * rs + pq)];
  };
  // Software pipeline over the steps: v is free once a step's transform sits
  // in LDS, so the next step's pencils are requested before this step's mask
  // values, and their latency and the mask loads' share one wait.
  bool lv[NQ];
  int s0 = seek(0, lv);
  float2 v[F::NV];
  if (s0 < ns && lv[quarter]) load_step(s0, v);
  while (s0 < ns) {
    const bool act = lv[quarter];
    float2 *pb = buf[quarter] + pq * PS;
    if (act) {
      F::phase1(v, j, tw);
      F::store_cols(v, pb, j);
    }
    wave_sync();  // the exchange region of a pencil is private to its R2 lanes
    if (act) {
      F::load_rows(v, pb, j);
      F::phase2(v);
    }
    wave_sync();
    if (act) {
#pragma unroll
      for (int uu = 0; uu < ROWS; ++uu)
        if (F::row_ok(j, uu)) {
#pragma unroll
          for (int qq = 0; qq < R2; ++qq) pb[(j * ROWS + uu) + R1 * qq] = v[uu * R2 + qq];
        }
    }
    __syncthreads();
    bool ln[NQ];
    const int s1 = seek(s0 + NQ, ln);
    if (s1 < ns && ln[quarter]) load_step(s1, v);
#pragma unroll
    for (int h = 0; h < NQ; ++h) {
      if (!lv[h]) continue;
      const int4 sl = slist[s0 + h];
      if (sl.y != gcur) {
        const float *Gk = ns_group_table(mv, sl.y, false) + kb;
#pragma unroll
        for (int m = 0; m < E; ++m) {
          const TE t(tid, m, c0, n3);
          gv[m] = in_tile(m) ? Gk[t.off] : 0.0f;
        }
        gcur = sl.y;
      }
      const PlaneRef wb = wedge_b_ref(mv, items[sl.z], k1, false);
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const TE t(tid, m, c0, n3);
        if (in_tile(m)) a[m] = fma2(buf[h][t.pp * PS + t.kp], gv[m] * wb.at(t.k2, t.k3), a[m]);
      }
    }
    __syncthreads();
    s0 = s1;
#pragma unroll
    for (int h = 0; h < NQ; ++h) lv[h] = ln[h];
  }
  if (OR == AX2) {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const TE t(tid, m, c0, n3);
      if (in_tile(m)) acc[kb + t.off] = a[m];
    }
    return;
  }
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const TE t(tid, m, c0, n3);
    if (in_tile(m)) buf[0][t.pp * PS + t.kp] = a[m];
  }
  __syncthreads();
  tile_fft<NP, 1, true>(buf[0] + pq * PS, tw, j, quarter == 0, v);
  if (quarter == 0) {
    float2 *dst = V + (size_t)k1 * n2 + c0;
    const int vs = h1 * n2;
#pragma unroll
    for (int uu = 0; uu < ROWS; ++uu)
      if (F::row_ok(j, uu)) {
#pragma unroll
        for (int qq = 0; qq < R2; ++qq)
          dst[(unsigned)(((j * ROWS + uu) + R1 * qq) * vs + pq)] = v[uu * R2 + qq];
      }
  }
}

// X[k1][k2][:] = FFT_3(T[:][k1][k2]) (T plane-major [i3][k1][k2]).
template <int NP>
__global__ __launch_bounds__(pencil_threads<NP>()) void kP_fwd(const float2 *__restrict__ T, float2 *__restrict__ X,
                                                               int h1, int n2, const float2 *__restrict__ twg) {
  constexpr int TP = tp_of<NP>(), NTH = pencil_threads<NP>(), PS = pencil_stride<NP>();
  using F = RegFFT<NP, -1>;
  constexpr int R1 = F::R1, R2 = F::R2, ROWS = F::ROWS, E = TP * NP / NTH;
  static_assert(E * NTH == TP * NP, "a tile's elements share out evenly");
  __shared__ float2 buf[TP * PS];
  __shared__ float2 tw[NP];
  const int tid = threadIdx.x, ntiles = n2 / TP;
  const int k1 = blockIdx.x / ntiles, c0 = (blockIdx.x % ntiles) * TP;
  RegFFT<NP, 1>::fill_twiddles(tw, twg, tid, NTH);
  __syncthreads();
  const int pq = tid % TP, j = tid / TP;
  const float2 *src = T + (size_t)k1 * n2 + c0;
  const int rs = h1 * n2;
  float2 v[F::NV];
#pragma unroll
  for (int r = 0; r < R1; ++r) v[r] = src[(unsigned)((j + R2 * r) * rs + pq)];
  F::phase1(v, j, tw);
  float2 *pb = buf + pq * PS;
  F::store_cols(v, pb, j);
  __syncthreads();
  F::load_rows(v, pb, j);
  F::phase2(v);
  __syncthreads();
#pragma unroll
  for (int uu = 0; uu < ROWS; ++uu)
    if (F::row_ok(j, uu)) {
#pragma unroll
      for (int qq = 0; qq < R2; ++qq) pb[(j * ROWS + uu) + R1 * qq] = v[uu * R2 + qq];
    }
  __syncthreads();
  float2 *dst = X + ((size_t)k1 * n2 + c0) * NP;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int e = tid + NTH * m, pp = e / NP, kp = e % NP;
    dst[e] = buf[pp * PS + kp];
  }
}

// ------------------------------------------------------------- plane passes
//
// Plane ip of a slab of orientation AX3 is i3 = ip with columns k2 (NC = n2);
// of AX2, i2 = ip with columns k3 (NC = n3).  In the volume, column ic of the
// plane starts at pbase + ic cstr: AX3 pbase = ip n1 n2, cstr = n1; AX2
// pbase = ip n1, cstr = n1 n2.  Plane a of a group is [k1][kc] in the pool
// for both orientations (pair (1,2) for AX3, (1,3) for AX2).

// Analysis plane pass, block (ip, job): for every a-index p in [p0, p1), the
// plane of subband out0 + (p - p0) out_step = scale * c2r_1(IFFT_c(U w_a,p)),
// U the slab plane.  Row k1 of the plane is transformed by the R2 threads of
// slot k1 (slot 0 carries rows 0 and n1/2, see below); the results go to LDS
// column-major (column stride CS) for the transform along k1, whose outputs
// are stored straight from registers.
template <int NC> constexpr int col_stride() { return NC >= 64 ? 68 : 36; }
// Analysis: row k1 < n1/2 of column kc lives at slot k1 ^ ana_row_swz(kc) of
// the column (row n1/2 unswizzled), so the eight columns that 16 lanes write
// at once spread over all banks.  kc and kc ^ 1 share the swizzle.
__device__ __forceinline__ int ana_row_swz(int kc) { return 2 * ((kc >> 2) & 3); }
// Synthesis: column kc of row k1 lives at slot kc ^ plane_col_swz(k1) of the
// row, so the eight rows that 16 lanes write at once spread over all banks.
// The swizzle is even: columns 2 pr and 2 pr + 1 stay adjacent.
__device__ __forceinline__ int plane_col_swz(int k1) { return 2 * ((k1 >> 1) & 7); }
template <int N1, int NC> constexpr int ana_slots() {
  constexpr int a = (N1 / 2) * plane_stride<NC>(), b = NC * col_stride<N1>(), c = (NC / 2) * plane_stride<N1>();
  constexpr int m = a > b ? a : b;
  return m > c ? m : c;
}
template <int N1, int NC> constexpr int ana_lds() {
  return 8 * ana_slots<N1, NC>() + 8 * (N1 + 2 * NC) + 4 * (N1 / 2 + 1);
}
template <int N1, int NC> constexpr int syn_lds() {
  return 8 * plane_slots<N1, NC>() + 8 * (N1 + 4 * NC) + 4 * (N1 / 2 + 1);
}
template <int N1, int NC> constexpr int ana_threads() {
  constexpr int a = (N1 / 2) * FourStep<NC>::R2, b = (NC / 2) * FourStep<N1>::R2;
  constexpr int m = a > b ? a : b;
  return (m + 63) / 64 * 64;
}

template <int N1, int NC, int OR>
__global__ __launch_bounds__((ana_threads<N1, NC>()), (plane_waves<ana_threads<N1, NC>(), ana_lds<N1, NC>()>())) void kQ_ana(
    const float2 *__restrict__ slabs, int64_t H, int n1n2, NSMaskView mv, const uint32_t *__restrict__ live,
    const PlaneJob *__restrict__ jobs, float *__restrict__ out, int64_t substride, int os, float scale,
    const float2 *__restrict__ tw1g, const float2 *__restrict__ twcg) {
  constexpr int tpw = tp_of<NC>();  // columns per live-tile bit
  constexpr int H1 = N1 / 2 + 1, HS = N1 / 2, PS = plane_stride<NC>(), PP = plane_stride<N1>();
  constexpr int CS = col_stride<N1>();
  constexpr int NTH = ana_threads<N1, NC>();
  using FC = RegFFT<NC, 1>;
  constexpr int RC1 = FC::R1, RC2 = FC::R2, RCR = FC::ROWS;
  using F = RegFFT<N1, 1>;
  constexpr int R1 = F::R1, R2 = F::R2, ROWS = F::ROWS;
  static_assert(R1 % 2 == 0 && R1 * R2 == 2 * HS, "row HS sits at element R1 / 2 of thread 0");
  __shared__ __align__(16) float2 L[ana_slots<N1, NC>()];
  __shared__ float2 tw1[N1], twc[NC], UW[NC];
  __shared__ uint32_t rl[H1];
  const int ip = blockIdx.x;
  const PlaneJob jb = jobs[blockIdx.y];
  const bool masked = jb.slab >= 0 && jb.group >= 0;
  const bool wunit = !masked || jb.pa0 < 0 || ns_unit_wedge(mv, true);
  const float2 *Ub = slabs + (size_t)(jb.slab < 0 ? 0 : jb.slab) * H + (size_t)ip * H1 * NC;
  const int pbase = OR == AX3 ? ip * n1n2 : ip * N1;
  const int cstr = OR == AX3 ? N1 : n1n2;
  // The slab plane is the same for every subband of the block (only plane a
  // changes): each thread loads its share of its row once and keeps it in
  // registers across the subbands; dead tiles are selected away below.
  // Row HS, scaled by its wedge-plane row per subband, goes to UW (thread
  // kc < NC takes column kc) for slot 0, which transforms z = row 0 + i row
  // HS: the inverse along kc of each row is real (their spectra are
  // Hermitian in kc), so the real part is row 0's result and the imaginary
  // part row HS's.
  // Only live tiles are read: the pencil pass writes nothing else, and the
  // row pass drops dead tiles whatever they hold.
  // When the row slots of a wave are exactly the column pairs of its column
  // pass (same stride, same lanes per line), a wave touches no LDS of another
  // wave between the barrier after the column pass's transposed reads and
  // the row pass's transposed writes, and row HS (UW) is staged by wave 0,
  // which alone reads it: no block barrier starts a subband but the first.
  constexpr bool OWN = PS == PP && RC2 == R2 && (NC / 2) * R2 == NTH && NC <= 128;
  constexpr int UHN = OWN ? (NC + 63) / 64 : 1;  // row HS columns per wave-0 lane
  float2 u[RC1];
  float2 uh[UHN];
#pragma unroll
  for (int h = 0; h < UHN; ++h) uh[h] = make_float2(0.0f, 0.0f);
  {
    const int tid = (int)threadIdx.x;
    const int slot = tid / RC2, j = tid % RC2;
    const int row = slot < HS ? slot : 0;
    const uint32_t lw = slot >= HS ? 0u : masked ? live[(size_t)jb.slab * H1 + row] : 0xffffffffu;
#pragma unroll
    for (int r = 0; r < RC1; ++r) {
      u[r] = make_float2(0.0f, 0.0f);
      if ((lw >> ((j + RC2 * r) / tpw)) & 1u) u[r] = Ub[(unsigned)(row * NC + j + RC2 * r)];
    }
    static_assert(NTH >= NC, "one thread per column of row HS");
    const uint32_t lh = masked ? live[(size_t)jb.slab * H1 + HS] : 0xffffffffu;
    if constexpr (OWN) {
#pragma unroll
      for (int h = 0; h < UHN; ++h) {
        const int kc = tid + 64 * h;
        if (tid < 64 && kc < NC && ((lh >> (kc / tpw)) & 1u)) uh[h] = Ub[(unsigned)(HS * NC + kc)];
      }
    } else {
      if (tid < NC && ((lh >> (tid / tpw)) & 1u)) uh[0] = Ub[(unsigned)(HS * NC + tid)];
    }
  }
  for (int p = jb.p0; p < jb.p1; ++p) {
    // Index math is redone per subband from an opaque copy of the thread
    // index, so that little but the loop counter and the slab row is live
    // across the loop.
    const int tid = opaque((int)threadIdx.x);
    const int slot = tid / RC2, j = tid % RC2;
    const bool onslot = slot < HS;
    const int row = onslot ? slot : 0;
    const float *wab = mv.planes + (wunit ? 0 : jb.pa0 + (p - jb.p0) * jb.pa_step);
    // Block-uniform branch, one offset per thread and immediate steps (no
    // per-element selects or 64-bit address arithmetic).
    float wa[RC1];
    if (wunit) {
#pragma unroll
      for (int r = 0; r < RC1; ++r) wa[r] = 1.0f;
    } else {
      const unsigned o0 = 4u * (unsigned)(row * NC + j);
#pragma unroll
      for (int r = 0; r < RC1; ++r) wa[r] = at_b(wab, o0 + 4u * RC2 * r);
    }
    if constexpr (OWN) {
#pragma unroll
      for (int h = 0; h < UHN; ++h) {
        const int kc = tid + 64 * h;
        if (tid < 64 && kc < NC) UW[kc] = wunit ? uh[h] : cscale(uh[h], wab[(unsigned)(HS * NC + kc)]);
      }
    } else {
      if (tid < NC) UW[tid] = wunit ? uh[0] : cscale(uh[0], wab[(unsigned)(HS * NC + tid)]);
    }
    if (p == jb.p0) {
      RegFFT<N1, 1>::fill_twiddles(tw1, tw1g, tid, NTH);
      RegFFT<NC, 1>::fill_twiddles(twc, twcg, tid, NTH);
      for (int k = tid; k < H1; k += NTH) rl[k] = masked ? live[(size_t)jb.slab * H1 + k] : 0xffffffffu;
    }
    if (!OWN || p == jb.p0)
      __syncthreads();  // also: the previous subband is done with L
    else
      wave_sync();
    const uint32_t lrow = onslot ? rl[slot] : 0u, lhs = slot == 0 ? rl[HS] : 0u;
    const bool act = (lrow | lhs) != 0u;
    float2 *pb = L + row * PS;
    {
      float2 v[FC::NV];
      if (act) {
        // Dead tiles hold +0: their slab values were never loaded (u = 0, uh
        // = 0) and the wedge values are finite and nonnegative.
        if (wunit) {
#pragma unroll
          for (int r = 0; r < RC1; ++r) v[r] = u[r];
        } else {
#pragma unroll
          for (int r = 0; r < RC1; ++r) v[r] = cscale(u[r], wa[r]);
        }
        if (slot == 0) {
#pragma unroll
          for (int r = 0; r < RC1; ++r) {
            const float2 b = UW[j + RC2 * r];
            v[r] = make_float2(v[r].x - b.y, v[r].y + b.x);
          }
        }
        FC::phase1(v, j, twc);
        FC::store_cols(v, pb, j);
      }
      wave_sync();  // the exchange region of a row is private to its R2 lanes
      if (act) {
        FC::load_rows(v, pb, j);
        FC::phase2(v);
      }
      __syncthreads();
      if (onslot) {
        // Rows without live tiles hold zeros; rows 0 and HS are real.
#pragma unroll
        for (int uu = 0; uu < RCR; ++uu)
#pragma unroll
          for (int qq = 0; qq < RC2; ++qq) {
            if (!FC::row_ok(j, uu)) continue;
            const float2 z = act ? v[uu * RC2 + qq] : make_float2(0.0f, 0.0f);
            const int kc = (j * RCR + uu) + RC1 * qq;
            float2 *lc = L + kc * CS;
            if (slot == 0) {
              lc[ana_row_swz(kc)] = make_float2(z.x, 0.0f);
              lc[HS] = make_float2(z.y, 0.0f);
            } else {
              lc[slot ^ ana_row_swz(kc)] = z;
            }
          }
      }
      __syncthreads();
    }
    // c2r along k1 on column pairs (2 pr, 2 pr + 1): z = a + i b through the
    // Hermitian extension; rows that hold no data are zero.
    const int pr = tid / R2, jj = tid % R2;
    const bool on = pr < NC / 2;
    float2 w[F::NV];
    if (on) {
      const float2 *ca = L + (2 * pr) * CS, *cb = ca + CS;
      const int gz = ana_row_swz(2 * pr);
#pragma unroll
      for (int r = 0; r < R1; ++r) {
        // Row t = jj + R2 r: below HS = R1 R2 / 2 for r < R1 / 2, above it
        // (the mirror N1 - t, conjugated) for r > R1 / 2; r = R1 / 2 holds
        // HS on thread jj = 0 and mirrors elsewhere.  Rows 0 and HS are real.
        const int t = jj + R2 * r;
        const bool mir = 2 * r > R1 || (2 * r == R1 && jj > 0);
        const bool real = (r == 0 || 2 * r == R1) && jj == 0;
        const int k1 = mir ? N1 - t : t;
        const int sk = 2 * r == R1 && jj == 0 ? HS : k1 ^ gz;  // row HS is not swizzled
        float2 A = ca[sk], B = cb[sk];
        if (mir) {
          A = conjf2(A);
          B = conjf2(B);
        }
        if (real) {
          A.y = 0.0f;
          B.y = 0.0f;
        }
        w[r] = make_float2(A.x - B.y, A.y + B.x);
      }
      F::phase1(w, jj, tw1);
    }
    __syncthreads();
    float2 *qb = L + (on ? pr : 0) * PP;
    if (on) F::store_cols(w, qb, jj);
    wave_sync();  // the exchange region of a column pair is private to its R2 lanes
    if (on) {
      F::load_rows(w, qb, jj);
      F::phase2(w);
      // Thread (pr, jj) holds columns 2 pr (.x) and 2 pr + 1 (.y) at
      // i1 = (jj ROWS + u) + R1 q.
      float *dst = out + (int64_t)(jb.out0 + (p - jb.p0) * jb.out_step) * substride + (int64_t)pbase * os;
      if (!F::row_ok(jj, 0)) {
        // no rows on this thread (R2 does not divide R1)
      } else if (ROWS == 2 && R1 % 2 == 0 && os == 1 && (reinterpret_cast<uintptr_t>(dst) & 7) == 0 &&
                 (cstr & 1) == 0) {
        // Block-uniform base, 32-bit byte offsets per thread.
        float2 *db = reinterpret_cast<float2 *>(dst);
        const unsigned o0 = 4u * (unsigned)((2 * pr) * cstr + jj * 2), o1 = o0 + 4u * (unsigned)cstr;
#pragma unroll
        for (int q = 0; q < R2; ++q) {
          const float2 z0 = w[q], z1 = w[R2 + q];
          at_b(db, o0 + 4u * R1 * q) = make_float2(z0.x * scale, z1.x * scale);
          at_b(db, o1 + 4u * R1 * q) = make_float2(z0.y * scale, z1.y * scale);
        }
      } else {
#pragma unroll
        for (int uu = 0; uu < ROWS; ++uu)
#pragma unroll
          for (int q = 0; q < R2; ++q) {
            if (!F::row_ok(jj, uu)) continue;
            const int i1 = (jj * ROWS + uu) + R1 * q;
            const float2 z = w[uu * R2 + q];
            dst[(int64_t)((2 * pr) * cstr + i1) * os] = z.x * scale;
            dst[(int64_t)((2 * pr + 1) * cstr + i1) * os] = z.y * scale;
          }
      }
    }
  }
}

// Element e = tid + NTH m of a real plane as column pairs: pair e / N1, row
// i1 = e % N1 of columns 2 pair (a) and 2 pair + 1 (b); a warp reads
// consecutive i1 of both columns.  sp: the plane's first sample.  NTH is a
// multiple of N1, so element m of a thread is its element 0 moved by the
// block-uniform m (NTH / N1) column pairs: one 32-bit offset per thread.
template <int N1, int NC, int NTH, int NLD>
__device__ __forceinline__ void plane_pairs_load(const float *__restrict__ sp, int tid, int cstr, int is,
                                                 float (&a)[NLD], float (&b)[NLD]) {
  static_assert(NTH % N1 == 0 && NLD * NTH == N1 * NC / 2, "plane load shape");
  const unsigned o = (unsigned)(((tid / N1) * 2 * cstr + tid % N1) * is) * 4u;
  const unsigned ob = o + (unsigned)(cstr * is) * 4u;
  const int64_t step = (int64_t)(NTH / N1) * 2 * cstr * is;
#pragma unroll
  for (int m = 0; m < NLD; ++m) {
    const float *spm = sp + m * step;
    a[m] = at_b(spm, o);
    b[m] = at_b(spm, ob);
  }
}

// Synthesis plane pass, block (ip, job): R = (init ? 0 : slab plane), then for
// p in [p0, p1): R = fma(FFT_c(r2c_1(in_p plane)), w_a,p, R); R -> slab plane
// (rows k1 <= n1/2, live tiles).  The transform along axis 1 runs on column
// pairs and is unpacked in registers (the mirror Z[-k] comes from a lane of
// the same pencil).  Rows 0 and n1/2 of r2c_1 are real, so slot 0 transforms
// z = row 0 + i row n1/2 and only parks the result Z in LDS (ZR); the split
// F(row 0) = (Z[kc] + conj Z[-kc]) / 2, F(row n1/2) = (Z[kc] - conj Z[-kc]) / 2i
// and the two masked sums (R0, R64, in LDS, natural column order) are done
// by threads kc < NC, one column each, at the start of the next subband
// while its plane loads are in flight (and once after the last subband).
// Slot 0's own lanes therefore do no more work than any other row's.
template <int N1, int NC, int OR>
__global__ __launch_bounds__((plane_threads<N1, NC>()), (plane_waves<plane_threads<N1, NC>(), syn_lds<N1, NC>()>())) void kQ_syn(
    const float *__restrict__ in, int64_t substride, int is, int n1n2, NSMaskView mv,
    const uint32_t *__restrict__ live, const PlaneJob *__restrict__ jobs, float2 *__restrict__ slabs, int64_t H,
    const float2 *__restrict__ tw1g, const float2 *__restrict__ twcg) {
  constexpr int tpw = tp_of<NC>();  // columns per live-tile bit
  constexpr int H1 = N1 / 2 + 1, HS = N1 / 2, PS = plane_stride<NC>(), PP = plane_stride<N1>();
  constexpr int NTH = plane_threads<N1, NC>();
  using FR = RegFFT<NC, -1>;  // along the tile axis
  using F1 = RegFFT<N1, -1>;  // along k1
  constexpr int RC1 = FR::R1, RC2 = FR::R2, RCR = FR::ROWS;
  constexpr int C1 = F1::R1, C2 = F1::R2, CR = F1::ROWS;
  static_assert(C2 % 2 == 0 && C1 * C2 == 2 * HS, "row HS sits at element (0, C2 / 2) of thread 0");
  static_assert(NTH >= NC, "one thread per column of rows 0 and HS");
  constexpr int NLD = (N1 * NC / 2 + NTH - 1) / NTH;  // column-pair elements per thread
  // Whether every column-pass lane is a thread (NTH = NC / 2 pairs x C2) and
  // a pair's rows split into whole 64-row runs: then each wave stages only
  // the pairs it transforms.
  constexpr bool WAVEIN = NTH == (NC / 2) * C2 && N1 % 64 == 0 && 64 % C2 == 0;
  __shared__ __align__(16) float2 L[plane_slots<N1, NC>()];
  __shared__ float2 tw1[N1], twc[NC], R0[NC], R64[NC], ZR[NC];
  __shared__ uint32_t rl[H1];
  const int tid = threadIdx.x, ip = blockIdx.x;
  const PlaneJob jb = jobs[blockIdx.y];
  const bool masked = jb.group >= 0;
  RegFFT<N1, 1>::fill_twiddles(tw1, tw1g, tid, NTH);
  RegFFT<NC, 1>::fill_twiddles(twc, twcg, tid, NTH);
  for (int k = tid; k < H1; k += NTH) rl[k] = masked ? live[(size_t)jb.slab * H1 + k] : 0xffffffffu;
  __syncthreads();
  const int slot = tid / RC2, j = tid % RC2;
  const bool onslot = slot < HS;
  // Slot 0 runs (block-uniform) when row 0 or row HS has a live tile.
  const bool act0 = (rl[0] | rl[HS]) != 0u;
  const uint32_t lrow = onslot ? (slot == 0 ? (act0 ? 1u : 0u) : rl[slot]) : 0u;
  const bool act = lrow != 0u;
  float2 *Sb = slabs + (size_t)jb.slab * H + (size_t)ip * H1 * NC;
  // Running sums of rows 1 .. HS - 1 in the row FFT's output order: element
  // (uu, qq) of thread (slot, j) is column kc = (j RCR + uu) + RC1 qq.  Rows 0
  // and HS sum in R0 and R64 (thread kc < NC owns column kc).
  float2 R[FR::NV];
#pragma unroll
  for (int uu = 0; uu < RCR; ++uu)
#pragma unroll
    for (int qq = 0; qq < RC2; ++qq) {
      const int kc = (j * RCR + uu) + RC1 * qq, m = uu * RC2 + qq;
      R[m] = make_float2(0.0f, 0.0f);
      if (!FR::row_ok(j, uu) || slot == 0) continue;
      if (act && !jb.init && ((rl[slot] >> (kc / tpw)) & 1u)) R[m] = Sb[(unsigned)(slot * NC + kc)];
    }
  if (act0 && tid < NC) {
    const bool ld = !jb.init;
    R0[tid] = ld && ((rl[0] >> (tid / tpw)) & 1u) ? Sb[(unsigned)tid] : make_float2(0.0f, 0.0f);
    R64[tid] = ld && ((rl[HS] >> (tid / tpw)) & 1u) ? Sb[(unsigned)(HS * NC + tid)] : make_float2(0.0f, 0.0f);
  }
  const int pbase = OR == AX3 ? ip * n1n2 : ip * N1;
  const int cstr = OR == AX3 ? N1 : n1n2;
  const bool wunit = !masked || jb.pa0 < 0 || ns_unit_wedge(mv, false);
  const int w64 = wave_base();
  // Rows 0 and HS of subband pp from ZR (written by slot 0 before the barrier
  // that closed pp's iteration; rewritten only two barriers into the next),
  // with w0, w1 = plane a of pp at (0, kc) and (HS, kc).
  auto split_w = [&](int pp, int kc, float &w0, float &w1) {
    const float *wa = mv.planes + (wunit ? 0 : jb.pa0 + (pp - jb.p0) * jb.pa_step);
    w0 = wunit ? 1.0f : wa[kc];
    w1 = wunit ? 1.0f : wa[HS * NC + kc];
  };
  auto split0 = [&](int kc, float w0, float w1) {
    const float2 Z = ZR[kc], Zc = conjf2(ZR[(NC - kc) % NC]);
    const float2 F0 = cscale(cadd(Z, Zc), 0.5f);
    const float2 d = csub(Z, Zc);
    const float2 F1 = make_float2(0.5f * d.y, -0.5f * d.x);
    R0[kc] = fma2(F0, w0, R0[kc]);
    R64[kc] = fma2(F1, w1, R64[kc]);
  };
  for (int p = jb.p0; p < jb.p1; ++p) {
    const int t = thread_index(w64);
    // The split of the previous subband: its wedge values are loaded ahead of
    // the plane, the split itself runs once the plane is in LDS.
    const bool sp = act0 && p > jb.p0 && t < NC;
    float w0 = 1.0f, w1 = 1.0f;
    if (sp) split_w(p - 1, t, w0, w1);
    {
      // Columns (2 pr, 2 pr + 1) of the real plane -> pair pr, pair-contiguous.
      const float *sp = in + (int64_t)(jb.out0 + (p - jb.p0) * jb.out_step) * substride + (int64_t)pbase * is;
      if constexpr (WAVEIN) {
        // Each wave loads the column pairs its own lanes transform (64 / C2
        // pairs, one 64-row run of one pair per load), so the staging in L
        // needs only a wave-level fence, not a block barrier.
        float a[C1], b[C1];
        const int lane = t - w64, pw = w64 / C2;
#pragma unroll
        for (int m = 0; m < C1; ++m) {
          const int pr = pw + (64 * m) / N1, r0 = (64 * m) % N1;
          const float *col = sp + (int64_t)(2 * pr * cstr + r0) * is;
          a[m] = at_b(col, 4u * (unsigned)(lane * is));
          b[m] = at_b(col, 4u * (unsigned)((cstr + lane) * is));
        }
#pragma unroll
        for (int m = 0; m < C1; ++m) L[(pw + (64 * m) / N1) * PP + (64 * m) % N1 + lane] = make_float2(a[m], b[m]);
      } else if constexpr (NTH % N1 == 0 && (N1 * NC / 2) % NTH == 0) {
        float a[NLD], b[NLD];
        plane_pairs_load<N1, NC, NTH, NLD>(sp, t, cstr, is, a, b);
        float2 *lp = L + (t / N1) * PP + t % N1;
#pragma unroll
        for (int m = 0; m < NLD; ++m) lp[m * (NTH / N1) * PP] = make_float2(a[m], b[m]);
      } else {
        // Block size not a multiple of n1: element e = t + NTH m is row
        // e % n1 of pair e / n1, each with its own offset.
        float a[NLD], b[NLD];
#pragma unroll
        for (int m = 0; m < NLD; ++m) {
          const int e = t + NTH * m, pr = e / N1, i1 = e % N1;
          const bool in_plane = m < NLD - 1 || e < N1 * NC / 2;
          const int o = in_plane ? (2 * pr * cstr + i1) * is : 0;
          a[m] = sp[o];
          b[m] = sp[o + cstr * is];
        }
#pragma unroll
        for (int m = 0; m < NLD; ++m) {
          const int e = t + NTH * m;
          if (m < NLD - 1 || e < N1 * NC / 2) L[(e / N1) * PP + e % N1] = make_float2(a[m], b[m]);
        }
      }
    }
    if (sp) split0(t, w0, w1);
    if constexpr (WAVEIN)
      wave_sync();
    else
      __syncthreads();
    {
      // FFT along k1 of z = col 2 pr + i col 2 pr + 1, then rows k <= HS:
      // col 2 pr gets A = (Z[k] + conj Z[-k]) / 2, col 2 pr + 1 gets
      // B = (Z[k] - conj Z[-k]) / 2i; rows 0 and HS (real) go to slot 0 as
      // row 0 + i row HS.
      const int pr = t / C2, jj = t % C2;
      const bool on = pr < NC / 2;
      float2 *qb = L + (on ? pr : 0) * PP;
      float2 v[F1::NV];
      // The exchange region of a column pair is private to its C2 lanes.
      if (on) {
#pragma unroll
        for (int r = 0; r < C1; ++r) v[r] = qb[jj + C2 * r];
        F1::phase1(v, jj, tw1);
      }
      wave_sync();
      if (on) F1::store_cols(v, qb, jj);
      wave_sync();
      if (on) {
        F1::load_rows(v, qb, jj);
        F1::phase2(v);
      }
      __syncthreads();
      {
        // Row k = (jj CR + u) + C1 q lies below HS = C1 C2 / 2 for every
        // q < C2 / 2; row HS is element (0, C2 / 2) of thread jj = 0.  Rows
        // without live tiles are written too: the row pass skips them.
#pragma unroll
        for (int u = 0; u < CR; ++u)
#pragma unroll
          for (int q = 0; q < C2 / 2; ++q) {
            const float2 Zm = fft_mirror<N1>(v, jj, u, q);
            if (on && F1::row_ok(jj, u)) {
              const float2 Z = v[u * C2 + q], Zc = conjf2(Zm);
              const float2 d = csub(Z, Zc);
              const int k1 = jj * CR + u + C1 * q;
              const float2 A = cscale(cadd(Z, Zc), 0.5f);
              *reinterpret_cast<float4 *>(L + k1 * PS + ((2 * pr) ^ plane_col_swz(k1))) =
                  make_float4(A.x, A.y, 0.5f * d.y, -0.5f * d.x);
            }
          }
        // Rows 0 and HS are real: slot 0 takes row 0 + i row HS (replacing
        // the row 0 pair written above by the same thread).
        if (on && jj == 0)  // row 0: no swizzle
          *reinterpret_cast<float4 *>(L + 2 * pr) = make_float4(v[0].x, v[C2 / 2].x, v[0].y, v[C2 / 2].y);
      }
    }
    __syncthreads();
    {
      const int sl = t / RC2, jr = t % RC2;
      float2 *pb = L + (onslot ? sl : 0) * PS;
      const int cz = plane_col_swz(sl);
      float2 v[FR::NV];
      // The exchange region of a row is private to its RC2 lanes.
      if (act) {
#pragma unroll
        for (int r = 0; r < RC1; ++r) v[r] = pb[(jr + RC2 * r) ^ cz];
        FR::phase1(v, jr, twc);
      }
      wave_sync();
      if (act) FR::store_cols(v, pb, jr);
      wave_sync();
      if (act) {
        FR::load_rows(v, pb, jr);
        FR::phase2(v);
      }
      // Plane a of subband p, from row sl and column jr RCR (unit: no table).
      const float *wrow = mv.planes + (wunit ? 0 : jb.pa0 + (p - jb.p0) * jb.pa_step + sl * NC + jr * RCR);
      if (sl == 0) {
        // Z = F(row 0) + i F(row HS), parked for the split.
        if (act) {
#pragma unroll
          for (int uu = 0; uu < RCR; ++uu)
#pragma unroll
            for (int qq = 0; qq < RC2; ++qq)
              if (FR::row_ok(jr, uu)) ZR[(jr * RCR + uu) + RC1 * qq] = v[uu * RC2 + qq];
        }
      } else if (act && wunit) {
#pragma unroll
        for (int m = 0; m < RCR * RC2; ++m)
          if (FR::row_ok(jr, m / RC2)) R[m] = cadd(v[m], R[m]);
      } else if (act) {
#pragma unroll
        for (int uu = 0; uu < RCR; ++uu)
#pragma unroll
          for (int qq = 0; qq < RC2; ++qq) {
            if (!FR::row_ok(jr, uu)) continue;
            const int m = uu * RC2 + qq;
            R[m] = fma2(v[m], wrow[uu + RC1 * qq], R[m]);
          }
      }
    }
    __syncthreads();  // L is reloaded by the next subband; ZR is complete
  }
  {
    const int t = thread_index(w64);
    if (act0 && t < NC) {
      float w0, w1;
      split_w(jb.p1 - 1, t, w0, w1);
      split0(t, w0, w1);
      if ((rl[0] >> (t / tpw)) & 1u) Sb[(unsigned)t] = R0[t];
      if ((rl[HS] >> (t / tpw)) & 1u) Sb[(unsigned)(HS * NC + t)] = R64[t];
    }
  }
  if (act && slot != 0) {
#pragma unroll
    for (int uu = 0; uu < RCR; ++uu)
#pragma unroll
      for (int qq = 0; qq < RC2; ++qq) {
        if (!FR::row_ok(j, uu)) continue;
        const int kc = (j * RCR + uu) + RC1 * qq, m = uu * RC2 + qq;
        if ((rl[slot] >> (kc / tpw)) & 1u) Sb[(unsigned)(slot * NC + kc)] = R[m];
      }
  }
}

// ------------------------------------------------------------------ host side

namespace {

template <int NP, int OR> struct PencilL {
  static void ana(const FusedCtx &c, const float2 *X, float2 *slabs, const int4 *units, int nunits, int) {
    if (nunits <= 0) return;
    const float2 *tw = OR == AX3 ? c.tw3 : c.tw2;
    hipLaunchKernelGGL((kP_ana<NP, OR>), dim3(nunits), dim3(pblock_threads<NP>()), 0, c.stream, X, slabs, c.H, c.n2,
                       c.n3, c.mv, c.items, c.groups, c.live, units, tw);
    SH_LAUNCHED();
  }
  static void syn(const FusedCtx &c, int, const float2 *slabs, const int4 *slist, int ns, float2 *acc,
                  int has_acc, float2 *V) {
    const int nt = OR == AX3 ? c.n2 : c.n3;
    const float2 *tw = OR == AX3 ? c.tw3 : c.tw2;
    hipLaunchKernelGGL((kP_syn<NP, OR>), dim3(c.h1 * (nt / tp_of<NP>())), dim3(syn_threads<NP>()), 0, c.stream,
                       slabs, c.H, c.n2, c.n3, c.mv, c.items, c.live, slist, ns, acc, has_acc, V, tw);
    SH_LAUNCHED();
  }
};

template <int NP> void launch_fwd(const FusedCtx &c, const float2 *T, float2 *X) {
  hipLaunchKernelGGL(kP_fwd<NP>, dim3(c.h1 * (c.n2 / tp_of<NP>())), dim3(pencil_threads<NP>()), 0, c.stream, T, X,
                     c.h1, c.n2, c.tw3);
  SH_LAUNCHED();
}

template <int N1, int NC, int OR> struct PlaneL {
  static void ana(const FusedCtx &c, int, const float2 *slabs, const PlaneJob *jobs, int njobs, float *out,
                  int64_t substride, int64_t os, float scale) {
    if (njobs <= 0) return;
    const int npl = OR == AX3 ? c.n3 : c.n2;
    const float2 *twc = OR == AX3 ? c.tw2 : c.tw3;
    hipLaunchKernelGGL((kQ_ana<N1, NC, OR>), dim3(npl, njobs), dim3(ana_threads<N1, NC>()), 0, c.stream, slabs, c.H,
                       c.n1 * c.n2, c.mv, c.live, jobs, out, substride, (int)os, scale, c.tw1, twc);
    SH_LAUNCHED();
  }
  static void syn(const FusedCtx &c, int, const float *in, int64_t substride, int64_t is, float2 *slabs,
                  const PlaneJob *jobs, int njobs) {
    if (njobs <= 0) return;
    const int npl = OR == AX3 ? c.n3 : c.n2;
    const float2 *twc = OR == AX3 ? c.tw2 : c.tw3;
    hipLaunchKernelGGL((kQ_syn<N1, NC, OR>), dim3(npl, njobs), dim3(plane_threads<N1, NC>()), 0, c.stream, in,
                       substride, (int)is, c.n1 * c.n2, c.mv, c.live, jobs, slabs, c.H, c.tw1, twc);
    SH_LAUNCHED();
  }
};

template <int N1, int NC, int OR> bool set_plane_n(FusedOps &o) {
  o.q_ana[OR] = &PlaneL<N1, NC, OR>::ana;
  o.q_syn[OR] = &PlaneL<N1, NC, OR>::syn;
  return true;
}
template <int N1, int OR> bool set_plane(int nc, FusedOps &o) {
  switch (nc) {
    case 32: return set_plane_n<N1, 32, OR>(o);
    case 64: return set_plane_n<N1, 64, OR>(o);
    case 80: return set_plane_n<N1, 80, OR>(o);
    case 96: return set_plane_n<N1, 96, OR>(o);
    case 112: return set_plane_n<N1, 112, OR>(o);
    case 128: return set_plane_n<N1, 128, OR>(o);
    default: return false;
  }
}

template <int NP, int OR> bool set_pencil_n(FusedOps &o) {
  o.p_ana[OR] = &PencilL<NP, OR>::ana;
  o.p_syn[OR] = &PencilL<NP, OR>::syn;
  return true;
}
template <int OR> bool set_pencil(int np, FusedOps &o) {
  switch (np) {
    case 32: return set_pencil_n<32, OR>(o);
    case 64: return set_pencil_n<64, OR>(o);
    case 80: return set_pencil_n<80, OR>(o);
    case 96: return set_pencil_n<96, OR>(o);
    case 112: return set_pencil_n<112, OR>(o);
    case 128: return set_pencil_n<128, OR>(o);
    default: return false;
  }
}

template <int N1> bool set_planes(int n2, int n3, FusedOps &o) {
  return set_plane<N1, AX3>(n2, o) && set_plane<N1, AX2>(n3, o);
}

}  // namespace

bool fused_ops_for(int n1, int n2, int n3, FusedOps &o) {
  o = FusedOps{};
  auto fits = [](int n) { return n == 32 || n == 64 || n == 80 || n == 96 || n == 112 || n == 128; };
  if (!fits(n1) || !fits(n2) || !fits(n3)) return false;
  bool ok = set_pencil<AX3>(n3, o) && set_pencil<AX2>(n2, o);
  switch (n1) {
    case 32: ok = ok && set_planes<32>(n2, n3, o); break;
    case 64: ok = ok && set_planes<64>(n2, n3, o); break;
    case 80: ok = ok && set_planes<80>(n2, n3, o); break;
    case 96: ok = ok && set_planes<96>(n2, n3, o); break;
    case 112: ok = ok && set_planes<112>(n2, n3, o); break;
    default: ok = ok && set_planes<128>(n2, n3, o); break;
  }
  switch (n3) {
    case 32: o.p_fwd = &launch_fwd<32>; break;
    case 64: o.p_fwd = &launch_fwd<64>; break;
    case 80: o.p_fwd = &launch_fwd<80>; break;
    case 96: o.p_fwd = &launch_fwd<96>; break;
    case 112: o.p_fwd = &launch_fwd<112>; break;
    default: o.p_fwd = &launch_fwd<128>; break;
  }
  o.tile = tp_of<32>();
  return ok;
}

}  // namespace surfhip
