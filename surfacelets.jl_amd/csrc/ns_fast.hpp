// Launch interface of the fast NSSurfacelet kernels (ns_fast.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "ns_tables.hpp"

namespace surfhip {

// Everything a launch needs: sizes, the device tables and twiddles.
struct FastCtx {
  hipStream_t stream;
  int n1, n2, n3, h1;
  int64_t H;              // rfft-grid size h1 n2 n3
  NSMaskView mv;          // device tables
  const NSItem *items;    // device copy of the S items
  const float2 *tw1, *tw2, *tw3;  // forward twiddles of n1, n2, n3
};

// The frequency support of a group's masks: every group mask carries the
// separable pyramid factor a1 a2 a3 of its level, which is exactly zero for
// k1 > K1 and for |k2| > K2 (the raised-cosine profiles vanish beyond their
// band edge).  Passes skip the pencils, rows and columns outside the box;
// K1 = h1 - 1 and K2 = n2 / 2 mean the full grid.
struct Support {
  int K1, K2;
};

// Host launchers, specialised for one (n1, n2, n3).  Spectra X / acc /
// accg are pencil-major [k1][k2][k3]; V / T are plane-major [i3][k1][k2],
// one H-sized slab per wedge.
struct FastOps {
  void (*a_fwd)(const FastCtx &, const float2 *T, float2 *X) = nullptr;
  // `cls` is the size of the item's whole group: a launch for one item of a
  // group (streaming) runs the same kernel instance as the whole group, so
  // both give bit-identical results.
  void (*a_ana)(const FastCtx &, const float2 *X, float2 *V, int item0, int nw, int cls, Support sp) = nullptr;
  void (*a_syn)(const FastCtx &, const float2 *T, int nw, float2 *acc, const float2 *gin, float2 *gout, float2 *V,
                int item0, int close, int finalize, int cls, Support sp) = nullptr;
  void (*a_inv)(const FastCtx &, const float2 *acc, float2 *V) = nullptr;
  void (*a_close)(const FastCtx &, float2 *acc, const float2 *accg, int item0, Support sp) = nullptr;
  void (*b_inv)(const FastCtx &, const float2 *V, int nw, float *out, size_t owstride, size_t os, float scale,
                Support sp) = nullptr;
  void (*b_fwd)(const FastCtx &, const float *in, size_t iwstride, size_t is, int nw, float2 *T,
                Support sp) = nullptr;
};

// Fill `o` for a grid the fast path covers; false otherwise.
bool fast_ops_for(int n1, int n2, int n3, FastOps &o);

}  // namespace surfhip
