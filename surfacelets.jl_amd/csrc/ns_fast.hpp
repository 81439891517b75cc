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

// Host launchers, specialised for one (n1, n2, n3).  Spectra X / acc /
// accg are pencil-major [k1][k2][k3]; V / T are plane-major [i3][k1][k2],
// one H-sized slab per wedge.
struct FastOps {
  void (*a_fwd)(const FastCtx &, const float2 *T, float2 *X) = nullptr;
  void (*a_ana)(const FastCtx &, const float2 *X, float2 *V, int item0, int nw) = nullptr;
  void (*a_syn)(const FastCtx &, const float2 *T, int nw, float2 *acc, const float2 *gin, float2 *gout, float2 *V,
                int item0, int close, int finalize) = nullptr;
  void (*a_inv)(const FastCtx &, const float2 *acc, float2 *V) = nullptr;
  void (*a_close)(const FastCtx &, float2 *acc, const float2 *accg, int item0) = nullptr;
  void (*b_inv)(const FastCtx &, const float2 *V, int nw, float *out, size_t owstride, size_t os,
                float scale) = nullptr;
  void (*b_fwd)(const FastCtx &, const float *in, size_t iwstride, size_t is, int nw, float2 *T) = nullptr;
};

// Fill `o` for a grid the fast path covers; false otherwise.
bool fast_ops_for(int n1, int n2, int n3, FastOps &o);

}  // namespace surfhip
