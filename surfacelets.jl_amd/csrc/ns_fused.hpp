// Launch interface of the factored NSSurfacelet kernels (ns_fused.hip).
//
// Every wedge mask is a product w_a w_b of two planes, each over the
// hourglass axis and one transverse axis (src/directional/ndfb_wedges.jl:35-43,
// src/transforms/ns_surfacelet.jl:376-427).  Plane a never involves one of the
// three axes -- k3 for hourglasses 1 and 2, k2 for hourglass 3 -- so a 1-D
// transform along that axis (the "pencil axis" of the group) commutes with it:
//   subband (p, q) = irfft3(X G w_a,p w_b,q)
//                  = IFFT_plane( w_a,p . IFFT_pencil( X G w_b,q ) ).
// The pencil pass therefore produces one intermediate slab per (group, q)
// instead of one per subband, and the plane pass applies w_a,p while it
// produces the nA subbands that share the slab.  Synthesis runs the same
// factorisation backwards: R_q = sum_p w_a,p FFT_plane(c_pq), then
// acc += G sum_q w_b,q FFT_pencil(R_q).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "ns_tables.hpp"

namespace surfhip {

// Pencil axis of a group: k3 (hourglasses 1 and 2, the lowpass) or k2
// (hourglass 3).  The tile axis is the other transverse axis.
enum { AX3 = 0, AX2 = 1 };

// One group on the device: the lowpass (one item) or a (scale, hourglass)
// pair of nA x nB items, item (p, q) at subband index first + p nB + q.
struct GroupDev {
  int32_t first, nA, nB, orient, slab0, lowpass;
};

// One plane-pass job: slab `slab`, b-index q of group `group` (-1: no mask,
// the whole grid), a-indices [p0, p1).  Subband p goes to (comes from) index
// out0 + (p - p0) out_step of the coefficient array; its plane a starts at
// pa0 + (p - p0) pa_step in the plane pool (pa0 < 0: no plane a).  `init`
// (synthesis) starts the sum over p from zero instead of from the slab.
struct PlaneJob {
  int32_t slab, group, q, p0, p1, init, out0, out_step, pa0, pa_step;
  int32_t pad[6];
};

struct FusedCtx {
  hipStream_t stream;
  int n1, n2, n3, h1;
  int64_t H;
  NSMaskView mv;
  const NSItem *items;
  const GroupDev *groups;
  const uint32_t *live;   // per slab, h1 words: bit t of row k1 = tile t holds nonzeros
  const float2 *tw1, *tw2, *tw3;
};

// Spectra X / acc are pencil-major [k1][k2][k3]; a slab of orientation AX3 is
// [i3][k1][k2], one of AX2 [i2][k1][k3] (H elements each).
struct FusedOps {
  int tile = 16;  // tile width (pencils per block) of both orientations
  // Analysis pencil pass over `nunits` units {group, k1, first tile column}
  // of orientation `orient`.
  void (*p_ana[2])(const FusedCtx &, const float2 *X, float2 *slabs, const int4 *units, int nunits,
                   int orient) = {nullptr, nullptr};
  // Analysis plane pass: jobs of one orientation -> subbands (or the volume).
  void (*q_ana[2])(const FusedCtx &, int orient, const float2 *slabs, const PlaneJob *jobs, int njobs, float *out,
                   int64_t substride, int64_t os, float scale) = {nullptr, nullptr};
  // Synthesis plane pass: subbands (or the volume) -> slabs.
  void (*q_syn[2])(const FusedCtx &, int orient, const float *in, int64_t substride, int64_t is, float2 *slabs,
                   const PlaneJob *jobs, int njobs) = {nullptr, nullptr};
  // Synthesis pencil pass over all tiles of one orientation for the slabs
  // slist[0..ns) = {slab, group, item of b-index q, 0}, summed in list order:
  // AX2 writes its sum into acc; AX3 adds acc (has_acc), closes the sum and
  // writes its inverse along k3 into V ([i3][k1][k2]).
  void (*p_syn[2])(const FusedCtx &, int orient, const float2 *slabs, const int4 *slist, int ns, float2 *acc,
                   int has_acc, float2 *V) = {nullptr, nullptr};
  // X = FFT along k3 of T ([i3][k1][k2] -> [k1][k2][k3]).
  void (*p_fwd)(const FusedCtx &, const float2 *T, float2 *X) = nullptr;
};

// Fill `o` for a grid the factored kernels cover; false otherwise.
bool fused_ops_for(int n1, int n2, int n3, FusedOps &o);

}  // namespace surfhip
