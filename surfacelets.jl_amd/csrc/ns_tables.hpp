// NSSurfacelet mask tables (built on the host in FP64, stored as FP32) and
// the per-voxel mask expression the kernels evaluate from them.
//
// Reference: src/transforms/ns_surfacelet.jl:98-123 (tables), :222-242 (the
// mask bank), :265-272 (bandpass and hourglass factors), :295-303 (group
// product), :350-358 (lowpass), :376-427 (wedge product).
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

#include "common.hpp"

namespace surfhip {

// One subband of the bank.  Subband 1 (index 0) is the lowpass.
struct NSItem {
  int32_t lowpass;    // 1 for the coarse lowpass
  int32_t level;      // pyramid level of the bandpass (0 = finest)
  int32_t hourglass;  // 1, 2, 3
  int32_t pa_off;     // offset of wedge plane a in the plane pool
  int32_t pa_pair;    // axes of plane a: 0 = (1,2), 1 = (1,3), 2 = (2,3)
  int32_t pb_off;
  int32_t pb_pair;
  int32_t group;      // (scale, hourglass) group, 0-based; -1 for the lowpass
};

// Raw table pointers (host or device) and the grid.
struct NSMaskView {
  // Group factors, one rfft-grid table per group and the lowpass last,
  // [g][k1][k2][k3] (k3 fastest, so a pencil along axis 3 reads contiguously):
  //   analysis: G = B_j W_i (the bandpass of the group's pyramid level times
  //   its hourglass weight, both square-rooted when tight); lowpass a1 a2 a3;
  //   synthesis: the same when tight, B_j alone otherwise.
  // Each value is formed in FP64 and rounded once.
  const float *gana, *gsyn;
  const float *planes;      // wedge plane pool (sqrt'd when tight), each [lower axis][higher axis]
  int32_t h1, n2, n3, ngroups, tight;
  // Mask swaps for the adjoints: every subband operator irfft . M . rfft is
  // self-adjoint (M real and even), so A^T is the accumulation with the
  // analysis masks and S^T the analysis with the synthesis masks.  Only a
  // bank that is not tight has different analysis and synthesis masks.
  int32_t syn_uses_ana = 0;  // "synthesis" masks are the analysis masks
  int32_t ana_uses_syn = 0;  // "analysis" masks are the synthesis masks
};

SH_HD inline float ns_plane(const NSMaskView &v, int32_t off, int32_t pair, int k1, int k2, int k3) {
  if (off < 0) return 1.0f;
  int idx = pair == 0 ? k1 * v.n2 + k2 : (pair == 1 ? k1 * v.n3 + k3 : k2 * v.n3 + k3);
  return v.planes[off + idx];
}

// Table index of group `g` of the item (the lowpass is the last group).
SH_HD inline int ns_group_index(const NSMaskView &v, const NSItem &it) {
  return it.lowpass ? v.ngroups - 1 : it.group;
}

// The group-factor table of group g for analysis or synthesis masks.
SH_HD inline const float *ns_group_table(const NSMaskView &v, int g, bool analysis) {
  const bool ana = analysis ? !v.ana_uses_syn : v.syn_uses_ana;
  return (ana ? v.gana : v.gsyn) + (size_t)g * ((size_t)v.h1 * v.n2 * v.n3);
}

// Group factor of subband `it` at rfft bin (k1, k2, k3): B_j times the
// hourglass factor (analysis, and synthesis of a tight bank), B_j alone (the
// synthesis of a bank that is not tight), or the lowpass a1 a2 a3
// (src/transforms/ns_surfacelet.jl:226-227,265-272,295-303,350-358).
SH_HD inline float ns_group(const NSMaskView &v, const NSItem &it, int k1, int k2, int k3, bool analysis) {
  return ns_group_table(v, ns_group_index(v, it), analysis)[((size_t)k1 * v.n2 + k2) * v.n3 + k3];
}

// Whether the wedge factor is the unit mask: the synthesis of a bank that is
// not tight (src/transforms/ns_surfacelet.jl:226-238).
SH_HD inline bool ns_unit_wedge(const NSMaskView &v, bool analysis) {
  return !v.tight && (analysis ? v.ana_uses_syn : !v.syn_uses_ana);
}

// The two wedge planes of subband `it`: w_a over (hourglass axis, first
// transverse axis), w_b over (hourglass axis, second transverse axis); 1 for
// the lowpass and for a unit wedge.
SH_HD inline float ns_wedge_a(const NSMaskView &v, const NSItem &it, int k1, int k2, int k3, bool analysis) {
  if (it.lowpass || ns_unit_wedge(v, analysis)) return 1.0f;
  return ns_plane(v, it.pa_off, it.pa_pair, k1, k2, k3);
}
SH_HD inline float ns_wedge_b(const NSMaskView &v, const NSItem &it, int k1, int k2, int k3, bool analysis) {
  if (it.lowpass || ns_unit_wedge(v, analysis)) return 1.0f;
  return ns_plane(v, it.pb_off, it.pb_pair, k1, k2, k3);
}

// Wedge factor w_a w_b (src/transforms/ns_surfacelet.jl:376-427); 1 for the
// lowpass and for the synthesis of a bank that is not tight (unit wedge).
SH_HD inline float ns_wedge(const NSMaskView &v, const NSItem &it, int k1, int k2, int k3, bool analysis) {
  if (it.lowpass || ns_unit_wedge(v, analysis)) return 1.0f;
  return ns_plane(v, it.pa_off, it.pa_pair, k1, k2, k3) * ns_plane(v, it.pb_off, it.pb_pair, k1, k2, k3);
}

// Full mask of subband `it`: group factor times wedge factor.  analysis =
// false gives the synthesis mask, which differs only when the bank is not
// tight (src/transforms/ns_surfacelet.jl:226-238).
SH_HD inline float ns_mask(const NSMaskView &v, const NSItem &it, int k1, int k2, int k3, bool analysis) {
  return ns_group(v, it, k1, k2, k3, analysis) * ns_wedge(v, it, k1, k2, k3, analysis);
}

// Host-side tables of one (transform, size).
struct NSTables {
  int32_t n1, n2, n3, h1, J, S, tight;
  int32_t ngroups;                 // 3 J groups and the lowpass
  std::vector<float> gana, gsyn;   // ngroups rfft-grid tables each; gsyn empty when tight (= gana)
  std::vector<float> planes;
  std::vector<NSItem> items;  // S entries, subband_layout order
  std::vector<int32_t> group_first, group_count;  // items of each group (subband ranges)
  std::vector<int32_t> group_na, group_nb;        // wedge counts of planes a and b per group
  NSMaskView view() const {
    const float *gs = gsyn.empty() ? gana.data() : gsyn.data();
    return NSMaskView{gana.data(), gs, planes.data(), h1, n2, n3, ngroups, tight, 0, 0};
  }
};

// Validate parameters (src/types/transforms.jl:93-104) and return S.
int32_t ns_check_params(const surfhip_ns_params &p);

// Build every table in FP64 and round once (src/transforms/ns_surfacelet.jl:98-123).
// pyramid = false builds the directional filter bank alone (one scale, B = 1,
// a zero lowpass): src/directional/ndfb.jl:91-241, hourglass.jl:149-420.
void ns_build_tables(const surfhip_ns_params &p, int n1, int n2, int n3, NSTables &out, bool pyramid = true);

// FP64 helpers shared with the decimated tables.
double halfband_split(double x, int m);
void rcos_axis(int n, long long wnum, long long wden, long long tnum, long long tden,
               std::vector<double> &prof, int &P, int &Q);

}  // namespace surfhip
