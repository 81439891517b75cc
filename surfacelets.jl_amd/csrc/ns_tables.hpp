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

// f0 of the halfband split (src/filters/bernstein.jl:171-190) in FP64:
// ((1+x)/2)^(m+1) sum_{k<=m} C(m+k, k) ((1-x)/2)^k, the powers by repeated
// multiplication and no contraction, so host and device round alike.
SH_HD inline double halfband_split_hd(double x, int m) {
#pragma clang fp contract(off)
  const double u = (1.0 + x) / 2, v = (1.0 - x) / 2;
  if (u == 0.0) return 0.0;
  double s = 0.0, c = 1.0, vk = 1.0, um = u;
  for (int k = 0; k <= m; ++k) {
    s += c * vk;
    vk *= v;
    c = c * (m + k + 1) / (k + 1);
  }
  for (int k = 0; k < m; ++k) um *= u;
  return um * s;
}

// Inputs of the group-factor tables, small enough to build on the host: the
// stage-1 mapping response contracted over axes 1 and 2 (t2[j3][k2][k1]),
// the axis-3 Bernstein table b3[j3][k3], the stage-2 splits hg2/hg2c[k2][k3]
// (square-rooted when tight), and the per-axis pyramid products acc[lev][.]
// (levels 0..J) and lowpasses lp[lev][.] (levels 0..J-1) over the
// concatenated axes (h1, n2, n3).
struct NSGroupInputs {
  const double *t2, *b3, *hg2, *hg2c, *acc, *lp;
  int32_t h1, n2, n3, np1, poly, J, tight, pyramid, lsum;
};

// Every group factor of voxel (k1, k2, k3) (src/transforms/ns_surfacelet.jl:
// 226-227,265-272,295-303,350-358), formed in FP64 and rounded once:
// ga[g V + q] = B_lev W_i (the lowpass: a1 a2 a3), gs[g V + q] = B_lev (when
// not tight; gs null otherwise).  The same code builds the tables on the host
// and on the device.
SH_HD inline void ns_group_voxel(const NSGroupInputs &in, int k1, int k2, int k3, float *ga, float *gs, size_t V) {
#pragma clang fp contract(off)
  const int h1 = in.h1, n2 = in.n2, n3 = in.n3;
  double col = 0.0;
  for (int j3 = 0; j3 < in.np1; ++j3) col += in.b3[(size_t)j3 * n3 + k3] * in.t2[((size_t)j3 * n2 + k2) * h1 + k1];
  const double f0 = halfband_split_hd(col, in.poly), f1 = halfband_split_hd(-col, in.poly);
  const double hg1 = in.tight ? sqrt(f0 > 0.0 ? f0 : 0.0) : f0;
  const double hg1c = in.tight ? sqrt(f1 > 0.0 ? f1 : 0.0) : f1;
  const size_t q = ((size_t)k1 * n2 + k2) * n3 + k3, q2 = (size_t)k2 * n3 + k3;
  const int G = 3 * in.J;
  for (int g = 0; g <= G; ++g) {
    const bool low = g == G;
    const int lev = low ? in.J : in.J - (g / 3 + 1);
    const int i = g % 3 + 1;
    const double *ax = in.acc + (size_t)lev * in.lsum;
    const double a3 = in.pyramid ? ax[k1] * ax[h1 + k2] * ax[h1 + n2 + k3] : (low ? 0.0 : 1.0);
    if (low) {
      ga[(size_t)g * V + q] = (float)a3;
      if (gs) gs[(size_t)g * V + q] = (float)a3;
      continue;
    }
    const double *lx = in.lp + (size_t)lev * in.lsum;
    const double L = in.pyramid ? lx[k1] * lx[h1 + k2] * lx[h1 + n2 + k3] : 0.0;
    const double r = 1.0 - L * L;
    const double B = a3 * sqrt(r > 0.0 ? r : 0.0);
    const double W = i == 1 ? hg1 : hg1c * (i == 2 ? in.hg2[q2] : in.hg2c[q2]);
    ga[(size_t)g * V + q] = (float)(B * W);
    if (gs) gs[(size_t)g * V + q] = (float)B;
  }
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
  // Inputs of the group tables (NSGroupInputs), kept when the tables are left
  // to the device (ns_build_tables with device_groups).
  std::vector<double> g_t2, g_b3, g_hg2, g_hg2c, g_acc, g_lp;
  NSGroupInputs group_inputs() const {
    return NSGroupInputs{g_t2.data(), g_b3.data(), g_hg2.data(), g_hg2c.data(), g_acc.data(), g_lp.data(),
                         h1, n2, n3, np1, poly, J, tight, pyramid, lsum};
  }
  int32_t np1 = 0, poly = 0, pyramid = 1, lsum = 0;
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
// device_groups = true leaves gana/gsyn empty and keeps their inputs in
// `out` for a device build (ns_group_voxel per voxel).
void ns_build_tables(const surfhip_ns_params &p, int n1, int n2, int n3, NSTables &out, bool pyramid = true,
                     bool device_groups = false);

// FP64 helpers shared with the decimated tables.
double halfband_split(double x, int m);
void rcos_axis(int n, long long wnum, long long wden, long long tnum, long long tden,
               std::vector<double> &prof, int &P, int &Q);

}  // namespace surfhip
