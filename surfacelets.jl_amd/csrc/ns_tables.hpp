// NSSurfacelet mask tables (built on the host in FP64, stored as FP32) and
// the per-voxel mask expression the kernels evaluate on the fly.
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
  // All tables are C-ordered over (k1, k2, k3) -- k3 fastest -- so that a
  // pencil along axis 3 reads them contiguously.
  const float *hg1, *hg1c;  // [k1][k2][k3]: stage-1 factor f0(M1) and f1(M1), sqrt'd when tight
  const float *hg2, *hg2c;  // [k2][k3]: stage-2 factor f0(M2) and f1(M2), sqrt'd when tight
  const float *a;           // (J+1) x (h1+n2+n3): prod_{i<j} l_i per axis
  const float *l;           // J x (h1+n2+n3): l_j per axis
  const float *c;           // J x (h1+n2+n3): 1 - l_j per axis (formed in FP64)
  const float *planes;      // wedge plane pool (sqrt'd when tight), each [lower axis][higher axis]
  int32_t h1, n2, n3, J, tight;
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

// The bandpass of pyramid level `lev` at (k1, k2, k3):
// B = a1 a2 a3 sqrt(max(1 - (l1 l2 l3)^2, 0)) (src/transforms/ns_surfacelet.jl:271-272).
// 1 - L is formed as c1 + l1 c2 + l1 l2 c3 from the FP64 complements, which
// is exact algebra and avoids the FP32 cancellation of 1 - L near L = 1.
SH_HD inline float ns_bandpass(const NSMaskView &v, int lev, int k1, int k2, int k3) {
  const int L = v.h1 + v.n2 + v.n3;
  const float *a = v.a + (size_t)lev * L;
  const float *l = v.l + (size_t)lev * L;
  const float *c = v.c + (size_t)lev * L;
  const int i2 = v.h1 + k2, i3 = v.h1 + v.n2 + k3;
  float l1 = l[k1], l2 = l[i2], l3 = l[i3];
  float om = c[k1] + l1 * c[i2] + (l1 * l2) * c[i3];
  float op = 1.0f + (l1 * l2) * l3;
  float r = om * op;
  return (a[k1] * a[i2] * a[i3]) * sqrtf(r > 0.0f ? r : 0.0f);
}

// Group factor of subband `it` at rfft bin (k1, k2, k3): B_j times the
// hourglass factor (analysis, and synthesis of a tight bank), B_j alone (the
// synthesis of a bank that is not tight), or the lowpass a1 a2 a3
// (src/transforms/ns_surfacelet.jl:226-227,265-272,295-303,350-358).
SH_HD inline float ns_group(const NSMaskView &v, const NSItem &it, int k1, int k2, int k3, bool analysis) {
  if (it.lowpass) {
    const int L = v.h1 + v.n2 + v.n3;
    const float *a = v.a + (size_t)v.J * L;
    return a[k1] * a[v.h1 + k2] * a[v.h1 + v.n2 + k3];
  }
  float B = ns_bandpass(v, it.level, k1, k2, k3);
  if (!v.tight && (analysis ? v.ana_uses_syn : !v.syn_uses_ana)) return B;
  size_t p3 = ((size_t)k1 * v.n2 + (size_t)k2) * v.n3 + k3;
  size_t p2 = (size_t)k2 * v.n3 + k3;
  float h = it.hourglass == 1 ? v.hg1[p3] : v.hg1c[p3] * (it.hourglass == 2 ? v.hg2[p2] : v.hg2c[p2]);
  return B * h;
}

// Wedge factor w_a w_b (src/transforms/ns_surfacelet.jl:376-427); 1 for the
// lowpass and for the synthesis of a bank that is not tight (unit wedge).
SH_HD inline float ns_wedge(const NSMaskView &v, const NSItem &it, int k1, int k2, int k3, bool analysis) {
  if (it.lowpass || (!v.tight && (analysis ? v.ana_uses_syn : !v.syn_uses_ana))) return 1.0f;
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
  std::vector<float> hg1, hg1c, hg2, hg2c, a, l, c, planes;
  std::vector<NSItem> items;  // S entries, subband_layout order
  std::vector<int32_t> group_first, group_count;  // items of each group (subband ranges)
  NSMaskView view() const {
    return NSMaskView{hg1.data(), hg1c.data(), hg2.data(), hg2c.data(), a.data(), l.data(),
                      c.data(), planes.data(), h1, n2, n3, J, tight, 0, 0};
  }
};

// Validate parameters (src/types/transforms.jl:93-104) and return S.
int32_t ns_check_params(const surfhip_ns_params &p);

// Build every table in FP64 and round once (src/transforms/ns_surfacelet.jl:98-123).
void ns_build_tables(const surfhip_ns_params &p, int n1, int n2, int n3, NSTables &out);

// FP64 helpers shared with the decimated tables.
double halfband_split(double x, int m);
void rcos_axis(int n, long long wnum, long long wden, long long tnum, long long tden,
               std::vector<double> &prof, int &P, int &Q);

}  // namespace surfhip
