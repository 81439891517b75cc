// FP64 construction of the NSSurfacelet tables (host), rounded once to FP32.
//
// Reference, per function below:
//   bernstein_table      src/filters/bernstein.jl:36-53
//   mapping_response     src/filters/bernstein.jl:95-151 (group (1,), 3-D and 2-D)
//   halfband_split       src/filters/bernstein.jl:171-190
//   rcos_axis            src/pyramid/fourier_pyramid.jl:42-61
//   rc_responses         src/pyramid/fourier_pyramid.jl:143-150
//   wedge planes         src/directional/ndfb_wedges.jl:14-113
//   hourglass tables     src/directional/hourglass.jl:172-189,289-303
//   workspace + bank     src/transforms/ns_surfacelet.jl:98-123,178-185,222-242
#include <algorithm>
#include <cmath>
#include <string>

#include "ns_tables.hpp"

namespace surfhip {

namespace {

double binom(int n, int k) {
  double r = 1.0;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return std::round(r);
}

// (N+1) x m table, entry [j * m + k].
std::vector<double> bernstein_table(int N, const std::vector<double> &omega) {
  const int m = (int)omega.size();
  std::vector<double> t((size_t)(N + 1) * m);
  for (int k = 0; k < m; ++k) {
    double s = std::sin(omega[k] / 2);
    double x = std::min(std::max(s * s, 0.0), 1.0);
    double y = 1.0 - x;
    for (int j = 0; j <= N; ++j) t[(size_t)j * m + k] = binom(N, j) * std::pow(x, j) * std::pow(y, N - j);
  }
  return t;
}

std::vector<double> dft_frequencies(int n, bool half) {
  int m = half ? n / 2 + 1 : n;
  std::vector<double> w(m);
  for (int k = 0; k < m; ++k) w[k] = 2.0 * M_PI * k / n;
  return w;
}

std::vector<double> folded_frequencies(int n, bool half) {
  if (half) return dft_frequencies(n, true);
  std::vector<double> w(n);
  for (int k = 0; k < n; ++k) {
    double v = 2.0 * M_PI * k / n;
    w[k] = v > M_PI ? v - 2.0 * M_PI : v;
  }
  return w;
}

double rcos(double x) { return x <= 0 ? 0.0 : (x >= 1 ? 1.0 : (1 - std::cos(M_PI * x)) / 2); }

long long floordiv(long long a, long long b) {
  long long q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
  return q;
}
long long ceildiv(long long a, long long b) { return -floordiv(-a, b); }

double soft_step(double s, double b, double w, int poly) {
  if (w <= 0) return s >= b ? 1.0 : 0.0;
  double u = (s - b) / w;
  u = std::min(std::max(u, -1.0), 1.0);
  return halfband_split(u, poly);
}

double wedge_weight(double s, int idx, int nw, double transition, int poly) {
  if (nw == 1) return 1.0;
  double delta = 2.0 / nw;
  double width = transition * delta;
  double lo = idx == 1 ? 1.0 : soft_step(s, -1.0 + (idx - 1) * delta, width, poly);
  double hi = idx == nw ? 0.0 : soft_step(s, -1.0 + idx * delta, width, poly);
  return lo - hi;
}

// Sign fixed by index, never by value: src/directional/ndfb_wedges.jl:78,97-106.
bool self_conjugate_bin(int k, int n) { return (2 * k) % n == 0; }

double wedge_ratio(double wr, double wt, bool ref_self, bool tr_self) {
  if (wr == 0) return 0.0;
  if (ref_self) return std::fabs(wt) / std::fabs(wr);
  if (tr_self) return wt / std::fabs(wr);
  return wt / wr;
}

void transverse_axes(int i, int &a, int &b) {
  if (i == 1) { a = 2; b = 3; }
  else if (i == 2) { a = 1; b = 3; }
  else { a = 1; b = 2; }
}

}  // namespace

double halfband_split(double x, int m) { return halfband_split_hd(x, m); }

void rcos_axis(int n, long long wnum, long long wden, long long tnum, long long tden,
               std::vector<double> &prof, int &P, int &Q) {
  // P = ceil(n/2 (w + t)), Q = floor(n/2 (w - t)), exact rationals.
  long long den = 2 * wden * tden;
  P = (int)ceildiv((long long)n * (wnum * tden + tnum * wden), den);
  Q = (int)floordiv((long long)n * (wnum * tden - tnum * wden), den);
  prof.assign(n, 0.0);
  for (int k = 0; k < n; ++k) {
    int f = std::abs(k <= n / 2 ? k : k - n);
    if (f <= P) prof[k] = std::sqrt(rcos((double)(P - f) / (double)(P - Q)));
  }
}

int32_t ns_check_params(const surfhip_ns_params &p) {
  if (p.scales < 1) fail(SURFHIP_ERR_ARG, fmt("scales must be ≥ 1, got %d", p.scales));
  if (p.scales > SURFHIP_MAX_SCALES) fail(SURFHIP_ERR_ARG, fmt("at most %d scales are supported, got %d", SURFHIP_MAX_SCALES, p.scales));
  for (int j = 0; j < p.scales; ++j) {
    if (p.direction_levels[j] < 0) {
      std::string s = "[";
      for (int k = 0; k < p.scales; ++k) s += (k ? ", " : "") + std::to_string(p.direction_levels[k]);
      fail(SURFHIP_ERR_ARG, "direction_levels entries must be ≥ 0, got " + s + "]");
    }
    if (p.direction_levels[j] > 12) fail(SURFHIP_ERR_ARG, fmt("direction_levels entries above 12 are not supported, got %d", p.direction_levels[j]));
  }
  if (p.bernstein_order < 1) fail(SURFHIP_ERR_ARG, fmt("bernstein_order must be ≥ 1, got %d", p.bernstein_order));
  if (p.halfband_order < 0) fail(SURFHIP_ERR_ARG, fmt("halfband_order must be ≥ 0, got %d", p.halfband_order));
  if (!(p.transition >= 0)) fail(SURFHIP_ERR_ARG, fmt("transition must be ≥ 0, got %g", p.transition));
  int32_t S = 1;
  for (int j = 0; j < p.scales; ++j) S += 3 * (1 << p.direction_levels[j]);
  return S;
}

void ns_build_tables(const surfhip_ns_params &p, int n1, int n2, int n3, NSTables &T, bool pyramid,
                     bool device_groups) {
  const int32_t S = ns_check_params(p);
  if (n1 < 1 || n2 < 1 || n3 < 1)
    fail(SURFHIP_ERR_ARG, fmt("volume size must be positive, got (%d, %d, %d)", n1, n2, n3));
  const int h1 = n1 / 2 + 1;
  const int J = p.scales;
  const int N = p.bernstein_order, poly = p.halfband_order;
  T.n1 = n1; T.n2 = n2; T.n3 = n3; T.h1 = h1; T.J = J; T.S = S; T.tight = p.tight ? 1 : 0;
  const int nb[3] = {h1, n2, n3};
  const int nn[3] = {n1, n2, n3};
  const int Lsum = h1 + n2 + n3;

  // Pyramid per-axis factors and the empty-scale refusal (ns_surfacelet.jl:107-118).
  std::vector<std::vector<double>> acc(J + 1, std::vector<double>(Lsum, 1.0));
  std::vector<std::vector<double>> lp(J, std::vector<double>(Lsum, 0.0));
  for (int d = 0, off = 0; d < 3; off += nb[d], ++d) {
    std::vector<double> prof;
    int P, Q;
    for (int j = 1; j <= J; ++j) {
      rcos_axis(nn[d], 1, 1LL << j, 1, 3LL << j, prof, P, Q);
      for (int k = 0; k < nb[d]; ++k) {
        lp[j - 1][off + k] = prof[k];
        acc[j][off + k] = acc[j - 1][off + k] * prof[k];
      }
    }
  }
  for (int j = 0; j < J && pyramid; ++j) {
    bool empty = true;
    for (int d = 0, off = 0; d < 3 && empty; off += nb[d], ++d)
      for (int k = 0; k < nb[d]; ++k)
        if (!(acc[j][off + k] == 0.0 || lp[j][off + k] == 1.0)) { empty = false; break; }
    if (empty)
      fail(SURFHIP_ERR_ARG, fmt("size (%d, %d, %d) is too small for %d scales: scale %d would be empty", n1, n2, n3, J, J - j));
  }
  // Hourglass tables: f0(M1) on the rfft grid, f0(M2) on (w2, w3).
  const int np1 = N + 1;
  auto b1 = bernstein_table(N, dft_frequencies(n1, true));
  auto b2 = bernstein_table(N, dft_frequencies(n2, false));
  auto b3 = bernstein_table(N, dft_frequencies(n3, false));
  // t1[(j2, j3), k1] = sum_j1 c[j1, j2, j3] b1[j1, k1]
  std::vector<double> t1((size_t)np1 * np1 * h1, 0.0);
  for (int j3 = 0; j3 < np1; ++j3)
    for (int j2 = 0; j2 < np1; ++j2) {
      double *dst = &t1[((size_t)j3 * np1 + j2) * h1];
      int mx = std::max(j2, j3);
      for (int j1 = 0; j1 < np1; ++j1) {
        double c = j1 > mx ? 1.0 : (j1 < mx ? -1.0 : 0.0);
        if (c == 0.0) continue;
        const double *src = &b1[(size_t)j1 * h1];
        for (int k1 = 0; k1 < h1; ++k1) dst[k1] += c * src[k1];
      }
    }
  // t2[j3, k2, k1] = sum_j2 t1[j2, j3, k1] b2[j2, k2]
  std::vector<double> t2((size_t)np1 * n2 * h1, 0.0);
  for (int j3 = 0; j3 < np1; ++j3)
    for (int k2 = 0; k2 < n2; ++k2) {
      double *dst = &t2[((size_t)j3 * n2 + k2) * h1];
      for (int j2 = 0; j2 < np1; ++j2) {
        double w = b2[(size_t)j2 * n2 + k2];
        const double *src = &t1[((size_t)j3 * np1 + j2) * h1];
        for (int k1 = 0; k1 < h1; ++k1) dst[k1] += w * src[k1];
      }
    }
  const size_t V = (size_t)h1 * n2 * n3;
  std::vector<double> hg2((size_t)n2 * n3), hg2c((size_t)n2 * n3);
  for (int k3 = 0; k3 < n3; ++k3)
    for (int k2 = 0; k2 < n2; ++k2) {
      double m = 0.0;
      for (int j2 = 0; j2 < np1; ++j2)
        for (int j3 = 0; j3 < np1; ++j3) {
          double c = j2 > j3 ? 1.0 : (j2 < j3 ? -1.0 : 0.0);
          if (c != 0.0) m += c * b2[(size_t)j2 * n2 + k2] * b3[(size_t)j3 * n3 + k3];
        }
      double f0 = halfband_split(m, poly), f1 = halfband_split(-m, poly);
      hg2[(size_t)k2 * n3 + k3] = T.tight ? std::sqrt(std::max(f0, 0.0)) : f0;
      hg2c[(size_t)k2 * n3 + k3] = T.tight ? std::sqrt(std::max(f1, 0.0)) : f1;
    }

  // Wedge planes, one pool; each plane is stored over its two axes in
  // increasing order (p < q), C-ordered: index kp * len_q + kq.
  T.planes.clear();
  T.items.assign(S, NSItem{});
  T.items[0] = NSItem{1, J, 0, -1, 0, -1, 0, -1};
  T.group_first.clear();
  T.group_count.clear();
  T.group_na.clear();
  T.group_nb.clear();
  int offset = 1;  // 0-based subband index of scale j's first subband
  for (int j = 1; j <= J; ++j) {
    const int Lj = p.direction_levels[j - 1];
    const int lv[2] = {(Lj + 1) / 2, Lj / 2};
    for (int i = 1; i <= 3; ++i) {
      int tr[2];
      transverse_axes(i, tr[0], tr[1]);
      int offs[2][64], pair[2] = {0, 0}, count[2];
      for (int s = 0; s < 2; ++s) {
        const int ref = i, t = tr[s];
        const int nw = 1 << lv[s];
        count[s] = nw;
        const int lo = std::min(ref, t), hi = std::max(ref, t);
        pair[s] = (lo == 1 && hi == 2) ? 0 : (lo == 1 ? 1 : 2);
        if (lv[s] == 0) { offs[s][0] = -1; continue; }
        auto wref = folded_frequencies(nn[ref - 1], ref == 1);
        auto wtr = folded_frequencies(nn[t - 1], t == 1);
        const int lenlo = nb[lo - 1], lenhi = nb[hi - 1];
        for (int k = 0; k < nw; ++k) {
          offs[s][k] = (int)T.planes.size();
          T.planes.resize(T.planes.size() + (size_t)lenlo * lenhi);
        }
        for (int q = 0; q < (int)wtr.size(); ++q)
          for (int pp = 0; pp < (int)wref.size(); ++pp) {
            double sr = wedge_ratio(wref[pp], wtr[q], self_conjugate_bin(pp, nn[ref - 1]),
                                    self_conjugate_bin(q, nn[t - 1]));
            int ilo = ref < t ? pp : q, ihi = ref < t ? q : pp;
            for (int k = 0; k < nw; ++k) {
              double w = wedge_weight(sr, k + 1, nw, p.transition, poly);
              if (p.tight) w = std::sqrt(std::max(w, 0.0));
              T.planes[(size_t)offs[s][k] + (size_t)ilo * lenhi + ihi] = (float)w;
            }
          }
      }
      const int g = 3 * (j - 1) + (i - 1);
      T.group_first.push_back(offset + (i - 1) * count[0] * count[1]);
      T.group_count.push_back(count[0] * count[1]);
      T.group_na.push_back(count[0]);
      T.group_nb.push_back(count[1]);
      for (int a = 0; a < count[0]; ++a)
        for (int b = 0; b < count[1]; ++b) {
          int k = offset + ((i - 1) * count[0] + a) * count[1] + b;  // _ns_index, 0-based
          T.items[k] = NSItem{0, J - j, i, offs[0][a], pair[0], offs[1][b], pair[1], g};
        }
    }
    offset += 3 * (1 << Lj);
  }

  // Group factors in FP64, rounded once (src/transforms/ns_surfacelet.jl:
  // 205-208,226-227,265-272): group g = 3 (j - 1) + (i - 1) has
  // B = a1 a2 a3 sqrt(max(1 - (l1 l2 l3)^2, 0)) at pyramid level J - j times
  // the hourglass weight W_i = (W1, W2, W3) = (f0(M1), f1(M1) f0(M2),
  // f1(M1) f1(M2)); the lowpass (last) is a1 a2 a3 at level J.  Without the
  // pyramid (the directional filter bank alone: src/directional/ndfb.jl:
  // 215-241, src/directional/hourglass.jl:400-420) B = 1 and the lowpass is
  // zero.
  const int G = 3 * J;
  T.ngroups = G + 1;
  T.np1 = np1;
  T.poly = poly;
  T.pyramid = pyramid ? 1 : 0;
  T.lsum = Lsum;
  T.g_t2 = std::move(t2);
  T.g_b3 = std::move(b3);
  T.g_hg2 = std::move(hg2);
  T.g_hg2c = std::move(hg2c);
  T.g_acc.clear();
  for (int lev = 0; lev <= J; ++lev) T.g_acc.insert(T.g_acc.end(), acc[lev].begin(), acc[lev].end());
  T.g_lp.clear();
  for (int lev = 0; lev < J; ++lev) T.g_lp.insert(T.g_lp.end(), lp[lev].begin(), lp[lev].end());
  if (T.g_lp.empty()) T.g_lp.push_back(0.0);
  T.gana.clear();
  T.gsyn.clear();
  if (device_groups) return;
  T.gana.assign((size_t)T.ngroups * V, 0.0f);
  if (!T.tight) T.gsyn.assign((size_t)T.ngroups * V, 0.0f);
  const NSGroupInputs in = T.group_inputs();
  for (int k1 = 0; k1 < h1; ++k1)
    for (int k2 = 0; k2 < n2; ++k2)
      for (int k3 = 0; k3 < n3; ++k3)
        ns_group_voxel(in, k1, k2, k3, T.gana.data(), T.tight ? nullptr : T.gsyn.data(), V);
}

}  // namespace surfhip
