// Host construction of the decimated Surfacelet's integer maps and FP64
// tables (rounded once to FP32).  Each function cites the reference it follows.
#include <algorithm>
#include <cmath>
#include <complex>
#include <string>

#include "dec_tables.hpp"
#include "ns_tables.hpp"

namespace surfhip {

namespace {

// The ladder taps t_q (q = 1, 3, ..., bo-1) of SurfBox's checkerboard
// filters: src/filters/checkerboard.jl:26-53 (a table that must be identical).
const std::vector<double> &cbd_ladder(int bo) {
  static const std::vector<double> t4 = {-0.6262169131293146, -0.2391935764853579};
  static const std::vector<double> t6 = {-0.6263429050738657, -0.18302382681237633, -0.10753046674888095};
  static const std::vector<double> t8 = {-0.627819441355453, -0.1868110527281001, -0.08821345831338342,
                                         -0.05299463507171143};
  static const std::vector<double> t10 = {-0.6290391996524649, -0.19023436487159193, -0.09316419099049718,
                                          -0.04787638089257991, -0.02729049477362672};
  static const std::vector<double> t12 = {-0.6299909650161403, -0.19294788893803155, -0.09719811683520764,
                                          -0.05257765509290852, -0.027224347419537415, -0.014421093056370137};
  static const std::vector<double> t14 = {-0.6307470366053054, -0.19509552613399547, -0.10042251574577052,
                                          -0.056418179282558024, -0.03113763249657735, -0.015819995468636647,
                                          -0.007750295209451851};
  static const std::vector<double> t16 = {-0.631356163548133, -0.19683521105934682, -0.10304471724680653,
                                          -0.05956342644490844, -0.03441276144847804, -0.018849533868796398,
                                          -0.009297418074817033, -0.004215553479603451};
  static const std::vector<double> t18 = {-0.6318535467876545, -0.19826135647461407, -0.10520738123671233,
                                          -0.06218709792109189, -0.037184557113685186, -0.021472356755448885,
                                          -0.011535270028700011, -0.00549634585408209, -0.0023119823375261405};
  switch (bo) {
    case 4: return t4;
    case 6: return t6;
    case 8: return t8;
    case 10: return t10;
    case 12: return t12;
    case 14: return t14;
    case 16: return t16;
    case 18: return t18;
  }
  fail(SURFHIP_ERR_ARG, fmt("checkerboard_order must be one of [4, 6, 8, 10, 12, 14, 16, 18], got %d", bo));
}

// src/pyramid/fourier_pyramid.jl:35-40, as exact rationals (num, den).
struct Mode {
  long long w[2][2], t[2][2];
  int D[2][2];
};
Mode pyramid_mode(double r) {
  if (r == 1.0) return Mode{{{1, 2}, {1, 4}}, {{1, 6}, {1, 12}}, {{1, 1}, {2, 1}}};
  if (r == 1.5) return Mode{{{1, 2}, {3, 8}}, {{1, 7}, {1, 9}}, {{3, 2}, {2, 1}}};
  if (r == 2.0) return Mode{{{1, 3}, {1, 3}}, {{1, 7}, {1, 7}}, {{2, 1}, {2, 1}}};
  std::string s = fmt("%g", r);
  if (r == std::floor(r) && std::fabs(r) < 1e15) s = fmt("%.1f", r);
  fail(SURFHIP_ERR_ARG, "pyramid_ratio must be 1, 1.5 or 2, got " + s);
}

std::string ratio_str(double r) {
  // Julia prints a Float64: 1.5, 2.0, 1.0
  if (r == std::floor(r)) return fmt("%.1f", r);
  return fmt("%g", r);
}

// b(w) = 2 sum t_q sin(q w): src/filters/checkerboard.jl:71-77.
double cbd_b(const std::vector<double> &t, double w) {
  double s = 0.0;
  for (size_t i = 0; i < t.size(); ++i) s += t[i] * std::sin((2.0 * i + 1.0) * w);
  return 2 * s;
}
std::complex<double> cbd_analysis(const std::vector<double> &t, int chan, double w1, double w2) {
  double B = -cbd_b(t, w1) * cbd_b(t, w2);
  double r = 1 / std::sqrt(2.0);
  if (chan == 0) return {r * (1 + B), 0.0};
  return std::polar(1.0, w2) * (r * (2 - B - B * B));
}
std::complex<double> cbd_synthesis(const std::vector<double> &t, int chan, double w1, double w2) {
  double B = -cbd_b(t, w1) * cbd_b(t, w2);
  double r = 1 / std::sqrt(2.0);
  if (chan == 0) return {r * (2 + B - B * B), 0.0};
  return std::polar(1.0, -w2) * (r * (1 - B));
}

// src/directional/ritf.jl:24-37.
double ritf_I0(double x) {
  x /= 2;
  double y = 1.0, z = 1.0, t = 1.0;
  int k = 1;
  while (t > 1.0e-10 * y) {
    z = z * x / k;
    k += 1;
    t = z * z;
    y += t;
  }
  return y;
}

double sinc(double x) { return x == 0 ? 1.0 : std::sin(M_PI * x) / (M_PI * x); }

// src/directional/ritf.jl:40-52.
std::vector<double> ritf_window(int n, double beta) {
  double den = ritf_I0(beta);
  int half = (n - 1) / 2;
  std::vector<double> wh(half + 1);
  for (int i = 0; i <= half; ++i) {
    double v = 2.0 * i / (n - 1) - 1;
    wh[i] = ritf_I0(beta * std::sqrt(1 - v * v)) / den;
  }
  std::vector<double> w(n);
  for (int i = 0; i <= half; ++i) w[i] = wh[i];
  for (int i = 0; i < half; ++i) w[half + 1 + i] = wh[half - 1 - i];
  for (int i = 0; i < n; ++i) w[i] *= sinc((double)(i - half) / 2);
  return w;
}

// src/directional/ritf.jl:58-73.
double ritf_diamond(const std::vector<double> &w, double a, double b) {
  int half = ((int)w.size() - 1) / 2;
  double al = (a + b) / 2, be = (a - b) / 2;
  double cea = 0, ceb = 0, coa = 0, cob = 0;
  for (int i = 0; i < (int)w.size(); ++i) {
    int u = i - half;
    if (u % 2 == 0) {
      cea += w[i] * std::cos(u * al);
      ceb += w[i] * std::cos(u * be);
    } else {
      coa += w[i] * std::cos(u * al);
      cob += w[i] * std::cos(u * be);
    }
  }
  return cea * ceb + coa * cob;
}

long long floordiv(long long a, long long b) {
  long long q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
  return q;
}
long long ceildiv(long long a, long long b) { return -floordiv(-a, b); }

int levels_max_axis(const surfhip_s_params &p, int s, int a) {
  int c = 0;
  for (int K = 0; K < 3; ++K) c = std::max(c, p.levels[s][K][a]);
  return c;
}

}  // namespace

void dec_check_params(const surfhip_s_params &p) {
  if (p.scales < 1) fail(SURFHIP_ERR_ARG, fmt("scales must be ≥ 1, got %d", p.scales));
  if (p.scales > SURFHIP_MAX_SCALES)
    fail(SURFHIP_ERR_ARG, fmt("at most %d scales are supported, got %d", SURFHIP_MAX_SCALES, p.scales));
  for (int j = 0; j < p.scales; ++j)
    for (int K = 0; K < 3; ++K)
      for (int m = 0; m < 3; ++m) {
        if (p.levels[j][K][m] < 0) fail(SURFHIP_ERR_ARG, "directional levels must be ≥ 0");
        if (p.levels[j][K][m] > 16) fail(SURFHIP_ERR_ARG, "directional levels above 16 are not supported");
      }
  pyramid_mode(p.pyramid_ratio);
  cbd_ladder(p.checkerboard_order);
  if (!(p.diamond_length % 2 == 1 && p.diamond_length >= 3))
    fail(SURFHIP_ERR_ARG, fmt("diamond_length must be an odd integer ≥ 3, got %d", p.diamond_length));
  if (!(p.kaiser_beta > 0)) fail(SURFHIP_ERR_ARG, fmt("kaiser_beta must be > 0, got %s", ratio_str(p.kaiser_beta).c_str()));
  if (!(p.diamond_power % 2 == 0 && p.diamond_power > 0))
    fail(SURFHIP_ERR_ARG, fmt("diamond_power must be a positive even integer, got %d", p.diamond_power));
}

std::string dec_axis_problem(const surfhip_s_params &p, int a, int n) {
  const int J = p.scales;
  Mode md = pyramid_mode(p.pyramid_ratio);
  for (int j = 1; j <= J; ++j) {
    int s = J - j + 1;
    std::string where = fmt("at scale %d of %d (1 the coarsest) the length is %d", s, J, n);
    if (n % 2) return where + ", which must be even";
    long long c = 1LL << levels_max_axis(p, s - 1, a);
    if (n % c) return where + fmt(", which the directional levels must split into %lld wedge blocks", c);
    int k = j == 1 ? 0 : 1;
    int Dn = md.D[k][0], Dd = md.D[k][1];
    if (((long long)n * Dd) % Dn)
      return where + fmt(", which resampling by %d/%d needs to be a multiple of %d", Dn, Dd, Dn);
    int nlow = (int)((long long)n * Dd / Dn);
    long long wn = md.w[k][0], wd = md.w[k][1], tn = md.t[k][0], td = md.t[k][1];
    int P = (int)ceildiv((long long)n * (wn * td + tn * wd), 2 * wd * td);
    if (2 * P > nlow) return where + ", too small for the pyramid's lowpass";
    n = nlow;
  }
  return "";
}

void dec_check_size(const surfhip_s_params &p, const Dims3 &dims) {
  dec_check_params(p);
  for (int a = 0; a < 3; ++a) {
    std::string problem = dec_axis_problem(p, a, dims[a]);
    if (problem.empty()) continue;
    int n = dims[a], near = 0;
    for (int d = 1; d <= 4 * n + 64 && near == 0; ++d) {
      for (int m : {n - d, n + d})
        if (m >= 1 && dec_axis_problem(p, a, m).empty()) {
          near = m;
          break;
        }
    }
    std::string hint = near > 0 ? fmt("; the nearest length that fits is %d", near) : "";
    fail(SURFHIP_ERR_ARG,
         fmt("a volume of size (%d, %d, %d) does not fit Surfacelet(scales = %d, pyramid_ratio = %s): axis %d: ",
             dims[0], dims[1], dims[2], p.scales, ratio_str(p.pyramid_ratio).c_str(), a + 1) +
             problem + hint);
  }
}

namespace {

// src/pyramid/fourier_pyramid.jl:68-103.
std::vector<PyrLevel> fourier_pyramid(const surfhip_s_params &p, const Dims3 &dims) {
  Mode md = pyramid_mode(p.pyramid_ratio);
  std::vector<PyrLevel> out;
  Dims3 n = dims;
  for (int j = 0; j < p.scales; ++j) {
    int k = j == 0 ? 0 : 1;
    PyrLevel lv;
    lv.n = n;
    lv.Dnum = md.D[k][0];
    lv.Dden = md.D[k][1];
    lv.D15 = std::pow((double)lv.Dnum / lv.Dden, 1.5);
    for (int a = 0; a < 3; ++a) {
      if (((long long)n[a] * lv.Dden) % lv.Dnum)
        fail(SURFHIP_ERR_ARG, fmt("a pyramid level with factor %d//%d cannot divide size (%d, %d, %d)", lv.Dnum,
                                  lv.Dden, n[0], n[1], n[2]));
      lv.nlow[a] = (int)((long long)n[a] * lv.Dden / lv.Dnum);
    }
    for (int a = 0; a < 3; ++a) {
      std::vector<double> full;
      int P, Q;
      rcos_axis(n[a], md.w[k][0], md.w[k][1], md.t[k][0], md.t[k][1], full, P, Q);
      if (2 * P > lv.nlow[a]) fail(SURFHIP_ERR_ARG, fmt("size %d is too small for the pyramid", n[a]));
      for (int f = 0; f <= P; ++f) {
        lv.idx[a].push_back(f);
        lv.idx_low[a].push_back(f);
      }
      for (int f = n[a] - P; f < n[a]; ++f) lv.idx[a].push_back(f);
      for (int f = lv.nlow[a] - P; f < lv.nlow[a]; ++f) lv.idx_low[a].push_back(f);
      for (int v : lv.idx[a]) lv.prof[a].push_back(full[v]);
    }
    out.push_back(lv);
    n = lv.nlow;
  }
  return out;
}

}  // namespace

std::vector<Dims3> dec_layout_sizes(const surfhip_s_params &p, const Dims3 &dims) {
  dec_check_size(p, dims);
  auto pyr = fourier_pyramid(p, dims);
  std::vector<Dims3> sizes = {pyr.back().nlow};
  for (int j = 1; j <= p.scales; ++j) {
    Dims3 n = pyr[p.scales - j].n;
    for (int K = 0; K < 3; ++K) {
      Dims3 counts, sb;
      for (int d = 0; d < 3; ++d) {
        counts[d] = 1 << p.levels[j - 1][K][d];
        if (n[d] % counts[d])
          fail(SURFHIP_ERR_ARG, fmt("scale %d of size (%d, %d, %d) cannot be split into (%d, %d, %d) wedge blocks in hourglass %d",
                                    j, n[0], n[1], n[2], counts[0], counts[1], counts[2], K + 1));
        sb[d] = d == K ? n[K] : n[d] / counts[d];
      }
      long long cnt = (long long)counts[0] * counts[1] * counts[2];
      for (long long i = 0; i < cnt; ++i) sizes.push_back(sb);
    }
  }
  return sizes;
}

void dec_build_tables(const surfhip_s_params &p, const Dims3 &dims, DecTables &T) {
  T.dims = dims;
  T.J = p.scales;
  T.sizes = dec_layout_sizes(p, dims);
  T.offsets.clear();
  int64_t off = 0;
  for (auto &s : T.sizes) {
    T.offsets.push_back(off);
    off += (int64_t)s[0] * s[1] * s[2];
  }
  T.ncoef = off;
  T.pyr = fourier_pyramid(p, dims);
  const auto &tq = cbd_ladder(p.checkerboard_order);
  const auto w = ritf_window(p.diamond_length, p.kaiser_beta);
  T.masks.clear();
  T.filt.clear();
  T.scales.clear();
  // Subband offsets of the scales (surfacelet.jl:78-81): coarse first.
  std::vector<int> counts(p.scales);
  for (int j = 0; j < p.scales; ++j) {
    const int s = p.scales - j - 1;  // levels index of pyramid level j
    int c = 0;
    for (int K = 0; K < 3; ++K) c += 1 << (p.levels[s][K][0] + p.levels[s][K][1] + p.levels[s][K][2]);
    counts[j] = c;
  }
  for (int j = 0; j < p.scales; ++j) {
    DecScaleT sc;
    sc.n = T.pyr[j].n;
    const int s = p.scales - j - 1;
    for (int K = 0; K < 3; ++K)
      for (int m = 0; m < 3; ++m) sc.lv[K][m] = K == m ? 0 : p.levels[s][K][m];
    int first = 1;
    for (int q = j + 1; q < p.scales; ++q) first += counts[q];
    sc.first_subband = first;
    const Dims3 n = sc.n;
    for (int d = 0; d < 3; ++d)
      if (n[d] % 2) fail(SURFHIP_ERR_ARG, fmt("the hourglass stage needs even dimensions, got (%d, %d, %d)", n[0], n[1], n[2]));
    // ritf masks V_K = sqrt(3 hg_K / sum hg) on the half of axis K (ritf.jl:76-108).
    std::array<std::array<std::vector<double>, 3>, 3> plane;  // plane[K][m]: [k + nK * l]
    for (int K = 0; K < 3; ++K)
      for (int m = 0; m < 3; ++m) {
        if (m == K) continue;
        auto &P = plane[K][m];
        P.resize((size_t)n[K] * n[m]);
        for (int l = 0; l < n[m]; ++l)
          for (int k = 0; k < n[K]; ++k) {
            double a = 2 * M_PI * k / n[K] + M_PI;
            double b = 2 * M_PI * l / n[m];
            P[(size_t)k + (size_t)n[K] * l] = std::pow(ritf_diamond(w, a, b), p.diamond_power);
          }
      }
    auto hg = [&](int K, int i1, int i2, int i3) {
      int ii[3] = {i1, i2, i3};
      double v = 1.0;
      for (int m = 0; m < 3; ++m)
        if (m != K) v *= plane[K][m][(size_t)ii[K] + (size_t)n[K] * ii[m]];
      return v;
    };
    for (int K = 0; K < 3; ++K) {
      Dims3 hd = half_dims(n, K);
      sc.mask_off[K] = (int)T.masks.size();
      T.masks.resize(T.masks.size() + (size_t)hd[0] * hd[1] * hd[2]);
      for (int i3 = 0; i3 < hd[2]; ++i3)
        for (int i2 = 0; i2 < hd[1]; ++i2)
          for (int i1 = 0; i1 < hd[0]; ++i1) {
            double tot = hg(0, i1, i2, i3) + hg(1, i1, i2, i3) + hg(2, i1, i2, i3);
            double v = std::sqrt(3 * hg(K, i1, i2, i3) / tot);
            T.masks[(size_t)sc.mask_off[K] + i1 + (size_t)hd[0] * (i2 + (size_t)hd[1] * i3)] = (float)v;
          }
      // IRC stages, axes m != K ascending (irc.jl:49-91).
      for (int m = 0; m < 3; ++m) {
        if (m == K) continue;
        const int L = sc.lv[K][m];
        if (hd[m] % (1 << L)) fail(SURFHIP_ERR_ARG, fmt("axis %d of length %d cannot be split %d times", m + 1, hd[m], L));
        for (int le = 1; le <= L; ++le) {
          IRCStageT st;
          st.m = m;
          st.b = hd[m] >> le;
          const int nKh = hd[K], nm = hd[m], W = nm >> (le - 1);
          for (int f = 0; f < 4; ++f) {
            st.off[f] = (int)T.filt.size();
            T.filt.resize(T.filt.size() + (size_t)nKh * nm);
            const int chan = f % 2;
            const bool ana = f < 2;
            const double scale = ana ? 0.5 : 1.0;
            for (int c = 0; c < nm; ++c)
              for (int k = 0; k < nKh; ++k) {
                int g = c / W, l = c % W;
                int sf = (1 << (le - 1)) - 1 - 2 * g;
                double wK = 2 * M_PI * k / n[K];
                double wm = 2 * M_PI * l / W + sf * wK;
                std::complex<double> v = scale * (ana ? cbd_analysis(tq, chan, wK, wm) : cbd_synthesis(tq, chan, wK, wm));
                T.filt[(size_t)st.off[f] + k + (size_t)nKh * c] = make_float2((float)v.real(), (float)v.imag());
              }
          }
          sc.stages[K].push_back(st);
        }
      }
    }
    T.scales.push_back(sc);
  }
}

}  // namespace surfhip
