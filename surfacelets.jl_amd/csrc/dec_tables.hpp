// Decimated Surfacelet: parameters, size rules, subband layout, pyramid index
// maps and the FP64 filter tables (ritf hourglass masks, IRC checkerboard
// slices), built on the host.
//
// Reference: src/types/transforms.jl:222-315, src/transforms/surfacelet.jl:
// 73-149,220-310, src/pyramid/fourier_pyramid.jl:26-127,
// src/directional/ritf.jl:24-108, src/directional/irc.jl:29-91,
// src/directional/surfacelet_scale.jl:12-60, src/filters/checkerboard.jl.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "common.hpp"

namespace surfhip {

using Dims3 = std::array<int, 3>;

struct PyrLevel {
  Dims3 n, nlow;
  int Dnum, Dden;
  double D15;                      // D^1.5
  std::array<std::vector<int>, 3> idx, idx_low;  // 0-based bins
  std::array<std::vector<double>, 3> prof;       // lowpass profile on idx
};

struct IRCStageT {
  int m;   // axis decimated (0-based)
  int b;   // block size after the stage
  int off[4];  // ana0, ana1, syn0, syn1 plane offsets (complex pool)
  int offT[4] = {-1, -1, -1, -1};  // the same planes transposed ([c + n_m k]), when the device keeps them
};

struct DecScaleT {
  Dims3 n;                               // size of the spectrum of this scale
  std::array<std::array<int, 3>, 3> lv;  // lv[K][m]
  std::array<int, 3> mask_off;           // V_K offsets in the real pool
  std::array<std::vector<IRCStageT>, 3> stages;
  int first_subband;                     // 0-based subband index of the scale's first block
};

struct DecTables {
  Dims3 dims;
  int J;
  std::vector<PyrLevel> pyr;             // finest first
  std::vector<DecScaleT> scales;         // pyramid order (finest first)
  std::vector<float> masks;              // real pool (ritf masks)
  std::vector<float2> filt;              // complex pool (IRC slices)
  std::vector<Dims3> sizes;              // layout order
  std::vector<int64_t> offsets;          // element offset of each subband
  int64_t ncoef = 0;
};

// Validate the transform parameters (src/types/transforms.jl:231-251).
void dec_check_params(const surfhip_s_params &p);
// First reason axis a of length n does not fit, or "" (surfacelet.jl:230-249).
std::string dec_axis_problem(const surfhip_s_params &p, int a, int n);
// Throws ArgumentError naming the nearest fitting length (surfacelet.jl:253-269).
void dec_check_size(const surfhip_s_params &p, const Dims3 &dims);
// Subband sizes in layout order (surfacelet.jl:275-290).
std::vector<Dims3> dec_layout_sizes(const surfhip_s_params &p, const Dims3 &dims);
// Every table of a workspace.
void dec_build_tables(const surfhip_s_params &p, const Dims3 &dims, DecTables &T);

// Helpers of the scale geometry (surfacelet_scale.jl:28-60).
inline Dims3 half_dims(const Dims3 &n, int K) {
  Dims3 h = n;
  h[K] = n[K] / 2 + 1;
  return h;
}
inline void scale_blocks(const DecScaleT &sc, int K, Dims3 &counts, Dims3 &bsize) {
  for (int d = 0; d < 3; ++d) {
    counts[d] = 1 << sc.lv[K][d];
    bsize[d] = d == K ? sc.n[K] / 2 + 1 : sc.n[d] / counts[d];
  }
}
inline Dims3 subband_size(const DecScaleT &sc, int K) {
  Dims3 c, b;
  scale_blocks(sc, K, c, b);
  b[K] = sc.n[K];
  return b;
}

}  // namespace surfhip
