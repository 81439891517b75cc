// Host planning of the LDS Stockham FFT: radix factorisation and FP64
// twiddles rounded once to FP32.
#include <hip/hip_runtime.h>

#include <cmath>

#include "common.hpp"
#include "fft.hpp"

namespace surfhip {

FftPlan1D make_fft_plan(int n, std::vector<float2> &pool) {
  if (n < 1) fail(SURFHIP_ERR_ARG, fmt("FFT length must be positive, got %d", n));
  FftPlan1D pl{};
  pl.n = n;
  pl.npass = 0;
  int m = n;
  // Largest hand-written radices first: fewer passes (96 = 16 x 6,
  // 80 = 16 x 5, 72 = 12 x 6, 112 = 16 x 7, 128 = 16 x 8).
  const int fixed[] = {16, 14, 12, 10, 9, 8, 7, 6, 5, 4, 3, 2};
  for (int r : fixed)
    while (m % r == 0) {
      if (pl.npass >= FFT_MAX_PASSES) fail(SURFHIP_ERR_ARG, fmt("FFT length %d has too many factors", n));
      pl.radix[pl.npass++] = r;
      m /= r;
    }
  for (int p = 11; m > 1; p += 2) {
    while (m % p == 0) {
      if (p > FFT_MAX_RADIX)
        fail(SURFHIP_ERR_ARG, fmt("axis length %d has the prime factor %d; factors up to %d are supported", n, p, FFT_MAX_RADIX));
      if (pl.npass >= FFT_MAX_PASSES) fail(SURFHIP_ERR_ARG, fmt("FFT length %d has too many factors", n));
      pl.radix[pl.npass++] = p;
      m /= p;
    }
  }
  pl.tw = (int32_t)pool.size();
  for (int k = 0; k < n; ++k) {
    double a = -2.0 * M_PI * (double)k / (double)n;
    pool.push_back(make_float2((float)std::cos(a), (float)std::sin(a)));
  }
  return pl;
}

}  // namespace surfhip
