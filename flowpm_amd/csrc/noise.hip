// Seeded Gaussian white noise generated in HBM (flowpm/utils.py:337-355
// draws N(0, sqrt(M)) with TF's RNG, whose stream cannot be reproduced).
//
// Counter-based: normal i comes from block i / 4 of the Philox4x64-10
// stream (Salmon et al. 2011, "Parallel random numbers: as easy as 1, 2,
// 3"), the generator numpy ships as numpy.random.Philox: block q is
// philox4x64_10(counter = q + 1, key), exactly the words
// numpy.random.Philox(key=key).random_raw() returns, so the same seed gives
// the same field on the GPU and on the host.  Each block's four words make
// two Box-Muller pairs: u1 = ((w0 >> 11) + 1) 2^-53 in (0, 1],
// u2 = (w1 >> 11) 2^-53, r = sqrt(-2 ln u1), (r cos 2 pi u2, r sin 2 pi u2),
// evaluated in float64 and scaled, then rounded once to float32.  No state
// is carried between threads, so any element range can be regenerated.
#include "pmx_internal.h"

namespace pmx {

struct Philox4x64 {
  unsigned long long c[4];
};

__device__ __forceinline__ Philox4x64 philox4x64_10(Philox4x64 x,
                                                    unsigned long long k0,
                                                    unsigned long long k1) {
  constexpr unsigned long long M0 = 0xD2E7470EE14C6C93ULL;
  constexpr unsigned long long M1 = 0xCA5A826395121157ULL;
  constexpr unsigned long long W0 = 0x9E3779B97F4A7C15ULL;
  constexpr unsigned long long W1 = 0xBB67AE8584CAA73BULL;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += W0;
      k1 += W1;
    }
    const unsigned long long hi0 = __umul64hi(M0, x.c[0]), lo0 = M0 * x.c[0];
    const unsigned long long hi1 = __umul64hi(M1, x.c[2]), lo1 = M1 * x.c[2];
    Philox4x64 y;
    y.c[0] = hi1 ^ x.c[1] ^ k0;
    y.c[1] = lo1;
    y.c[2] = hi0 ^ x.c[3] ^ k1;
    y.c[3] = lo0;
    x = y;
  }
  return x;
}

__global__ __launch_bounds__(256) void k_white_noise(
    float* __restrict__ out, long long n, unsigned long long k0,
    unsigned long long k1, unsigned long long ctr0, double scale) {
  const long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (4 * q >= n) return;
  // 256-bit counter ctr0 + q + 1 (carry into the second word)
  Philox4x64 x;
  const unsigned long long lo = ctr0 + (unsigned long long)q + 1ULL;
  x.c[0] = lo;
  x.c[1] = lo < ctr0 ? 1ULL : 0ULL;
  x.c[2] = 0;
  x.c[3] = 0;
  x = philox4x64_10(x, k0, k1);
  const double two_pi = 6.283185307179586;
  const double inv53 = 1.0 / 9007199254740992.0;
  double z[4];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const double u1 = (double)((x.c[2 * h] >> 11) + 1ULL) * inv53;
    const double u2 = (double)(x.c[2 * h + 1] >> 11) * inv53;
    const double r = sqrt(-2.0 * log(u1));
    z[2 * h] = r * cos(two_pi * u2);
    z[2 * h + 1] = r * sin(two_pi * u2);
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (4 * q + u < n) out[4 * q + u] = (float)(z[u] * scale);
}

int launch_white_noise(float* out, long long n, unsigned long long k0,
                       unsigned long long k1, unsigned long long ctr0,
                       double scale, hipStream_t s) {
  if (n <= 0) return kOk;
  const long long nq = (n + 3) / 4;
  hipLaunchKernelGGL(k_white_noise, dim3((unsigned)((nq + 255) / 256)),
                     dim3(256), 0, s, out, n, k0, k1, ctr0, scale);
  PMX_LAUNCH_CHECK();
  return kOk;
}

}  // namespace pmx
