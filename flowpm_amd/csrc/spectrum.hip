// Binned power spectrum of a half spectrum (flowpm/pk.py:7-103).
//
// The reference transforms the real field to the FULL complex spectrum and
// weights every mode by W[..., kz] = 1 for the first and last kz index and
// 2 otherwise (pk.py:16-19, a half-spectrum weighting applied to a full
// spectrum), bins |k| with numpy.digitize against kedges and sums
// W |F|^2 per bin.  Here only the half spectrum (kz = 0..nz/2) exists: the
// full-spectrum mode (kx, ky, nz - j) is conj F(-kx, -ky, j), with the same
// |F|^2 and the same |k| (fftfreq is odd, so the squares are equal), so half
// column j carries the weights of both full columns:
//   j = 0:          W(0)
//   0 < j < nz/2:   W(j) + W(nz - j)
//   j = nz/2:       W(nz/2)
// which reproduces the reference's sums exactly.  With count = 1 the same
// pass sums W alone: the reference's weighted mode count Nsum (pk.py:32-35,
// a bincount over the full spectrum), exact in float64.
//
// |k| = sqrt((kx^2 + ky^2) + kz^2) is evaluated in float64 without
// contraction from the same per-axis tables numpy builds, so the bin of
// every mode is the one numpy.digitize finds.  One thread walks one (x, y)
// row, keeping a running sum while the bin does not change; per-block bins
// live in LDS and are added to the float64 output once per block.
#include "pmx_internal.h"

namespace pmx {

struct PkArgs {
  const float2* spec;
  const double* kx;
  const double* ky;
  const double* kz;
  const double* edges;
  int nedges;
  double* psum;  // [B][nedges + 1]
  PlanDev d;
};

__device__ __forceinline__ int pk_w(int idx, int nz) {
  return (idx == 0 || idx == nz - 1) ? 1 : 2;
}

// numpy.digitize(x, edges), increasing edges, right=False: the number of
// edges <= x
__device__ __forceinline__ int digitize(double x, const double* e, int n) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (e[mid] <= x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__global__ void k_zero_d(double* p, long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

template <bool COUNT>
__global__ __launch_bounds__(256) void k_pk_bin(PkArgs a) {
  // numpy rounds every product and sum: no fused multiply-adds here
#pragma clang fp contract(off)
  extern __shared__ double bins[];
  const int nb = a.nedges + 1;
  const int b = blockIdx.y;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) bins[i] = 0.0;
  __syncthreads();
  const PlanDev& d = a.d;
  const long long rows = (long long)d.nx * d.ny;
  const int H = d.H;
  for (long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x;
       r < rows; r += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(r / d.ny), y = (int)(r % d.ny);
    const double kxy = a.kx[x] * a.kx[x] + a.ky[y] * a.ky[y];
    const long long row = (long long)b * rows + r;
    const float2* src = a.spec + row * H;
    int cur = -1;
    double acc = 0.0;
    for (int j = 0; j <= H; ++j) {
      float2 v = make_float2(0.f, 0.f);
      if constexpr (!COUNT) v = j < H ? src[j] : a.spec[d.nyq0 + row];
      const double kk = kxy + a.kz[j] * a.kz[j];
      const int bin = digitize(sqrt(kk), a.edges, a.nedges);
      int w = pk_w(j, d.nz);
      if (j > 0 && j < H) w += pk_w(d.nz - j, d.nz);
      // float32 as the reference forms it: real(F conj F) times W
      const float p2 = v.x * v.x + v.y * v.y;
      const float pw = COUNT ? (float)w : (float)w * p2;
      if (bin != cur) {
        if (cur >= 0) atomicAdd(&bins[cur], acc);
        cur = bin;
        acc = 0.0;
      }
      acc += (double)pw;
    }
    if (cur >= 0) atomicAdd(&bins[cur], acc);
  }
  __syncthreads();
  double* out = a.psum + (long long)b * nb;
  for (int i = threadIdx.x; i < nb; i += blockDim.x)
    if (bins[i] != 0.0) atomicAdd(out + i, bins[i]);
}

int launch_power_bins(const Plan& p, const float2* spec, const double* kx,
                      const double* ky, const double* kz, const double* edges,
                      int nedges, double* psum, hipStream_t s, int count) {
  PMX_REQUIRE(nedges >= 1 && nedges + 1 <= kMaxPkBins,
              "power spectrum: 1 <= number of edges < 8192");
  const int nb_members = count ? 1 : p.d.B;
  const long long nout = (long long)nb_members * (nedges + 1);
  hipLaunchKernelGGL(k_zero_d, dim3((unsigned)((nout + 255) / 256)),
                     dim3(256), 0, s, psum, nout);
  PMX_LAUNCH_CHECK();
  PkArgs a;
  a.spec = spec;
  a.kx = kx;
  a.ky = ky;
  a.kz = kz;
  a.edges = edges;
  a.nedges = nedges;
  a.psum = psum;
  a.d = p.d;
  const long long rows = (long long)p.d.nx * p.d.ny;
  const unsigned g = (unsigned)std::min<long long>((rows + 255) / 256, 1024);
  hipLaunchKernelGGL(count ? k_pk_bin<true> : k_pk_bin<false>,
                     dim3(g, nb_members), dim3(256),
                     sizeof(double) * (nedges + 1), s, a);
  PMX_LAUNCH_CHECK();
  return kOk;
}

// Adjoint of the binning (the gradient of pk.py:40-103 w.r.t. the field):
// for bin sums P_b = sum_{modes in b} W |F|^2, d L / d x =
// 2 Re IDFT_unnorm(sum_k g_b(k) W_k F_k) over the full spectrum.  W is not
// symmetric under k -> -k, but the real part keeps only the symmetric part
// (W_k + W_-k) / 2, which on the half spectrum is the binning weight of the
// column (as in k_pk_bin) halved where the column stands for two full
// modes.  So spec[mode] *= g[b][bin] * Wh(j) * (0 < j < nz/2 ? 1/2 : 1) in
// place; a c2r with scale 2 then gives the field gradient.  One thread per
// mode (coalesced), bins by the same float64 digitize as k_pk_bin.
struct PkAdjArgs {
  float2* spec;
  const double* kx;
  const double* ky;
  const double* kz;
  const double* edges;
  int nedges;
  const double* g;  // [B][nedges + 1]
  PlanDev d;
};

__global__ __launch_bounds__(256) void k_pk_bin_adj(PkAdjArgs a) {
#pragma clang fp contract(off)
  const PlanDev& d = a.d;
  const int b = blockIdx.y;
  const long long rows = (long long)d.nx * d.ny;
  const int H = d.H;
  const long long nmain = rows * H;
  const double* g = a.g + (long long)b * (a.nedges + 1);
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
       e < nmain + rows; e += (long long)gridDim.x * blockDim.x) {
    long long r;
    int j;
    float2* dst;
    if (e < nmain) {
      r = e / H;
      j = (int)(e % H);
      dst = a.spec + ((long long)b * rows + r) * H + j;
    } else {
      r = e - nmain;
      j = H;
      dst = a.spec + d.nyq0 + (long long)b * rows + r;
    }
    const int x = (int)(r / d.ny), y = (int)(r % d.ny);
    const double kk =
        (a.kx[x] * a.kx[x] + a.ky[y] * a.ky[y]) + a.kz[j] * a.kz[j];
    const int bin = digitize(sqrt(kk), a.edges, a.nedges);
    double w = (double)pk_w(j, d.nz);
    if (j > 0 && j < H) w = 0.5 * (w + (double)pk_w(d.nz - j, d.nz));
    const float m = (float)(g[bin] * w);
    const float2 v = *dst;
    *dst = make_float2(v.x * m, v.y * m);
  }
}

int launch_power_bins_adj(const Plan& p, float2* spec, const double* kx,
                          const double* ky, const double* kz,
                          const double* edges, int nedges, const double* g,
                          hipStream_t s) {
  PMX_REQUIRE(nedges >= 1 && nedges + 1 <= kMaxPkBins,
              "power spectrum: 1 <= number of edges < 8192");
  PkAdjArgs a;
  a.spec = spec;
  a.kx = kx;
  a.ky = ky;
  a.kz = kz;
  a.edges = edges;
  a.nedges = nedges;
  a.g = g;
  a.d = p.d;
  const long long n = (long long)p.d.nx * p.d.ny * (p.d.H + 1);
  const unsigned gx = (unsigned)std::min<long long>((n + 255) / 256, 16384);
  hipLaunchKernelGGL(k_pk_bin_adj, dim3(gx, p.d.B), dim3(256), 0, s, a);
  PMX_LAUNCH_CHECK();
  return kOk;
}

}  // namespace pmx
