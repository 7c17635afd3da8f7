// Gradient of the potential in real space (flowpm/kernels.py:51-78).
//
// The reference multiplies the potential's spectrum by
//   i g(k_d),  g = (8 sin w - sin 2w) / 6,  w = 2 pi k_d / n_d,
// once per axis and inverse-transforms each product.  Under the forward
// transform sum_n x(n) exp(-2 pi i k n / N) that multiplier is exactly the
// transfer function of the periodic central difference
//   (D x)(n) = 2/3 (x(n+1) - x(n-1)) - 1/12 (x(n+2) - x(n-2)),
// (the reference's kernel is the fourth-order finite-difference gradient
// written in Fourier space; at the Nyquist mode both are zero), so the three
// force meshes are D_x, D_y, D_z of ONE inverse-transformed potential: one
// transform chain instead of three, and this pass writes the interleaved
// (F_x, F_y, F_z, -) mesh the particle passes gather.  D is antisymmetric,
// so the adjoint map from the three mesh cotangents to the potential's
// cotangent is -(D_x g_x + D_y g_y + D_z g_z), also one pass.
//
// One thread takes 4 consecutive z cells: one float4 and two float2 loads
// cover z - 2 .. z + 5, four float4 loads the y and the x neighbours.
#include "pmx_internal.h"

namespace pmx {

struct GridDims {
  int B, nx, ny, nz, nq;  // nq = nz / 4
};

static __device__ __forceinline__ float4 ld4(const float* p) {
  return *reinterpret_cast<const float4*>(p);
}
static __device__ __forceinline__ float2 ld2(const float* p) {
  return *reinterpret_cast<const float2*>(p);
}
static __device__ __forceinline__ float dcoef(float m2, float m1, float p1,
                                       float p2) {
  return (2.0f / 3.0f) * (p1 - m1) - (1.0f / 12.0f) * (p2 - m2);
}
static __device__ __forceinline__ float4 dvec(float4 m2, float4 m1, float4 p1,
                                       float4 p2) {
  return make_float4(dcoef(m2.x, m1.x, p1.x, p2.x),
                     dcoef(m2.y, m1.y, p1.y, p2.y),
                     dcoef(m2.z, m1.z, p1.z, p2.z),
                     dcoef(m2.w, m1.w, p1.w, p2.w));
}

// cell index of thread i: (b, x, y, 4 q)
struct Cell {
  int b, x, y, z0;
  bool ok;
};
static __device__ __forceinline__ Cell cell_of(const GridDims& g) {
  const long long tot = (long long)g.B * g.nx * g.ny * g.nq;
  const long long i =
      xcd_block(blockIdx.x, gridDim.x) * (long long)blockDim.x + threadIdx.x;
  Cell c;
  c.ok = i < tot;
  const long long j = c.ok ? i : 0;
  const int q = (int)(j % g.nq);
  const long long r = j / g.nq;
  c.y = (int)(r % g.ny);
  const long long r2 = r / g.ny;
  c.x = (int)(r2 % g.nx);
  c.b = (int)(r2 / g.nx);
  c.z0 = 4 * q;
  return c;
}

// D_z of the 4 cells z0 .. z0 + 3 of the row at p (row length nz)
static __device__ __forceinline__ float4 dz_row(const float* row, int z0, int nz) {
  const int zm = z0 == 0 ? nz - 2 : z0 - 2;
  const int zp = z0 + 4 == nz ? 0 : z0 + 4;
  const float2 a = ld2(row + zm);  // z0 - 2, z0 - 1
  const float4 c = ld4(row + z0);  // z0 .. z0 + 3
  const float2 e = ld2(row + zp);  // z0 + 4, z0 + 5
  return make_float4(dcoef(a.x, a.y, c.y, c.z), dcoef(a.y, c.x, c.z, c.w),
                     dcoef(c.x, c.y, c.w, e.x), dcoef(c.y, c.z, e.x, e.y));
}

static __device__ __forceinline__ int wrap(int v, int n) {
  return v < 0 ? v + n : (v >= n ? v - n : v);
}

// (b, x, y) row offsets of the y and x neighbours at distance 1 and 2
struct Nbr {
  long long ym2, ym1, yp1, yp2, xm2, xm1, xp1, xp2, c;
};
static __device__ __forceinline__ Nbr nbr_of(const Cell& c, const GridDims& g) {
  const long long nz = g.nz;
  const long long plane = (long long)g.ny * nz;
  const long long xb = ((long long)c.b * g.nx + c.x) * plane;
  const long long yb = (long long)c.y * nz;
  Nbr n;
  n.c = xb + yb;
  n.ym2 = xb + wrap(c.y - 2, g.ny) * nz;
  n.ym1 = xb + wrap(c.y - 1, g.ny) * nz;
  n.yp1 = xb + wrap(c.y + 1, g.ny) * nz;
  n.yp2 = xb + wrap(c.y + 2, g.ny) * nz;
  const long long bb = (long long)c.b * g.nx * plane;
  n.xm2 = bb + wrap(c.x - 2, g.nx) * plane + yb;
  n.xm1 = bb + wrap(c.x - 1, g.nx) * plane + yb;
  n.xp1 = bb + wrap(c.x + 1, g.nx) * plane + yb;
  n.xp2 = bb + wrap(c.x + 2, g.nx) * plane + yb;
  return n;
}

// f4[cell] = (D_x phi, D_y phi, D_z phi, 0).  A thread's 4 cells are
// consecutive, so its own 4 float4 stores would leave each store
// instruction spread over 64-byte strides; they go through LDS instead and
// every store instruction writes one contiguous 1 KB run per wave (measured
// at 256^3: 78 us against 103).
__global__ __launch_bounds__(256) void k_grad_f4(const float* __restrict__ phi,
                                                 float4* __restrict__ f4,
                                                 GridDims g) {
  const Cell c = cell_of(g);
  float4 gx = make_float4(0.f, 0.f, 0.f, 0.f), gy = gx, gz = gx;
  if (c.ok) {
    const Nbr n = nbr_of(c, g);
    const int z0 = c.z0;
    gz = dz_row(phi + n.c, z0, g.nz);
    gy = dvec(ld4(phi + n.ym2 + z0), ld4(phi + n.ym1 + z0),
              ld4(phi + n.yp1 + z0), ld4(phi + n.yp2 + z0));
    gx = dvec(ld4(phi + n.xm2 + z0), ld4(phi + n.xm1 + z0),
              ld4(phi + n.xp1 + z0), ld4(phi + n.xp2 + z0));
  }
  __shared__ float4 st[256 * 4];
  const int t = threadIdx.x;
  st[4 * t] = make_float4(gx.x, gy.x, gz.x, 0.0f);
  st[4 * t + 1] = make_float4(gx.y, gy.y, gz.y, 0.0f);
  st[4 * t + 2] = make_float4(gx.z, gy.z, gz.z, 0.0f);
  st[4 * t + 3] = make_float4(gx.w, gy.w, gz.w, 0.0f);
  __syncthreads();
  // the block's threads cover 4 * 256 consecutive cells from its first one
  const long long tot = (long long)g.B * g.nx * g.ny * g.nq;
  const long long i0 =
      xcd_block(blockIdx.x, gridDim.x) * (long long)blockDim.x;
  const long long ncell = 4 * min((long long)blockDim.x, tot - i0);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int q = t + 256 * k;
    if (q < ncell) f4[4 * i0 + q] = st[q];
  }
}

// out = -(D_x gx + D_y gy + D_z gz)
__global__ __launch_bounds__(256) void k_grad_adj(const float* __restrict__ gx,
                                                  const float* __restrict__ gy,
                                                  const float* __restrict__ gz,
                                                  float* __restrict__ out,
                                                  GridDims g) {
  const Cell c = cell_of(g);
  if (!c.ok) return;
  const Nbr n = nbr_of(c, g);
  const int z0 = c.z0;
  const float4 a = dz_row(gz + n.c, z0, g.nz);
  const float4 b = dvec(ld4(gy + n.ym2 + z0), ld4(gy + n.ym1 + z0),
                        ld4(gy + n.yp1 + z0), ld4(gy + n.yp2 + z0));
  const float4 d = dvec(ld4(gx + n.xm2 + z0), ld4(gx + n.xm1 + z0),
                        ld4(gx + n.xp1 + z0), ld4(gx + n.xp2 + z0));
  *reinterpret_cast<float4*>(out + n.c + z0) =
      make_float4(-((d.x + b.x) + a.x), -((d.y + b.y) + a.y),
                  -((d.z + b.z) + a.z), -((d.w + b.w) + a.w));
}

static GridDims grid_dims(const Plan& p) {
  return GridDims{p.d.B, p.d.nx, p.d.ny, p.d.nz, p.d.nz / 4};
}

static unsigned nblocks(const GridDims& g) {
  const long long tot = (long long)g.B * g.nx * g.ny * g.nq;
  return (unsigned)((tot + 255) / 256);
}

bool grad_stencil_ok(const Plan& p) {
  return p.d.nz % 4 == 0 && p.d.nx >= 4 && p.d.ny >= 4 && p.d.nz >= 8;
}

int launch_grad_f4(const Plan& p, const float* phi, float* f4,
                   hipStream_t s) {
  PMX_REQUIRE(grad_stencil_ok(p), "gradient stencil: unsupported mesh");
  const GridDims g = grid_dims(p);
  ProfScope prof_(kProfGrad, s);
  hipLaunchKernelGGL(k_grad_f4, dim3(nblocks(g)), dim3(256), 0, s, phi,
                     reinterpret_cast<float4*>(f4), g);
  PMX_LAUNCH_CHECK();
  return kOk;
}

int launch_grad_adj(const Plan& p, const float* gx, const float* gy,
                    const float* gz, float* out, hipStream_t s) {
  PMX_REQUIRE(grad_stencil_ok(p), "gradient stencil: unsupported mesh");
  const GridDims g = grid_dims(p);
  ProfScope prof_(kProfGradAdj, s);
  hipLaunchKernelGGL(k_grad_adj, dim3(nblocks(g)), dim3(256), 0, s, gx, gy,
                     gz, out, g);
  PMX_LAUNCH_CHECK();
  return kOk;
}

}  // namespace pmx
