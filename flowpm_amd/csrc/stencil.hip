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
// (F_x, F_y, F_z[, -]) mesh the particle passes gather.  D is antisymmetric,
// so the adjoint map from the three mesh cotangents to the potential's
// cotangent is -(D_x g_x + D_y g_y + D_z g_z), also one pass.
//
// One thread takes 4 consecutive z cells: one float4 and two float2 loads
// cover z - 2 .. z + 5, four float4 loads the y and the x neighbours.
#include "pmx_internal.h"

namespace pmx {

// Threads visit the cells in y bands: band of yb rows, then x, then band,
// then member (yb = ny is plain C order).  A row's x neighbours are read
// again 2 x steps later, i.e. 2 * yb rows later in this order; with whole
// planes at 512^3 (1 MB per plane and mesh) that reuse distance no longer
// fits one XCD's 4 MB L2.
struct GridDims {
  int B, nx, ny, nz, nq;  // nq = nz / 4
  int yb;                 // band height in rows, divides ny
  unsigned tot;           // B * nx * ny * nq threads (< 2^31, host-checked)
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
// index arithmetic in 32 bits: the launcher checks tot < 2^31
static __device__ __forceinline__ Cell cell_at(const GridDims& g, unsigned i) {
  Cell c;
  c.ok = i < g.tot;
  const unsigned j = c.ok ? i : 0u;
  const unsigned q = j % (unsigned)g.nq;
  unsigned r = j / (unsigned)g.nq;  // row in band order
  const unsigned yy = r % (unsigned)g.yb;
  r /= (unsigned)g.yb;
  c.x = (int)(r % (unsigned)g.nx);
  r /= (unsigned)g.nx;
  const unsigned nband = (unsigned)(g.ny / g.yb);
  c.y = (int)((r % nband) * (unsigned)g.yb + yy);
  c.b = (int)(r / nband);
  c.z0 = (int)(4 * q);
  return c;
}
static __device__ __forceinline__ unsigned block_first(void) {
  return (unsigned)xcd_block(blockIdx.x, gridDim.x) * blockDim.x;
}
static __device__ __forceinline__ Cell cell_of(const GridDims& g) {
  return cell_at(g, block_first() + threadIdx.x);
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

// NC = 4: fm[cell] = (D_x phi, D_y phi, D_z phi, 0) as one float4;
// NC = 3: (D_x phi, D_y phi, D_z phi) packed, 12 bytes per cell (the
// forward step's mesh: a quarter fewer bytes written here and gathered by
// the readout).  A thread's 4 cells are consecutive, so its own stores
// would leave each store instruction spread over 4 * NC-float strides; they
// go through LDS instead and every store instruction writes one contiguous
// 1 KB run per wave (measured at 256^3 with NC = 4: 78 us against 103).
template <int NC>
__global__ __launch_bounds__(256) void k_grad_mesh(const float* __restrict__ phi,
                                                   float* __restrict__ fm,
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
  __shared__ float4 st[256 * NC];
  const int t = threadIdx.x;
  if (NC == 4) {
    st[4 * t] = make_float4(gx.x, gy.x, gz.x, 0.0f);
    st[4 * t + 1] = make_float4(gx.y, gy.y, gz.y, 0.0f);
    st[4 * t + 2] = make_float4(gx.z, gy.z, gz.z, 0.0f);
    st[4 * t + 3] = make_float4(gx.w, gy.w, gz.w, 0.0f);
  } else {
    // cells 0..3 of the thread: x0 y0 z0 x1 | y1 z1 x2 y2 | z2 x3 y3 z3
    st[3 * t] = make_float4(gx.x, gy.x, gz.x, gx.y);
    st[3 * t + 1] = make_float4(gy.y, gz.y, gx.z, gy.z);
    st[3 * t + 2] = make_float4(gz.z, gx.w, gy.w, gz.w);
  }
  __syncthreads();
  // the block's threads cover 4 * 256 cells consecutive in memory from its
  // first one (the host keeps every block inside one band row group):
  // NC * 256 float4, 16-byte aligned (the first cell's offset is a multiple
  // of 4 * 256, or of the row length when rows are longer)
  const unsigned i0 = block_first();
  const Cell c0 = cell_at(g, i0);
  const long long cell0 =
      (((long long)c0.b * g.nx + c0.x) * g.ny + c0.y) * g.nz + c0.z0;
  const int nv = NC * (int)min(blockDim.x, g.tot - i0);
  float4* out = reinterpret_cast<float4*>(fm + NC * cell0);
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    const int q = t + 256 * k;
    if (q < nv) out[q] = st[q];
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

// Band height for nmesh input meshes: the largest divisor of ny whose
// reuse distance (2 x steps either way: 4 * yb rows of each input mesh)
// stays within 1 MB, among the heights that keep every 256-thread block
// inside one (band, x) row group, so each block's cells stay consecutive in
// memory; whole planes when they already fit.
static GridDims grid_dims(const Plan& p, int nmesh) {
  GridDims g{p.d.B, p.d.nx, p.d.ny, p.d.nz, p.d.nz / 4, p.d.ny, 0u};
  g.tot = (unsigned)((long long)g.B * g.nx * g.ny * g.nq);
  const long long budget = 1 << 20;
  auto dist = [&](long long yb) { return 4 * yb * p.d.nz * 4LL * nmesh; };
  if (dist(g.ny) <= budget) return g;
  int rpb = 0;  // rows per block
  if (g.nq >= 256 && g.nq % 256 == 0) rpb = 1;
  else if (g.nq < 256 && 256 % g.nq == 0) rpb = 256 / g.nq;
  if (rpb == 0 || g.ny % rpb != 0) return g;
  int yb = rpb;  // smallest legal band if none fits the budget
  for (int h = rpb; h < g.ny; h += rpb)
    if (g.ny % h == 0 && dist(h) <= budget) yb = h;
  g.yb = yb;
  return g;
}

static unsigned nblocks(const GridDims& g) { return (g.tot + 255u) / 256u; }

bool grad_stencil_ok(const Plan& p) {
  return p.d.nz % 4 == 0 && p.d.nx >= 4 && p.d.ny >= 4 && p.d.nz >= 8 &&
         (long long)p.d.B * p.d.nx * p.d.ny * (p.d.nz / 4) < (1LL << 31);
}

int launch_grad_mesh(const Plan& p, const float* phi, float* fm, int nc,
                     hipStream_t s) {
  PMX_REQUIRE(grad_stencil_ok(p),
              "gradient stencil: nz % 4 != 0, a side below 4 (nz below 8) "
              "or B * cells >= 2^33");
  PMX_REQUIRE(nc == 3 || nc == 4, "gradient stencil: 3 or 4 floats per cell");
  const GridDims g = grid_dims(p, 1);
  ProfScope prof_(kProfGrad, s);
  if (nc == 4)
    hipLaunchKernelGGL(k_grad_mesh<4>, dim3(nblocks(g)), dim3(256), 0, s, phi,
                       fm, g);
  else
    hipLaunchKernelGGL(k_grad_mesh<3>, dim3(nblocks(g)), dim3(256), 0, s, phi,
                       fm, g);
  PMX_LAUNCH_CHECK();
  return kOk;
}

int launch_grad_adj(const Plan& p, const float* gx, const float* gy,
                    const float* gz, float* out, hipStream_t s) {
  PMX_REQUIRE(grad_stencil_ok(p),
              "gradient stencil: nz % 4 != 0, a side below 4 (nz below 8) "
              "or B * cells >= 2^33");
  const GridDims g = grid_dims(p, 3);
  ProfScope prof_(kProfGradAdj, s);
  hipLaunchKernelGGL(k_grad_adj, dim3(nblocks(g)), dim3(256), 0, s, gx, gy,
                     gz, out, g);
  PMX_LAUNCH_CHECK();
  return kOk;
}

}  // namespace pmx
