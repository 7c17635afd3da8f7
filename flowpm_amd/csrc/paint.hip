// Atomic-free CIC paint of a Lagrangian particle grid ("brick paint").
//
// Same deposit as flowpm/utils.py:32-91 (8 corners, weight (wx*wy)*wz times
// the particle weight, periodic wrap), organised for gfx950:
//
// phase 1 (k_brick_deposit): one workgroup per Lagrangian brick of
//   particles (e.g. 8x16x16).  Particles that start close stay close, so
//   the brick's deposits fit a small box.  The workgroup finds the box,
//   accumulates all 8*C deposits per particle with LDS float atomics, and
//   stores the dense box to a scratch arena (bump-allocated, plain stores).
//   A brick whose box does not fit the LDS budget or the arena falls back
//   to device-scope float atomics straight into the mesh.
// phase 2 (k_brick_gather): one workgroup per 16^3 mesh tile sums, in a
//   fixed brick order, the parts of every box that overlap it and writes
//   the tile with plain stores.
//
// Device-scope float atomics cannot complete inside one XCD's L2 (the eight
// L2s are not coherent with each other), so the plain scatter pays a far
// memory transaction per deposit; here the mesh sees none in the common
// case.  C = 1 is the density paint, C = 3 the weighted paint of the force
// VJP (three mesh cotangents at once).
#include <climits>

#include "pmx_internal.h"

namespace pmx {

constexpr int kBoxFloats = 16384;  // LDS budget of one brick box (64 KB)
constexpr int kTile = 16;          // gather tile side

struct BrickArgs {
  float* mesh[3];        // C output meshes [B][nx][ny][nz] (zeroed)
  float* arena;          // box arena
  long long arena_cap;   // floats
  int* meta;             // 8 ints per brick: lo[3], ext[3], off lo/hi
  int* ctr;              // [0..1] arena bump (64-bit), [2] overflow count,
                         // [3] box radius
  const float* X;        // positions [B][N][3]
  const float* w;        // weights [B][N][wstride] or null
  int wstride;
  float wscale;
  float ps;              // position scale (pm_nc_factor)
  MeshDims d;            // mesh dims and N
  BrickGrid g;
};

__global__ void k_zero_ints(int* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

__device__ __forceinline__ bool ok_pos(float x, float y, float z) {
  return fabsf(x) < 1.0e9f && fabsf(y) < 1.0e9f && fabsf(z) < 1.0e9f;
}

__device__ __forceinline__ int imod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

template <int C>
__global__ __launch_bounds__(256) void k_brick_deposit(BrickArgs a) {
  __shared__ float box[kBoxFloats];
  __shared__ int red[4][6];
  __shared__ int sh[10];
  const BrickGrid& g = a.g;
  const int bid = blockIdx.x, b = blockIdx.y;
  const int tid = threadIdx.x;
  const int kz = bid % g.nbz, ky = (bid / g.nbz) % g.nby,
            kx = bid / (g.nbz * g.nby);
  const int q0x = kx * g.ex, q0y = ky * g.ey, q0z = kz * g.ez;
  const int cx = min(g.ex, g.gx - q0x), cy = min(g.ey, g.gy - q0y),
            cz = min(g.ez, g.gz - q0z);
  const int nloc = cx * cy * cz;
  const long long pb = (long long)b * a.d.N;
  const float ps = a.ps;

  int mn[3] = {INT_MAX, INT_MAX, INT_MAX}, mx[3] = {INT_MIN, INT_MIN, INT_MIN};
  for (int t = tid; t < nloc; t += 256) {
    const int lz = t % cz, r = t / cz, ly = r % cy, lx = r / cy;
    const long long p =
        pb + ((long long)(q0x + lx) * g.gy + (q0y + ly)) * g.gz + q0z + lz;
    const float x = a.X[3 * p] * ps, y = a.X[3 * p + 1] * ps,
                z = a.X[3 * p + 2] * ps;
    if (!ok_pos(x, y, z)) continue;
    const int f[3] = {(int)floorf(x), (int)floorf(y), (int)floorf(z)};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      mn[d] = min(mn[d], f[d]);
      mx[d] = max(mx[d], f[d]);
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    mn[d] = wave_min(mn[d]);
    mx[d] = wave_max(mx[d]);
  }
  const int wv = tid >> 6;
  if ((tid & 63) == 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      red[wv][d] = mn[d];
      red[wv][3 + d] = mx[d];
    }
  }
  __syncthreads();
  if (tid == 0) {
    int lo[3], hi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lo[d] = min(min(red[0][d], red[1][d]), min(red[2][d], red[3][d]));
      hi[d] = max(max(red[0][3 + d], red[1][3 + d]),
                  max(red[2][3 + d], red[3][3 + d]));
    }
    const int n[3] = {a.d.nx, a.d.ny, a.d.nz};
    int ext[3] = {0, 0, 0};
    bool any = lo[0] <= hi[0];
    long long vol = 0;
    bool small = true;
    if (any) {
      for (int d = 0; d < 3; ++d) {
        // long long: positions may span up to 2e9 cells
        const long long e = (long long)hi[d] - lo[d] + 2;
        small = small && e <= n[d];
        ext[d] = small ? (int)e : 0;
      }
      if (small) vol = (long long)ext[0] * ext[1] * ext[2];
    }
    int mode = 0;  // 0: nothing, 1: LDS box, 2: device atomics
    long long off = 0;
    if (any) {
      bool fits = small && vol * C <= kBoxFloats;
      if (fits) {
        // bump-allocate the arena (one atomic per workgroup)
        off = (long long)atomicAdd(reinterpret_cast<unsigned long long*>(
                                       a.ctr),
                                   (unsigned long long)(vol * C));
        fits = off + vol * C <= a.arena_cap;
      }
      mode = fits ? 1 : 2;
      if (mode == 2) atomicAdd(a.ctr + 2, 1);
      // box radius around the brick's own mesh interval, for the gather
      const int q0[3] = {q0x, q0y, q0z}, cc[3] = {cx, cy, cz};
      if (mode == 1) {
        int R = 0;
        for (int d = 0; d < 3; ++d) {
          const int f = g.f;
          R = max(R, q0[d] * f - lo[d]);
          R = max(R, lo[d] + ext[d] - (q0[d] + cc[d]) * f);
        }
        atomicMax(a.ctr + 3, R);
      }
    }
    int* m = a.meta + 8LL * ((long long)b * g.nbricks + bid);
    m[0] = lo[0];
    m[1] = lo[1];
    m[2] = lo[2];
    m[3] = mode == 1 ? ext[0] : 0;
    m[4] = mode == 1 ? ext[1] : 0;
    m[5] = mode == 1 ? ext[2] : 0;
    m[6] = (int)(off & 0x7fffffff);
    m[7] = (int)(off >> 31);
    sh[0] = mode;
    sh[1] = lo[0];
    sh[2] = lo[1];
    sh[3] = lo[2];
    sh[4] = ext[0];
    sh[5] = ext[1];
    sh[6] = ext[2];
    sh[7] = (int)vol;
    sh[8] = m[6];
    sh[9] = m[7];
  }
  __syncthreads();
  const int mode = sh[0];
  if (mode == 0) return;
  const int lo0 = sh[1], lo1 = sh[2], lo2 = sh[3];
  const int e1 = sh[5], e2 = sh[6], vol = sh[7];
  if (mode == 1) {
    for (int i = tid; i < vol * C; i += 256) box[i] = 0.0f;
    __syncthreads();
  }
  constexpr int OX[8] = {0, 1, 0, 0, 1, 1, 0, 1};
  constexpr int OY[8] = {0, 0, 1, 0, 1, 0, 1, 1};
  constexpr int OZ[8] = {0, 0, 0, 1, 0, 1, 1, 1};
  for (int t = tid; t < nloc; t += 256) {
    const int lz = t % cz, r = t / cz, ly = r % cy, lx = r / cy;
    const long long p =
        pb + ((long long)(q0x + lx) * g.gy + (q0y + ly)) * g.gz + q0z + lz;
    const float x = a.X[3 * p] * ps, y = a.X[3 * p + 1] * ps,
                z = a.X[3 * p + 2] * ps;
    if (!ok_pos(x, y, z)) continue;
    const float flx = floorf(x), fly = floorf(y), flz = floorf(z);
    const float w0x = 1.0f - fabsf(x - flx), w1x = 1.0f - fabsf(x - (flx + 1.0f));
    const float w0y = 1.0f - fabsf(y - fly), w1y = 1.0f - fabsf(y - (fly + 1.0f));
    const float w0z = 1.0f - fabsf(z - flz), w1z = 1.0f - fabsf(z - (flz + 1.0f));
    float wc[C];
#pragma unroll
    for (int c = 0; c < C; ++c)
      wc[c] = a.w ? a.w[p * a.wstride + c] * a.wscale : a.wscale;
    const int ix = (int)flx, iy = (int)fly, iz = (int)flz;
    if (mode == 1) {
      const int bx = ix - lo0, by = iy - lo1, bz = iz - lo2;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const float kw = ((OX[k] ? w1x : w0x) * (OY[k] ? w1y : w0y)) *
                         (OZ[k] ? w1z : w0z);
        const int li = ((bx + OX[k]) * e1 + (by + OY[k])) * e2 + bz + OZ[k];
#pragma unroll
        for (int c = 0; c < C; ++c) atomicAdd(&box[c * vol + li], wc[c] * kw);
      }
    } else {
      const MeshDims& d = a.d;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const float kw = ((OX[k] ? w1x : w0x) * (OY[k] ? w1y : w0y)) *
                         (OZ[k] ? w1z : w0z);
        const long long cell =
            (((long long)b * d.nx + imod(ix + OX[k], d.nx)) * d.ny +
             imod(iy + OY[k], d.ny)) * d.nz + imod(iz + OZ[k], d.nz);
#pragma unroll
        for (int c = 0; c < C; ++c) atomicAdd(a.mesh[c] + cell, wc[c] * kw);
      }
    }
  }
  if (mode == 1) {
    __syncthreads();
    const long long off = (long long)sh[8] | ((long long)sh[9] << 31);
    float* dst = a.arena + off;
    for (int i = tid; i < vol * C; i += 256) dst[i] = box[i];
  }
}

// overlap of the wrapped box interval [lo, lo+ext) (ext <= n) with the tile
// interval [t0, t0+T): up to two pieces (start in tile coords, box offset,
// length)
struct Pieces {
  int n;
  int ts[2], bo[2], len[2];
};
__device__ __forceinline__ Pieces overlap(int lo, int ext, int t0, int T,
                                          int n) {
  Pieces P;
  P.n = 0;
  const int a = imod(lo, n);
  // piece 1: wrapped [a, min(a+ext, n)), box offset 0
  // piece 2: wrapped [0, a+ext-n),       box offset n-a
  const int s[2] = {a, 0};
  const int e[2] = {min(a + ext, n), max(a + ext - n, 0)};
  const int o[2] = {0, n - a};
  for (int k = 0; k < 2; ++k) {
    const int b0 = max(s[k], t0), b1 = min(e[k], t0 + T);
    if (b1 > b0) {
      P.ts[P.n] = b0 - t0;
      P.bo[P.n] = o[k] + (b0 - s[k]);
      P.len[P.n] = b1 - b0;
      ++P.n;
    }
  }
  return P;
}

// does brick k (mesh interval [k*E, k*E + cnt*f) widened by R) touch the
// tile [t0, t0+T) modulo n?
__device__ __forceinline__ bool brick_hits(int k, int e, int gdim, int f,
                                           int R, int t0, int T, int n) {
  const int cnt = min(e, gdim - k * e);
  if ((long long)cnt * f + 2LL * R >= n) return true;
  const int a0 = k * e * f - R;
  const int len = cnt * f + 2 * R;
  const int s = imod(a0, n);
  const int s1 = s + len;
  return (s < t0 + T && s1 > t0) || (s < t0 + T + n && s1 > t0 + n);
}

template <int C>
__global__ __launch_bounds__(256) void k_brick_gather(BrickArgs a) {
  __shared__ float acc[C * kTile * kTile * kTile];
  __shared__ short cand[3][1024];
  __shared__ int ncand[3];
  const BrickGrid& g = a.g;
  const MeshDims& d = a.d;
  const int tid = threadIdx.x;
  const int b = blockIdx.y;
  const int ntx = (d.nx + kTile - 1) / kTile, nty = (d.ny + kTile - 1) / kTile,
            ntz = (d.nz + kTile - 1) / kTile;
  const int tt = (int)xcd_block(blockIdx.x, gridDim.x);
  const int tz = tt % ntz, ty = (tt / ntz) % nty, tx = tt / (ntz * nty);
  const int t0[3] = {tx * kTile, ty * kTile, tz * kTile};
  const int n[3] = {d.nx, d.ny, d.nz};
  const int T[3] = {min(kTile, d.nx - t0[0]), min(kTile, d.ny - t0[1]),
                    min(kTile, d.nz - t0[2])};
  for (int i = tid; i < C * kTile * kTile * kTile; i += 256) acc[i] = 0.0f;
  const int R = a.ctr[3];
  // candidate bricks per dimension (wave d builds list d)
  const int wv = tid >> 6, ln = tid & 63;
  if (wv < 3) {
    const int e = wv == 0 ? g.ex : (wv == 1 ? g.ey : g.ez);
    const int gd = wv == 0 ? g.gx : (wv == 1 ? g.gy : g.gz);
    const int nb = wv == 0 ? g.nbx : (wv == 1 ? g.nby : g.nbz);
    int base = 0;
    for (int k0 = 0; k0 < nb; k0 += 64) {
      const int k = k0 + ln;
      const bool hit =
          k < nb && brick_hits(k, e, gd, g.f, R, t0[wv], T[wv], n[wv]);
      const unsigned long long mask = __ballot(hit);
      if (hit) {
        const int pos = base + __popcll(mask & ((1ULL << ln) - 1ULL));
        if (pos < 1024) cand[wv][pos] = (short)k;
      }
      base += __popcll(mask);
    }
    if (ln == 0) ncand[wv] = min(base, 1024);
  }
  __syncthreads();
  const int nc0 = ncand[0], nc1 = ncand[1], nc2 = ncand[2];
  for (int i0 = 0; i0 < nc0; ++i0)
    for (int i1 = 0; i1 < nc1; ++i1)
      for (int i2 = 0; i2 < nc2; ++i2) {
        const int bid =
            (cand[0][i0] * g.nby + cand[1][i1]) * g.nbz + cand[2][i2];
        const int* m = a.meta + 8LL * ((long long)b * g.nbricks + bid);
        const int ex0 = m[3], ex1 = m[4], ex2 = m[5];
        if (ex0 == 0) continue;
        const Pieces px = overlap(m[0], ex0, t0[0], T[0], n[0]);
        if (px.n == 0) continue;
        const Pieces py = overlap(m[1], ex1, t0[1], T[1], n[1]);
        if (py.n == 0) continue;
        const Pieces pz = overlap(m[2], ex2, t0[2], T[2], n[2]);
        if (pz.n == 0) continue;
        const long long off = (long long)m[6] | ((long long)m[7] << 31);
        const float* src = a.arena + off;
        const int vol = ex0 * ex1 * ex2;
        for (int jx = 0; jx < px.n; ++jx)
          for (int jy = 0; jy < py.n; ++jy)
            for (int jz = 0; jz < pz.n; ++jz) {
              const int lx = px.len[jx], ly = py.len[jy], lz = pz.len[jz];
              const int cnt = lx * ly * lz;
              for (int i = tid; i < cnt; i += 256) {
                const int iz = i % lz, r = i / lz, iy = r % ly, ix = r / ly;
                const int ti = ((px.ts[jx] + ix) * kTile + py.ts[jy] + iy) *
                                   kTile + pz.ts[jz] + iz;
                const int bi = ((px.bo[jx] + ix) * ex1 + py.bo[jy] + iy) * ex2 +
                               pz.bo[jz] + iz;
#pragma unroll
                for (int c = 0; c < C; ++c)
                  acc[c * kTile * kTile * kTile + ti] += src[c * vol + bi];
              }
            }
        __syncthreads();
      }
  __syncthreads();
  const bool rmw = a.ctr[2] != 0;  // some bricks went straight to the mesh
  const int cnt = T[0] * T[1] * T[2];
  for (int i = tid; i < cnt; i += 256) {
    const int iz = i % T[2], r = i / T[2], iy = r % T[1], ix = r / T[1];
    const long long cell =
        (((long long)b * d.nx + t0[0] + ix) * d.ny + t0[1] + iy) * d.nz +
        t0[2] + iz;
    const int ti = (ix * kTile + iy) * kTile + iz;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const float v = acc[c * kTile * kTile * kTile + ti];
      a.mesh[c][cell] = rmw ? a.mesh[c][cell] + v : v;
    }
  }
}

BrickGrid make_brick_grid(int gx, int gy, int gz, int f, int C) {
  BrickGrid g;
  g.gx = gx;
  g.gy = gy;
  g.gz = gz;
  g.f = f;
  // mesh extent of a brick ~ 8 x 16 x 16 cells (C = 1) or 8 x 8 x 16 (C = 3)
  const int bx = 8, by = C == 1 ? 16 : 8, bz = 16;
  g.ex = std::max(1, std::min(gx, bx / f));
  g.ey = std::max(1, std::min(gy, by / f));
  g.ez = std::max(1, std::min(gz, bz / f));
  g.nbx = (gx + g.ex - 1) / g.ex;
  g.nby = (gy + g.ey - 1) / g.ey;
  g.nbz = (gz + g.ez - 1) / g.ez;
  g.nbricks = g.nbx * g.nby * g.nbz;
  return g;
}

bool brick_paint_ok(const MeshDims& d, const BrickGrid& g) {
  return g.gx * g.f == d.nx && g.gy * g.f == d.ny && g.gz * g.f == d.nz &&
         (long long)g.gx * g.gy * g.gz == d.N && g.nbx <= 1024 &&
         g.nby <= 1024 && g.nbz <= 1024;
}

long long brick_work_floats(const MeshDims& d, const BrickGrid& g, int C) {
  const long long M = (long long)d.B * d.nx * d.ny * d.nz;
  const long long arena = 3 * M * C;
  const long long meta = 8LL * d.B * g.nbricks;
  return arena + meta + 16;
}

int launch_brick_paint(float* const* mesh, int C, const float* X,
                       const float* w, int wstride, float wscale, float ps,
                       const MeshDims& d, const BrickGrid& g, float* work,
                       hipStream_t s) {
  if ((long long)d.B * d.N == 0) return kOk;
  const long long M = (long long)d.B * d.nx * d.ny * d.nz;
  BrickArgs a{};
  for (int c = 0; c < C; ++c) a.mesh[c] = mesh[c];
  a.arena = work;
  a.arena_cap = 3 * M * C;
  a.meta = reinterpret_cast<int*>(work + a.arena_cap);
  a.ctr = a.meta + 8LL * d.B * g.nbricks;
  a.X = X;
  a.w = w;
  a.wstride = wstride;
  a.wscale = wscale;
  a.ps = ps;
  a.d = d;
  a.g = g;
  hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, a.ctr, 8);
  PMX_LAUNCH_CHECK();
  {
    ProfScope prof_(kProfPaint, s);
    dim3 grid(g.nbricks, d.B);
    if (C == 1)
      hipLaunchKernelGGL(k_brick_deposit<1>, grid, dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL(k_brick_deposit<3>, grid, dim3(256), 0, s, a);
    PMX_LAUNCH_CHECK();
  }
  {
    ProfScope prof_(kProfPaintGather, s);
    const int nt = ((d.nx + kTile - 1) / kTile) * ((d.ny + kTile - 1) / kTile) *
                   ((d.nz + kTile - 1) / kTile);
    dim3 grid(nt, d.B);
    if (C == 1)
      hipLaunchKernelGGL(k_brick_gather<1>, grid, dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL(k_brick_gather<3>, grid, dim3(256), 0, s, a);
    PMX_LAUNCH_CHECK();
  }
  return kOk;
}

}  // namespace pmx
