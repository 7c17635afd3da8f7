// Index arithmetic of the Lagrangian-brick paint (paint.hip), shared by the
// kernels and by the CPU emulation that checks it (tests/cpu), so the
// emulation runs this exact code.  Host-only builds (g++) see plain inline
// functions; hipcc builds see __host__ __device__ ones.
#pragma once

#include <climits>

#ifdef __HIPCC__
#define PMX_HD __host__ __device__ __forceinline__
#else
#define PMX_HD inline
#endif

namespace pmx {

struct BrickGrid {
  int gx, gy, gz;     // particle grid (N = gx*gy*gz per batch member)
  int ex, ey, ez;     // brick size in particles
  int nbx, nby, nbz;  // bricks per dimension
  int nbricks;
  int f;              // mesh cells per particle spacing (pm_nc_factor)
};

constexpr int kBoxFloats = 16384;  // largest brick box: its arena slot (64 KB)
constexpr int kTile = 16;          // gather tile side

PMX_HD int imin(int a, int b) { return a < b ? a : b; }
PMX_HD int imax(int a, int b) { return a > b ? a : b; }

PMX_HD int imod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

// c - a modulo n for 0 <= c, a < n: the offset of mesh coordinate c inside
// a wrapped box that starts at a
PMX_HD int wrap_off(int c, int a, int n) {
  const int o = c - a;
  return o < 0 ? o + n : o;
}

// Brick grid of a particle grid gx*gy*gz on a mesh f times finer: bricks of
// about 8 x 16 x 16 mesh cells (fewer particles along an axis shorter than
// that).
PMX_HD BrickGrid make_brick_grid(int gx, int gy, int gz, int f) {
  BrickGrid g;
  g.gx = gx;
  g.gy = gy;
  g.gz = gz;
  g.f = f;
  const int bx = 8, by = 16, bz = 16;
  g.ex = imax(1, imin(gx, bx / f));
  g.ey = imax(1, imin(gy, by / f));
  g.ez = imax(1, imin(gz, bz / f));
  g.nbx = (gx + g.ex - 1) / g.ex;
  g.nby = (gy + g.ey - 1) / g.ey;
  g.nbz = (gz + g.ez - 1) / g.ez;
  g.nbricks = g.nbx * g.nby * g.nbz;
  return g;
}

// Particle block of brick bid: first particle (q0) and counts (cnt) per
// axis.
PMX_HD void brick_block(const BrickGrid& g, int bid, int q0[3], int cnt[3]) {
  const int kz = bid % g.nbz, ky = (bid / g.nbz) % g.nby,
            kx = bid / (g.nbz * g.nby);
  q0[0] = kx * g.ex;
  q0[1] = ky * g.ey;
  q0[2] = kz * g.ez;
  cnt[0] = imin(g.ex, g.gx - q0[0]);
  cnt[1] = imin(g.ey, g.gy - q0[1]);
  cnt[2] = imin(g.ez, g.gz - q0[2]);
}

// The deposit box of a brick from the floor bounds lo..hi of its particles'
// (scaled) positions: extent hi - lo + 2 per axis (the CIC corners), mode 1
// (an LDS box) when every extent fits the mesh, the volume fits the budget
// and the weights are finite; mode 2 (device atomics) otherwise; mode 0 for
// a brick without particles.  R: how far the box reaches beyond the
// brick's own mesh interval [q0 f, (q0 + cnt) f) (mode 1 only).
struct BrickBox {
  int ext[3];
  long long vol;
  int mode;
  int R;
};
PMX_HD BrickBox brick_box(const int lo[3], const int hi[3], const int n[3],
                          const int q0[3], const int cnt[3], int f, bool wfin,
                          long long cap) {
  BrickBox b;
  b.ext[0] = b.ext[1] = b.ext[2] = 0;
  b.vol = 0;
  b.mode = 0;
  b.R = 0;
  if (lo[0] > hi[0]) return b;  // no particles
  bool small = true;
  for (int d = 0; d < 3; ++d) {
    // long long: positions may span up to 2e9 cells
    const long long e = (long long)hi[d] - lo[d] + 2;
    small = small && e <= n[d];
    b.ext[d] = small ? (int)e : 0;
  }
  if (small) b.vol = (long long)b.ext[0] * b.ext[1] * b.ext[2];
  b.mode = (small && b.vol <= cap && wfin) ? 1 : 2;
  if (b.mode == 1) {
    for (int d = 0; d < 3; ++d) {
      b.R = imax(b.R, q0[d] * f - lo[d]);
      b.R = imax(b.R, lo[d] + b.ext[d] - (q0[d] + cnt[d]) * f);
    }
  } else {
    b.ext[0] = b.ext[1] = b.ext[2] = 0;
  }
  return b;
}

// does brick k (mesh interval [k*E, k*E + cnt*f) widened by R) touch the
// tile [t0, t0+T) modulo n?
PMX_HD bool brick_hits(int k, int e, int gdim, int f, int R, int t0, int T,
                       int n) {
  const int cnt = imin(e, gdim - k * e);
  if ((long long)cnt * f + 2LL * R >= n) return true;
  const int a0 = k * e * f - R;
  const int len = cnt * f + 2 * R;
  const int s = imod(a0, n);
  const int s1 = s + len;
  return (s < t0 + T && s1 > t0) || (s < t0 + T + n && s1 > t0 + n);
}

// does the wrapped box [am, am + ext) (0 <= am < n, ext <= n) meet the tile
// [t0, t0 + T) (t0 + T <= n)?
PMX_HD bool box_meets(int am, int ext, int t0, int T, int n) {
  return (am < t0 + T && am + ext > t0) || (am + ext - n > t0);
}

}  // namespace pmx
