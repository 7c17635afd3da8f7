// CPU emulation of the Lagrangian-brick paint (flowpm_amd/csrc/paint.hip)
// built on the kernels' own index code (brick_index.h): every brick's
// particles are deposited into its box (brick_block, brick_box), the boxes
// are gathered tile by tile through the same candidate tests (brick_hits,
// box_meets) and wrapped offsets (wrap_off), bricks that do not fit are
// added directly (the spill path), and the mesh is compared with a direct
// periodic CIC deposit (flowpm/utils.py:32-91) of the same float weights.
// Three weighted channels (the force-VJP paint) are checked the same way.
// Exits non-zero on any mismatch or out-of-box index.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "../../flowpm_amd/csrc/brick_index.h"

using namespace pmx;

// accumulation window of the emulated deposit, in floats
constexpr int kEmuWindow = 3200;

struct Case {
  const char* name;
  int gx, gy, gz, f;
  double amp;       // Gaussian displacement (particle spacings)
  double jump_p;    // probability of an extra jump
  double jump;      // jump size (mesh sides)
  unsigned seed;
  double shift;     // bulk displacement of every particle (spacings)
};

// the kernel's per-axis weights: float arithmetic as in k_brick_deposit
static void weights(float x, int* i0, float* w0, float* w1) {
  const float fl = floorf(x);
  *i0 = (int)fl;
  *w0 = 1.0f - fabsf(x - fl);
  *w1 = 1.0f - fabsf(x - (fl + 1.0f));
}

static int run(const Case& cs, int C) {
  const int n[3] = {cs.gx * cs.f, cs.gy * cs.f, cs.gz * cs.f};
  const BrickGrid g = make_brick_grid(cs.gx, cs.gy, cs.gz, cs.f);
  const long long N = (long long)cs.gx * cs.gy * cs.gz;
  const long long M = (long long)n[0] * n[1] * n[2];
  std::mt19937 rng(cs.seed);
  std::normal_distribution<double> gauss(0.0, cs.amp);
  std::uniform_real_distribution<double> uni(0.0, 1.0);
  // scaled positions x = X * f (float, as the kernel forms them)
  std::vector<float> pos(3 * N);
  for (long long p = 0; p < N; ++p) {
    const int q[3] = {(int)(p / ((long long)cs.gy * cs.gz)),
                      (int)((p / cs.gz) % cs.gy), (int)(p % cs.gz)};
    for (int d = 0; d < 3; ++d) {
      double X = q[d] + cs.shift + gauss(rng);
      if (uni(rng) < cs.jump_p) X += (uni(rng) - 0.5) * 2.0 * cs.jump * n[d];
      pos[3 * p + d] = (float)X * (float)cs.f;
    }
  }
  // channel weights: 1 for the density, random for the three VJP channels
  std::vector<float> wch(C * N, 1.0f);
  if (C > 1)
    for (auto& v : wch) v = (float)(2.0 * uni(rng) - 1.0);
  auto cell = [&](int x, int y, int z) {
    return ((long long)imod(x, n[0]) * n[1] + imod(y, n[1])) * n[2] +
           imod(z, n[2]);
  };
  // the 8 CIC corners of particle p, weighted by channel c
  auto deposit = [&](long long p, int c, auto&& add) {
    int i[3];
    float w[3][2];
    for (int d = 0; d < 3; ++d) weights(pos[3 * p + d], &i[d], &w[d][0], &w[d][1]);
    for (int k = 0; k < 8; ++k) {
      const int ox = k & 1, oy = (k >> 1) & 1, oz = k >> 2;
      add(i[0] + ox, i[1] + oy, i[2] + oz,
          (double)wch[C * p + c] * ((w[0][ox] * w[1][oy]) * w[2][oz]));
    }
  };
  // direct CIC of every channel
  std::vector<double> direct[3];
  for (int c = 0; c < C; ++c) {
    direct[c].assign(M, 0.0);
    for (long long p = 0; p < N; ++p)
      deposit(p, c, [&](int x, int y, int z, double v) { direct[c][cell(x, y, z)] += v; });
  }

  // phase 1: boxes
  struct Box {
    int lo[3], ext[3], mode, R;
    std::vector<double> v[3];
  };
  std::vector<Box> boxes(g.nbricks);
  int Rmax = 0, nspill = 0, nwrap = 0, nslab = 0;  // nslab: boxes in slabs
  for (int bid = 0; bid < g.nbricks; ++bid) {
    int q0[3], cnt[3];
    brick_block(g, bid, q0, cnt);
    int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
    auto for_particles = [&](auto&& fn) {
      for (int lx = 0; lx < cnt[0]; ++lx)
        for (int ly = 0; ly < cnt[1]; ++ly)
          for (int lz = 0; lz < cnt[2]; ++lz)
            fn(((long long)(q0[0] + lx) * g.gy + (q0[1] + ly)) * g.gz + q0[2] + lz);
    };
    for_particles([&](long long p) {
      for (int d = 0; d < 3; ++d) {
        const int fl = (int)floorf(pos[3 * p + d]);
        lo[d] = imin(lo[d], fl);
        hi[d] = imax(hi[d], fl);
      }
    });
    const BrickBox bx = brick_box(lo, hi, n, q0, cnt, g.f, true, kBoxFloats);
    const BrickBox& ob = bx;
    Box& B = boxes[bid];
    for (int d = 0; d < 3; ++d) {
      B.lo[d] = lo[d];
      B.ext[d] = ob.ext[d];
    }
    B.mode = ob.mode;
    B.R = ob.R;
    if (ob.mode == 1) {
      if (bx.vol > kBoxFloats || bx.vol != (long long)bx.ext[0] * bx.ext[1] * bx.ext[2] ||
          ob.vol != (long long)B.ext[0] * B.ext[1] * B.ext[2]) {
        printf("%s: brick %d box volume %lld\n", cs.name, bid, bx.vol);
        return 1;
      }
      Rmax = imax(Rmax, ob.R);
      for (int d = 0; d < 3; ++d)
        if (B.lo[d] < 0 || B.lo[d] + B.ext[d] > n[d]) {
          ++nwrap;
          break;
        }
      // deposit boxes, one per channel, through an accumulation window of
      // whole x planes as k_brick_deposit does (a window much smaller than
      // the kernel's, so that most boxes here take several slabs)
      bool bad = false;
      const int plane = bx.ext[1] * bx.ext[2];
      const int hx = imax(1, kEmuWindow / plane);
      for (int c = 0; c < C; ++c) {
        B.v[c].assign(bx.vol, 0.0);
        for (int x0 = 0; x0 < bx.ext[0]; x0 += hx) {
          const int nxw = imin(hx, bx.ext[0] - x0);
          std::vector<double> win((size_t)nxw * plane, 0.0);
          for_particles([&](long long p) {
            deposit(p, c, [&](int x, int y, int z, double v) {
              const int o[3] = {x - lo[0], y - lo[1], z - lo[2]};
              for (int d = 0; d < 3; ++d) bad = bad || o[d] < 0 || o[d] >= bx.ext[d];
              const int xx = o[0] - x0;
              if (bad || xx < 0 || xx >= nxw) return;  // another slab's corner
              win[((size_t)xx * bx.ext[1] + o[1]) * bx.ext[2] + o[2]] += v;
            });
          });
          for (size_t i = 0; i < win.size(); ++i)
            B.v[c][(size_t)x0 * plane + i] = win[i];
        }
        if (c == 0 && hx < bx.ext[0]) ++nslab;  // a box of several slabs
      }
      if (bad) {
        printf("%s: brick %d deposit outside its box\n", cs.name, bid);
        return 1;
      }
    } else if (ob.mode == 2) {
      ++nspill;
    }
  }

  // phase 2: every mesh cell from the boxes that meet its tile
  std::vector<double> mesh[3];
  for (int c = 0; c < C; ++c) mesh[c].assign(M, 0.0);
  std::vector<int> owner_hits(M, 0);
  // gather tiles: 16^3, 4 x 16 x 16 for three channels (as the kernel)
  const int tl[3] = {C == 1 ? kTile : kTile / 4, kTile, kTile};
  const int nt[3] = {(n[0] + tl[0] - 1) / tl[0], (n[1] + tl[1] - 1) / tl[1],
                     (n[2] + tl[2] - 1) / tl[2]};
  const int e[3] = {g.ex, g.ey, g.ez}, gd[3] = {g.gx, g.gy, g.gz},
            nb[3] = {g.nbx, g.nby, g.nbz};
  long long used = 0;
  for (int tx = 0; tx < nt[0]; ++tx)
    for (int ty = 0; ty < nt[1]; ++ty)
      for (int tz = 0; tz < nt[2]; ++tz) {
        const int t0[3] = {tx * tl[0], ty * tl[1], tz * tl[2]};
        int T[3];
        for (int d = 0; d < 3; ++d) T[d] = imin(tl[d], n[d] - t0[d]);
        std::vector<int> cand[3];
        for (int d = 0; d < 3; ++d)
          for (int k = 0; k < nb[d]; ++k)
            if (brick_hits(k, e[d], gd[d], g.f, Rmax, t0[d], T[d], n[d]))
              cand[d].push_back(k);
        std::vector<int> lst;
        for (int i0 : cand[0])
          for (int i1 : cand[1])
            for (int i2 : cand[2]) {
              const int bid = (i0 * g.nby + i1) * g.nbz + i2;
              const Box& B = boxes[bid];
              if (B.mode != 1) continue;
              bool meets = true;
              for (int d = 0; d < 3; ++d)
                meets = meets && box_meets(imod(B.lo[d], n[d]), B.ext[d], t0[d], T[d], n[d]);
              if (meets) lst.push_back(bid);
            }
        for (int x = t0[0]; x < t0[0] + T[0]; ++x)
          for (int y = t0[1]; y < t0[1] + T[1]; ++y)
            for (int z = t0[2]; z < t0[2] + T[2]; ++z) {
              const long long c = ((long long)x * n[1] + y) * n[2] + z;
              ++owner_hits[c];
              double acc[3] = {0.0, 0.0, 0.0};
              for (int bid : lst) {
                const Box& B = boxes[bid];
                const int ox = wrap_off(x, imod(B.lo[0], n[0]), n[0]);
                const int oy = wrap_off(y, imod(B.lo[1], n[1]), n[1]);
                const int oz = wrap_off(z, imod(B.lo[2], n[2]), n[2]);
                if (ox < B.ext[0] && oy < B.ext[1] && oz < B.ext[2]) {
                  for (int ch = 0; ch < C; ++ch)
                    acc[ch] += B.v[ch][((long long)ox * B.ext[1] + oy) * B.ext[2] + oz];
                  ++used;
                }
              }
              for (int ch = 0; ch < C; ++ch) mesh[ch][c] = acc[ch];
            }
      }
  // every box cell is read exactly once (a box never wraps onto itself:
  // ext <= n per axis)
  long long box_cells = 0;
  for (const Box& B : boxes)
    if (B.mode == 1) box_cells += (long long)B.v[0].size();
  if (used != box_cells) {
    printf("%s: %lld box cells read, %lld exist\n", cs.name, used, box_cells);
    return 1;
  }
  for (long long c = 0; c < M; ++c)
    if (owner_hits[c] != 1) {
      printf("%s: cell %lld written %d times\n", cs.name, c, owner_hits[c]);
      return 1;
    }
  // spill: the overflowing bricks' particles directly
  for (int bid = 0; bid < g.nbricks; ++bid) {
    if (boxes[bid].mode != 2) continue;
    int q0[3], cnt[3];
    brick_block(g, bid, q0, cnt);
    for (int lx = 0; lx < cnt[0]; ++lx)
      for (int ly = 0; ly < cnt[1]; ++ly)
        for (int lz = 0; lz < cnt[2]; ++lz) {
          const long long p =
              ((long long)(q0[0] + lx) * g.gy + (q0[1] + ly)) * g.gz + q0[2] + lz;
          for (int c = 0; c < C; ++c)
            deposit(p, c, [&](int x, int y, int z, double v) {
              mesh[c][cell(x, y, z)] += v;
            });
        }
  }
  double err = 0.0, mx = 0.0, tot = 0.0;
  for (int c = 0; c < C; ++c)
    for (long long q = 0; q < M; ++q) {
      err = fmax(err, fabs(mesh[c][q] - direct[c][q]));
      mx = fmax(mx, fabs(direct[c][q]));
      tot += mesh[c][q];
    }
  printf("%-24s C=%d mesh %dx%dx%d f=%d bricks %d (wrapping %d, spilled %d, "
         "in slabs %d) R %d max err %.2e sum %.6f\n", cs.name, C, n[0], n[1], n[2],
         cs.f, g.nbricks, nwrap, nspill, nslab, Rmax, err, tot);
  return err <= 1e-9 * fmax(1.0, mx) ? 0 : 1;
}

int main() {
  const Case cases[] = {
      {"small displacements", 48, 48, 48, 1, 0.8, 0.0, 0.0, 1},
      {"f=2", 32, 32, 32, 2, 0.5, 0.0, 0.0, 2},
      {"rectangular", 32, 64, 48, 1, 1.5, 0.0, 0.0, 3},
      {"jumps, spilled bricks", 48, 32, 32, 1, 0.7, 0.0005, 1.5, 4},
      {"grid smaller than brick", 8, 8, 8, 1, 0.3, 0.0, 0.0, 5},
      {"partial bricks", 40, 24, 40, 1, 1.0, 0.0, 0.0, 6},
      {"partial bricks f=2", 20, 12, 20, 2, 0.6, 0.0, 0.0, 7},
      {"boxes over the budget", 32, 32, 32, 1, 6.0, 0.0, 0.0, 8},
      {"many wraps", 16, 16, 16, 1, 40.0, 0.0, 0.0, 9},
      {"f=4", 16, 16, 16, 4, 0.3, 0.0, 0.0, 10},
      {"bulk shift down", 48, 32, 32, 1, 0.3, 0.0, 0.0, 11, -3.5},
      {"bulk shift up", 48, 32, 32, 1, 0.3, 0.0, 0.0, 12, 2.5},
  };
  int bad = 0;
  for (const Case& c : cases) bad += run(c, 1);
  for (const Case& c : cases) bad += run(c, 3);
  if (bad) {
    printf("FAILED %d case(s)\n", bad);
    return 1;
  }
  printf("ok\n");
  return 0;
}
