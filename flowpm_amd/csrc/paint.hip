// Atomic-free CIC paint of a Lagrangian particle grid ("brick paint").
//
// Same deposit as flowpm/utils.py:32-91 (8 corners, weight (wx*wy)*wz times
// the particle weight, periodic wrap), organised for gfx950:
//
// phase 1 (k_brick_deposit): one workgroup per Lagrangian brick of
//   particles (e.g. 8x16x16).  Particles that start close stay close, so
//   the brick's deposits fit a small box.  The workgroup finds the box,
//   accumulates the 8 deposits per particle in LDS (in x slabs when the box
//   exceeds the 39.5 KB window) as 32-bit fixed point
//   (integer LDS atomics run several times faster than float ones on
//   gfx950, and integer sums do not depend on their order), and stores the
//   dense box, converted back to float, to its own arena slot; the C
//   channels are deposited one after the other through the same box.  The
//   fixed-point scale is a power of two chosen per brick and channel from
//   the sum S of the channel's |weights| so no cell can overflow: the scale
//   is 2^(30 - ex) with S = m 2^ex, m in [0.5, 1), so one deposit is rounded
//   by at most half a unit, 2^(ex - 31) <= 2^-30 S.
//   A brick whose box exceeds its 64 KB arena slot falls back to
//   device-scope float atomics into the finished mesh (k_brick_spill, after
//   phase 2).
// phase 2 (k_brick_gather): one workgroup per mesh tile (16^3; 4 x 16 x 16
//   for the three channels) evaluates the
//   candidate bricks in parallel, compacts those whose boxes meet the tile
//   (in candidate order), and every thread adds them, in that order, to the
//   cells of the x row it owns; the tile is written with plain stores, no barrier per
//   brick, so the mesh is never cleared first.
//
// Device-scope float atomics cannot complete inside one XCD's L2 (the eight
// L2s are not coherent with each other), so the plain scatter pays a far
// memory transaction per deposit; here the mesh sees none in the common
// case.  C = 1 is the density paint, C = 3 the weighted paint of the force
// VJP (three mesh cotangents at once).
#include "pmx_internal.h"

namespace pmx {

struct BrickArgs {
  float* mesh[3];        // C output meshes [B][nx][ny][nz] (overwritten)
  float* arena;          // box arena
  long long arena_cap;   // floats
  long long slot;        // arena floats per brick
  int accum;             // gather adds to the mesh instead of overwriting
  int* meta;             // 8 ints per brick: lo[3], ext[3], R (how far the
                         // box reaches past the brick), overflow flag
  int* ctr;              // [2] overflow count, [3] box radius (written by
                         // k_brick_reduce between the two phases)
  const float* X;        // positions [B][N][3]
  const float* w;        // weights [B][N][wstride] or null
  int wstride;
  float wscale;
  float ps;              // position scale (pm_nc_factor)
  MeshDims d;            // mesh dims and N
  BrickGrid g;
};

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

// deposit workgroup: 512 threads (8 waves; four workgroups per CU give 32
// waves to hide the loads and the LDS atomics), 4 particles per thread;
// bricks hold <= 2048 particles.  A box (<= kBoxFloats, its arena slot) is
// accumulated through an LDS window of kWinFloats ints: four windows and
// the workgroups' other LDS fit a CU's 160 KB; a larger box is deposited
// in slabs of whole x planes, one pass over the particles per slab.
constexpr int kDT = 512;
constexpr int kWinFloats = 10112;
// gather tile x lengths of the density and of the three-channel paint
constexpr int kGatherTx1 = 16, kGatherTx3 = 4;
constexpr int kDW = kDT / 64;
constexpr int kPPT = 2048 / kDT;

// one brick's particles in registers, kPPT per thread (raw positions and
// weights; lanes past the brick's end hold ok = false)
template <int C>
struct BrickRegs {
  float px[kPPT], py[kPPT], pz[kPPT], wc[kPPT][C];
  bool ok[kPPT];
};

// every load issued up front, branch-free: lanes past the brick's end load
// its first particle and drop it; the weight test is one uniform branch
template <int C>
__device__ __forceinline__ void load_brick(const BrickArgs& a, int bid, int b,
                                           int tid, BrickRegs<C>& r) {
  const BrickGrid& g = a.g;
  int q0[3], cnt[3];
  brick_block(g, bid, q0, cnt);
  const int cy = cnt[1], cz = cnt[2];
  const int nloc = cnt[0] * cy * cz;
  const long long pb = (long long)b * a.d.N;
  long long pidx[kPPT];
#pragma unroll
  for (int u = 0; u < kPPT; ++u) {
    const int t = tid + u * kDT;
    r.ok[u] = t < nloc;
    const int tt = r.ok[u] ? t : 0;
    const int lz = tt % cz, q = tt / cz, ly = q % cy, lx = q / cy;
    pidx[u] = pb + ((long long)(q0[0] + lx) * g.gy + (q0[1] + ly)) * g.gz +
              q0[2] + lz;
    r.px[u] = a.X[3 * pidx[u]];
    r.py[u] = a.X[3 * pidx[u] + 1];
    r.pz[u] = a.X[3 * pidx[u] + 2];
  }
  if (a.w) {
#pragma unroll
    for (int u = 0; u < kPPT; ++u)
#pragma unroll
      for (int c = 0; c < C; ++c) r.wc[u][c] = a.w[pidx[u] * a.wstride + c];
  }
}

template <int C>
// (launch bounds: at least 8 waves per SIMD, i.e. at most 64 VGPRs, so that
// four 512-thread workgroups share a CU; their LDS windows fit its 160 KB)
__global__ __launch_bounds__(kDT, 8) void k_brick_deposit(BrickArgs a) {
  __shared__ int box[kWinFloats];  // fixed-point accumulators
  __shared__ int red[kDW][6];
  __shared__ float redw[kDW][C];
  __shared__ int sh[10];
  __shared__ float shs[C][2];
  const BrickGrid& g = a.g;
  // XCD-aware: each XCD deposits one contiguous run of bricks, so the
  // cache lines that z-neighbouring bricks share are fetched into one L2
  const int bid = (int)xcd_block(blockIdx.x, gridDim.x);
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  const float ps = a.ps;
  BrickRegs<C> cur;
  load_brick<C>(a, bid, b, tid, cur);
  {
    int q0[3], cnt[3];
    brick_block(g, bid, q0, cnt);
    // the brick's registers, modified in place
    float(&px)[kPPT] = cur.px;
    float(&py)[kPPT] = cur.py;
    float(&pz)[kPPT] = cur.pz;
    float(&wc)[kPPT][C] = cur.wc;
    bool(&ok)[kPPT] = cur.ok;
#pragma unroll
    for (int u = 0; u < kPPT; ++u) {
      px[u] = ok[u] ? px[u] : 0.0f;
      py[u] = ok[u] ? py[u] : 0.0f;
      pz[u] = ok[u] ? pz[u] : 0.0f;
#pragma unroll
      for (int c = 0; c < C; ++c)
        wc[u][c] = ok[u] ? (a.w ? wc[u][c] * a.wscale : a.wscale) : 0.0f;
    }
    int mn[3] = {INT_MAX, INT_MAX, INT_MAX},
        mx[3] = {INT_MIN, INT_MIN, INT_MIN};
    float sw[C];  // per channel: sum over particles of |weight|
#pragma unroll
    for (int c = 0; c < C; ++c) sw[c] = 0.0f;
#pragma unroll
    for (int u = 0; u < kPPT; ++u) {
      px[u] *= ps;
      py[u] *= ps;
      pz[u] *= ps;
      ok[u] = ok[u] && ok_pos(px[u], py[u], pz[u]);
      if (ok[u]) {
        const int f[3] = {(int)floorf(px[u]), (int)floorf(py[u]),
                          (int)floorf(pz[u])};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          mn[d] = min(mn[d], f[d]);
          mx[d] = max(mx[d], f[d]);
        }
#pragma unroll
        for (int c = 0; c < C; ++c) sw[c] += fabsf(wc[u][c]);
      }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      mn[d] = wave_min(mn[d]);
      mx[d] = wave_max(mx[d]);
    }
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) sw[c] += __shfl_xor(sw[c], o);
    const int wv = tid >> 6;
    if ((tid & 63) == 0) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        red[wv][d] = mn[d];
        red[wv][3 + d] = mx[d];
      }
#pragma unroll
      for (int c = 0; c < C; ++c) redw[wv][c] = sw[c];
    }
    __syncthreads();
    if (tid == 0) {
      int lo[3], hi[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        lo[d] = red[0][d];
        hi[d] = red[0][3 + d];
        for (int w = 1; w < kDW; ++w) {
          lo[d] = min(lo[d], red[w][d]);
          hi[d] = max(hi[d], red[w][3 + d]);
        }
      }
      const int n[3] = {a.d.nx, a.d.ny, a.d.nz};
      const long long off = ((long long)b * g.nbricks + bid) * a.slot;
      // fixed-point scale 2^e with (sum of |weights|) * 2^e < 2^30: no cell
      // total (nor a partial sum) can leave the int32 range
      bool wfin = true;
      for (int c = 0; c < C; ++c) {
        float swt = redw[0][c];
        for (int w = 1; w < kDW; ++w) swt += redw[w][c];
        int e = 0;
        wfin = wfin && swt <= 3.0e38f;  // false for inf / nan weights
        if (swt <= 3.0e38f && swt > 0.0f) {
          int ex;
          frexpf(swt, &ex);  // swt = m 2^ex, m in [0.5, 1)
          e = min(max(30 - ex, -120), 120);
        }
        shs[c][0] = ldexpf(1.0f, e);
        shs[c][1] = ldexpf(1.0f, -e);
      }
      // every brick owns a fixed arena slot: no shared allocator, and no
      // shared counter either (k_brick_reduce takes the radius maximum)
      const BrickBox bx = brick_box(lo, hi, n, q0, cnt, g.f, wfin, kBoxFloats);
      const int mode = bx.mode;
      const int* ext = bx.ext;
      const long long vol = bx.vol;
      int* m = a.meta + 8LL * ((long long)b * g.nbricks + bid);
      m[0] = lo[0];
      m[1] = lo[1];
      m[2] = lo[2];
      m[3] = ext[0];
      m[4] = ext[1];
      m[5] = ext[2];
      m[6] = bx.R;
      m[7] = mode == 2 ? 1 : 0;
      sh[0] = mode;
      sh[1] = lo[0];
      sh[2] = lo[1];
      sh[3] = lo[2];
      sh[4] = ext[0];
      sh[5] = ext[1];
      sh[6] = ext[2];
      sh[7] = (int)vol;
      sh[8] = (int)(off & 0x7fffffff);
      sh[9] = (int)(off >> 31);
    }
    __syncthreads();
    const int mode = sh[0];
    const int lo0 = sh[1], lo1 = sh[2], lo2 = sh[3];
    const int e0 = sh[4], e1 = sh[5], e2 = sh[6], vol = sh[7];
    const long long off = (long long)sh[8] | ((long long)sh[9] << 31);
    // mode 0: no particles; mode 2: k_brick_spill adds it after the gather
    if (mode != 1) return;
    constexpr int OX[8] = {0, 1, 0, 0, 1, 1, 0, 1};
    constexpr int OY[8] = {0, 0, 1, 0, 1, 0, 1, 1};
    constexpr int OZ[8] = {0, 0, 0, 1, 0, 1, 1, 1};
    // x planes of the box per LDS window: boxes larger than the window are
    // deposited in several passes over x slabs
    const int plane = e1 * e2;
    const int hx = kWinFloats / plane;  // >= 1: plane <= vol / 2 <= 8192
    const bool clip = hx < e0;
    // clears the window and deposits channel c of the brick's particles
    // whose corners fall in x planes [x0, x0 + nxw) of the box
    auto deposit = [&](int c, float fxs, int x0, int nxw) {
      const int wv = nxw * plane;
      for (int i = tid; i < wv; i += kDT) box[i] = 0;
      __syncthreads();
      // the positions are opaque here, so the corner weights are computed
      // per pass instead of being kept live across the passes and channels
#pragma unroll
      for (int u = 0; u < kPPT; ++u)
        asm volatile("" : "+v"(px[u]), "+v"(py[u]), "+v"(pz[u]));
#pragma unroll
      for (int u = 0; u < kPPT; ++u) {
        if (!ok[u]) continue;
        const float x = px[u], y = py[u], z = pz[u];
        const float flx = floorf(x), fly = floorf(y), flz = floorf(z);
        const float w0x = 1.0f - fabsf(x - flx), w1x = 1.0f - fabsf(x - (flx + 1.0f));
        const float w0y = 1.0f - fabsf(y - fly), w1y = 1.0f - fabsf(y - (fly + 1.0f));
        const float w0z = 1.0f - fabsf(z - flz), w1z = 1.0f - fabsf(z - (flz + 1.0f));
        const int bx = (int)flx - lo0 - x0, by = (int)fly - lo1, bz = (int)flz - lo2;
        const float wq = wc[u][c] * fxs;
        if (!clip) {  // the whole box in one window: no x test
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            const float kw = ((OX[k] ? w1x : w0x) * (OY[k] ? w1y : w0y)) *
                             (OZ[k] ? w1z : w0z);
            const int li =
                ((bx + OX[k]) * e1 + (by + OY[k])) * e2 + bz + OZ[k];
            atomicAdd(&box[li], __float2int_rn(wq * kw));
          }
        } else {
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            const int xx = bx + OX[k];
            if ((unsigned)xx >= (unsigned)nxw) continue;
            const float kw = ((OX[k] ? w1x : w0x) * (OY[k] ? w1y : w0y)) *
                             (OZ[k] ? w1z : w0z);
            const int li = (xx * e1 + (by + OY[k])) * e2 + bz + OZ[k];
            atomicAdd(&box[li], __float2int_rn(wq * kw));
          }
        }
      }
      __syncthreads();
    };
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const float fxinv = shs[c][1];
      for (int x0 = 0; x0 < e0; x0 += hx) {
        const int nxw = min(hx, e0 - x0);
        deposit(c, shs[c][0], x0, nxw);
        float* dst = a.arena + off + (long long)c * vol + (long long)x0 * plane;
        const int wv = nxw * plane;
        for (int i = tid; i < wv; i += kDT) dst[i] = (float)box[i] * fxinv;
        if (c + 1 < C || x0 + hx < e0) __syncthreads();  // window reused
      }
    }
  }
}

// Radius maximum and overflow count over all bricks (one workgroup), read by
// the gather: ctr[2] = bricks that went to device atomics, ctr[3] = max R.
__global__ __launch_bounds__(1024) void k_brick_reduce(const int* __restrict__ meta,
                                                       long long nb,
                                                       int* __restrict__ ctr) {
  __shared__ int smax[16], ssum[16];
  int r = 0, o = 0;
  // (R, overflow) of 8 bricks per thread and round, all loads issued before
  // the first is used (indices past the end reload the last brick and are
  // dropped)
  constexpr int U = 8;
  for (long long i0 = threadIdx.x; i0 < nb; i0 += 1024LL * U) {
    int vr[U], vo[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long i = min(i0 + 1024LL * u, nb - 1);
      vr[u] = meta[8 * i + 6];
      vo[u] = meta[8 * i + 7];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool in = i0 + 1024LL * u < nb;
      r = max(r, in ? vr[u] : 0);
      o += in ? vo[u] : 0;
    }
  }
#pragma unroll
  for (int k = 32; k > 0; k >>= 1) {
    r = max(r, __shfl_xor(r, k));
    o += __shfl_xor(o, k);
  }
  if ((threadIdx.x & 63) == 0) {
    smax[threadIdx.x >> 6] = r;
    ssum[threadIdx.x >> 6] = o;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int rr = 0, oo = 0;
    for (int w = 0; w < 16; ++w) {
      rr = max(rr, smax[w]);
      oo += ssum[w];
    }
    ctr[2] = oo;
    ctr[3] = rr;
  }
}

// TX: the tile's x length, the cells each thread owns (16 for the density;
// 4 for the three channels, whose 3 x 16 accumulators would take 124 VGPRs
// and half the waves: 281 us at 256^3 with 16, 249 with 8, 235 with 4)
template <int C, int TX>
__global__ __launch_bounds__(256) void k_brick_gather(BrickArgs a) {
  constexpr int PER = TX;  // owned cells per thread: (k, iy, iz), k < TX
  __shared__ short cand[3][1024];
  __shared__ int ncand[3];
  __shared__ int lst[256][8];  // overlapping bricks of one chunk
  __shared__ int wcnt[4];
  const BrickGrid& g = a.g;
  const MeshDims& d = a.d;
  const int tid = threadIdx.x;
  const int b = blockIdx.y;
  const int ntx = (d.nx + TX - 1) / TX, nty = (d.ny + kTile - 1) / kTile,
            ntz = (d.nz + kTile - 1) / kTile;
  const int tt = (int)xcd_block(blockIdx.x, gridDim.x);
  const int tz = tt % ntz, ty = (tt / ntz) % nty, tx = tt / (ntz * nty);
  const int t0x = tx * TX, t0y = ty * kTile, t0z = tz * kTile;
  const int Tx = min(TX, d.nx - t0x), Ty = min(kTile, d.ny - t0y),
            Tz = min(kTile, d.nz - t0z);
  const int R = a.ctr[3];
  // candidate bricks per dimension (wave w builds list w)
  const int wv = tid >> 6, ln = tid & 63;
  if (wv < 3) {
    const int e = wv == 0 ? g.ex : (wv == 1 ? g.ey : g.ez);
    const int gd = wv == 0 ? g.gx : (wv == 1 ? g.gy : g.gz);
    const int nb = wv == 0 ? g.nbx : (wv == 1 ? g.nby : g.nbz);
    const int t0 = wv == 0 ? t0x : (wv == 1 ? t0y : t0z);
    const int T = wv == 0 ? Tx : (wv == 1 ? Ty : Tz);
    const int n = wv == 0 ? d.nx : (wv == 1 ? d.ny : d.nz);
    int base = 0;
    for (int k0 = 0; k0 < nb; k0 += 64) {
      const int k = k0 + ln;
      const bool hit = k < nb && brick_hits(k, e, gd, g.f, R, t0, T, n);
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
  const int total = nc0 * nc1 * nc2;
  // owned cells: (x = t0x + k, y = t0y + iy, z = t0z + iz), k = 0..15
  const int iy = tid / kTile, iz = tid % kTile;
  const int cy = t0y + iy, cz = t0z + iz;
  const bool yz_ok = iy < Ty && iz < Tz;
  float acc[C][PER];
#pragma unroll
  for (int c = 0; c < C; ++c)
#pragma unroll
    for (int k = 0; k < PER; ++k) acc[c][k] = 0.0f;
  for (int base = 0; base < total; base += 256) {
    // 1. evaluate one candidate per thread, compact the overlapping ones in
    //    candidate order (ballot within a wave, wave counts across)
    const int ci = base + tid;
    bool ok = false;
    int rec[7];
    if (ci < total) {
      const int i2 = ci % nc2, r = ci / nc2, i1 = r % nc1, i0 = r / nc1;
      const int bid = (cand[0][i0] * g.nby + cand[1][i1]) * g.nbz + cand[2][i2];
      const int* m = a.meta + 8LL * ((long long)b * g.nbricks + bid);
      const int e0 = m[3], e1 = m[4], e2 = m[5];
      if (e0 > 0) {
        const int a0 = imod(m[0], d.nx), a1 = imod(m[1], d.ny),
                  a2 = imod(m[2], d.nz);
        ok = box_meets(a0, e0, t0x, Tx, d.nx) &&
             box_meets(a1, e1, t0y, Ty, d.ny) &&
             box_meets(a2, e2, t0z, Tz, d.nz);
        rec[0] = bid;
        rec[1] = a0;
        rec[2] = a1;
        rec[3] = a2;
        rec[4] = e0;
        rec[5] = e1;
        rec[6] = e2;
      }
    }
    const unsigned long long mask = __ballot(ok);
    if (ln == 0) wcnt[wv] = __popcll(mask);
    __syncthreads();
    int off = 0, nl = 0;
    for (int w = 0; w < 4; ++w) {
      if (w < wv) off += wcnt[w];
      nl += wcnt[w];
    }
    if (ok) {
      const int pos = off + __popcll(mask & ((1ULL << ln) - 1ULL));
#pragma unroll
      for (int q = 0; q < 7; ++q) lst[pos][q] = rec[q];
    }
    __syncthreads();
    // 2. every thread adds the bricks, in list order, to its own cells
    if (yz_ok) {
      for (int j = 0; j < nl; ++j) {
        const int e1 = lst[j][5], e2 = lst[j][6];
        const int oy = wrap_off(cy, lst[j][2], d.ny);
        const int oz = wrap_off(cz, lst[j][3], d.nz);
        if (oy >= e1 || oz >= e2) continue;
        const int e0 = lst[j][4];
        const int vol = e0 * e1 * e2;
        const float* src =
            a.arena + ((long long)b * g.nbricks + lst[j][0]) * a.slot;
        const int a0 = lst[j][1];
        // branch-free: every load is issued (out-of-box ones read box
        // element 0) before the first is consumed, so the 16 x C loads of
        // one brick are in flight together instead of one round trip each
        float v[C][PER];
        bool in[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
          const int ox = wrap_off(t0x + k, a0, d.nx);
          in[k] = k < Tx && ox < e0;
          const int bi = in[k] ? (ox * e1 + oy) * e2 + oz : 0;
#pragma unroll
          for (int c = 0; c < C; ++c) v[c][k] = src[c * vol + bi];
        }
#pragma unroll
        for (int k = 0; k < PER; ++k)
#pragma unroll
          for (int c = 0; c < C; ++c) acc[c][k] += in[k] ? v[c][k] : 0.0f;
      }
    }
    __syncthreads();  // lst is rewritten by the next chunk
  }
  if (!yz_ok) return;
  // every cell of the mesh is written here (the mesh needs no clearing);
  // with accum the painted values are added to the mesh's own
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    if (k < Tx) {
      const long long cell =
          (((long long)b * d.nx + t0x + k) * d.ny + cy) * d.nz + cz;
#pragma unroll
      for (int c = 0; c < C; ++c)
        a.mesh[c][cell] = a.accum ? a.mesh[c][cell] + acc[c][k] : acc[c][k];
    }
  }
}

// Bricks whose box did not fit (box over its 64 KB arena slot, positions spread
// over more than the mesh, non-finite weights) are added to the gathered
// mesh with device-scope float atomics, one thread per particle.  Without
// such bricks (k_brick_reduce counted none) every workgroup returns after
// one load.
template <int C>
__global__ __launch_bounds__(256) void k_brick_spill(BrickArgs a) {
  if (a.ctr[2] == 0) return;
  const BrickGrid& g = a.g;
  const MeshDims& d = a.d;
  const long long nq = (long long)d.B * g.nbricks;
  for (long long q = blockIdx.x; q < nq; q += gridDim.x) {
    if (a.meta[8 * q + 7] == 0) continue;
    const int b = (int)(q / g.nbricks), bid = (int)(q % g.nbricks);
    int q0[3], cnt[3];
    brick_block(g, bid, q0, cnt);
    const int cy = cnt[1], cz = cnt[2];
    const int nloc = cnt[0] * cy * cz;
    for (int t = threadIdx.x; t < nloc; t += blockDim.x) {
      const int lz = t % cz, r = t / cz, ly = r % cy, lx = r / cy;
      const long long pi = (long long)b * d.N +
                           ((long long)(q0[0] + lx) * g.gy + (q0[1] + ly)) *
                               g.gz +
                           q0[2] + lz;
      const float x = a.X[3 * pi] * a.ps, y = a.X[3 * pi + 1] * a.ps,
                  z = a.X[3 * pi + 2] * a.ps;
      if (!ok_pos(x, y, z)) continue;
      float wc[C];
#pragma unroll
      for (int c = 0; c < C; ++c)
        wc[c] = a.w ? a.w[pi * a.wstride + c] * a.wscale : a.wscale;
      const float flx = floorf(x), fly = floorf(y), flz = floorf(z);
      const float wx[2] = {1.0f - fabsf(x - flx), 1.0f - fabsf(x - (flx + 1.0f))};
      const float wy[2] = {1.0f - fabsf(y - fly), 1.0f - fabsf(y - (fly + 1.0f))};
      const float wz[2] = {1.0f - fabsf(z - flz), 1.0f - fabsf(z - (flz + 1.0f))};
      const int ix = (int)flx, iy = (int)fly, iz = (int)flz;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int ox = k & 1, oy = (k >> 1) & 1, oz = k >> 2;
        const float kw = (wx[ox] * wy[oy]) * wz[oz];
        const long long cell =
            (((long long)b * d.nx + imod(ix + ox, d.nx)) * d.ny +
             imod(iy + oy, d.ny)) * d.nz + imod(iz + oz, d.nz);
#pragma unroll
        for (int c = 0; c < C; ++c) atomicAdd(a.mesh[c] + cell, wc[c] * kw);
      }
    }
  }
}

bool brick_paint_ok(const MeshDims& d, const BrickGrid& g) {
  return g.ex * g.ey * g.ez <= kDT * kPPT && g.gx * g.f == d.nx &&
         g.gy * g.f == d.ny && g.gz * g.f == d.nz &&
         (long long)g.gx * g.gy * g.gz == d.N && g.nbx <= 1024 &&
         g.nby <= 1024 && g.nbz <= 1024;
}

long long brick_work_floats(const MeshDims& d, const BrickGrid& g, int C) {
  // one arena slot of C boxes per brick
  const long long arena = (long long)d.B * g.nbricks * C * kBoxFloats;
  const long long meta = 8LL * d.B * g.nbricks;
  return arena + meta + 16;
}

int launch_brick_paint(float* const* mesh, int C, const float* X,
                       const float* w, int wstride, float wscale, float ps,
                       const MeshDims& d, const BrickGrid& g, float* work,
                       hipStream_t s, bool accumulate) {
  if ((long long)d.B * d.N == 0) return kOk;
  BrickArgs a{};
  a.accum = accumulate ? 1 : 0;
  for (int c = 0; c < C; ++c) a.mesh[c] = mesh[c];
  a.slot = (long long)C * kBoxFloats;
  a.arena = work;
  a.arena_cap = (long long)d.B * g.nbricks * a.slot;
  a.meta = reinterpret_cast<int*>(work + a.arena_cap);
  a.ctr = a.meta + 8LL * d.B * g.nbricks;
  a.X = X;
  a.w = w;
  a.wstride = wstride;
  a.wscale = wscale;
  a.ps = ps;
  a.d = d;
  a.g = g;
  {
    ProfScope prof_(C == 1 ? kProfPaint : kProfPaintW, s);
    dim3 grid(g.nbricks, d.B);
    if (C == 1)
      hipLaunchKernelGGL(k_brick_deposit<1>, grid, dim3(kDT), 0, s, a);
    else
      hipLaunchKernelGGL(k_brick_deposit<3>, grid, dim3(kDT), 0, s, a);
    PMX_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_brick_reduce, dim3(1), dim3(1024), 0, s, a.meta,
                     (long long)d.B * g.nbricks, a.ctr);
  PMX_LAUNCH_CHECK();
  {
    ProfScope prof_(C == 1 ? kProfPaintGather : kProfPaintWGather, s);
    const int tx = C == 1 ? kGatherTx1 : kGatherTx3;
    const int nt = ((d.nx + tx - 1) / tx) * ((d.ny + kTile - 1) / kTile) *
                   ((d.nz + kTile - 1) / kTile);
    dim3 grid(nt, d.B);
    if (C == 1)
      hipLaunchKernelGGL((k_brick_gather<1, kGatherTx1>), grid, dim3(256), 0,
                         s, a);
    else
      hipLaunchKernelGGL((k_brick_gather<3, kGatherTx3>), grid, dim3(256), 0,
                         s, a);
    PMX_LAUNCH_CHECK();
    const long long nq = (long long)d.B * g.nbricks;
    const dim3 sgrid((unsigned)std::min<long long>(nq, 256));
    if (C == 1)
      hipLaunchKernelGGL(k_brick_spill<1>, sgrid, dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL(k_brick_spill<3>, sgrid, dim3(256), 0, s, a);
    PMX_LAUNCH_CHECK();
  }
  return kOk;
}

}  // namespace pmx
