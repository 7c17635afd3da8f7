// Cloud-in-cell paint / readout and their adjoints, for gfx950.
//
// Semantics follow flowpm/utils.py:32-91 (cic_paint) and :163-210
// (cic_readout): neighbour = floor(x) + {0,1}^3 in the order
// (000,100,010,001,110,101,011,111); weight = prod_d (1 - |x_d - n_d|)
// evaluated as (wx*wy)*wz; periodic wrap by int floormod; positions are
// never wrapped themselves.  Non-finite positions are skipped (the reference
// raises on CPU and silently drops out-of-range indices on GPU).
//
// Adjoints use TF's conventions: d|u|/du = sign(u) with sign(0) = 0, floor
// has no gradient.  For the lower corner d/dx (1 - (x - fl)) = -1 when
// x > fl and 0 when x == fl; for the upper corner the derivative is +1.
#include "pmx_internal.h"

namespace pmx {

struct Cic1 {
  int i0, i1;    // wrapped indices of the two neighbours
  float w0, w1;  // weights
  float d0, d1;  // derivatives of the weights w.r.t. x
};

__device__ __forceinline__ int wrapi(int i, int n) {
  int r = i % n;
  return r < 0 ? r + n : r;
}

__device__ __forceinline__ Cic1 cic1(float x, int n) {
  Cic1 c;
  const float fl = floorf(x);
  const int fi = (int)fl;
  c.w0 = 1.0f - fabsf(x - fl);
  c.w1 = 1.0f - fabsf(x - (fl + 1.0f));
  c.d0 = (x - fl > 0.0f) ? -1.0f : 0.0f;
  c.d1 = 1.0f;
  // floormod: a mask for power-of-two sides (two's complement makes it
  // exact for negative fi too), the division otherwise
  c.i0 = (n & (n - 1)) == 0 ? (fi & (n - 1)) : wrapi(fi, n);
  c.i1 = c.i0 + 1 == n ? 0 : c.i0 + 1;
  return c;
}

// force mesh layouts: kMeshSep three meshes, kMeshF3 packed (x, y, z) per
// cell, kMeshF4 float4 (x, y, z, -) per cell
constexpr int kMeshSep = 0, kMeshF3 = 3, kMeshF4 = 4;
template <int L>
__device__ __forceinline__ float3 load_cell(const Mesh3& m, long long cell) {
  if (L == kMeshF4) {
    const float4 v = m.f4[cell];
    return make_float3(v.x, v.y, v.z);
  }
  if (L == kMeshF3) {
    const float* q = m.f3 + 3 * cell;
    return make_float3(q[0], q[1], q[2]);
  }
  return make_float3(m.x[cell], m.y[cell], m.z[cell]);
}
template <bool I4>
__device__ __forceinline__ float3 load3(const Mesh3& m, long long cell) {
  return load_cell<I4 ? kMeshF4 : kMeshSep>(m, cell);
}

__device__ __forceinline__ bool pos_ok(float x, float y, float z) {
  // beyond 2^30 the int conversion of floor would overflow
  return fabsf(x) < 1.0e9f && fabsf(y) < 1.0e9f && fabsf(z) < 1.0e9f;
}

// cell offsets within one mesh are 32-bit (sides <= 1024: < 2^30 cells);
// the batch member's base is added once in 64 bits
#define PMX_CORNERS(BODY)                                                 \
  {                                                                       \
    constexpr int OX[8] = {0, 1, 0, 0, 1, 1, 0, 1};                       \
    constexpr int OY[8] = {0, 0, 1, 0, 1, 0, 1, 1};                       \
    constexpr int OZ[8] = {0, 0, 0, 1, 0, 1, 1, 1};                       \
    const long long cb_ = (long long)b * ((long long)d.nx * d.ny * d.nz); \
    const int rx0_ = cx.i0 * d.ny, rx1_ = cx.i1 * d.ny;                   \
    const int rw_[4] = {(rx0_ + cy.i0) * d.nz, (rx0_ + cy.i1) * d.nz,     \
                        (rx1_ + cy.i0) * d.nz, (rx1_ + cy.i1) * d.nz};    \
    _Pragma("unroll") for (int cc = 0; cc < 8; ++cc) {                    \
      const float wx = OX[cc] ? cx.w1 : cx.w0;                            \
      const float wy = OY[cc] ? cy.w1 : cy.w0;                            \
      const float wz = OZ[cc] ? cz.w1 : cz.w0;                            \
      const float dx = OX[cc] ? cx.d1 : cx.d0;                            \
      const float dy = OY[cc] ? cy.d1 : cy.d0;                            \
      const float dz = OZ[cc] ? cz.d1 : cz.d0;                            \
      const long long cell =                                              \
          cb_ + (rw_[2 * OX[cc] + OY[cc]] + (OZ[cc] ? cz.i1 : cz.i0));    \
      (void)dx; (void)dy; (void)dz;                                       \
      BODY                                                                \
    }                                                                     \
  }

__global__ void k_fill(float* __restrict__ p, long long n, float v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// Plain scatter: one thread per particle, 8 hardware float atomics.
__global__ __launch_bounds__(256) void k_paint(
    float* __restrict__ mesh, const float* __restrict__ pos,
    const float* __restrict__ w, float wscale, MeshDims d, float ps) {
  const long long tot = (long long)d.B * d.N;
  const long long p =
      xcd_block(blockIdx.x, gridDim.x) * (long long)blockDim.x + threadIdx.x;
  if (p >= tot) return;
  const int b = (int)(p / d.N);
  const float x = pos[3 * p] * ps, y = pos[3 * p + 1] * ps,
              z = pos[3 * p + 2] * ps;
  if (!pos_ok(x, y, z)) return;
  const float wt = w ? w[p] * wscale : wscale;
  const Cic1 cx = cic1(x, d.nx), cy = cic1(y, d.ny), cz = cic1(z, d.nz);
  PMX_CORNERS({
    const float k = (wx * wy) * wz;
    atomicAdd(mesh + cell, wt * k);
  })
}

__global__ __launch_bounds__(256) void k_readout(
    float* __restrict__ out, const float* __restrict__ mesh,
    const float* __restrict__ pos, MeshDims d, float ps) {
  const long long tot = (long long)d.B * d.N;
  const long long p =
      xcd_block(blockIdx.x, gridDim.x) * (long long)blockDim.x + threadIdx.x;
  if (p >= tot) return;
  const int b = (int)(p / d.N);
  const float x = pos[3 * p] * ps, y = pos[3 * p + 1] * ps,
              z = pos[3 * p + 2] * ps;
  if (!pos_ok(x, y, z)) {
    out[p] = 0.0f;
    return;
  }
  const Cic1 cx = cic1(x, d.nx), cy = cic1(y, d.ny), cz = cic1(z, d.nz);
  float acc = 0.0f;
  PMX_CORNERS({ acc += mesh[cell] * ((wx * wy) * wz); })
  out[p] = acc;
}

// d(pos) = gscale * w * sum_c g[c] dK_c/dx ; d(w) = sum_c g[c] K_c
__global__ __launch_bounds__(256) void k_paint_adj(
    float* __restrict__ dpos, float* __restrict__ dw,
    const float* __restrict__ g, const float* __restrict__ pos,
    const float* __restrict__ w, MeshDims d, float ps, float gscale,
    int accumulate) {
  const long long tot = (long long)d.B * d.N;
  const long long p =
      xcd_block(blockIdx.x, gridDim.x) * (long long)blockDim.x + threadIdx.x;
  if (p >= tot) return;
  const int b = (int)(p / d.N);
  const float x = pos[3 * p] * ps, y = pos[3 * p + 1] * ps,
              z = pos[3 * p + 2] * ps;
  float gx = 0.f, gy = 0.f, gz = 0.f, gw = 0.f;
  if (pos_ok(x, y, z)) {
    const Cic1 cx = cic1(x, d.nx), cy = cic1(y, d.ny), cz = cic1(z, d.nz);
    PMX_CORNERS({
      const float v = g[cell];
      gw += v * ((wx * wy) * wz);
      gx += v * ((dx * wy) * wz);
      gy += v * ((wx * dy) * wz);
      gz += v * ((wx * wy) * dz);
    })
  }
  const float wt = (w ? w[p] : 1.0f) * gscale * ps;
  if (dpos) {
    if (accumulate) {
      dpos[3 * p] += wt * gx;
      dpos[3 * p + 1] += wt * gy;
      dpos[3 * p + 2] += wt * gz;
    } else {
      dpos[3 * p] = wt * gx;
      dpos[3 * p + 1] = wt * gy;
      dpos[3 * p + 2] = wt * gz;
    }
  }
  if (dw) dw[p] = accumulate ? dw[p] + gw * gscale : gw * gscale;
}

// readout adjoint: gmesh[c] += gout_p K_c ; dpos = gout_p sum_c m[c] dK_c/dx
__global__ __launch_bounds__(256) void k_readout_adj(
    float* __restrict__ gmesh, float* __restrict__ dpos,
    const float* __restrict__ gout, const float* __restrict__ mesh,
    const float* __restrict__ pos, MeshDims d, float ps, int accumulate) {
  const long long tot = (long long)d.B * d.N;
  const long long p =
      xcd_block(blockIdx.x, gridDim.x) * (long long)blockDim.x + threadIdx.x;
  if (p >= tot) return;
  const int b = (int)(p / d.N);
  const float x = pos[3 * p] * ps, y = pos[3 * p + 1] * ps,
              z = pos[3 * p + 2] * ps;
  const float go = gout[p];
  float gx = 0.f, gy = 0.f, gz = 0.f;
  if (pos_ok(x, y, z)) {
    const Cic1 cx = cic1(x, d.nx), cy = cic1(y, d.ny), cz = cic1(z, d.nz);
    PMX_CORNERS({
      if (gmesh) atomicAdd(gmesh + cell, go * ((wx * wy) * wz));
      if (dpos) {
        const float v = mesh[cell];
        gx += v * ((dx * wy) * wz);
        gy += v * ((wx * dy) * wz);
        gz += v * ((wx * wy) * dz);
      }
    })
  }
  if (dpos) {
    const float s = go * ps;
    if (accumulate) {
      dpos[3 * p] += s * gx;
      dpos[3 * p + 1] += s * gy;
      dpos[3 * p + 2] += s * gz;
    } else {
      dpos[3 * p] = s * gx;
      dpos[3 * p + 1] = s * gy;
      dpos[3 * p + 2] = s * gz;
    }
  }
}

// Fused force readout + kick (+ next kick + drift), one thread per particle.
// F_d = (readout(mesh_d, f X) * fac) / f          tfpm.py:352-356
// P  += k2 F                                      tfpm.py:284-291
// P  += k1 F ; X += d P                           tfpm.py:306-313
// The (B,N,3) state rows are moved through LDS so that every global load
// and store of X, P and F is a contiguous, coalesced dword stream instead
// of three 12-byte-strided accesses per particle.
// Particle tiles of the fused readout: with a particle grid, a workgroup
// takes 4 (y) x 64 (z) particles at one x and consecutive workgroups step x
// first, so the mesh rows that neighbouring tiles gather are shared in L1 /
// L2 (late-time positions are far from their grid cells: one particle row
// per workgroup made each mesh line be fetched from L2 ~4 times); the
// 64-particle runs keep the particle streams in 768-byte contiguous pieces.
constexpr int kTileY = 4, kTileZ = 64;

__host__ __device__ __forceinline__ bool grid_tiled(int gx, int gy, int gz,
                                                    long long N) {
  return gx > 0 && gy % kTileY == 0 && gz % kTileZ == 0 &&
         (long long)gx * gy * gz == N;
}

// first particle of logical workgroup lb (its 4 runs are gz apart)
// Consecutive workgroups: every z tile of G = 8 (or the largest power of two
// <= 8 dividing gy / 4) neighbouring y tiles at one x, i.e. a 32 x gz (y, z)
// slab, then the next x.  The workgroups in flight on one XCD then cover a
// few such planes, whose mesh rows stay in its L2 while neighbouring planes
// gather them.  Measured at 256^3 (40 steps from 2LPT ICs): readout 273 us
// against 297 with 16 x 64 columns and z tiles outermost, 289 with G = 16
// and 309 with whole-y planes.
__device__ __forceinline__ long long grid_tile(long long lb, int gx, int gy,
                                               int gz, long long N) {
  // 32-bit: a launch has < 2^31 workgroups and a member < 2^31 particles
  const unsigned per_b = (unsigned)gx * (unsigned)(gy / kTileY) *
                         (unsigned)(gz / kTileZ);
  const unsigned l = (unsigned)lb;
  const unsigned b = l / per_b, r0 = l % per_b;
  const unsigned nty = (unsigned)(gy / kTileY);
  const unsigned G = nty % 8 == 0 ? 8 : (nty % 4 == 0 ? 4 : (nty % 2 == 0 ? 2 : 1));
  const unsigned ntz = (unsigned)(gz / kTileZ);
  const unsigned sub = r0 % (G * ntz), r = r0 / (G * ntz);
  const unsigned x = r % (unsigned)gx, grp = r / (unsigned)gx;
  const unsigned ty = grp * G + sub / ntz, tz = sub % ntz;
  return (long long)b * N +
         (long long)((x * (unsigned)gy + ty * kTileY) * (unsigned)gz +
                     tz * kTileZ);
}

// float offset of staged element i (0..767) of a 256-particle workgroup:
// tiled, 4 runs of 192 floats gz particles apart; else contiguous
__device__ __forceinline__ long long stage_off(bool tiled, long long p0,
                                               int gz, int i) {
  if (!tiled) return 3 * p0 + i;
  const int r = i / (3 * kTileZ), o = i % (3 * kTileZ);
  return 3 * (p0 + (long long)r * gz) + o;
}

// CK: the step also writes the state between its two kicks (X, P after
// the closing kick, F; the layout of a state (3, B, N, 3)) to c.ckpt, the
// checkpoint a recomputing reverse sweep restarts from.
template <int L, bool P2F, bool CK = false>
__global__ __launch_bounds__(256) void k_readout_kick_drift(
    float* __restrict__ X, float* __restrict__ P, float* __restrict__ F,
    Mesh3 m, MeshDims d, KickDrift c) {
  __shared__ float sx[768], sp[768], sf[768];
  __shared__ float sck[CK ? 768 : 1];
  const long long tot = (long long)d.B * d.N;
  const long long lb = xcd_block(blockIdx.x, gridDim.x);
  const bool tiled = c.gx > 0;  // the launcher checked the grid
  const long long p0 =
      tiled ? grid_tile(lb, c.gx, c.gy, c.gz, d.N) : lb * 256;
  const int np = tiled ? 256 : (int)min((long long)256, tot - p0);
  const int nf = 3 * np;
  const bool kick = c.do_k2 || c.do_k1d;
  float* const Xo = c.xout ? c.xout : X;  // where the drifted X goes
  const float* const Xl = c.xin ? c.xin : X;  // where X and P are read
  const float* const Pl = c.pin ? c.pin : P;
  const int t = threadIdx.x;
  auto goff = [&](int i) { return stage_off(tiled, p0, c.gz, i); };
  // particle state is streamed with non-temporal loads / stores so it does
  // not evict the force mesh, which neighbouring blocks re-read, from L2
  if (np == 256) {
    // full block: all loads in flight before the first LDS write
    float vx[3], vp[3];
    if (kick) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const long long o = goff(t + 256 * k);
        vx[k] = __builtin_nontemporal_load(Xl + o);
        vp[k] = __builtin_nontemporal_load(Pl + o);
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        sx[t + 256 * k] = vx[k];
        sp[t + 256 * k] = vp[k];
        // the checkpoint's X is the force's position, as loaded
        if (CK) __builtin_nontemporal_store(vx[k], c.ckpt + goff(t + 256 * k));
      }
    } else {
#pragma unroll
      for (int k = 0; k < 3; ++k)
        vx[k] = __builtin_nontemporal_load(Xl + goff(t + 256 * k));
#pragma unroll
      for (int k = 0; k < 3; ++k) sx[t + 256 * k] = vx[k];
    }
  } else {
    for (int i = t; i < nf; i += 256) {
      sx[i] = __builtin_nontemporal_load(Xl + 3 * p0 + i);
      if (kick) sp[i] = __builtin_nontemporal_load(Pl + 3 * p0 + i);
      if (CK) __builtin_nontemporal_store(sx[i], c.ckpt + 3 * p0 + i);
    }
  }
  __syncthreads();
  if (t < np) {
    // a tile never spans batch members; a linear block may
    const int b = d.B == 1 ? 0 : (int)((tiled ? p0 : p0 + t) / d.N);
    float px = sx[3 * t], py = sx[3 * t + 1], pz = sx[3 * t + 2];
    const float x = px * c.f, y = py * c.f, z = pz * c.f;
    float ax = 0.f, ay = 0.f, az = 0.f;
    if (pos_ok(x, y, z)) {
      const Cic1 cx = cic1(x, d.nx), cy = cic1(y, d.ny), cz = cic1(z, d.nz);
      float3 v[8];
      PMX_CORNERS({ v[cc] = load_cell<L>(m, cell); })
      PMX_CORNERS({
        const float k = (wx * wy) * wz;
        ax += v[cc].x * k;
        ay += v[cc].y * k;
        az += v[cc].z * k;
      })
    }
    // (v fac) / f as tfpm.py:361-365; for a power-of-two f this equals
    // v (fac / f) exactly (scaling by 2^-k commutes with rounding), which
    // saves the division sequence
    const float Fx = P2F ? ax * (c.fac / c.f) : (ax * c.fac) / c.f;
    const float Fy = P2F ? ay * (c.fac / c.f) : (ay * c.fac) / c.f;
    const float Fz = P2F ? az * (c.fac / c.f) : (az * c.fac) / c.f;
    sf[3 * t] = Fx;
    sf[3 * t + 1] = Fy;
    sf[3 * t + 2] = Fz;
    if (kick) {
      float vx = sp[3 * t], vy = sp[3 * t + 1], vz = sp[3 * t + 2];
      if (c.do_k2) {
        vx = vx + c.k2 * Fx;
        vy = vy + c.k2 * Fy;
        vz = vz + c.k2 * Fz;
      }
      if (CK) {
        sck[3 * t] = vx;
        sck[3 * t + 1] = vy;
        sck[3 * t + 2] = vz;
      }
      if (c.do_k1d) {
        vx = vx + c.k1 * Fx;
        vy = vy + c.k1 * Fy;
        vz = vz + c.k1 * Fz;
        px = px + c.d * vx;
        py = py + c.d * vy;
        pz = pz + c.d * vz;
        sx[3 * t] = px;
        sx[3 * t + 1] = py;
        sx[3 * t + 2] = pz;
      }
      sp[3 * t] = vx;
      sp[3 * t + 1] = vy;
      sp[3 * t + 2] = vz;
    }
  }
  __syncthreads();
  if (np == 256) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int i = t + 256 * k;
      const long long o = goff(i);
      if (!c.skip_f) __builtin_nontemporal_store(sf[i], F + o);
      if (kick) __builtin_nontemporal_store(sp[i], P + o);
      if (c.do_k1d) __builtin_nontemporal_store(sx[i], Xo + o);
      if (c.do_k1d && c.xcopy) __builtin_nontemporal_store(sx[i], c.xcopy + o);
      if (CK) {
        __builtin_nontemporal_store(sck[i], c.ckpt + 3 * tot + o);
        __builtin_nontemporal_store(sf[i], c.ckpt + 6 * tot + o);
      }
    }
  } else {
    for (int i = t; i < nf; i += 256) {
      if (!c.skip_f) __builtin_nontemporal_store(sf[i], F + 3 * p0 + i);
      if (kick) __builtin_nontemporal_store(sp[i], P + 3 * p0 + i);
      if (c.do_k1d) __builtin_nontemporal_store(sx[i], Xo + 3 * p0 + i);
      if (c.do_k1d && c.xcopy)
        __builtin_nontemporal_store(sx[i], c.xcopy + 3 * p0 + i);
      if (CK) {
        __builtin_nontemporal_store(sck[i], c.ckpt + 3 * tot + 3 * p0 + i);
        __builtin_nontemporal_store(sf[i], c.ckpt + 6 * tot + 3 * p0 + i);
      }
    }
  }
}

// Contraction of a particle cotangent with a readout: sum_p sum_c
// G[p][c] readout(m_c, X_p f), reduced in float64 (the PGD parameter
// gradients).
template <bool I4>
__global__ __launch_bounds__(256) void k_readout_dot(
    const float* __restrict__ X, const float* __restrict__ G, Mesh3 m,
    MeshDims d, float f, double* __restrict__ out, double scale) {
  __shared__ double red[4];
  const long long tot = (long long)d.B * d.N;
  const long long p =
      xcd_block(blockIdx.x, gridDim.x) * (long long)blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (p < tot) {
    const int b = d.B == 1 ? 0 : (int)(p / d.N);
    const float x = X[3 * p] * f, y = X[3 * p + 1] * f, z = X[3 * p + 2] * f;
    if (pos_ok(x, y, z)) {
      const Cic1 cx = cic1(x, d.nx), cy = cic1(y, d.ny), cz = cic1(z, d.nz);
      float3 v[8];
      PMX_CORNERS({ v[cc] = load3<I4>(m, cell); })
      float ax = 0.f, ay = 0.f, az = 0.f;
      PMX_CORNERS({
        const float k = (wx * wy) * wz;
        ax += v[cc].x * k;
        ay += v[cc].y * k;
        az += v[cc].z * k;
      })
      acc = (double)G[3 * p] * ax + (double)G[3 * p + 1] * ay +
            (double)G[3 * p + 2] * az;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0)
    atomicAdd(out, scale * ((red[0] + red[1]) + (red[2] + red[3])));
}

// Scalar cotangents of the time loop's coefficients (the cosmology-parameter
// gradients of nbody): for the force F = (fac / f) readout(m4, f X) of one
// step, a = sum_p (gP - dstep gX) . F and, with Xprev, b = sum_p
// gX . (Xd - Xprev) / dstep (the momentum of the drift that produced Xd
// from Xprev; Xd is X itself unless a PGD correction moved X after the
// drift).  gP and gX are the cotangents after that step's update (gP - dstep
// gX restores the momentum cotangent before the drift's adjoint added
// dstep gX).  One (a, b) pair of float64 partial sums per workgroup.
__global__ __launch_bounds__(256) void k_coef_dot(
    const float* __restrict__ X, const float* __restrict__ Xd,
    const float* __restrict__ Xprev, const float* __restrict__ gX,
    const float* __restrict__ gP, const float4* __restrict__ m4, MeshDims d,
    float f, float fac, float dstep, double* __restrict__ part) {
  __shared__ double red[2][4];
  const long long tot = (long long)d.B * d.N;
  const long long blk = xcd_block(blockIdx.x, gridDim.x);
  const long long p = blk * 256 + threadIdx.x;
  double a = 0.0, bb = 0.0;
  if (p < tot) {
    const int b = d.B == 1 ? 0 : (int)(p / d.N);
    const float x0 = X[3 * p], y0 = X[3 * p + 1], z0 = X[3 * p + 2];
    const float x = x0 * f, y = y0 * f, z = z0 * f;
    float ax = 0.f, ay = 0.f, az = 0.f;
    if (pos_ok(x, y, z)) {
      const Cic1 cx = cic1(x, d.nx), cy = cic1(y, d.ny), cz = cic1(z, d.nz);
      float4 v[8];
      PMX_CORNERS({ v[cc] = m4[cell]; })
      PMX_CORNERS({
        const float k = (wx * wy) * wz;
        ax += v[cc].x * k;
        ay += v[cc].y * k;
        az += v[cc].z * k;
      })
    }
    const double Fx = (double)((ax * fac) / f), Fy = (double)((ay * fac) / f),
                 Fz = (double)((az * fac) / f);
    const double gx = gX[3 * p], gy = gX[3 * p + 1], gz = gX[3 * p + 2];
    const double ds = dstep;
    a = ((double)gP[3 * p] - ds * gx) * Fx +
        ((double)gP[3 * p + 1] - ds * gy) * Fy +
        ((double)gP[3 * p + 2] - ds * gz) * Fz;
    if (Xprev) {
      const float* xd = Xd ? Xd : X;
      bb = (gx * ((double)xd[3 * p] - (double)Xprev[3 * p]) +
            gy * ((double)xd[3 * p + 1] - (double)Xprev[3 * p + 1]) +
            gz * ((double)xd[3 * p + 2] - (double)Xprev[3 * p + 2])) / ds;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o);
    bb += __shfl_xor(bb, o);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = a;
    red[1][threadIdx.x >> 6] = bb;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * blk] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    part[2 * blk + 1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
  }
}

// out[ia] += sum of the a partials, out[ib] += sum of the b partials (ia or
// ib < 0: not wanted); one workgroup, in a fixed order (deterministic).
__global__ __launch_bounds__(1024) void k_coef_reduce(
    const double* __restrict__ part, long long nb, double* __restrict__ out,
    int ia, int ib) {
  __shared__ double red[2][16];
  double a = 0.0, bb = 0.0;
  constexpr int U = 8;
  for (long long i0 = threadIdx.x; i0 < nb; i0 += 1024LL * U) {
    double va[U], vb[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long i = min(i0 + 1024LL * u, nb - 1);
      va[u] = part[2 * i];
      vb[u] = part[2 * i + 1];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool in = i0 + 1024LL * u < nb;
      a += in ? va[u] : 0.0;
      bb += in ? vb[u] : 0.0;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o);
    bb += __shfl_xor(bb, o);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = a;
    red[1][threadIdx.x >> 6] = bb;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sa = 0.0, sb = 0.0;
    for (int w = 0; w < 16; ++w) {
      sa += red[0][w];
      sb += red[1][w];
    }
    if (ia >= 0) out[ia] += sa;
    if (ib >= 0) out[ib] += sb;
  }
}

// PGD displacement (tfpm.py:404-421): X += c * readout(m, X * f), with
// c = alpha / pm_nc_factor and m the three PGD meshes.
template <bool I4>
__global__ __launch_bounds__(256) void k_pgd_shift(float* __restrict__ X,
                                                   Mesh3 m, MeshDims d,
                                                   float f, float c) {
  const long long tot = (long long)d.B * d.N;
  const long long p =
      xcd_block(blockIdx.x, gridDim.x) * (long long)blockDim.x + threadIdx.x;
  if (p >= tot) return;
  const int b = d.B == 1 ? 0 : (int)(p / d.N);
  float xv[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) xv[k] = X[3 * p + k];
  const float x = xv[0] * f, y = xv[1] * f, z = xv[2] * f;
  if (!pos_ok(x, y, z)) return;
  const Cic1 cx = cic1(x, d.nx), cy = cic1(y, d.ny), cz = cic1(z, d.nz);
  float3 v[8];
  PMX_CORNERS({ v[cc] = load3<I4>(m, cell); })
  float ax = 0.f, ay = 0.f, az = 0.f;
  PMX_CORNERS({
    const float k = (wx * wy) * wz;
    ax += v[cc].x * k;
    ay += v[cc].y * k;
    az += v[cc].z * k;
  })
  X[3 * p] = xv[0] + c * ax;
  X[3 * p + 1] = xv[1] + c * ay;
  X[3 * p + 2] = xv[2] + c * az;
}

// Kick / drift from the stored force (no readout).
// Xs, Ps: the state read; Xd, Pd: where the kicked / drifted state goes
// (the same arrays in place; each element is read before it is written)
__global__ __launch_bounds__(256) void k_kick_drift_only(
    const float* Xs, const float* Ps, const float* __restrict__ F, float* Xd,
    float* Pd, long long n3, KickDrift c) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n3) return;
  float v = Ps[i];
  if (c.do_k2) v = v + c.k2 * F[i];
  if (c.do_k1d) {
    v = v + c.k1 * F[i];
    Xd[i] = Xs[i] + c.d * v;
  } else if (Xd != Xs) {
    Xd[i] = Xs[i];
  }
  Pd[i] = v;
}

// Force VJP, particle side.  With F_d = c * readout(mesh_d, f X):
//   gmesh_d[cell] += c gF_d K_cell           (mesh cotangents, 24 atomics)
//   dX += c f sum_d gF_d sum_cell mesh_d[cell] dK_cell/dx
template <bool I4>
__global__ __launch_bounds__(256) void k_force_vjp_particles(
    float* __restrict__ gmx, float* __restrict__ gmy, float* __restrict__ gmz,
    float* __restrict__ dX, const float* __restrict__ gF,
    const float* __restrict__ X, Mesh3 m, MeshDims d, float f, float cf,
    int scatter) {
  const long long tot = (long long)d.B * d.N;
  const long long p =
      xcd_block(blockIdx.x, gridDim.x) * (long long)blockDim.x + threadIdx.x;
  if (p >= tot) return;
  const int b = (int)(p / d.N);
  const float x = X[3 * p] * f, y = X[3 * p + 1] * f, z = X[3 * p + 2] * f;
  const float gfx = gF[3 * p] * cf, gfy = gF[3 * p + 1] * cf,
              gfz = gF[3 * p + 2] * cf;
  float sx = 0.f, sy = 0.f, sz = 0.f;
  if (pos_ok(x, y, z)) {
    const Cic1 cx = cic1(x, d.nx), cy = cic1(y, d.ny), cz = cic1(z, d.nz);
    PMX_CORNERS({
      if (scatter) {
        const float k = (wx * wy) * wz;
        atomicAdd(gmx + cell, gfx * k);
        atomicAdd(gmy + cell, gfy * k);
        atomicAdd(gmz + cell, gfz * k);
      }
      const float3 fm = load3<I4>(m, cell);
      const float v = gfx * fm.x + gfy * fm.y + gfz * fm.z;
      sx += v * ((dx * wy) * wz);
      sy += v * ((wx * dy) * wz);
      sz += v * ((wx * wy) * dz);
    })
  }
  dX[3 * p] += f * sx;
  dX[3 * p + 1] += f * sy;
  dX[3 * p + 2] += f * sz;
}

// Last particle pass of one reverse step (force VJP + KDK cotangents):
//   gX += f [ sum_d (cf gF_d) dR_d + dPaint(rbar) ]     (readout and paint
//         position gradients: dR_d = sum_cell mesh_d[cell] dK_cell/dx)
//   then, by mode: 1: gP += d gX ; gF = k1 gP
//                  2: as 1, plus gF += k2prev gP
//                  3: gF = 0 (the jump-start force replaced F)
// After an update gF is a multiple of gP, so the sweep may leave it unstored
// (u.implicit_gf) and the next pass reads its cotangent as u.gin_scale * gP:
// one particle stream less to read and one less to write per step.
// With the interleaved mesh (I4) rbar is its fourth component and the rbar
// argument is unused.
// gin is the readout's cotangent (gF for the force, gX itself for the PGD
// displacement: gX and gin may alias, so neither is __restrict__; every
// thread loads its inputs before it stores).
// CG (the step-coefficient cotangents, with an update): the block's partial
// sums a = sum_p (gP' - d gX') . F_p and b = sum_p gX' . (X_p - Xprev_p) / d
// of the updated cotangents gX', gP' and the force F_p = cf * (readout of
// the mesh) go to part[2 * block], part[2 * block + 1] in float64; the
// separate dot-product pass (k_coef_dot) is then not needed.
template <bool I4, bool CG>
__global__ __launch_bounds__(256) void k_force_vjp_finish(
    float* gX, float* __restrict__ gP, float* gF, const float* gin,
    const float* __restrict__ X, Mesh3 m, const float* __restrict__ rbar,
    MeshDims d, float f, float cf, AdjUpdate u,
    const float* __restrict__ Xprev, double* __restrict__ part) {
  __shared__ float sh[CG ? 5 : 4][768];
  __shared__ double red[2][4];
  const long long tot = (long long)d.B * d.N;
  const long long lb = xcd_block(blockIdx.x, gridDim.x);
  const bool tiled = u.gx > 0;  // the launcher checked the grid
  const long long p0 =
      tiled ? grid_tile(lb, u.gx, u.gy, u.gz, d.N) : lb * 256;
  const int t = threadIdx.x;
  const long long p = p0 + t;  // (linear order only)
  const bool full = tiled || p0 + 256 <= tot;
  const bool upd = u.mode == 1 || u.mode == 2;
  const bool gin_gp = u.gin_scale != 0.0f;  // readout cotangent from gP
  const bool ldgp = upd || gin_gp;
  // particle streams: non-temporal (read once, kept out of L2 where the
  // gathered mesh is re-read by neighbouring blocks); a full block stages
  // them through LDS so every load and store is a contiguous run
  float xv[3], gfv[3], gxv[3], gpv[3], xpv[3];
  constexpr int NQ = CG ? 5 : 4;
  if (full) {
    float v[NQ][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const long long o = stage_off(tiled, p0, u.gz, t + 256 * k);
      v[0][k] = __builtin_nontemporal_load(X + o);
      v[1][k] = gin_gp ? 0.0f : __builtin_nontemporal_load(gin + o);
      v[2][k] = __builtin_nontemporal_load(gX + o);
      v[3][k] = ldgp ? __builtin_nontemporal_load(gP + o) : 0.0f;
      if (CG) v[NQ - 1][k] = __builtin_nontemporal_load(Xprev + o);
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int k = 0; k < 3; ++k) sh[q][t + 256 * k] = v[q][k];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      xv[c] = sh[0][3 * t + c];
      gfv[c] = sh[1][3 * t + c];
      gxv[c] = sh[2][3 * t + c];
      gpv[c] = sh[3][3 * t + c];
      xpv[c] = CG ? sh[NQ - 1][3 * t + c] : 0.0f;
    }
  } else {
    // the block reduction of CG needs every thread: a thread past the end
    // computes on the last particle and contributes nothing
    if (!CG && p >= tot) return;  // no barrier follows on this path
    const long long q = p < tot ? p : tot - 1;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      xv[c] = X[3 * q + c];
      gfv[c] = gin_gp ? 0.0f : gin[3 * q + c];
      gxv[c] = gX[3 * q + c];
      gpv[c] = ldgp ? gP[3 * q + c] : 0.0f;
      xpv[c] = CG ? Xprev[3 * q + c] : 0.0f;
    }
  }
  if (gin_gp) {
#pragma unroll
    for (int c = 0; c < 3; ++c) gfv[c] = gpv[c] * u.gin_scale;
  }
  // a tile never spans batch members; a linear block may
  const int b = d.B == 1 ? 0 : (int)((tiled ? p0 : p) / d.N);
  const float x = xv[0] * f, y = xv[1] * f, z = xv[2] * f;
  const float gfx = gfv[0] * cf, gfy = gfv[1] * cf, gfz = gfv[2] * cf;
  float sx = 0.f, sy = 0.f, sz = 0.f, rx = 0.f, ry = 0.f, rz = 0.f;
  float ax = 0.f, ay = 0.f, az = 0.f;  // readout of the force mesh (CG)
  if (pos_ok(x, y, z)) {
    const Cic1 cx = cic1(x, d.nx), cy = cic1(y, d.ny), cz = cic1(z, d.nz);
    // the 8 (or 16) gathers are all issued before the first is used
    float3 fm[8];
    float rb[8];
    PMX_CORNERS({
      if (I4) {
        // interleaved: (force x, y, z, density cotangent)
        const float4 v = m.f4[cell];
        fm[cc] = make_float3(v.x, v.y, v.z);
        rb[cc] = v.w;
      } else {
        fm[cc] = load3<I4>(m, cell);
        rb[cc] = rbar[cell];
      }
    })
    PMX_CORNERS({
      const float v = gfx * fm[cc].x + gfy * fm[cc].y + gfz * fm[cc].z;
      const float kx = (dx * wy) * wz;
      const float ky = (wx * dy) * wz;
      const float kz = (wx * wy) * dz;
      sx += v * kx;
      sy += v * ky;
      sz += v * kz;
      rx += rb[cc] * kx;
      ry += rb[cc] * ky;
      rz += rb[cc] * kz;
      if (CG) {
        const float k = (wx * wy) * wz;
        ax += fm[cc].x * k;
        ay += fm[cc].y * k;
        az += fm[cc].z * k;
      }
    })
  }
  float g[3], gp[3], gf[3];
  const float s3[3] = {sx, sy, sz}, r3[3] = {rx, ry, rz};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    g[c] = gxv[c] + f * s3[c];
    g[c] = g[c] + f * r3[c];
    gp[c] = gpv[c] + u.d * g[c];
    gf[c] = u.k1 * gp[c];
    if (u.mode == 2) gf[c] = gf[c] + u.k2prev * gp[c];
    if (u.mode == 3) gf[c] = 0.0f;
  }
  const bool wgf = (upd && !u.implicit_gf) || u.mode == 3;
  if (full) {
    __syncthreads();  // every thread has read its inputs from sh
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      sh[0][3 * t + c] = g[c];
      sh[1][3 * t + c] = gp[c];
      sh[2][3 * t + c] = gf[c];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const long long o = stage_off(tiled, p0, u.gz, t + 256 * k);
      __builtin_nontemporal_store(sh[0][t + 256 * k], gX + o);
      if (upd) __builtin_nontemporal_store(sh[1][t + 256 * k], gP + o);
      if (wgf) __builtin_nontemporal_store(sh[2][t + 256 * k], gF + o);
    }
  } else if (p < tot) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      gX[3 * p + c] = g[c];
      if (upd) gP[3 * p + c] = gp[c];
      if (wgf) gF[3 * p + c] = gf[c];
    }
  }
  if (CG) {
    const float av[3] = {ax, ay, az};
    double a = 0.0, bb = 0.0;
    if (full || p < tot) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double gx = g[c];
        a += ((double)gp[c] - (double)u.d * gx) * (double)(av[c] * cf);
        bb += gx * ((double)xv[c] - (double)xpv[c]);
      }
      bb /= (double)u.d;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      a += __shfl_xor(a, o);
      bb += __shfl_xor(bb, o);
    }
    if ((t & 63) == 0) {
      red[0][t >> 6] = a;
      red[1][t >> 6] = bb;
    }
    __syncthreads();
    if (t == 0) {
      part[2 * blockIdx.x] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
      part[2 * blockIdx.x + 1] =
          (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    }
  }
}

// ------------------------------------------------------------ launchers --
static inline unsigned nblk(long long n, int t) {
  return (unsigned)((n + t - 1) / t);
}

int launch_fill(float* p, long long n, float v, hipStream_t s) {
  if (n <= 0) return kOk;
  const unsigned g = (unsigned)std::min<long long>(nblk(n, 256), 8192);
  ProfScope prof_(kProfFill, s);
  hipLaunchKernelGGL(k_fill, dim3(g), dim3(256), 0, s, p, n, v);
  PMX_LAUNCH_CHECK();
  return kOk;
}

int launch_paint(float* mesh, const float* pos, const float* w, float wscale,
                 const MeshDims& d, float ps, int mode, hipStream_t s) {
  (void)mode;
  const long long tot = (long long)d.B * d.N;
  if (tot == 0) return kOk;
  ProfScope prof_(kProfPaint, s);
  hipLaunchKernelGGL(k_paint, dim3(nblk(tot, 256)), dim3(256), 0, s, mesh,
                     pos, w, wscale, d, ps);
  PMX_LAUNCH_CHECK();
  return kOk;
}

int launch_readout(float* out, const float* mesh, const float* pos,
                   const MeshDims& d, float ps, hipStream_t s) {
  const long long tot = (long long)d.B * d.N;
  if (tot == 0) return kOk;
  ProfScope prof_(kProfReadout, s);
  hipLaunchKernelGGL(k_readout, dim3(nblk(tot, 256)), dim3(256), 0, s, out,
                     mesh, pos, d, ps);
  PMX_LAUNCH_CHECK();
  return kOk;
}

int launch_paint_adj(float* dpos, float* dw, const float* g, const float* pos,
                     const float* w, const MeshDims& d, float ps, float gscale,
                     int accumulate, hipStream_t s) {
  const long long tot = (long long)d.B * d.N;
  if (tot == 0) return kOk;
  ProfScope prof_(kProfAdjCic, s);
  hipLaunchKernelGGL(k_paint_adj, dim3(nblk(tot, 256)), dim3(256), 0, s, dpos,
                     dw, g, pos, w, d, ps, gscale, accumulate);
  PMX_LAUNCH_CHECK();
  return kOk;
}

int launch_readout_adj(float* gmesh, float* dpos, const float* gout,
                       const float* mesh, const float* pos, const MeshDims& d,
                       float ps, int accumulate, hipStream_t s) {
  const long long tot = (long long)d.B * d.N;
  if (tot == 0) return kOk;
  ProfScope prof_(kProfAdjCic, s);
  hipLaunchKernelGGL(k_readout_adj, dim3(nblk(tot, 256)), dim3(256), 0, s,
                     gmesh, dpos, gout, mesh, pos, d, ps, accumulate);
  PMX_LAUNCH_CHECK();
  return kOk;
}

int launch_readout_kick_drift(float* X, float* P, float* F, const Mesh3& m,
                              const MeshDims& d, const KickDrift& c,
                              hipStream_t s) {
  const long long tot = (long long)d.B * d.N;
  if (tot == 0) return kOk;
  ProfScope prof_(kProfReadoutKD, s);
  KickDrift cc = c;
  if (!grid_tiled(c.gx, c.gy, c.gz, d.N)) cc.gx = cc.gy = cc.gz = 0;
  int ex;
  const bool p2 = c.f > 0.0f && std::frexp(c.f, &ex) == 0.5f;
  if (c.ckpt)
    PMX_REQUIRE(m.f3 && c.do_k2 && c.do_k1d && !c.xout && !c.xcopy,
                "readout: a checkpoint is written from an untaped step with "
                "both kicks and the packed force mesh");
  auto kern = c.ckpt ? (p2 ? k_readout_kick_drift<kMeshF3, true, true>
                           : k_readout_kick_drift<kMeshF3, false, true>)
              : m.f4 ? (p2 ? k_readout_kick_drift<kMeshF4, true>
                         : k_readout_kick_drift<kMeshF4, false>)
              : m.f3 ? (p2 ? k_readout_kick_drift<kMeshF3, true>
                           : k_readout_kick_drift<kMeshF3, false>)
                     : (p2 ? k_readout_kick_drift<kMeshSep, true>
                           : k_readout_kick_drift<kMeshSep, false>);
  hipLaunchKernelGGL(kern, dim3(nblk(tot, 256)), dim3(256), 0, s, X, P, F, m,
                     d, cc);
  PMX_LAUNCH_CHECK();
  return kOk;
}

int launch_coef_dot(const float* X, const float* Xd, const float* Xprev,
                    const float* gX, const float* gP, const float4* m4,
                    const MeshDims& d, float f, float fac, float dstep,
                    double* part, double* out, int ia, int ib,
                    hipStream_t s) {
  const long long tot = (long long)d.B * d.N;
  if (tot == 0) return kOk;
  const long long nb = nblk(tot, 256);
  ProfScope prof_(kProfAdjUpdate, s);
  hipLaunchKernelGGL(k_coef_dot, dim3((unsigned)nb), dim3(256), 0, s, X, Xd,
                     Xprev, gX, gP, m4, d, f, fac, dstep, part);
  PMX_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_coef_reduce, dim3(1), dim3(1024), 0, s, part, nb, out,
                     ia, ib);
  PMX_LAUNCH_CHECK();
  return kOk;
}

int launch_readout_dot(const float* X, const float* G, const Mesh3& m,
                       const MeshDims& d, float f, double* out, double scale,
                       hipStream_t s) {
  const long long tot = (long long)d.B * d.N;
  if (tot == 0) return kOk;
  ProfScope prof_(kProfAdjCic, s);
  PMX_REQUIRE(!m.f3, "readout_dot: no packed-mesh variant");
  if (m.f4)
    hipLaunchKernelGGL(k_readout_dot<true>, dim3(nblk(tot, 256)), dim3(256),
                       0, s, X, G, m, d, f, out, scale);
  else
    hipLaunchKernelGGL(k_readout_dot<false>, dim3(nblk(tot, 256)), dim3(256),
                       0, s, X, G, m, d, f, out, scale);
  PMX_LAUNCH_CHECK();
  return kOk;
}

int launch_pgd_shift(float* X, const Mesh3& m, const MeshDims& d, float f,
                     float c, hipStream_t s) {
  const long long tot = (long long)d.B * d.N;
  if (tot == 0) return kOk;
  ProfScope prof_(kProfReadout, s);
  PMX_REQUIRE(!m.f3, "pgd_shift: no packed-mesh variant");
  if (m.f4)
    hipLaunchKernelGGL(k_pgd_shift<true>, dim3(nblk(tot, 256)), dim3(256), 0,
                       s, X, m, d, f, c);
  else
    hipLaunchKernelGGL(k_pgd_shift<false>, dim3(nblk(tot, 256)), dim3(256), 0,
                       s, X, m, d, f, c);
  PMX_LAUNCH_CHECK();
  return kOk;
}

int launch_kick_drift_only(float* X, float* P, const float* F,
                           const MeshDims& d, const KickDrift& c,
                           hipStream_t s) {
  return launch_kick_drift_from(X, P, F, X, P, d, c, s);
}

int launch_kick_drift_from(const float* Xs, const float* Ps, const float* F,
                           float* Xd, float* Pd, const MeshDims& d,
                           const KickDrift& c, hipStream_t s) {
  const long long n3 = (long long)d.B * d.N * 3;
  if (n3 == 0) return kOk;
  ProfScope prof_(kProfKickDrift, s);
  hipLaunchKernelGGL(k_kick_drift_only, dim3(nblk(n3, 256)), dim3(256), 0, s,
                     Xs, Ps, F, Xd, Pd, n3, c);
  PMX_LAUNCH_CHECK();
  return kOk;
}

int launch_force_vjp_finish(float* gX, float* gP, float* gF, const float* gin,
                            const float* X, const Mesh3& m, const float* rbar,
                            const MeshDims& d, float f, float cf,
                            const AdjUpdate& u, hipStream_t s,
                            const CoefDot* cd) {
  const long long tot = (long long)d.B * d.N;
  if (tot == 0) return kOk;
  AdjUpdate uu = u;
  if (!grid_tiled(u.gx, u.gy, u.gz, d.N)) uu.gx = uu.gy = uu.gz = 0;
  PMX_REQUIRE(!m.f3, "force_vjp_finish: no packed-mesh variant");
  const unsigned nb = nblk(tot, 256);
  if (cd) {
    PMX_REQUIRE(m.f4 && (u.mode == 1 || u.mode == 2) && cd->xprev &&
                    cd->part && cd->out,
                "force_vjp_finish: coefficient dots need the interleaved "
                "mesh, an update and Xprev / part / out");
    {
      ProfScope prof_(kProfAdjCic, s);
      hipLaunchKernelGGL((k_force_vjp_finish<true, true>), dim3(nb), dim3(256),
                         0, s, gX, gP, gF, gin ? gin : gF, X, m, rbar, d, f,
                         cf, uu, cd->xprev, cd->part);
      PMX_LAUNCH_CHECK();
    }
    ProfScope prof_(kProfAdjUpdate, s);
    hipLaunchKernelGGL(k_coef_reduce, dim3(1), dim3(1024), 0, s, cd->part,
                       (long long)nb, cd->out, cd->ia, cd->ib);
    PMX_LAUNCH_CHECK();
    return kOk;
  }
  ProfScope prof_(kProfAdjCic, s);
  if (m.f4)
    hipLaunchKernelGGL((k_force_vjp_finish<true, false>), dim3(nb), dim3(256),
                       0, s, gX, gP, gF, gin ? gin : gF, X, m, rbar, d, f, cf,
                       uu, nullptr, nullptr);
  else
    hipLaunchKernelGGL((k_force_vjp_finish<false, false>), dim3(nb),
                       dim3(256), 0, s, gX, gP, gF, gin ? gin : gF, X, m,
                       rbar, d, f, cf, uu, nullptr, nullptr);
  PMX_LAUNCH_CHECK();
  return kOk;
}

int launch_force_vjp_particles(float* gmx, float* gmy, float* gmz, float* dX,
                               const float* gF, const float* X,
                               const Mesh3& m, const MeshDims& d, float f,
                               float c, int scatter, hipStream_t s) {
  const long long tot = (long long)d.B * d.N;
  if (tot == 0) return kOk;
  ProfScope prof_(kProfAdjCic, s);
  PMX_REQUIRE(!m.f3, "force_vjp_particles: no packed-mesh variant");
  if (m.f4)
    hipLaunchKernelGGL(k_force_vjp_particles<true>, dim3(nblk(tot, 256)),
                       dim3(256), 0, s, gmx, gmy, gmz, dX, gF, X, m, d, f, c,
                       scatter);
  else
    hipLaunchKernelGGL(k_force_vjp_particles<false>, dim3(nblk(tot, 256)),
                       dim3(256), 0, s, gmx, gmy, gmz, dX, gF, X, m, d, f, c,
                       scatter);
  PMX_LAUNCH_CHECK();
  return kOk;
}

}  // namespace pmx
