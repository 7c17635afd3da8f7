// Batched 3-D real<->complex FFT for the PM force, written for gfx950.
//
// Replaces tf.signal.fft3d / ifft3d as used by flowpm/utils.py:213-272
// (r2c3d: fft of the real field cast to complex, times 1/norm; c2r3d: real
// part of ifft times norm, i.e. an unnormalised inverse).  Only the half
// spectrum kz = 0..nz/2 is kept (the reference's full spectrum of a real
// field is Hermitian; every multiplied kernel preserves that symmetry).
//
// Structure: three line passes, each a workgroup-per-tile kernel that stages
// a tile in LDS, runs a Stockham autosort FFT there (radix 16/8/4/2/3/5 stages,
// register-staged, in place), and writes the tile back.
//   Z pass  contiguous rows; two real rows are packed into one complex line.
//   Y pass  column tiles: all ky for 16 consecutive kz (128-byte segments).
//   X pass  column tiles: all kx for 16 consecutive kz.  The Fourier-space
//           Poisson x gradient multiply (kernels.py:28-102, tfpm.py:183-219)
//           is fused here between the forward and the inverse x transform,
//           so the force pipeline never materialises the 3-D spectrum.
#include "pmx_internal.h"

namespace pmx {

// --------------------------------------------------------------- complex --
__device__ __forceinline__ float2 cadd(float2 a, float2 b) {
  return make_float2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ float2 csub(float2 a, float2 b) {
  return make_float2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) {
  return make_float2(a.x * s, a.y * s);
}
// multiply by (S * i), S = +1 or -1
template <int S>
__device__ __forceinline__ float2 mul_i(float2 a) {
  return S > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}
// multiply by i^E
__device__ __forceinline__ float2 rot_i(float2 v, int E) {
  switch (E & 3) {
    case 1: return make_float2(-v.y, v.x);
    case 2: return make_float2(-v.x, -v.y);
    case 3: return make_float2(v.y, -v.x);
    default: return v;
  }
}

// ------------------------------------------------------- radix codelets --
// D = -1: forward exp(-2 pi i jk/R); D = +1: inverse (unnormalised).
template <int R, int D>
struct Dft;

template <int D>
struct Dft<2, D> {
  static __device__ __forceinline__ void run(float2* a) {
    float2 t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
  }
};

template <int D>
struct Dft<4, D> {
  static __device__ __forceinline__ void run(float2* a) {
    float2 t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
    float2 t2 = cadd(a[1], a[3]), t3 = mul_i<D>(csub(a[1], a[3]));
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
  }
};

template <int D>
struct Dft<8, D> {
  static __device__ __forceinline__ void run(float2* a) {
    constexpr float c = 0.70710678118654752440f;
    float2 e[4] = {a[0], a[2], a[4], a[6]};
    float2 o[4] = {a[1], a[3], a[5], a[7]};
    Dft<4, D>::run(e);
    Dft<4, D>::run(o);
    // o[m] *= W8^m, W8 = exp(D 2 pi i / 8) = c (1 + D i)
    o[1] = make_float2(c * (o[1].x - D * o[1].y), c * (o[1].y + D * o[1].x));
    o[2] = mul_i<D>(o[2]);
    o[3] = make_float2(c * (-o[3].x - D * o[3].y), c * (-o[3].y + D * o[3].x));
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      a[m] = cadd(e[m], o[m]);
      a[m + 4] = csub(e[m], o[m]);
    }
  }
};

template <int D>
struct Dft<16, D> {
  // two 8-point DFTs of the even and odd inputs, odd ones twiddled by
  // W16^m = cos(2 pi m / 16) + D i sin(2 pi m / 16)
  static __device__ __forceinline__ void run(float2* a) {
    constexpr float C[8] = {1.0f,
                            0.92387953251128675613f,
                            0.70710678118654752440f,
                            0.38268343236508977173f,
                            0.0f,
                            -0.38268343236508977173f,
                            -0.70710678118654752440f,
                            -0.92387953251128675613f};
    constexpr float S[8] = {0.0f,
                            0.38268343236508977173f,
                            0.70710678118654752440f,
                            0.92387953251128675613f,
                            1.0f,
                            0.92387953251128675613f,
                            0.70710678118654752440f,
                            0.38268343236508977173f};
    float2 e[8], o[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      e[m] = a[2 * m];
      o[m] = a[2 * m + 1];
    }
    Dft<8, D>::run(e);
    Dft<8, D>::run(o);
#pragma unroll
    for (int m = 1; m < 8; ++m)
      o[m] = cmul(o[m], make_float2(C[m], D * S[m]));
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      a[m] = cadd(e[m], o[m]);
      a[m + 8] = csub(e[m], o[m]);
    }
  }
};

template <int D>
struct Dft<3, D> {
  static __device__ __forceinline__ void run(float2* a) {
    constexpr float h = 0.86602540378443864676f;  // sqrt(3)/2
    float2 s = cadd(a[1], a[2]), d = csub(a[1], a[2]);
    float2 m = make_float2(a[0].x - 0.5f * s.x, a[0].y - 0.5f * s.y);
    float2 r = cscale(mul_i<D>(d), h);
    a[0] = cadd(a[0], s);
    a[1] = cadd(m, r);
    a[2] = csub(m, r);
  }
};

template <int D>
struct Dft<5, D> {
  // X_m = sum_j a_j W^(jm), W = exp(D 2 pi i / 5).  With p_j = a_j + a_(5-j),
  // m_j = a_j - a_(5-j) (j = 1, 2), c_j = cos(2 pi j / 5) and
  // s_j = sin(2 pi j / 5):
  //   X_1, X_4 = a_0 + c1 p_1 + c2 p_2 +- D i (s1 m_1 + s2 m_2)
  //   X_2, X_3 = a_0 + c2 p_1 + c1 p_2 +- D i (s2 m_1 - s1 m_2)
  static __device__ __forceinline__ void run(float2* a) {
    constexpr float c1 = 0.30901699437494742410f;   // cos(2 pi / 5)
    constexpr float c2 = -0.80901699437494742410f;  // cos(4 pi / 5)
    constexpr float s1 = 0.95105651629515357212f;   // sin(2 pi / 5)
    constexpr float s2 = 0.58778525229247312917f;   // sin(4 pi / 5)
    const float2 p1 = cadd(a[1], a[4]), m1 = csub(a[1], a[4]);
    const float2 p2 = cadd(a[2], a[3]), m2 = csub(a[2], a[3]);
    const float2 r1 = make_float2(a[0].x + c1 * p1.x + c2 * p2.x,
                                  a[0].y + c1 * p1.y + c2 * p2.y);
    const float2 r2 = make_float2(a[0].x + c2 * p1.x + c1 * p2.x,
                                  a[0].y + c2 * p1.y + c1 * p2.y);
    const float2 q1 = mul_i<D>(make_float2(s1 * m1.x + s2 * m2.x,
                                           s1 * m1.y + s2 * m2.y));
    const float2 q2 = mul_i<D>(make_float2(s2 * m1.x - s1 * m2.x,
                                           s2 * m1.y - s1 * m2.y));
    a[0] = cadd(a[0], cadd(p1, p2));
    a[1] = cadd(r1, q1);
    a[4] = csub(r1, q1);
    a[2] = cadd(r2, q2);
    a[3] = csub(r2, q2);
  }
};

// radix plan: 16 first (fewer LDS round trips: 256 = 16 x 16 is two
// stages instead of three), then 8, 4, 2, 3, 5
__host__ __device__ constexpr int pick_radix(int n) {
  return (n % 16 == 0) ? 16 : (n % 8 == 0) ? 8 : (n % 4 == 0) ? 4
       : (n % 2 == 0) ? 2 : (n % 3 == 0) ? 3 : (n % 5 == 0) ? 5 : n;
}

// LDS tile: element (line index i, column c) at s[i * P + c], P = TW + 1.
template <int TW>
struct Pitch {
  static constexpr int P = TW + 1;
};

// One Stockham stage (Bainville's formulation): butterfly j reads
// in[j + r N/R], twiddles by W_{NS R}^{(j mod NS) r}, and writes
// out[(j / NS) NS R + (j mod NS) + r NS].  In place: all reads are staged in
// registers before the barrier, all writes follow it.
template <int N, int R, int NS, int TW, int T, int D>
__device__ __forceinline__ void fft_stage(float2* __restrict__ s,
                                          const float2* __restrict__ tw) {
  constexpr int P = Pitch<TW>::P;
  constexpr int NB = (N / R) * TW;
  constexpr int BPT = (NB + T - 1) / T;
  float2 a[BPT][R];
#pragma unroll
  for (int u = 0; u < BPT; ++u) {
    const int b = threadIdx.x + u * T;
    if (NB % T == 0 || b < NB) {
      const int j = b / TW, c = b % TW;
#pragma unroll
      for (int r = 0; r < R; ++r) a[u][r] = s[(j + r * (N / R)) * P + c];
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < BPT; ++u) {
    const int b = threadIdx.x + u * T;
    if (NB % T == 0 || b < NB) {
      const int j = b / TW, c = b % TW;
      const int k = j % NS;
      if (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[k * r * (N / (NS * R))];
          if (D > 0) w.y = -w.y;
          a[u][r] = cmul(a[u][r], w);
        }
      }
      Dft<R, D>::run(a[u]);
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) s[(base + r * NS) * P + c] = a[u][r];
    }
  }
  __syncthreads();
}

template <int N, int NS, int TW, int T, int D>
struct FftStages {
  static __device__ __forceinline__ void run(float2* s, const float2* tw) {
    if constexpr (NS < N) {
      constexpr int R = pick_radix(N / NS);
      static_assert(R == 2 || R == 3 || R == 4 || R == 5 || R == 8 ||
                        R == 16,
                    "unsupported radix");
      fft_stage<N, R, NS, TW, T, D>(s, tw);
      FftStages<N, NS * R, TW, T, D>::run(s, tw);
    }
  }
};

template <int N, int TW, int T, int D>
__device__ __forceinline__ void fft_tile(float2* s, const float2* tw) {
  FftStages<N, 1, TW, T, D>::run(s, tw);
}

template <int N, int T>
__device__ __forceinline__ void load_twiddles(float2* tw_s,
                                              const float2* __restrict__ tw) {
  for (int t = threadIdx.x; t < N; t += T) tw_s[t] = tw[t];
}

// Fourier multiplier for element (kx, ky, kz).
__device__ __forceinline__ float2 apply_term(float2 v, const Term& t, int kx,
                                             int ky, int kz,
                                             const PlanDev& p) {
  float c = t.scale;
  if (t.lap || t.pgd) {
    const double kk = p.ax[0].k2[kx] + p.ax[1].k2[ky] + p.ax[2].k2[kz];
    // the reference sets the zero mode's kk to 1 before its mask is formed,
    // so the mask keeps everything (kernels.py:43-47, :128-133)
    const double kk1 = kk == 0.0 ? 1.0 : kk;
    if (t.lap) {
      c *= (float)(1.0 / kk1);
      // long-range filter as the reference forms it: its float64 exp is
      // cast to complex64 and multiplies the complex64 Laplacian
      // (tfpm.py:205-207).  Evaluated in float (expf of the float64
      // exponent): a float64 exp here costs the 256-point x pass 19 VGPRs,
      // above the 128 that let two 512-thread workgroups share a CU.
      if (t.lr_r2 > 0.0) c *= expf((float)(-kk * t.lr_r2));
    }
    if (t.pgd) {
      // float32 as TF evaluates it: exp(-kl^2 / kk) exp(-kk^2 / ks^4)
      const float kkf = (float)kk1, kk2f = (float)(kk1 * kk1);
      const float e1 = expf(-t.pgd_kl2 / kkf);
      const float e2 = expf(-kk2f / t.pgd_ks4);
      float w = e1 * e2;
      if (t.pgd == 2) w *= -2.0f * t.pgd_kl / kkf;            // d/dkl
      if (t.pgd == 3) w *= 4.0f * kk2f / (t.pgd_ks4 * t.pgd_ks);  // d/dks
      c *= w;
    }
  }
  if (t.cic) {
    // (sinc_x sinc_y sinc_z)^-2 in float64, rounded once (kernels.py:137-152)
    const double sp = p.ax[0].sinc[kx] * p.ax[1].sinc[ky] * p.ax[2].sinc[kz];
    c *= (float)(1.0 / (sp * sp));
  }
  const int kidx[3] = {kx, ky, kz};
  int E = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int e = t.e[d];
    if (e) {
      const float gd = p.ax[d].g[kidx[d]];
      c *= (e == 2) ? gd * gd : gd;
      E += e;
    }
  }
  return rot_i(cscale(v, c), E);
}

// ------------------------------------------------------------ Z pass -----
// Rows of the real mesh are contiguous; a workgroup takes 2*TW consecutive
// rows (one contiguous block), packs row pairs as re/im of TW complex lines,
// transforms, and splits X1(k) = (Z(k) + Z*(N-k))/2, X2(k) = (Z(k) -
// Z*(N-k))/(2i) for k = 0..N/2.
template <int N, int TW, int T>
__global__ __launch_bounds__(T) void k_zpass_r2c(ZJobs jobs, PlanDev p) {
  constexpr int P = Pitch<TW>::P;
  constexpr int H = N / 2;
  __shared__ float2 s[N * P];
  __shared__ float2 tw[N];
  const ZJob J = blockIdx.y == 0 ? jobs.j[0]
                 : (blockIdx.y == 1 ? jobs.j[1] : jobs.j[2]);
  load_twiddles<N, T>(tw, p.ax[2].tw);
  const long long r0 = (long long)blockIdx.x * (2 * TW);
  const int nr = (int)min((long long)(2 * TW), p.rows - r0);
  const float* __restrict__ src = J.rin + r0 * N;
  float* sf = reinterpret_cast<float*>(s);
  {
    constexpr int NE = 2 * TW * N;
    constexpr int E = (NE + T - 1) / T;
    float v[E];
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int q = threadIdx.x + u * T;
      v[u] = ((NE % T == 0 || q < NE) && q / N < nr) ? src[q] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int q = threadIdx.x + u * T;
      const int row = q / N, z = q % N;
      if (NE % T == 0 || q < NE) sf[2 * (z * P + (row >> 1)) + (row & 1)] = v[u];
    }
  }
  __syncthreads();
  fft_tile<N, TW, T, -1>(s, tw);
  float2* __restrict__ dst = J.cout + r0 * H;
  const float sc = J.scale;
  for (int q = threadIdx.x; q < 2 * TW * H; q += T) {
    const int row = q / H, k = q % H;
    if (row < nr) {
      const float2 zk = s[k * P + (row >> 1)];
      const float2 zn = s[((N - k) % N) * P + (row >> 1)];
      float2 v;
      if (row & 1)
        v = make_float2(0.5f * (zk.y + zn.y), -0.5f * (zk.x - zn.x));
      else
        v = make_float2(0.5f * (zk.x + zn.x), 0.5f * (zk.y - zn.y));
      dst[q] = cscale(v, sc);
    }
  }
  for (int row = threadIdx.x; row < nr; row += T) {
    const float2 zk = s[H * P + (row >> 1)];
    float2 v = (row & 1) ? make_float2(zk.y, 0.0f) : make_float2(zk.x, 0.0f);
    J.cout[p.nyq0 + r0 + row] = cscale(v, sc);
  }
}

// Inverse: two half-spectrum rows -> Hermitian-extended Z = X1 + i X2 ->
// inverse FFT -> row pair (re, im).  The imaginary parts of the kz=0 and
// kz=N/2 entries are dropped (real part of the inverse, utils.py:270-271).
// A thread that loads (X1[k], X2[k]) writes both Z[k] = X1 + i X2 and
// Z[N-k] = conj(X1) + i conj(X2), so the half rows never need their own LDS
// staging buffer.
// Stages job J's rows [r0, r0 + nr) of the half spectrum into the LDS tile
// as packed complex lines: line l = rows 2l and 2l + 1.
template <int N, int TW, int T>
__device__ __forceinline__ void c2r_load(float2* s, const ZJob& J,
                                         const PlanDev& p, long long r0,
                                         int nr) {
  constexpr int P = Pitch<TW>::P;
  constexpr int H = N / 2;
  const float2* __restrict__ src = J.cin + r0 * H;
  const float* __restrict__ gz = p.ax[2].g;
  const int ez = J.ez;
  // element q = (line l, k), k = 0..H (k = H is the Nyquist plane)
  constexpr int NE = TW * (H + 1);
  constexpr int E = (NE + T - 1) / T;
  float2 x1[E], x2[E];
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int q = threadIdx.x + u * T;
    const int l = q / (H + 1), k = q % (H + 1);
    const int ra = 2 * l, rb = 2 * l + 1;
    const bool in = NE % T == 0 || q < NE;
    x1[u] = make_float2(0.f, 0.f);
    x2[u] = make_float2(0.f, 0.f);
    if (in && ra < nr)
      x1[u] = k < H ? src[ra * H + k] : J.cin[p.nyq0 + r0 + ra];
    if (in && rb < nr)
      x2[u] = k < H ? src[rb * H + k] : J.cin[p.nyq0 + r0 + rb];
  }
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int q = threadIdx.x + u * T;
    if (NE % T == 0 || q < NE) {
      const int l = q / (H + 1), k = q % (H + 1);
      float2 a = x1[u], b = x2[u];
      if (ez) {
        const float g = gz[k];
        const float c = ez == 2 ? g * g : g;
        a = rot_i(cscale(a, c), ez);
        b = rot_i(cscale(b, c), ez);
      }
      if (k == 0 || k == H) {
        a.y = 0.0f;
        b.y = 0.0f;
      }
      s[k * P + l] = make_float2(a.x - b.y, a.y + b.x);
      if (k > 0 && k < H)  // conj(X1) + i conj(X2)
        s[(N - k) * P + l] = make_float2(a.x + b.y, -a.y + b.x);
    }
  }
}

template <int N, int TW, int T>
__global__ __launch_bounds__(T) void k_zpass_c2r(ZJobs jobs, PlanDev p) {
  constexpr int P = Pitch<TW>::P;
  __shared__ float2 s[N * P];
  __shared__ float2 tw[N];
  // 1-D grid: the jobs of one row chunk are consecutive logical blocks on
  // one XCD, so interleaved outputs (ostride > 1) merge in that L2
  const long long lb = xcd_block(blockIdx.x, gridDim.x);
  const int jb = (int)(lb % jobs.n);
  const ZJob J = jb == 0 ? jobs.j[0] : (jb == 1 ? jobs.j[1] : jobs.j[2]);
  load_twiddles<N, T>(tw, p.ax[2].tw);
  const long long r0 = (lb / jobs.n) * (2 * TW);
  const int nr = (int)min((long long)(2 * TW), p.rows - r0);
  c2r_load<N, TW, T>(s, J, p, r0, nr);
  __syncthreads();
  fft_tile<N, TW, T, +1>(s, tw);
  // output element (row, z) at rout[((r0 + row) * N + z) * os + oc]
  const int os = J.ostride > 0 ? J.ostride : 1;
  float* __restrict__ dst = J.rout + r0 * N * os + J.ocomp;
  const float sc = J.scale;
  for (int q = threadIdx.x; q < 2 * TW * N; q += T) {
    const int row = q / N, z = q % N;
    if (row < nr) {
      const float2 v = s[z * P + (row >> 1)];
      const float r = sc * ((row & 1) ? v.y : v.x);
      float* o = dst + (long long)q * os;
      *o = J.accum ? *o + r : r;
    }
  }
}

// ------------------------------------------------------- column tiles ----
struct Tile {
  long long base;
  long long ls, cs;  // line stride, column stride (one of them is 1)
  int ncol;
  int nyq;           // 0: columns are kz; 1: Nyquist-plane tile
  int fix;           // main: fixed x (Y pass) or y (X pass); nyq: first col
};

// AX = 1 (y lines) or 0 (x lines)
template <int AX, int TW>
__device__ __forceinline__ Tile make_tile(const PlanDev& p, int t) {
  Tile tl;
  const int H = p.H;
  const int nkt = (H + TW - 1) / TW;
  if (AX == 1) {
    const int tiles_main = p.B * p.nx * nkt;
    if (t < tiles_main) {
      const int kt = t % nkt, bx = t / nkt;
      tl.base = (long long)bx * p.ny * H + kt * TW;
      tl.ls = H;
      tl.cs = 1;
      tl.ncol = min(TW, H - kt * TW);
      tl.nyq = 0;
      tl.fix = bx % p.nx;
    } else {
      const int nxt = (p.nx + TW - 1) / TW;
      const int u = t - tiles_main;
      const int xt = u % nxt, b = u / nxt;
      tl.base = p.nyq0 + ((long long)b * p.nx + xt * TW) * p.ny;
      tl.ls = 1;
      tl.cs = p.ny;
      tl.ncol = min(TW, p.nx - xt * TW);
      tl.nyq = 1;
      tl.fix = xt * TW;
    }
  } else {
    const int tiles_main = p.B * p.ny * nkt;
    if (t < tiles_main) {
      const int kt = t % nkt, by = t / nkt;
      const int b = by / p.ny, y = by % p.ny;
      tl.base = ((long long)b * p.nx * p.ny + y) * H + kt * TW;
      tl.ls = (long long)p.ny * H;
      tl.cs = 1;
      tl.ncol = min(TW, H - kt * TW);
      tl.nyq = 0;
      tl.fix = y;
    } else {
      const int nyt = (p.ny + TW - 1) / TW;
      const int u = t - tiles_main;
      const int yt = u % nyt, b = u / nyt;
      tl.base = p.nyq0 + (long long)b * p.nx * p.ny + yt * TW;
      tl.ls = p.ny;
      tl.cs = 1;
      tl.ncol = min(TW, p.ny - yt * TW);
      tl.nyq = 1;
      tl.fix = yt * TW;
    }
  }
  return tl;
}

template <int AX, int TW>
__device__ __forceinline__ int tile_count(const PlanDev& p) {
  const int nkt = (p.H + TW - 1) / TW;
  if (AX == 1) return p.B * p.nx * nkt + p.B * ((p.nx + TW - 1) / TW);
  return p.B * p.ny * nkt + p.B * ((p.ny + TW - 1) / TW);
}

// kz of column c of a main tile: the tile's first kz is base % H.
__device__ __forceinline__ int main_kz0(const Tile& tl, const PlanDev& p) {
  return (int)(tl.base % p.H);
}

// Tile moves between HBM and LDS.  The per-thread element count is a
// compile-time constant and the loops are unrolled in two phases (all global
// loads into registers, then all LDS stores), so every thread keeps E loads
// in flight instead of one load-wait-store chain per element.
template <int N, int TW, int T>
__device__ __forceinline__ void load_tile(float2* __restrict__ s,
                                          const float2* __restrict__ g,
                                          const Tile& tl) {
  constexpr int P = Pitch<TW>::P;
  constexpr int NE = N * TW;
  constexpr int E = (NE + T - 1) / T;
  float2 v[E];
  if (tl.cs == 1) {
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int q = threadIdx.x + u * T;
      const int i = q / TW, c = q % TW;
      v[u] = ((NE % T == 0 || q < NE) && c < tl.ncol)
                 ? g[tl.base + i * tl.ls + c]
                 : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int q = threadIdx.x + u * T;
      if (NE % T == 0 || q < NE) s[(q / TW) * P + q % TW] = v[u];
    }
  } else {
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int q = threadIdx.x + u * T;
      const int i = q % N, c = q / N;
      v[u] = ((NE % T == 0 || q < NE) && c < tl.ncol)
                 ? g[tl.base + c * tl.cs + i]
                 : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int q = threadIdx.x + u * T;
      if (NE % T == 0 || q < NE) s[(q % N) * P + q / N] = v[u];
    }
  }
}

template <int N, int TW, int T>
__device__ __forceinline__ void store_tile(float2* __restrict__ g,
                                           const float2* __restrict__ s,
                                           const Tile& tl, float scale) {
  constexpr int P = Pitch<TW>::P;
  constexpr int NE = N * TW;
  constexpr int E = (NE + T - 1) / T;
  float2 v[E];
  if (tl.cs == 1) {
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int q = threadIdx.x + u * T;
      if (NE % T == 0 || q < NE) v[u] = s[(q / TW) * P + q % TW];
    }
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int q = threadIdx.x + u * T;
      const int i = q / TW, c = q % TW;
      if ((NE % T == 0 || q < NE) && c < tl.ncol)
        g[tl.base + i * tl.ls + c] = cscale(v[u], scale);
    }
  } else {
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int q = threadIdx.x + u * T;
      if (NE % T == 0 || q < NE) v[u] = s[(q % N) * P + q / N];
    }
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int q = threadIdx.x + u * T;
      const int i = q % N, c = q / N;
      if ((NE % T == 0 || q < NE) && c < tl.ncol)
        g[tl.base + c * tl.cs + i] = cscale(v[u], scale);
    }
  }
}

// Y pass.  Forward: in place (out[0] may equal in).  Inverse: each output o
// gets IFFT_y((i g_y)^ey[o] * in) * scale[o].
// FAN = false: every job has one output and no multiplier (one LDS tile,
// twice the workgroups per CU); FAN = true: fan-out with multipliers.
template <int N, int TW, int T, int D, bool FAN>
__global__ __launch_bounds__(T) void k_ypass(YJobs jobs, PlanDev p) {
  constexpr int P = Pitch<TW>::P;
  __shared__ float2 a[N * P];
  __shared__ float2 w[FAN ? N * P : 1];
  __shared__ float2 tw[N];
  const YJob J = blockIdx.y == 0 ? jobs.j[0]
                 : (blockIdx.y == 1 ? jobs.j[1] : jobs.j[2]);
  load_twiddles<N, T>(tw, p.ax[1].tw);
  // XCD-aware tile order: neighbouring column tiles (which share cache
  // lines when a tile row is 64 bytes) run on one XCD, so its L2 fetches a
  // line once instead of each XCD fetching it (8-wide tiles read 2x the
  // bytes without it; y passes at 512^3 230 -> 203 us)
  const Tile tl = make_tile<1, TW>(p, (int)xcd_block(blockIdx.x, gridDim.x));
  load_tile<N, TW, T>(a, J.in, tl);
  __syncthreads();
  if (!FAN || (J.nout == 1 && J.ey[0] == 0)) {
    fft_tile<N, TW, T, D>(a, tw);
    store_tile<N, TW, T>(J.out[0], a, tl, J.scale[0]);
    return;
  }
  const float* __restrict__ gy = p.ax[1].g;
#pragma unroll
  for (int o = 0; o < kMaxJobs; ++o) {
    if (o >= J.nout) break;
    const int e = J.ey[o];
    for (int q = threadIdx.x; q < N * TW; q += T) {
      const int i = q / TW, c = q % TW;
      float2 v = a[i * P + c];
      if (e) {
        const float g = gy[i];
        v = rot_i(cscale(v, e == 2 ? g * g : g), e);
      }
      w[i * P + c] = v;
    }
    __syncthreads();
    fft_tile<N, TW, T, D>(w, tw);
    store_tile<N, TW, T>(J.out[o], w, tl, J.scale[o]);
    __syncthreads();
  }
}

// X pass: [forward FFT_x] -> multiply by Fourier terms -> [inverse FFT_x].
// FAN = false: one input and one output (the time loop's potential), the
// multiply in place on one LDS tile, which lets twice the workgroups share a
// CU; FAN = true: several inputs summed, or several outputs.
template <int N, int TW, int T, bool FAN>
__global__ __launch_bounds__(T) void k_xpass(XJob J, PlanDev p) {
  constexpr int P = Pitch<TW>::P;
  __shared__ float2 a[N * P];
  __shared__ float2 w[FAN ? N * P : 1];
  __shared__ float2 tw[N];
  load_twiddles<N, T>(tw, p.ax[0].tw);
  // XCD-aware tile order, as in k_ypass (x pass at 512^3 423 -> 392 us)
  const Tile tl = make_tile<0, TW>(p, (int)xcd_block(blockIdx.x, gridDim.x));
  const int kz0 = tl.nyq ? p.H : main_kz0(tl, p);
  // per element (i = kx, c): ky, kz
  auto kyz = [&](int c, int& ky, int& kz) {
    if (tl.nyq) {
      ky = min(tl.fix + c, p.ny - 1);
      kz = p.H;
    } else {
      ky = tl.fix;
      kz = min(kz0 + c, p.H);
    }
  };
  if constexpr (!FAN) {
    load_tile<N, TW, T>(a, J.in[0], tl);
    __syncthreads();
    if (J.do_fwd) fft_tile<N, TW, T, -1>(a, tw);
    const Term tm = J.term[0][0];
    for (int q = threadIdx.x; q < N * TW; q += T) {
      const int i = q / TW, c = q % TW;
      int ky, kz;
      kyz(c, ky, kz);
      a[i * P + c] = apply_term(a[i * P + c], tm, i, ky, kz, p);
    }
    __syncthreads();
    if (J.do_inv) fft_tile<N, TW, T, +1>(a, tw);
    store_tile<N, TW, T>(J.out[0], a, tl, 1.0f);
    return;
  } else if (J.nin == 1) {
    load_tile<N, TW, T>(a, J.in[0], tl);
    __syncthreads();
    if (J.do_fwd) fft_tile<N, TW, T, -1>(a, tw);
    for (int o = 0; o < J.nout; ++o) {
      const Term tm = J.term[0][o];
      for (int q = threadIdx.x; q < N * TW; q += T) {
        const int i = q / TW, c = q % TW;
        int ky, kz;
        kyz(c, ky, kz);
        w[i * P + c] = apply_term(a[i * P + c], tm, i, ky, kz, p);
      }
      __syncthreads();
      if (J.do_inv) fft_tile<N, TW, T, +1>(w, tw);
      store_tile<N, TW, T>(J.out[o], w, tl, 1.0f);
      __syncthreads();
    }
  } else {
    for (int q = threadIdx.x; q < N * TW; q += T) {
      const int i = q / TW, c = q % TW;
      w[i * P + c] = make_float2(0.f, 0.f);
    }
    for (int in = 0; in < J.nin; ++in) {
      __syncthreads();
      load_tile<N, TW, T>(a, J.in[in], tl);
      __syncthreads();
      if (J.do_fwd) fft_tile<N, TW, T, -1>(a, tw);
      const Term tm = J.term[in][0];
      for (int q = threadIdx.x; q < N * TW; q += T) {
        const int i = q / TW, c = q % TW;
        int ky, kz;
        kyz(c, ky, kz);
        w[i * P + c] = cadd(w[i * P + c],
                            apply_term(a[i * P + c], tm, i, ky, kz, p));
      }
    }
    __syncthreads();
    if (J.do_inv) fft_tile<N, TW, T, +1>(w, tw);
    store_tile<N, TW, T>(J.out[0], w, tl, 1.0f);
  }
}

// -------------------------------------------- full <-> half spectrum -----
// full[b][x][y][z] (the reference's complex64 layout) from the half spectrum.
__global__ void k_spec_to_full(const float2* __restrict__ spec,
                               float2* __restrict__ full, PlanDev p) {
  const long long M = (long long)p.nx * p.ny * p.nz;
  const long long tot = (long long)p.B * M;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x) {
    const int z = (int)(e % p.nz);
    const long long r = e / p.nz;
    const int y = (int)(r % p.ny);
    const long long r2 = r / p.ny;
    const int x = (int)(r2 % p.nx);
    const int b = (int)(r2 / p.nx);
    int xs = x, ys = y, zs = z;
    bool cj = false;
    if (z > p.H) {
      xs = (p.nx - x) % p.nx;
      ys = (p.ny - y) % p.ny;
      zs = p.nz - z;
      cj = true;
    }
    const long long row = ((long long)b * p.nx + xs) * p.ny + ys;
    float2 v = (zs == p.H) ? spec[p.nyq0 + row] : spec[row * p.H + zs];
    if (cj) v.y = -v.y;
    full[e] = v;
  }
}

// Hermitian projection of an arbitrary full spectrum, kept on the half grid:
// real(ifft(c)) == C2R((c(k) + conj(c(-k)))/2).
__global__ void k_full_to_spec(const float2* __restrict__ full,
                               float2* __restrict__ spec, PlanDev p) {
  const long long tot = (long long)p.rows * (p.H + 1);
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x) {
    const int kz = (int)(e % (p.H + 1));
    const long long row = e / (p.H + 1);
    const int y = (int)(row % p.ny);
    const long long r2 = row / p.ny;
    const int x = (int)(r2 % p.nx);
    const int b = (int)(r2 / p.nx);
    const int xm = (p.nx - x) % p.nx, ym = (p.ny - y) % p.ny,
              zm = (p.nz - kz) % p.nz;
    const long long M = (long long)p.nx * p.ny * p.nz;
    const float2 u = full[b * M + ((long long)x * p.ny + y) * p.nz + kz];
    const float2 v = full[b * M + ((long long)xm * p.ny + ym) * p.nz + zm];
    const float2 hv = make_float2(0.5f * (u.x + v.x), 0.5f * (u.y - v.y));
    if (kz == p.H)
      spec[p.nyq0 + row] = hv;
    else
      spec[row * p.H + kz] = hv;
  }
}

// ------------------------------------------------------------ dispatch ---
// 2^a, 3 * 2^a, 5 * 2^a, 15 * 2^a and 25 * 2^a up to 1024 (multiples of 4
// from 8, as the time loop's z stencil needs).  PMX_FFT_ONLY_16 keeps one
// line length, for the quick host-only checking build of the C ABI.
#ifdef PMX_FFT_ONLY_16
#define PMX_FFT_SIZES(X) X(16)
#else
#define PMX_FFT_SIZES(X)                                                    \
  X(8) X(12) X(16) X(20) X(24) X(32) X(40) X(48) X(60) X(64) X(80) X(96)    \
  X(100) X(120) X(128) X(160) X(192) X(200) X(240) X(256) X(320) X(384)     \
  X(400) X(480) X(512) X(640) X(768) X(800) X(960) X(1024)
#endif

bool fft_size_supported(int n) {
  switch (n) {
#define PMX_CASE(n) case n:
    PMX_FFT_SIZES(PMX_CASE)
#undef PMX_CASE
    return true;
    default:
      return false;
  }
}

// threads per block and tile width per line length
template <int N>
struct Cfg {
  // 512 threads on 256-point rows: a third of the registers per thread of
  // the 256-thread version (97 -> 34 VGPRs for the r2c pass)
  static constexpr int T = (N * 16 >= 4096) ? 512 : 128;
  static constexpr int ZTW = (N >= 1024) ? 4 : (N >= 512 ? 8 : 16);
};

template <int N>
static int zr2c_n(const Plan& pl, const ZJobs& jobs, hipStream_t s) {
  constexpr int TW = Cfg<N>::ZTW, T = Cfg<N>::T;
  dim3 grid((unsigned)((pl.d.rows + 2 * TW - 1) / (2 * TW)), jobs.n);
  ProfScope prof_(kProfZr2c, s);
  hipLaunchKernelGGL((k_zpass_r2c<N, TW, T>), grid, dim3(T), 0, s, jobs,
                     pl.d);
  PMX_LAUNCH_CHECK();
  return kOk;
}
template <int N>
static int zc2r_n(const Plan& pl, const ZJobs& jobs, hipStream_t s) {
  constexpr int TW = Cfg<N>::ZTW, T = Cfg<N>::T;
  ProfScope prof_(kProfZc2r, s);
  dim3 grid((unsigned)((pl.d.rows + 2 * TW - 1) / (2 * TW)) * jobs.n);
  hipLaunchKernelGGL((k_zpass_c2r<N, TW, T>), grid, dim3(T), 0, s, jobs,
                     pl.d);
  PMX_LAUNCH_CHECK();
  return kOk;
}
// Threads per column-tile workgroup: 512 for the large tiles (fewer
// registers per thread: a 256 x 16 x-pass tile needs 106 VGPRs at 512
// threads against 169 at 256, which doubles the waves a CU holds).
template <int N, int TW>
struct ColT {
  static constexpr int T = (N * TW >= 4096) ? 512 : (N * TW >= 2048 ? 256 : 128);
};

template <int N, int TW>
static int y_nt(const Plan& pl, const YJobs& jobs, int inv, hipStream_t s) {
  constexpr int T = ColT<N, TW>::T;
  const PlanDev& p = pl.d;
  const int nkt = (p.H + TW - 1) / TW;
  const int tiles = p.B * p.nx * nkt + p.B * ((p.nx + TW - 1) / TW);
  dim3 grid(tiles, jobs.n);
  ProfScope prof_(inv ? kProfYinv : kProfYfwd, s);
  bool fan = false;
  for (int j = 0; j < jobs.n; ++j)
    fan = fan || jobs.j[j].nout != 1 || jobs.j[j].ey[0] != 0;
  if (inv && fan)
    hipLaunchKernelGGL((k_ypass<N, TW, T, +1, true>), grid, dim3(T), 0, s,
                       jobs, p);
  else if (inv)
    hipLaunchKernelGGL((k_ypass<N, TW, T, +1, false>), grid, dim3(T), 0, s,
                       jobs, p);
  else if (fan)
    hipLaunchKernelGGL((k_ypass<N, TW, T, -1, true>), grid, dim3(T), 0, s,
                       jobs, p);
  else
    hipLaunchKernelGGL((k_ypass<N, TW, T, -1, false>), grid, dim3(T), 0, s,
                       jobs, p);
  PMX_LAUNCH_CHECK();
  return kOk;
}
template <int N, int TW>
static int x_nt(const Plan& pl, const XJob& job, hipStream_t s) {
  constexpr int T = ColT<N, TW>::T;
  const PlanDev& p = pl.d;
  const int nkt = (p.H + TW - 1) / TW;
  const int tiles = p.B * p.ny * nkt + p.B * ((p.ny + TW - 1) / TW);
  ProfScope prof_(kProfX, s);
  if (job.nin == 1 && job.nout == 1)
    hipLaunchKernelGGL((k_xpass<N, TW, T, false>), dim3(tiles), dim3(T), 0, s,
                       job, p);
  else
    hipLaunchKernelGGL((k_xpass<N, TW, T, true>), dim3(tiles), dim3(T), 0, s,
                       job, p);
  PMX_LAUNCH_CHECK();
  return kOk;
}
// Column-tile width: 16 float2 (128-byte rows) or 8 (more workgroups per CU,
// 64-byte rows); lines longer than 512 only fit LDS with 8 (two tiles of
// N x 17 float2 in the fan-out passes).
template <int N>
static int y_n(const Plan& pl, const YJobs& jobs, int inv, hipStream_t s) {
  if constexpr (N > 512) {
    return y_nt<N, 8>(pl, jobs, inv, s);
  } else {
    return pl.tw_y == 8 ? y_nt<N, 8>(pl, jobs, inv, s)
                        : y_nt<N, 16>(pl, jobs, inv, s);
  }
}
template <int N>
static int x_n(const Plan& pl, const XJob& job, hipStream_t s) {
  if constexpr (N > 512) {
    return x_nt<N, 8>(pl, job, s);
  } else {
    return pl.tw_x == 8 ? x_nt<N, 8>(pl, job, s) : x_nt<N, 16>(pl, job, s);
  }
}

int launch_zpass_r2c(const Plan& p, const ZJobs& jobs, hipStream_t s) {
  switch (p.d.nz) {
#define PMX_CASE(n) \
  case n:           \
    return zr2c_n<n>(p, jobs, s);
    PMX_FFT_SIZES(PMX_CASE)
#undef PMX_CASE
  }
  set_error("unsupported nz");
  return kUnsupported;
}
int launch_zpass_c2r(const Plan& p, const ZJobs& jobs, hipStream_t s) {
  switch (p.d.nz) {
#define PMX_CASE(n) \
  case n:           \
    return zc2r_n<n>(p, jobs, s);
    PMX_FFT_SIZES(PMX_CASE)
#undef PMX_CASE
  }
  set_error("unsupported nz");
  return kUnsupported;
}
int launch_ypass(const Plan& p, const YJobs& jobs, int inverse,
                 hipStream_t s) {
  switch (p.d.ny) {
#define PMX_CASE(n) \
  case n:           \
    return y_n<n>(p, jobs, inverse, s);
    PMX_FFT_SIZES(PMX_CASE)
#undef PMX_CASE
  }
  set_error("unsupported ny");
  return kUnsupported;
}
int launch_xpass(const Plan& p, const XJob& job, hipStream_t s) {
  switch (p.d.nx) {
#define PMX_CASE(n) \
  case n:           \
    return x_n<n>(p, job, s);
    PMX_FFT_SIZES(PMX_CASE)
#undef PMX_CASE
  }
  set_error("unsupported nx");
  return kUnsupported;
}

int launch_spec_to_full(const Plan& p, const float2* spec, float2* full,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_spec_to_full, dim3(2048), dim3(256), 0, s, spec, full,
                     p.d);
  PMX_LAUNCH_CHECK();
  return kOk;
}
int launch_full_to_spec(const Plan& p, const float2* full, float2* spec,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_full_to_spec, dim3(2048), dim3(256), 0, s, full, spec,
                     p.d);
  PMX_LAUNCH_CHECK();
  return kOk;
}

}  // namespace pmx
