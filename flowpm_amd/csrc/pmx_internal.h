// Internal declarations shared by the pmx HIP sources (gfx950 only).
//
// Data layout in HBM
//   real mesh      [B][nx][ny][nz] float, z fastest (flowpm/utils.py:32-91)
//   particles      [B][N][3] float, Lagrangian C-order (flowpm/tfpm.py:161-164)
//   state          [3][B][N][3] float = X, P, F (flowpm/tfpm.py:179-180)
//   spectrum       half spectrum split in two blocks of one allocation:
//                    main [B][nx][ny][nz/2] float2   (kz = 0 .. nz/2-1)
//                    nyq  [B][nx][ny]       float2   (kz = nz/2)
//                  Every line transform then reads/writes whole 128-byte
//                  segments: y and x passes tile 16 consecutive kz columns,
//                  the Nyquist plane is tiled along its own contiguous axis.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "brick_index.h"

namespace pmx {

// ------------------------------------------------------------ errors ------
void set_error(const std::string& msg);
const char* last_error();

enum Status : int {
  kOk = 0,
  kBadArg = -1,
  kUnsupported = -2,
  kHipError = -3,
  kNoMemory = -4,
};

#define PMX_HIP_CHECK(expr)                                                 \
  do {                                                                      \
    hipError_t _e = (expr);                                                 \
    if (_e != hipSuccess) {                                                 \
      ::pmx::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));  \
      return ::pmx::kHipError;                                              \
    }                                                                       \
  } while (0)

#define PMX_LAUNCH_CHECK()                                                  \
  do {                                                                      \
    hipError_t _e = hipGetLastError();                                      \
    if (_e != hipSuccess) {                                                 \
      ::pmx::set_error(std::string("kernel launch: ") +                    \
                       hipGetErrorString(_e));                             \
      return ::pmx::kHipError;                                              \
    }                                                                       \
  } while (0)

#define PMX_REQUIRE(cond, msg)                                              \
  do {                                                                      \
    if (!(cond)) {                                                          \
      ::pmx::set_error(msg);                                                \
      return ::pmx::kBadArg;                                                \
    }                                                                       \
  } while (0)

#define PMX_TRY(expr)                                                       \
  do {                                                                      \
    int _s = (expr);                                                        \
    if (_s != 0) return _s;                                                 \
  } while (0)

// --------------------------------------------------- live kernel timing --
// When enabled (pmx_profile_start), every launch is bracketed by a pair of
// hipEvents on its own stream; pmx_profile_stop sums the durations by kind.
enum ProfKind : int {
  kProfFill = 0,
  kProfPaint,
  kProfPaintGather,
  kProfZr2c,
  kProfYfwd,
  kProfX,
  kProfYinv,
  kProfZc2r,
  kProfReadoutKD,
  kProfKickDrift,
  kProfReadout,
  kProfAdjCic,
  kProfPaintW,        // weighted three-channel brick deposit (force VJP)
  kProfPaintWGather,  // its gather
  kProfAdjUpdate,     // KDK cotangent updates and checkpoint copies
  kProfGrad,          // real-space gradient: potential -> force mesh
  kProfGradAdj,       // its adjoint: three mesh cotangents -> potential
  kProfOther,
  kProfKinds
};
extern bool g_prof_on;
void prof_record(int kind, hipStream_t s, hipEvent_t* ev, bool end);
struct ProfScope {
  int kind;
  hipStream_t s;
  hipEvent_t a = nullptr;
  ProfScope(int k, hipStream_t st) : kind(k), s(st) {
    if (g_prof_on) prof_record(kind, s, &a, false);
  }
  ~ProfScope() {
    if (g_prof_on && a) prof_record(kind, s, &a, true);
  }
};

// XCD-aware block order: the hardware deals workgroups round-robin to the 8
// XCDs (block i -> XCD i % 8).  Return a logical block index such that each
// XCD walks one contiguous chunk of the grid, so neighbouring particle rows
// (which gather the same mesh rows) share one L2.
__device__ __forceinline__ long long xcd_block(long long i, long long nb) {
  const long long q = nb / 8, r = nb % 8;
  const long long c = i % 8, slot = i / 8;
  return c * q + (c < r ? c : r) + slot;
}

// ------------------------------------------------------ plan (device) ------
// Per-axis tables, computed on the host in float64 and uploaded once.
struct AxisTab {
  const float2* tw;    // [n]   exp(-2 pi i t / n)
  const double* k2;    // [n]   (2 pi fftfreq(n))^2          kernels.py:10-25
  const float* g;      // [n]   (8 sin w - sin 2w) / 6       kernels.py:51-78
  const double* sinc;  // [n]   numpy.sinc(fftfreq(n))       kernels.py:137-152
};

struct PlanDev {
  int B, nx, ny, nz, H;  // H = nz / 2
  long long S;           // float2 elements in one spectral buffer
  long long nyq0;        // offset of the Nyquist block = B*nx*ny*H
  long long rows;        // B*nx*ny
  AxisTab ax[3];
};

struct Plan {
  PlanDev d;
  void* dev_mem = nullptr;  // all tables in one allocation
  int tw_y = 16, tw_x = 16, tw_z = 16;  // column-tile widths (autotuned)
};

// Multiplier applied in Fourier space:
//   scale * [lap(k) if lap] * prod_d (i g_d(k_d))^e_d
// lap(k) = 1/(kx^2+ky^2+kz^2) evaluated in float64 and rounded to float,
// with the zero mode set to 1 as in kernels.py:42-47.
struct Term {
  float scale;
  int lap;
  int e[3];
  int cic;                 // x (sinc_x sinc_y sinc_z)^-2 (cic_compensation)
  // x PGD_kernel(kl, ks) (kernels.py:105-134) when 1; x its derivative
  // w.r.t. kl when 2, w.r.t. ks when 3 (the PGD parameter gradients)
  int pgd;
  float pgd_kl, pgd_ks, pgd_kl2, pgd_ks4;  // kl, ks, kl^2, ks^4
  // x longrange_kernel = exp(-k^2 r_split^2) (kernels.py:81-102) when > 0,
  // with k^2 r_split^2 in float64 and the exp in float32
  double lr_r2;
};

constexpr int kMaxJobs = 3;

// Z pass (contiguous rows): real -> half spectrum, or back.
struct ZJob {
  const float* rin;   // r2c input
  float2* cout;       // r2c output
  const float2* cin;  // c2r input
  float* rout;        // c2r output
  float scale;
  int ez;             // c2r: exponent of (i g_z) applied before the inverse
  int accum;          // c2r: rout += result instead of rout = result
  int ostride;        // c2r: output element stride (0 or 1: contiguous)
  int ocomp;          // c2r: offset inside the stride (interleaved output)
};
struct ZJobs {
  ZJob j[kMaxJobs];
  int n;
};

// Column pass along y: forward in place, or inverse with fan-out.
struct YJob {
  const float2* in;
  float2* out[kMaxJobs];
  float scale[kMaxJobs];
  int ey[kMaxJobs];
  int nout;
};
struct YJobs {
  YJob j[kMaxJobs];
  int n;
};

// Column pass along x, fused with the Fourier-space multiply:
//   out_o = [IFFT_x] sum_q Term(q,o) * [FFT_x] in_q     (nin == 1 or nout == 1)
struct XJob {
  const float2* in[kMaxJobs];
  float2* out[kMaxJobs];
  Term term[kMaxJobs][kMaxJobs];  // [q][o]
  int nin, nout;
  int do_fwd, do_inv;
};

// Host launchers (fft.hip).  All take the caller's stream.
int launch_zpass_r2c(const Plan& p, const ZJobs& jobs, hipStream_t s);
int launch_zpass_c2r(const Plan& p, const ZJobs& jobs, hipStream_t s);
int launch_ypass(const Plan& p, const YJobs& jobs, int inverse, hipStream_t s);
int launch_xpass(const Plan& p, const XJob& job, hipStream_t s);
int launch_spec_to_full(const Plan& p, const float2* spec, float2* full,
                        hipStream_t s);
int launch_full_to_spec(const Plan& p, const float2* full, float2* spec,
                        hipStream_t s);
bool fft_size_supported(int n);

// CIC launchers (cic.hip).
struct MeshDims {
  int B, nx, ny, nz;
  long long N;  // particles per batch member
};
int launch_fill(float* p, long long n, float v, hipStream_t s);
// Fourth-order central-difference gradient of a real mesh (stencil.hip):
// fm = (D_x phi, D_y phi, D_z phi, 0) per cell as float4 (nc = 4) or the
// first three packed (nc = 3), and the adjoint
// out = -(D_x gx + D_y gy + D_z gz).  Needs nz % 4 == 0.
bool grad_stencil_ok(const Plan& p);
int launch_grad_mesh(const Plan& p, const float* phi, float* fm, int nc,
                     hipStream_t s);
int launch_grad_adj(const Plan& p, const float* gx, const float* gy,
                    const float* gz, float* out, hipStream_t s);
int launch_paint(float* mesh, const float* pos, const float* w, float wscale,
                 const MeshDims& d, float pscale, int mode, hipStream_t s);
int launch_readout(float* out, const float* mesh, const float* pos,
                   const MeshDims& d, float pscale, hipStream_t s);
int launch_paint_adj(float* dpos, float* dw, const float* gmesh,
                     const float* pos, const float* w, const MeshDims& d,
                     float pscale, float gscale, int accumulate,
                     hipStream_t s);
int launch_readout_adj(float* gmesh, float* dpos, const float* gout,
                       const float* mesh, const float* pos, const MeshDims& d,
                       float pscale, int accumulate, hipStream_t s);

// Fused particle update: F = readout(meshes)*fac/f ; P += k2 F ;
// then (optional) P += k1 F ; X += d P      (tfpm.py:273-366, 475-498)
struct KickDrift {
  float fac;      // 1.5 Omega_m
  float f;        // pm_nc_factor
  float k2;       // closing kick of this step
  float k1;       // opening kick of the next step
  float d;        // drift of the next step
  int do_k2, do_k1d;
  int skip_f;     // F is consumed in registers by the next kick: not stored
  // particle grid (N = gx*gy*gz per batch member, C order) or 0: lets the
  // fused readout visit 4 x 64 (y, z) particle tiles, x fastest
  int gx, gy, gz;
  float* xout;    // if set (with do_k1d): drifted X stored here, not in X
  float* xcopy;   // if set (with do_k1d): drifted X also stored here
  // if set (both kicks, no xout / xcopy, packed mesh): the state between
  // the two kicks (X, P after the closing kick, F), laid out as a state
  // (3, B, N, 3), is also stored here (a checkpoint)
  float* ckpt;
  // if set: X / P are read from here instead (the results still go to X
  // or xout and to P)
  const float* xin;
  const float* pin;
};
// The three force meshes: separate (x, y, z), interleaved float4 (f4) or
// interleaved packed (x, y, z) floats (f3).
struct Mesh3 {
  const float* x;
  const float* y;
  const float* z;
  const float4* f4;
  const float* f3;
};
int launch_readout_kick_drift(float* X, float* P, float* F, const Mesh3& m,
                              const MeshDims& d, const KickDrift& c,
                              hipStream_t s);

// out += scale * sum_p sum_c G[p][c] readout(m_c, X_p * f) (float64)
int launch_readout_dot(const float* X, const float* G, const Mesh3& m,
                       const MeshDims& d, float f, double* out, double scale,
                       hipStream_t s);

// Scalar cotangents of one step's coefficients (cic.hip, k_coef_dot):
// out[ia] += sum_p (gP - dstep gX) . F, F = (fac / f) readout(m4, f X), and
// (Xprev set) out[ib] += sum_p gX . (Xd - Xprev) / dstep (Xd null: X);
// part: scratch of 2 * ceil(B N / 256) doubles.
int launch_coef_dot(const float* X, const float* Xd, const float* Xprev,
                    const float* gX, const float* gP, const float4* m4,
                    const MeshDims& d, float f, float fac, float dstep,
                    double* part, double* out, int ia, int ib,
                    hipStream_t s);

// X += c * readout(m, X * f) (the PGD displacement)
int launch_pgd_shift(float* X, const Mesh3& m, const MeshDims& d, float f,
                     float c, hipStream_t s);

int launch_kick_drift_only(float* X, float* P, const float* F,
                           const MeshDims& d, const KickDrift& c,
                           hipStream_t s);
// the same from the state (Xs, Ps, F) into (Xd, Pd): a copy of the state
// and the kick / drift in one pass (either pair may be the same arrays)
int launch_kick_drift_from(const float* Xs, const float* Ps, const float* F,
                           float* Xd, float* Pd, const MeshDims& d,
                           const KickDrift& c, hipStream_t s);

// Philox4x64-10 Gaussian white noise (noise.hip).
int launch_white_noise(float* out, long long n, unsigned long long k0,
                       unsigned long long k1, unsigned long long ctr0,
                       double scale, hipStream_t s);

// Binned power spectrum of a half spectrum (spectrum.hip).
constexpr int kMaxPkBins = 8192;
int launch_power_bins(const Plan& p, const float2* spec, const double* kx,
                      const double* ky, const double* kz, const double* edges,
                      int nedges, double* psum, hipStream_t s, int count = 0);
// spec *= g[b][bin] x the column weight (adjoint of launch_power_bins)
int launch_power_bins_adj(const Plan& p, float2* spec, const double* kx,
                          const double* ky, const double* kz,
                          const double* edges, int nedges, const double* g,
                          hipStream_t s);

// Lagrangian-brick paint (paint.hip).
bool brick_paint_ok(const MeshDims& d, const BrickGrid& g);
long long brick_work_floats(const MeshDims& d, const BrickGrid& g, int C);
// mesh[c] = sum_p w[p*wstride+c]*wscale * CIC(X*ps) (every cell written;
// accumulate: mesh[c] += instead)
int launch_brick_paint(float* const* mesh, int C, const float* X,
                       const float* w, int wstride, float wscale, float ps,
                       const MeshDims& d, const BrickGrid& g, float* work,
                       hipStream_t s, bool accumulate = false);

// Force VJP pieces.
// KDK cotangent update applied after a force VJP (see k_force_vjp_finish)
struct AdjUpdate {
  float d, k1, k2prev;
  int mode;  // 0 none, 1 update, 2 update + k2prev, 3 zero gF
  int gx, gy, gz;  // particle grid (tiled particle order) or 0
  // gin_scale != 0: the readout cotangent is gin_scale * gP, the force
  // cotangent an earlier update left unstored (gin is not read)
  float gin_scale;
  // 1: the update does not store gF = (k1 [+ k2prev]) gP; the next force
  // VJP reads it as gin_scale * gP
  int implicit_gf;
};
// Step-coefficient cotangents taken in the same pass: out[ia] += a and
// out[ib] += b (see launch_coef_dot, with Xd = X), via part (2 doubles per
// 256 particles).
struct CoefDot {
  const float* xprev;
  double* part;
  double* out;
  int ia, ib;
};
// m.f4 set: the density cotangent is m.f4[cell].w and rbar is unused.
// gin: the cotangent of the readout (gF for the force; the position
// cotangent itself for the PGD displacement); gF is the update target.
// cd (optional): the coefficient dot products of the updated cotangents
// (needs m.f4 and an update, mode 1 or 2).
int launch_force_vjp_finish(float* gX, float* gP, float* gF, const float* gin,
                            const float* X, const Mesh3& m, const float* rbar,
                            const MeshDims& d, float f, float cf,
                            const AdjUpdate& u, hipStream_t s,
                            const CoefDot* cd = nullptr);

// scatter = 0 skips the mesh-cotangent atomics (done by the brick paint)
int launch_force_vjp_particles(float* gmx, float* gmy, float* gmz,
                               float* dX, const float* gF, const float* X,
                               const Mesh3& m, const MeshDims& d, float f,
                               float c, int scatter, hipStream_t s);

}  // namespace pmx
