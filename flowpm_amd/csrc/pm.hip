// Plans, PM pipelines (force, LPT, adjoints), the nbody time loop and the
// C ABI of libpmx_hip.so.  See include/pmx.h for the contract of each entry
// point and the reference function it replaces.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "pmx_internal.h"
#include "../../include/pmx.h"

namespace pmx {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }

static inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ------------------------------------------------------ kernel timing ----
bool g_prof_on = false;
struct ProfRec {
  int kind;
  hipEvent_t a, b;
};
static std::vector<ProfRec> g_prof;
static std::mutex g_prof_mu;

void prof_record(int kind, hipStream_t s, hipEvent_t* ev, bool end) {
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return;
  (void)hipEventRecord(e, s);
  if (!end) {
    *ev = e;
    return;
  }
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof.push_back({kind, *ev, e});
}

static const char* kProfNames[kProfKinds] = {
    "fill", "paint", "paint_gather", "fft_z_r2c", "fft_y_fwd", "fft_x_fused", "fft_y_inv",
    "fft_z_c2r", "readout_kick_drift", "kick_drift", "readout", "cic_adjoint",
    "paint_w3", "paint_w3_gather", "adjoint_update", "grad", "grad_adjoint",
    "other"};

// ------------------------------------------------------------------ plan --
static int plan_create(Plan** out, int B, int nx, int ny, int nz) {
  PMX_REQUIRE(out != nullptr, "plan_create: null output");
  PMX_REQUIRE(B >= 1, "plan_create: batch must be >= 1");
  PMX_REQUIRE(fft_size_supported(nx) && fft_size_supported(ny) &&
                  fft_size_supported(nz),
              "plan_create: mesh sides must be in {8,12,16,20,24,32,40,48,"
              "60,64,80,96,100,120,128,160,192,200,240,256,320,384,400,480,"
              "512,640,768,800,960,1024}");
  Plan* p = new Plan();
  PlanDev& d = p->d;
  d.B = B;
  d.nx = nx;
  d.ny = ny;
  d.nz = nz;
  d.H = nz / 2;
  d.rows = (long long)B * nx * ny;
  d.nyq0 = d.rows * d.H;
  d.S = d.rows * (d.H + 1);
  const int n[3] = {nx, ny, nz};
  size_t bytes = 0;
  size_t off_tw[3], off_k2[3], off_g[3], off_sinc[3];
  for (int a = 0; a < 3; ++a) {
    off_k2[a] = bytes;
    bytes += sizeof(double) * n[a];
    off_sinc[a] = bytes;
    bytes += sizeof(double) * n[a];
    off_tw[a] = bytes;
    bytes += sizeof(float2) * n[a];
    off_g[a] = bytes;
    bytes += sizeof(float) * n[a];
    bytes = (bytes + 255) & ~size_t(255);
  }
  std::vector<unsigned char> host(bytes, 0);
  for (int a = 0; a < 3; ++a) {
    double* k2 = reinterpret_cast<double*>(host.data() + off_k2[a]);
    float2* tw = reinterpret_cast<float2*>(host.data() + off_tw[a]);
    float* g = reinterpret_cast<float*>(host.data() + off_g[a]);
    double* sn = reinterpret_cast<double*>(host.data() + off_sinc[a]);
    for (int i = 0; i < n[a]; ++i) {
      // numpy.fft.fftfreq: i/n for i < (n+1)/2 (n even: i < n/2), else
      // (i - n)/n; times 2 pi (kernels.py:16-17)
      const int fi = (i < (n[a] + 1) / 2) ? i : i - n[a];
      const double w = (double)fi / (double)n[a] * 2.0 * M_PI;
      k2[i] = w * w;
      // numpy.sinc(w / 2 pi): sin(pi x) / (pi x), x = 0 -> 1e-20
      const double x = w / (2.0 * M_PI);
      const double y = M_PI * (x == 0.0 ? 1.0e-20 : x);
      sn[i] = std::sin(y) / y;
      g[i] = (float)((8.0 * std::sin(w) - std::sin(2.0 * w)) / 6.0);
      const double ang = -2.0 * M_PI * (double)i / (double)n[a];
      tw[i] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
  }
  hipError_t e = hipMalloc(&p->dev_mem, bytes);
  if (e != hipSuccess) {
    delete p;
    set_error(std::string("plan_create: hipMalloc: ") + hipGetErrorString(e));
    return kNoMemory;
  }
  e = hipMemcpy(p->dev_mem, host.data(), bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(p->dev_mem);
    delete p;
    set_error(std::string("plan_create: hipMemcpy: ") + hipGetErrorString(e));
    return kHipError;
  }
  unsigned char* base = static_cast<unsigned char*>(p->dev_mem);
  for (int a = 0; a < 3; ++a) {
    d.ax[a].k2 = reinterpret_cast<const double*>(base + off_k2[a]);
    d.ax[a].tw = reinterpret_cast<const float2*>(base + off_tw[a]);
    d.ax[a].g = reinterpret_cast<const float*>(base + off_g[a]);
    d.ax[a].sinc = reinterpret_cast<const double*>(base + off_sinc[a]);
  }
  *out = p;
  return kOk;
}

static inline Term term(float scale, int lap, int ex, int ey, int ez) {
  Term t{};
  t.scale = scale;
  t.lap = lap;
  t.e[0] = ex;
  t.e[1] = ey;
  t.e[2] = ez;
  return t;
}

static inline ZJob zr2c(const float* in, float2* out, float scale) {
  ZJob j{};
  j.rin = in;
  j.cout = out;
  j.scale = scale;
  return j;
}
static inline ZJob zc2r(const float2* in, float* out, float scale, int ez) {
  ZJob j{};
  j.cin = in;
  j.rout = out;
  j.scale = scale;
  j.ez = ez;
  return j;
}
static inline YJob yfwd(float2* buf) {
  YJob j{};
  j.in = buf;
  j.out[0] = buf;
  j.scale[0] = 1.0f;
  j.nout = 1;
  return j;
}

// ------------------------------------------------------------ pipelines --
// Forward 3-D transform up to (not including) the x pass: real -> spec.
static int r2c_zy(const Plan& p, const float* const* in, float2* const* spec,
                  int n, float scale, hipStream_t s) {
  ZJobs zj{};
  YJobs yj{};
  zj.n = yj.n = n;
  for (int i = 0; i < n; ++i) {
    zj.j[i] = zr2c(in[i], spec[i], scale);
    yj.j[i] = yfwd(spec[i]);
  }
  PMX_TRY(launch_zpass_r2c(p, zj, s));
  PMX_TRY(launch_ypass(p, yj, 0, s));
  return kOk;
}

// From a spectrum whose x pass is pending (fwd = 1) or done (fwd = 0):
// out_d = IFFT3(spec * lap * (i g_d) * scale), d = x, y, z.
// w0, w1, w2: spec buffers (w0 may alias spec when fwd = 1).
// f4 (optional): write the three meshes interleaved as (x, y, z, -) float4
// instead of to fx / fy / fz.
// pgd (optional): the PGD kernel exp(-kl^2/k^2) exp(-k^4/ks^4) multiplies
// the Laplacian (apply_pgd, tfpm.py:222-270) instead of the long-range 1.
// mode 1: the kernel; 2 / 3: its derivative w.r.t. kl / ks.
struct PgdTerm {
  float kl, ks;
  int mode;
};
static inline Term with_pgd(Term t, const PgdTerm* pgd) {
  if (pgd) {
    t.pgd = pgd->mode;
    t.pgd_kl = pgd->kl;
    t.pgd_ks = pgd->ks;
    t.pgd_kl2 = pgd->kl * pgd->kl;
    t.pgd_ks4 = (float)std::pow((double)pgd->ks, 4.0);
  }
  return t;
}

static int poisson_grad(const Plan& p, const float2* spec, int fwd,
                        float2* w0, float2* w1, float2* w2, float* fx,
                        float* fy, float* fz, float scale, hipStream_t s,
                        const PgdTerm* pgd = nullptr, double lr_r2 = 0.0) {
  XJob xj{};
  xj.nin = 1;
  xj.nout = 2;
  xj.in[0] = spec;
  xj.out[0] = w0;  // lap
  xj.out[1] = w1;  // lap * i gx
  xj.term[0][0] = with_pgd(term(scale, 1, 0, 0, 0), pgd);
  xj.term[0][1] = with_pgd(term(scale, 1, 1, 0, 0), pgd);
  xj.term[0][0].lr_r2 = lr_r2;
  xj.term[0][1].lr_r2 = lr_r2;
  xj.do_fwd = fwd;
  xj.do_inv = 1;
  PMX_TRY(launch_xpass(p, xj, s));
  YJobs yj{};
  yj.n = 2;
  yj.j[0].in = w0;
  yj.j[0].nout = 2;
  yj.j[0].out[0] = w2;  // for Fy: i gy
  yj.j[0].ey[0] = 1;
  yj.j[0].scale[0] = 1.0f;
  yj.j[0].out[1] = w0;  // for Fz
  yj.j[0].ey[1] = 0;
  yj.j[0].scale[1] = 1.0f;
  yj.j[1].in = w1;
  yj.j[1].nout = 1;
  yj.j[1].out[0] = w1;
  yj.j[1].ey[0] = 0;
  yj.j[1].scale[0] = 1.0f;
  PMX_TRY(launch_ypass(p, yj, 1, s));
  ZJobs zj{};
  zj.n = 3;
  zj.j[0] = zc2r(w1, fx, 1.0f, 0);
  zj.j[1] = zc2r(w2, fy, 1.0f, 0);
  zj.j[2] = zc2r(w0, fz, 1.0f, 1);
  PMX_TRY(launch_zpass_c2r(p, zj, s));
  return kOk;
}

// The interleaved force mesh from one potential: phi = c2r(lap * r2c(rho)
// [* PGD]) into the real scratch mesh phi, then the real-space gradient
// (the reference's i g_d multiplier is the fourth-order central difference,
// stencil.hip): one transform chain instead of three.  nc = 4: float4
// cells (x, y, z, -), whose fourth slot the adjoint uses; nc = 3: packed.
static int force_mesh_f4(const Plan& p, const float* rho, float2* w,
                         float* phi, float* f4, float rho_scale,
                         hipStream_t s, const PgdTerm* pgd, int nc = 4) {
  PMX_TRY(r2c_zy(p, &rho, &w, 1, rho_scale, s));
  XJob xj{};
  xj.nin = 1;
  xj.nout = 1;
  xj.in[0] = w;
  xj.out[0] = w;
  xj.term[0][0] = with_pgd(term(1.0f, 1, 0, 0, 0), pgd);
  xj.do_fwd = 1;
  xj.do_inv = 1;
  PMX_TRY(launch_xpass(p, xj, s));
  YJobs yj{};
  yj.n = 1;
  yj.j[0].in = w;
  yj.j[0].nout = 1;
  yj.j[0].out[0] = w;
  yj.j[0].scale[0] = 1.0f;
  PMX_TRY(launch_ypass(p, yj, 1, s));
  ZJobs zj{};
  zj.n = 1;
  zj.j[0] = zc2r(w, phi, 1.0f, 0);
  PMX_TRY(launch_zpass_c2r(p, zj, s));
  return launch_grad_mesh(p, phi, f4, nc, s);
}

// Three separate force meshes (the drop-in force / apply_longrange path).
static int force_meshes(const Plan& p, const float* rho, float2* w, float* fx,
                        float* fy, float* fz, float rho_scale, hipStream_t s,
                        const PgdTerm* pgd = nullptr) {
  float2* w0 = w;
  float2* w1 = w + p.d.S;
  float2* w2 = w + 2 * p.d.S;
  PMX_TRY(r2c_zy(p, &rho, &w0, 1, rho_scale, s));
  return poisson_grad(p, w0, 1, w0, w1, w2, fx, fy, fz, 1.0f, s, pgd);
}

// Adjoint of the map (g_x, g_y, g_z) <- force_meshes(rho):
// grho = IFFT3(sum_d conj(lap i g_d) FFT3(g_d) * rho_scale).
// f4w (optional): write grho into component 3 of this interleaved
// (x, y, z, -) mesh instead of to grho.
// phibar (optional): a real scratch mesh; with it the three cotangents are
// first contracted by the adjoint real-space gradient,
// phibar = -(D_x gx + D_y gy + D_z gz), and one transform chain follows.
static int force_meshes_adj(const Plan& p, const float* gx, const float* gy,
                            const float* gz, float2* w, float* grho,
                            float rho_scale, hipStream_t s,
                            float* f4w = nullptr,
                            const PgdTerm* pgd = nullptr,
                            float* phibar = nullptr) {
  if (phibar && grad_stencil_ok(p)) {
    PMX_TRY(launch_grad_adj(p, gx, gy, gz, phibar, s));
    const float* pb = phibar;
    PMX_TRY(r2c_zy(p, &pb, &w, 1, rho_scale, s));
    XJob xj{};
    xj.nin = 1;
    xj.nout = 1;
    xj.in[0] = w;
    xj.out[0] = w;
    xj.term[0][0] = with_pgd(term(1.0f, 1, 0, 0, 0), pgd);
    xj.do_fwd = 1;
    xj.do_inv = 1;
    PMX_TRY(launch_xpass(p, xj, s));
    YJobs yj{};
    yj.n = 1;
    yj.j[0].in = w;
    yj.j[0].nout = 1;
    yj.j[0].out[0] = w;
    yj.j[0].scale[0] = 1.0f;
    PMX_TRY(launch_ypass(p, yj, 1, s));
    ZJobs zj{};
    zj.n = 1;
    zj.j[0] = zc2r(w, f4w ? f4w : grho, 1.0f, 0);
    if (f4w) {
      zj.j[0].ostride = 4;
      zj.j[0].ocomp = 3;
    }
    return launch_zpass_c2r(p, zj, s);
  }
  float2* ws[3] = {w, w + p.d.S, w + 2 * p.d.S};
  const float* in[3] = {gx, gy, gz};
  PMX_TRY(r2c_zy(p, in, ws, 3, rho_scale, s));
  XJob xj{};
  xj.nin = 3;
  xj.nout = 1;
  for (int d = 0; d < 3; ++d) {
    xj.in[d] = ws[d];
    xj.term[d][0] = with_pgd(term(-1.0f, 1, d == 0, d == 1, d == 2), pgd);
  }
  xj.out[0] = ws[0];
  xj.do_fwd = 1;
  xj.do_inv = 1;
  PMX_TRY(launch_xpass(p, xj, s));
  YJobs yj{};
  yj.n = 1;
  yj.j[0].in = ws[0];
  yj.j[0].nout = 1;
  yj.j[0].out[0] = ws[0];
  yj.j[0].scale[0] = 1.0f;
  PMX_TRY(launch_ypass(p, yj, 1, s));
  ZJobs zj{};
  zj.n = 1;
  zj.j[0] = zc2r(ws[0], f4w ? f4w : grho, 1.0f, 0);
  if (f4w) {
    zj.j[0].ostride = 4;
    zj.j[0].ocomp = 3;
  }
  PMX_TRY(launch_zpass_c2r(p, zj, s));
  return kOk;
}

// Adjoint of poisson_grad(spec, fwd=0): gdk = sum_d conj(lap i g_d) FFT3(g_d)
// * scale (fully transformed half spectrum).
// With pgd / lr_r2 the multiplier carries the same extra real factor as
// the forward (PGD kernel, long-range filter).
static int poisson_grad_adj(const Plan& p, const float* gx, const float* gy,
                            const float* gz, float2* w, float2* gdk,
                            float scale, hipStream_t s,
                            const PgdTerm* pgd = nullptr, double lr_r2 = 0.0) {
  float2* ws[3] = {w, w + p.d.S, w + 2 * p.d.S};
  const float* in[3] = {gx, gy, gz};
  PMX_TRY(r2c_zy(p, in, ws, 3, 1.0f, s));
  XJob xj{};
  xj.nin = 3;
  xj.nout = 1;
  for (int d = 0; d < 3; ++d) {
    xj.in[d] = ws[d];
    xj.term[d][0] =
        with_pgd(term(-scale, 1, d == 0, d == 1, d == 2), pgd);
    xj.term[d][0].lr_r2 = lr_r2;
  }
  xj.out[0] = gdk;
  xj.do_fwd = 1;
  xj.do_inv = 0;
  PMX_TRY(launch_xpass(p, xj, s));
  return kOk;
}

// out (+)= IFFT3(sum_q Term_q FFT3(in_q) * in_scale), nin <= 3.
static int fourier_apply(const Plan& p, int nin, const float* const* in,
                         const Term* terms, float in_scale, float2* w,
                         float* out, int accum, hipStream_t s) {
  float2* ws[3] = {w, w + p.d.S, w + 2 * p.d.S};
  PMX_TRY(r2c_zy(p, in, ws, nin, in_scale, s));
  XJob xj{};
  xj.nin = nin;
  xj.nout = 1;
  for (int q = 0; q < nin; ++q) {
    xj.in[q] = ws[q];
    xj.term[q][0] = terms[q];
  }
  xj.out[0] = ws[0];
  xj.do_fwd = 1;
  xj.do_inv = 1;
  PMX_TRY(launch_xpass(p, xj, s));
  YJobs yj{};
  yj.n = 1;
  yj.j[0].in = ws[0];
  yj.j[0].nout = 1;
  yj.j[0].out[0] = ws[0];
  yj.j[0].scale[0] = 1.0f;
  PMX_TRY(launch_ypass(p, yj, 1, s));
  ZJobs zj{};
  zj.n = 1;
  zj.j[0] = zc2r(ws[0], out, 1.0f, 0);
  zj.j[0].accum = accum;
  return launch_zpass_c2r(p, zj, s);
}

// ---------------------------------------------------------- LPT kernels --
__global__ void k_lpt2_combine(float* __restrict__ out,
                               const float* __restrict__ xx,
                               const float* __restrict__ yy,
                               const float* __restrict__ zz,
                               const float* __restrict__ xy,
                               const float* __restrict__ xz,
                               const float* __restrict__ yz, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    // tfpm.py:130-143: D1 = [1,2,0], D2 = [2,0,1]
    float s = 0.0f;
    s = s + yy[i] * zz[i];
    s = s + zz[i] * xx[i];
    s = s + xx[i] * yy[i];
    s = s - yz[i] * yz[i];
    s = s - xz[i] * xz[i];
    s = s - xy[i] * xy[i];
    out[i] = s * (3.0f / 7.0f);
  }
}

__global__ void k_lpt_assemble(float* __restrict__ state,
                               const float* __restrict__ d1x,
                               const float* __restrict__ d1y,
                               const float* __restrict__ d1z,
                               const float* __restrict__ d2x,
                               const float* __restrict__ d2y,
                               const float* __restrict__ d2z, int B, int nx,
                               int ny, int nz, float c0, float c1, float c2,
                               float c3, float c4, float c5) {
  const long long N = (long long)nx * ny * nz;
  const long long tot = (long long)B * N;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < tot;
       i += (long long)gridDim.x * blockDim.x) {
    const long long q = i % N;
    const int qz = (int)(q % nz);
    const int qy = (int)((q / nz) % ny);
    const int qx = (int)(q / ((long long)nz * ny));
    const float Q[3] = {(float)qx, (float)qy, (float)qz};
    const float a1[3] = {d1x[i], d1y[i], d1z[i]};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      float DX = c0 * a1[d];
      float P = c1 * DX;
      float F = c2 * DX;
      if (d2x) {
        const float a2 = d == 0 ? d2x[i] : (d == 1 ? d2y[i] : d2z[i]);
        const float DX2 = c3 * a2;
        P = P + c4 * DX2;
        F = F + c5 * DX2;
        DX = DX + DX2;
      }
      state[3 * i + d] = DX + Q[d];
      state[3 * tot + 3 * i + d] = P;
      state[6 * tot + 3 * i + d] = F;
    }
  }
}

// Cotangent update of the KDK scheme run backwards (see nbody_adjoint).
__global__ void k_adj_update(float* __restrict__ gX, float* __restrict__ gP,
                             float* __restrict__ gF, long long n, float d,
                             float k1, float k2prev, int mode) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    if (mode == 0) {  // closing kick of the last step: gF += k2 gP
      gF[i] = gF[i] + k2prev * gP[i];
    } else {  // after the force VJP of a step
      const float gp = gP[i] + d * gX[i];
      gP[i] = gp;
      float gf = k1 * gp;
      if (mode == 2) gf = gf + k2prev * gp;
      gF[i] = gf;
    }
  }
}

// Adjoint of k_lpt_assemble: per-component cotangent meshes of dx1 / dx2.
__global__ void k_lpt_assemble_adj(const float* __restrict__ gstate,
                                   float* __restrict__ g1,
                                   float* __restrict__ g2, long long tot,
                                   float c0, float c1, float c2, float c3,
                                   float c4, float c5) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < tot;
       i += (long long)gridDim.x * blockDim.x) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float gX = gstate[3 * i + d];
      const float gP = gstate[3 * tot + 3 * i + d];
      const float gF = gstate[6 * tot + 3 * i + d];
      g1[d * tot + i] = c0 * (gX + c1 * gP + c2 * gF);
      if (g2) g2[d * tot + i] = c3 * (gX + c4 * gP + c5 * gF);
    }
  }
}

// Adjoint of k_lpt2_combine: cotangents of the six phi_ij from that of
// the source.  phi order: xx, yy, zz, xy, xz, yz (contiguous meshes).
__global__ void k_lpt2_combine_adj(const float* __restrict__ phi,
                                   const float* __restrict__ gs,
                                   float* __restrict__ gphi, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float g = gs[i] * (3.0f / 7.0f);
    const float xx = phi[i], yy = phi[n + i], zz = phi[2 * n + i];
    gphi[i] = g * (yy + zz);
    gphi[n + i] = g * (zz + xx);
    gphi[2 * n + i] = g * (xx + yy);
    gphi[3 * n + i] = -2.0f * g * phi[3 * n + i];
    gphi[4 * n + i] = -2.0f * g * phi[4 * n + i];
    gphi[5 * n + i] = -2.0f * g * phi[5 * n + i];
  }
}

// spec *= amp (real, batch-independent, half-spectrum layout of one mesh)
__global__ void k_spec_mul_real(float2* __restrict__ spec,
                                const float* __restrict__ amp, PlanDev p) {
  const long long per_main = (long long)p.nx * p.ny * p.H;
  const long long per_nyq = (long long)p.nx * p.ny;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
       e < p.S; e += (long long)gridDim.x * blockDim.x) {
    float a;
    if (e < p.nyq0)
      a = amp[e % per_main];
    else
      a = amp[per_main + (e - p.nyq0) % per_nyq];
    spec[e] = make_float2(spec[e].x * a, spec[e].y * a);
  }
}

// amp[mode] = table[mx^2 + my^2 + mz^2], m_d the signed frequency index
// of the mode (numpy fftfreq order), over the half-spectrum layout of one
// mesh (main then Nyquist plane): P(k) evaluated once per distinct |k| of a
// box with equal sides instead of once per mode.
__global__ void k_amp_from_table(const float* __restrict__ table,
                                 long long ntable, float* __restrict__ amp,
                                 PlanDev p) {
  const long long per_main = (long long)p.nx * p.ny * p.H;
  const long long tot = per_main + (long long)p.nx * p.ny;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
       e < tot; e += (long long)gridDim.x * blockDim.x) {
    int x, y, z;
    if (e < per_main) {
      z = (int)(e % p.H);
      const long long r = e / p.H;
      y = (int)(r % p.ny);
      x = (int)(r / p.ny);
    } else {
      const long long r = e - per_main;
      y = (int)(r % p.ny);
      x = (int)(r / p.ny);
      z = p.H;
    }
    const long long mx = x < (p.nx + 1) / 2 ? x : x - p.nx;
    const long long my = y < (p.ny + 1) / 2 ? y : y - p.ny;
    const long long mz = z < (p.nz + 1) / 2 ? z : z - p.nz;
    const long long sq = mx * mx + my * my + mz * mz;
    amp[e] = sq < ntable ? table[sq] : __int_as_float(0x7fc00000);
  }
}

static inline unsigned gridfor(long long n) {
  return (unsigned)std::max<long long>(
      1, std::min<long long>((n + 255) / 256, 16384));
}

// ---------------------------------------------------------------- nbody --
// Particles on the Lagrangian grid of (n / f)^3 (the nbody state) let the
// particle passes visit tiles; 0 otherwise.
static void particle_grid(const Plan& p, long long N, float f, int* gx,
                          int* gy, int* gz) {
  *gx = *gy = *gz = 0;
  const int fi = (int)f;
  if ((float)fi == f && fi >= 1 && p.d.nx % fi == 0 && p.d.ny % fi == 0 &&
      p.d.nz % fi == 0 &&
      (long long)(p.d.nx / fi) * (p.d.ny / fi) * (p.d.nz / fi) == N) {
    *gx = p.d.nx / fi;
    *gy = p.d.ny / fi;
    *gz = p.d.nz / fi;
  }
}

static inline Mesh3 m3(const float* f4) {
  return Mesh3{nullptr, nullptr, nullptr, reinterpret_cast<const float4*>(f4),
               nullptr};
}
// force mesh of nc floats per cell (3: packed, 4: float4)
static inline Mesh3 mesh_nc(const float* fm, int nc) {
  return nc == 4 ? m3(fm) : Mesh3{nullptr, nullptr, nullptr, nullptr, fm};
}

struct NbodyWork {
  float* rho;
  float* f4;  // force meshes, interleaved (x, y, z, -) per cell
  float2* w;
  float* brick;  // brick-paint arena (C = 1 and C = 3 share it)
};
// Brick paint needs the particle grid: the mesh divided by pm_nc_factor.
static MeshDims paint_dims(const Plan& p, long long N) {
  return MeshDims{p.d.B, p.d.nx, p.d.ny, p.d.nz, N};
}
static long long brick_floats(const Plan& p) {
  // worst case over f in {1, 2, 4} and C in {1, 3}
  long long m = 0;
  for (int f = 1; f <= 4; f *= 2) {
    if (p.d.nx % f || p.d.ny % f || p.d.nz % f) continue;
    const long long N = (long long)(p.d.nx / f) * (p.d.ny / f) * (p.d.nz / f);
    const MeshDims d = paint_dims(p, N);
    for (int C = 1; C <= 3; C += 2) {
      const BrickGrid g =
          make_brick_grid(p.d.nx / f, p.d.ny / f, p.d.nz / f, f);
      m = std::max(m, brick_work_floats(d, g, C));
    }
  }
  return m;
}
static NbodyWork nbody_work(const Plan& p, float* work) {
  NbodyWork nw;
  const long long M = p.d.rows * p.d.nz;
  nw.rho = work;
  nw.f4 = work + M;
  nw.w = reinterpret_cast<float2*>(work + 5 * M);
  nw.brick = work + 5 * M + 2 * 3 * p.d.S;
  return nw;
}
static long long nbody_work_floats(const Plan& p) {
  const long long M = p.d.rows * p.d.nz;
  return 5 * M + 2 * 3 * p.d.S + brick_floats(p);
}

// Paint C meshes (overwritten) with the brick paint when the particles form
// the grid mesh/f, else with the plain atomic scatter into cleared meshes.
static int paint_meshes(const Plan& p, float* const* mesh, int C,
                        const float* X, const float* w, int wstride,
                        float wscale, long long N, float f, float* brick,
                        hipStream_t s) {
  const long long M = p.d.rows * p.d.nz;
  const int fi = (int)f;
  const MeshDims md = paint_dims(p, N);
  if ((float)fi == f && fi >= 1 && p.d.nx % fi == 0 && p.d.ny % fi == 0 &&
      p.d.nz % fi == 0) {
    const BrickGrid g =
        make_brick_grid(p.d.nx / fi, p.d.ny / fi, p.d.nz / fi, fi);
    if (brick_paint_ok(md, g) && brick_work_floats(md, g, C) <= brick_floats(p))
      return launch_brick_paint(mesh, C, X, w, wstride, wscale, f, md, g,
                                brick, s);
  }
  if (C != 1) return kUnsupported;  // C = 3 only arises on the grid path
  PMX_TRY(launch_fill(mesh[0], M, 0.0f, s));
  return launch_paint(mesh[0], X, w, wscale, md, f, 0, s);
}

// force meshes at X (positions * f), density scaled by 1/nbar and 1/M
// (with pgd: the PGD meshes of the same density); nc floats per cell.
// phi_out (optional): where the potential goes (else the density buffer).
static int force_at(const Plan& p, const float* X, long long N, float f,
                    const NbodyWork& nw, hipStream_t s,
                    const PgdTerm* pgd = nullptr, int nc = 4,
                    float* phi_out = nullptr) {
  const long long M = p.d.rows * p.d.nz;
  const long long Mb = (long long)p.d.nx * p.d.ny * p.d.nz;
  (void)M;
  {
    float* r[1] = {nw.rho};
    PMX_TRY(paint_meshes(p, r, 1, X, nullptr, 1, 1.0f, N, f, nw.brick, s));
  }
  // rho * (1/nbar) with nbar = N / Mb (tfpm.py:346-351), then r2c / Mb
  const double inv_nbar = (double)Mb / (double)N;
  const float sc = (float)(inv_nbar / (double)Mb);
  // the density buffer takes the potential once it is transformed
  PMX_REQUIRE(grad_stencil_ok(p),
              "force: nz must be a multiple of 4, sides >= 4 (nz >= 8), "
              "B * cells < 2^33");
  return force_mesh_f4(p, nw.rho, nw.w, phi_out ? phi_out : nw.rho, nw.f4,
                       sc, s, pgd, nc);
}

// Where a run leaves what the reverse sweep needs of each force evaluation
// (any pointer may be null): the positions (B*N*3 floats) and the
// interleaved force meshes (4M floats) of the jump-start force, and of the
// force of step i at xs / fs + (i - first) slots.
struct Tape {
  float* xj = nullptr;
  float* fj = nullptr;
  float* xs = nullptr;
  float* fs = nullptr;
  // with the PGD correction: the drifted positions before the correction
  // and the PGD meshes at them, per step (same slot numbering)
  float* xp = nullptr;
  float* fp = nullptr;
  int first = 0;
  // Potentials of the run's force evaluations, M floats per slot: slot 0
  // the jump-start force, slot 1 + i the force of step i (absolute step
  // numbers).  phi: the run stores them there; phi_in: the run takes its
  // force meshes from them (the gradient stencil) instead of painting and
  // transforming, as a checkpointed reverse sweep does when it recomputes
  // a segment whose potentials the forward run kept.
  float* phi = nullptr;
  const float* phi_in = nullptr;
  // the caller reads the run's last positions from their slot: the state's
  // X is left as it is (a reverse sweep's scratch state)
  bool x_from_tape = false;
};

// pgd (optional): (alpha, kl, ks) of every step; after the drift of step
// i the positions move by the PGD displacement (tfpm.py:487-488).
// src (optional, not with pgd, at least one step): the run starts from
// this state (X, P, F) instead of the one in state, which it only writes:
// the jump-start force and readout, or the first kick and drift, read src,
// so no copy of it is needed.
// ck_out (optional, with ck_every >= 1; no PGD, no kept positions or force
// meshes): the state after the closing kick of every step i with
// (i + 1 - b0) % ck_every == 0 and i + 1 < b1 (the state a run ending at
// step i + 1 would leave) goes to slot (i + 1 - b0) / ck_every - 1 of
// ck_out, 3 B N 3 floats each, written by that step's readout.
static int nbody_run(const Plan& p, float* state, long long N, float f,
                     float fac, const double* coef, int b0, int b1,
                     int jump, float* work, hipStream_t s,
                     const Tape* tp = nullptr, const float* pgd = nullptr,
                     const float* src = nullptr, float* ck_out = nullptr,
                     int ck_every = 0) {
  PMX_REQUIRE(!(src && (pgd || b0 >= b1)),
              "nbody: a run from a separate state needs at least one step "
              "and no PGD");
  PMX_REQUIRE(!(ck_out && (ck_every < 1 || pgd ||
                           (tp && (tp->xs || tp->fs || tp->xp)))),
              "nbody: checkpoints are written by untaped runs without PGD "
              "(every >= 1)");
  PMX_REQUIRE(!(pgd && tp && (!tp->xp || !tp->fp)),
              "nbody: a PGD run keeps its pre-correction positions and PGD "
              "meshes in the tape");
  PMX_REQUIRE(!(pgd && tp && (tp->phi || tp->phi_in)),
              "nbody: kept potentials are not supported with PGD");
  const long long BN3 = (long long)p.d.B * N * 3;
  const long long M = p.d.rows * p.d.nz;
  const Tape tape = tp ? *tp : Tape{};
  auto keep_x = [&](float* dst, const float* X) -> int {
    if (!dst) return kOk;
    ProfScope prof_(kProfAdjUpdate, s);
    PMX_HIP_CHECK(hipMemcpyAsync(dst, X, sizeof(float) * BN3,
                                 hipMemcpyDeviceToDevice, s));
    return kOk;
  };
  float* X = state;
  float* P = state + BN3;
  float* F = state + 2 * BN3;
  MeshDims md{p.d.B, p.d.nx, p.d.ny, p.d.nz, N};
  const NbodyWork nw0 = nbody_work(p, work);
  auto at = [&](float* dst) {
    NbodyWork w = nw0;
    if (dst)
      w.f4 = dst;
    return w;
  };
  KickDrift kd{};
  kd.fac = fac;
  kd.f = f;
  particle_grid(p, N, f, &kd.gx, &kd.gy, &kd.gz);
  // slot of step i in the tape of the positions the drift produces (the
  // pre-correction ones with PGD), null when not kept
  auto xslot = [&](int i) -> float* {
    float* base = pgd ? tape.xp : tape.xs;
    return base ? base + (long long)(i - tape.first) * BN3 : nullptr;
  };
  bool x_kept = false;  // X of the coming force is already in its slot
  // Without PGD a taped run keeps the drifted positions only in the tape:
  // the readout writes them into the next slot (xout) and the next force
  // reads them there, and one copy at the end returns the last positions
  // to the state.  With PGD the drifted positions are corrected in place,
  // so the tape gets a second copy (xcopy).
  float* Xc = X;  // the current positions: the state's or a tape slot
  auto drift_to = [&](int i) {
    float* slot = xslot(i);
    kd.xout = pgd ? nullptr : slot;
    kd.xcopy = pgd ? slot : nullptr;
    x_kept = slot != nullptr;
  };
  // the force meshes of potential slot k (0: jump start, 1 + i: step i)
  // at Xf: from the kept potential, or painted and transformed (keeping
  // the potential when asked)
  auto force_k = [&](int k, const float* Xf, const NbodyWork& nw,
                     int nc) -> int {
    if (tape.phi_in)
      return launch_grad_mesh(p, tape.phi_in + (long long)k * M, nw.f4, nc,
                              s);
    return force_at(p, Xf, N, f, nw, s, nullptr, nc,
                    tape.phi ? tape.phi + (long long)k * M : nullptr);
  };
  // a force mesh the tape keeps is float4 (the reverse sweep stores the
  // density cotangent in its fourth slot); one only the readout reads is
  // packed, 12 bytes per cell
  if (jump) {
    const float* Xj = src ? src : X;  // the jump-start force's positions
    PMX_TRY(keep_x(tape.xj, Xj));
    const NbodyWork nw = at(tape.fj);
    const int nc = tape.fj ? 4 : 3;
    PMX_TRY(force_k(0, Xj, nw, nc));
    kd.do_k2 = 0;
    kd.do_k1d = (b0 < b1);
    if (kd.do_k1d) {
      kd.k1 = (float)coef[3 * b0];
      kd.d = (float)coef[3 * b0 + 1];
      drift_to(b0);
    }
    kd.xin = src;
    kd.pin = src ? src + BN3 : nullptr;
    PMX_TRY(launch_readout_kick_drift(X, P, F, mesh_nc(nw.f4, nc), md,
                                      kd, s));
    kd.xin = kd.pin = nullptr;
    if (kd.do_k1d && kd.xout) Xc = kd.xout;
  } else if (b0 < b1) {
    // opening kick + drift of the first step from the stored force
    KickDrift c{};
    c.do_k1d = 1;
    c.k1 = (float)coef[3 * b0];
    c.d = (float)coef[3 * b0 + 1];
    if (src) {
      // the drifted positions go straight into their tape slot when the
      // run keeps them (the state's F is written by the run's last step)
      float* xd = xslot(b0);
      PMX_TRY(launch_kick_drift_from(src, src + BN3, src + 2 * BN3,
                                     xd ? xd : X, P, md, c, s));
      if (xd) {
        Xc = xd;
        x_kept = true;
      }
    } else {
      PMX_TRY(launch_kick_drift_only(X, P, F, md, c, s));
    }
  }
  for (int i = b0; i < b1; ++i) {
    const long long slot = i - tape.first;
    if (pgd) {
      if (!x_kept) PMX_TRY(keep_x(xslot(i), Xc));
      const float alpha = pgd[3 * i], kl = pgd[3 * i + 1],
                  ks = pgd[3 * i + 2];
      const PgdTerm pt{kl, ks, 1};
      const NbodyWork nwp = at(tape.fp ? tape.fp + slot * 4 * M : nullptr);
      PMX_TRY(force_at(p, Xc, N, f, nwp, s, &pt));
      PMX_TRY(launch_pgd_shift(Xc, m3(nwp.f4), md, f, alpha / f, s));
      // the corrected positions are the force's
      PMX_TRY(keep_x(tape.xs ? tape.xs + slot * BN3 : nullptr, Xc));
    } else if (!x_kept) {
      PMX_TRY(keep_x(xslot(i), Xc));
    }
    const NbodyWork nw = at(tape.fs ? tape.fs + slot * 4 * M : nullptr);
    const int nc = tape.fs ? 4 : 3;
    PMX_TRY(force_k(1 + i, Xc, nw, nc));
    kd.do_k2 = 1;
    kd.k2 = (float)coef[3 * i + 2];
    kd.do_k1d = (i + 1 < b1);
    kd.xout = kd.xcopy = nullptr;
    x_kept = false;
    if (kd.do_k1d) {
      kd.k1 = (float)coef[3 * (i + 1)];
      kd.d = (float)coef[3 * (i + 1) + 1];
      drift_to(i + 1);
    }
    // only the last force of the run has to stay in the state
    kd.skip_f = kd.do_k1d;
    kd.ckpt = (ck_out && kd.do_k1d && (i + 1 - b0) % ck_every == 0)
                  ? ck_out + (long long)((i + 1 - b0) / ck_every - 1) * 3 * BN3
                  : nullptr;
    PMX_TRY(launch_readout_kick_drift(Xc, P, F, mesh_nc(nw.f4, nc), md,
                                      kd, s));
    if (kd.do_k1d && kd.xout) Xc = kd.xout;
  }
  if (Xc != X && !tape.x_from_tape) {
    ProfScope prof_(kProfAdjUpdate, s);
    PMX_HIP_CHECK(hipMemcpyAsync(X, Xc, sizeof(float) * BN3,
                                 hipMemcpyDeviceToDevice, s));
  }
  return kOk;
}

// X-cotangent contribution of F = force(X) for cotangent gF, followed by
// the KDK cotangent update u.  fm: the force meshes at X from the forward
// pass (recomputed when null).  With pgd the same for the PGD displacement
// S = (alpha / f) readout(PGD meshes): fac is alpha, gin (the cotangent of
// S) is gX itself, and gF is only the target of the update u.
// gF = scale * gP: the force cotangent an earlier update left unstored
__global__ void k_scale_copy(float* __restrict__ dst,
                             const float* __restrict__ src, float scale,
                             long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    dst[i] = scale * src[i];
}

static int store_implicit_gf(float* gF, const float* gP, float* gf_scale,
                             long long BN3, hipStream_t s) {
  if (!gf_scale || *gf_scale == 0.0f) return kOk;
  ProfScope prof_(kProfAdjUpdate, s);
  hipLaunchKernelGGL(k_scale_copy, dim3(gridfor(BN3)), dim3(256), 0, s, gF,
                     gP, *gf_scale, BN3);
  PMX_LAUNCH_CHECK();
  *gf_scale = 0.0f;
  return kOk;
}

// gf_scale (optional, in / out): 0 when gF holds the force cotangent, else
// the factor s with force cotangent s * gP (gF not stored).  On return it
// says the same of the gF this update leaves: the grid path of the plain
// (no PGD) force leaves it unstored.
static int force_vjp(const Plan& p, const float* X, float* gF, float* gX,
                     float* gP, long long N, float f, float fac,
                     const NbodyWork& nw, float* gm, float* fm,
                     const AdjUpdate& u, hipStream_t s,
                     const PgdTerm* pgd = nullptr,
                     const float* gin = nullptr,
                     float* gf_scale = nullptr,
                     const CoefDot* cd = nullptr, bool* cd_done = nullptr) {
  const long long M = p.d.rows * p.d.nz;
  const long long Mb = (long long)p.d.nx * p.d.ny * p.d.nz;
  const long long BN3 = (long long)p.d.B * N * 3;
  if (gin || pgd) PMX_TRY(store_implicit_gf(gF, gP, gf_scale, BN3, s));
  const float gin_scale = gf_scale ? *gf_scale : 0.0f;
  // the readout's cotangent: gin, gF, or gin_scale * gP
  const float* gc = gin ? gin : (gin_scale != 0.0f ? gP : gF);
  const float wsc = gin_scale != 0.0f ? gin_scale * (fac / f) : fac / f;
  MeshDims md{p.d.B, p.d.nx, p.d.ny, p.d.nz, N};
  float* f4 = fm ? fm : nw.f4;
  const Mesh3 m = m3(f4);
  if (!fm) PMX_TRY(force_at(p, X, N, f, nw, s, pgd));
  float* gms[3] = {gm, gm + M, gm + 2 * M};
  const int rc3 = paint_meshes(p, gms, 3, X, gc, 3, wsc, N, f, nw.brick, s);
  if (rc3 != kOk && rc3 != kUnsupported) return rc3;
  const bool grid = rc3 == kOk;
  const double inv_nbar = (double)Mb / (double)N;
  const float sc = (float)(inv_nbar / (double)Mb);
  if (grid) {
    // mesh cotangents by the brick paint; one fused particle pass at the end
    // the density cotangent goes to the unused fourth component of the
    // force mesh, so the finish kernel gathers one float4 per corner
    // nw.rho is free here: it takes the potential's cotangent
    PMX_TRY(force_meshes_adj(p, gm, gm + M, gm + 2 * M, nw.w, nullptr, sc,
                             s, f4, pgd, nw.rho));
    AdjUpdate uu = u;
    uu.gin_scale = gin ? 0.0f : gin_scale;
    const bool upd = u.mode == 1 || u.mode == 2;
    uu.implicit_gf = (gf_scale && !pgd && !gin && upd) ? 1 : 0;
    // the step-coefficient dot products ride along (they need the update)
    const bool fuse = cd && m.f4 && upd;
    PMX_TRY(launch_force_vjp_finish(gX, gP, gF, gc, X, m, nullptr, md, f,
                                    fac / f, uu, s, fuse ? cd : nullptr));
    if (cd_done) *cd_done = fuse;
    if (gf_scale)
      *gf_scale = uu.implicit_gf ? (u.mode == 2 ? u.k1 + u.k2prev : u.k1)
                                 : 0.0f;
    return kOk;
  }
  PMX_REQUIRE(!pgd, "PGD adjoint: particles must form the Lagrangian grid");
  PMX_TRY(store_implicit_gf(gF, gP, gf_scale, BN3, s));
  for (int c = 0; c < 3; ++c) PMX_TRY(launch_fill(gms[c], M, 0.0f, s));
  PMX_TRY(launch_force_vjp_particles(gm, gm + M, gm + 2 * M, gX, gF, X, m,
                                     md, f, fac / f, 1, s));
  PMX_TRY(force_meshes_adj(p, gm, gm + M, gm + 2 * M, nw.w, nw.rho, sc, s));
  PMX_TRY(launch_paint_adj(gX, nullptr, nw.rho, X, nullptr, md, f, 1.0f, 1,
                           s));
  if (u.mode == 3) return launch_fill(gF, BN3, 0.0f, s);
  if (u.mode == 1 || u.mode == 2) {
    ProfScope prof_(kProfAdjUpdate, s);
    hipLaunchKernelGGL(k_adj_update, dim3(gridfor(BN3)), dim3(256), 0, s, gX,
                       gP, gF, BN3, u.d, u.k1, u.k2prev, u.mode);
    PMX_LAUNCH_CHECK();
  }
  return kOk;
}

// Adjoint workspace: nbody work and 3 cotangent meshes; with recomputation
// (every >= 1) also a scratch state, X of every step of a segment and the
// force meshes of every step of a segment plus the jump-start force.
// every = 0: the taped reverse sweep (no recomputation).
static long long nbody_adj_work_floats(const Plan& p, long long N, int every) {
  const long long M = p.d.rows * p.d.nz;
  const long long BN3 = 3LL * p.d.B * N;
  const long long base = nbody_work_floats(p) + 3 * M;
  if (every == 0) return base;
  return base + 3 * BN3 + (long long)every * BN3 +
         (long long)(every + 1) * 4 * M;
}

// Tape of a whole run: X and force meshes of the jump-start force and of
// every step, laid out X slots (nsteps + 1) then force slots (nsteps + 1).
static long long tape_floats(const Plan& p, long long N, int nsteps) {
  const long long M = p.d.rows * p.d.nz;
  const long long BN3 = 3LL * p.d.B * N;
  return (long long)(nsteps + 1) * (BN3 + 4 * M);
}

static Tape tape_of(const Plan& p, float* tape, long long N, int nsteps,
                    bool pgd = false) {
  const long long BN3 = 3LL * p.d.B * N;
  const long long M4 = 4 * p.d.rows * p.d.nz;
  Tape t;
  t.xj = tape;
  t.xs = tape + BN3;
  t.fj = tape + (long long)(nsteps + 1) * BN3;
  t.fs = t.fj + M4;
  if (pgd) {
    t.xp = tape + (long long)(nsteps + 1) * (BN3 + M4);
    t.fp = t.xp + (long long)nsteps * BN3;
  }
  t.first = 0;
  return t;
}

// Compact tape of a whole run: X of the jump-start force and of every step,
// then the potential of every force evaluation (slot 0 the jump start,
// 1 + i step i) instead of its float4 force mesh: (nsteps + 1)(BN3 + M)
// floats against (nsteps + 1)(BN3 + 4M).  The forward run keeps the packed
// force meshes of an untaped run; the reverse sweep takes each step's
// float4 mesh from its potential with the gradient stencil.
static long long compact_tape_floats(const Plan& p, long long N, int nsteps) {
  const long long M = p.d.rows * p.d.nz;
  const long long BN3 = 3LL * p.d.B * N;
  return (long long)(nsteps + 1) * (BN3 + M);
}
static Tape compact_tape_of(const Plan& p, float* tape, long long N,
                            int nsteps) {
  const long long BN3 = 3LL * p.d.B * N;
  Tape t;
  t.xj = tape;
  t.xs = tape + BN3;
  t.phi = tape + (long long)(nsteps + 1) * BN3;
  t.first = 0;
  return t;
}

// PGD run: the tape also keeps, per step, the pre-correction positions and
// the PGD meshes
static long long pgd_tape_floats(const Plan& p, long long N, int nsteps) {
  const long long M = p.d.rows * p.d.nz;
  const long long BN3 = 3LL * p.d.B * N;
  return tape_floats(p, N, nsteps) + (long long)nsteps * (BN3 + 4 * M);
}

__global__ void k_zero_dbl(double* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// out[0] += sum_i a[i] b[i] in float64: block partials, one atomic each
__global__ __launch_bounds__(256) void k_dot(const float* __restrict__ a,
                                             const float* __restrict__ b,
                                             long long n, double* out) {
  __shared__ double part[4];
  double acc = 0.0;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n;
       i += (long long)gridDim.x * 256)
    acc += (double)a[i] * (double)b[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(out, (part[0] + part[1]) + (part[2] + part[3]));
}


// Reverse sweep over steps [b0, b1) whose X and force meshes are in tp.
// Gradients of <G, S> w.r.t. the PGD parameters of one step, S = alpha / f
// readout(PGD meshes at Xpre): out[0] += <G, S> / alpha, out[1] / out[2] +=
// alpha / f <G, readout(meshes of d PGD / d kl or d ks)>.  fpm holds the
// PGD meshes at Xpre (x, y, z components are read).
static int pgd_param_grads(const Plan& p, const float* Xpre, const float* G,
                           const float* fpm, float alpha, float kl, float ks,
                           long long N, float f, const NbodyWork& nw,
                           double* out, hipStream_t s) {
  MeshDims md{p.d.B, p.d.nx, p.d.ny, p.d.nz, N};
  PMX_TRY(launch_readout_dot(Xpre, G, m3(fpm), md, f, out, 1.0 / f, s));
  for (int mode = 2; mode <= 3; ++mode) {
    const PgdTerm pt{kl, ks, mode};
    PMX_TRY(force_at(p, Xpre, N, f, nw, s, &pt));
    PMX_TRY(launch_readout_dot(Xpre, G, m3(nw.f4), md, f, out + mode - 1,
                               (double)alpha / f, s));
  }
  return kOk;
}

// pgd / gpgd (optional): the PGD parameters of every step and where their
// gradients are accumulated (3 float64 per step).
// gsc: the force cotangent's state across calls (see force_vjp's gf_scale)
// Where the reverse sweep accumulates the scalar cotangents of the time
// loop's coefficients (null: not wanted): for a run of nsteps steps,
// out[i] = sum_p gP . F_i over the force F_i of step i (gP the momentum
// cotangent at the kicks that use F_i), out[nsteps + i] = sum_p gX . P_i
// over the momentum P_i of the drift of step i, and out[2 nsteps] the
// first of these for the jump-start force.  xb0: the positions before the
// drift of the sweep's first step.
struct CoefGrad {
  double* out = nullptr;
  int nsteps = 0;
  const float* xb0 = nullptr;
};
static double* coef_scratch(const NbodyWork& nw) {
  const uintptr_t a = reinterpret_cast<uintptr_t>(nw.brick);
  return reinterpret_cast<double*>((a + 7) & ~uintptr_t(7));
}

static int reverse_steps(const Plan& p, const Tape& tp, int b0, int b1,
                         long long N, float f, float fac, const double* coef,
                         float* gstate, const NbodyWork& nw, float* gm,
                         hipStream_t s, float* gsc,
                         const float* pgd = nullptr,
                         double* gpgd = nullptr,
                         const CoefGrad& cg = CoefGrad{}) {
  const long long M = p.d.rows * p.d.nz;
  const long long BN3 = (long long)p.d.B * N * 3;
  float* gX = gstate;
  float* gP = gstate + BN3;
  float* gF = gstate + 2 * BN3;
  const MeshDims md{p.d.B, p.d.nx, p.d.ny, p.d.nz, N};
  for (int i = b1 - 1; i >= b0; --i) {
    AdjUpdate u{};
    u.d = (float)coef[3 * i + 1];
    u.k1 = (float)coef[3 * i];
    u.k2prev = i > 0 ? (float)coef[3 * (i - 1) + 2] : 0.f;
    u.mode = i > 0 ? 2 : 1;
    particle_grid(p, N, f, &u.gx, &u.gy, &u.gz);
    const long long slot = i - tp.first;
    if (!pgd) {
      // the step's float4 force mesh: taped, or (compact tape) from its
      // potential into the force-mesh scratch
      float* fm = tp.fs ? tp.fs + slot * 4 * M : nw.f4;
      if (!tp.fs)
        PMX_TRY(launch_grad_mesh(p, tp.phi_in + (long long)(1 + i) * M, fm,
                                 4, s));
      const float* xprev = i > b0 ? tp.xs + (slot - 1) * BN3 : cg.xb0;
      const CoefDot cd{xprev, coef_scratch(nw), cg.out, i, cg.nsteps + i};
      bool fused = false;
      PMX_TRY(force_vjp(p, tp.xs + slot * BN3, gF, gX, gP, N, f, fac, nw, gm,
                        fm, u, s, nullptr, nullptr, gsc,
                        cg.out ? &cd : nullptr, &fused));
      if (cg.out && !fused)
        PMX_TRY(launch_coef_dot(
            tp.xs + slot * BN3, nullptr, xprev, gX, gP,
            reinterpret_cast<const float4*>(fm), md, f, fac, u.d,
            coef_scratch(nw), cg.out, i, cg.nsteps + i, s));
      continue;
    }
    PMX_TRY(store_implicit_gf(gF, gP, gsc, BN3, s));
    // force at the corrected positions (no update yet), then the PGD
    // displacement, whose cotangent is the position cotangent itself, then
    // the drift / kick update
    AdjUpdate u0 = u;
    u0.mode = 0;
    PMX_TRY(force_vjp(p, tp.xs + slot * BN3, gF, gX, gP, N, f, fac, nw, gm,
                      tp.fs + slot * 4 * M, u0, s));
    const float alpha = pgd[3 * i], kl = pgd[3 * i + 1], ks = pgd[3 * i + 2];
    float* xpre = tp.xp + slot * BN3;
    float* fpm = tp.fp + slot * 4 * M;
    if (gpgd)
      PMX_TRY(pgd_param_grads(p, xpre, gX, fpm, alpha, kl, ks, N, f, nw,
                              gpgd + 3 * i, s));
    const PgdTerm pt{kl, ks, 1};
    PMX_TRY(force_vjp(p, xpre, gF, gX, gP, N, f, alpha, nw, gm, fpm, u, s,
                      &pt, gX));
    if (cg.out) {
      // the force at the corrected positions, the drift to the positions
      // before the correction
      const float* xprev = i > b0 ? tp.xs + (slot - 1) * BN3 : cg.xb0;
      PMX_TRY(launch_coef_dot(
          tp.xs + slot * BN3, xpre, xprev, gX, gP,
          reinterpret_cast<const float4*>(tp.fs + slot * 4 * M), md, f, fac,
          u.d, coef_scratch(nw), cg.out, i, cg.nsteps + i, s));
    }
  }
  return kOk;
}

// First update of the reverse sweep: the closing kick of the last step.
static int adjoint_open(const Plan& p, long long N, int nsteps,
                        const double* coef, float* gstate, hipStream_t s) {
  if (nsteps <= 0) return kOk;
  const long long BN3 = (long long)p.d.B * N * 3;
  ProfScope prof_(kProfAdjUpdate, s);
  hipLaunchKernelGGL(k_adj_update, dim3(gridfor(BN3)), dim3(256), 0, s,
                     gstate, gstate + BN3, gstate + 2 * BN3, BN3, 0.f, 0.f,
                     (float)coef[3 * (nsteps - 1) + 2], 0);
  PMX_LAUNCH_CHECK();
  return kOk;
}

// The recomputation tape of a checkpointed reverse sweep (inside its
// workspace): X and the force meshes of the steps of one segment, and the
// jump-start force.
static Tape segment_tape(const Plan& p, long long N, int every, float* work,
                         int b0) {
  const long long M = p.d.rows * p.d.nz;
  const long long BN3 = (long long)p.d.B * N * 3;
  float* gm = work + nbody_work_floats(p);
  float* st = gm + 3 * M;            // scratch state (3, B, N, 3)
  Tape tp;
  tp.xs = st + 3 * BN3;              // every * BN3
  tp.fs = tp.xs + (long long)every * BN3;  // every * 4M
  tp.fj = tp.fs + (long long)every * 4 * M;
  tp.first = b0;
  return tp;
}

// last_taped: the forward run already recorded the last segment's tape in
// work (nbody_run_last_segment), so that segment is not recomputed.
// phi (optional): the potentials of every force of the forward run (slot 0
// the jump start, 1 + i step i); the recomputation then takes its force
// meshes from them and neither paints nor transforms.
static int nbody_adjoint(const Plan& p, const float* ckpt, int every,
                         int nsteps, long long N, float f, float fac,
                         const double* coef, float* gstate, float* work,
                         hipStream_t s, bool last_taped = false,
                         const float* phi = nullptr,
                         double* gcoef = nullptr) {
  PMX_REQUIRE(every >= 1, "nbody_adjoint: every must be >= 1");
  const long long M = p.d.rows * p.d.nz;
  const long long BN3 = (long long)p.d.B * N * 3;
  NbodyWork nw = nbody_work(p, work);
  float* gm = work + nbody_work_floats(p);
  float* st = gm + 3 * M;            // scratch state (3, B, N, 3)
  const int nseg = (nsteps + every - 1) / every;
  float gsc = 0.0f;  // 0: gF stored; else the force cotangent is gsc * gP
  PMX_TRY(adjoint_open(p, N, nsteps, coef, gstate, s));
  for (int sg = nseg - 1; sg >= 0; --sg) {
    const int b0 = sg * every, b1 = std::min(nsteps, b0 + every);
    Tape tp = segment_tape(p, N, every, work, b0);
    tp.phi_in = phi;
    tp.x_from_tape = true;  // the sweep reads every position from the tape
    if (!(last_taped && sg == nseg - 1)) {
      const float* ck = ckpt + (long long)sg * 3 * BN3;
      if (b0 >= b1) {
        // no step: the jump start reads the state in place
        ProfScope prof_(kProfAdjUpdate, s);
        PMX_HIP_CHECK(hipMemcpyAsync(st, ck, sizeof(float) * 3 * BN3,
                                     hipMemcpyDeviceToDevice, s));
      }
      // recompute the segment from its checkpoint (read in place), keeping
      // X and the force meshes of every step
      PMX_TRY(nbody_run(p, st, N, f, fac, coef, b0, b1, sg == 0, work, s,
                        &tp, nullptr, b0 >= b1 ? nullptr : ck));
    }
    CoefGrad cg;
    cg.out = gcoef;
    cg.nsteps = nsteps;
    cg.xb0 = ckpt + (long long)sg * 3 * BN3;  // X before the segment
    PMX_TRY(reverse_steps(p, tp, b0, b1, N, f, fac, coef, gstate, nw, gm, s,
                          &gsc, nullptr, nullptr, cg));
  }
  float* fjump = segment_tape(p, N, every, work, 0).fj;
  // jump-start force at the initial positions; the initial F is discarded
  AdjUpdate u0{0.f, 0.f, 0.f, 3, 0, 0, 0};
  particle_grid(p, N, f, &u0.gx, &u0.gy, &u0.gz);
  PMX_TRY(force_vjp(p, ckpt, gstate + 2 * BN3, gstate, gstate + BN3, N, f,
                    fac, nw, gm, nsteps > 0 ? fjump : nullptr, u0, s, nullptr,
                    nullptr, &gsc));
  if (gcoef && nsteps > 0)
    PMX_TRY(launch_coef_dot(ckpt, nullptr, nullptr, gstate, gstate + BN3,
                            reinterpret_cast<const float4*>(fjump),
                            MeshDims{p.d.B, p.d.nx, p.d.ny, p.d.nz, N}, f,
                            fac, 0.0f, coef_scratch(nw), gcoef, 2 * nsteps,
                            -1, s));
  return kOk;
}

// Reverse sweep over a tape recorded by the forward run (no recomputation).
static int nbody_adjoint_taped(const Plan& p, float* tape, int nsteps,
                               long long N, float f, float fac,
                               const double* coef, float* gstate,
                               float* work, hipStream_t s,
                               double* gcoef = nullptr,
                               bool compact = false) {
  const long long BN3 = (long long)p.d.B * N * 3;
  NbodyWork nw = nbody_work(p, work);
  float* gm = work + nbody_work_floats(p);
  Tape tp = compact ? compact_tape_of(p, tape, N, nsteps)
                    : tape_of(p, tape, N, nsteps);
  if (compact) {
    // the reverse sweep reads the potentials, and forms each force mesh
    tp.phi_in = tp.phi;
    tp.phi = nullptr;
    PMX_REQUIRE(grad_stencil_ok(p),
                "compact tape: nz must be a multiple of 4, sides >= 4");
  }
  float gsc = 0.0f;  // 0: gF stored; else the force cotangent is gsc * gP
  PMX_TRY(adjoint_open(p, N, nsteps, coef, gstate, s));
  CoefGrad cg;
  cg.out = gcoef;
  cg.nsteps = nsteps;
  cg.xb0 = tp.xj;
  PMX_TRY(reverse_steps(p, tp, 0, nsteps, N, f, fac, coef, gstate, nw, gm,
                        s, &gsc, nullptr, nullptr, cg));
  AdjUpdate u0{0.f, 0.f, 0.f, 3, 0, 0, 0};
  particle_grid(p, N, f, &u0.gx, &u0.gy, &u0.gz);
  float* fj = tp.fj;
  if (compact) {
    fj = nw.f4;
    PMX_TRY(launch_grad_mesh(p, tp.phi_in, fj, 4, s));
  }
  PMX_TRY(force_vjp(p, tp.xj, gstate + 2 * BN3, gstate, gstate + BN3, N, f,
                    fac, nw, gm, fj, u0, s, nullptr, nullptr, &gsc));
  if (gcoef && nsteps > 0)
    PMX_TRY(launch_coef_dot(tp.xj, nullptr, nullptr, gstate, gstate + BN3,
                            reinterpret_cast<const float4*>(fj),
                            MeshDims{p.d.B, p.d.nx, p.d.ny, p.d.nz, N}, f,
                            fac, 0.0f, coef_scratch(nw), gcoef, 2 * nsteps,
                            -1, s));
  return kOk;
}

static int nbody_adjoint_pgd_taped(const Plan& p, float* tape, int nsteps,
                                   long long N, float f, float fac,
                                   const double* coef, const float* pgd,
                                   float* gstate, double* gpgd, float* work,
                                   hipStream_t s, double* gcoef = nullptr) {
  const long long BN3 = (long long)p.d.B * N * 3;
  NbodyWork nw = nbody_work(p, work);
  float* gm = work + nbody_work_floats(p);
  const Tape tp = tape_of(p, tape, N, nsteps, true);
  if (gpgd && nsteps > 0) {
    hipLaunchKernelGGL(k_zero_dbl, dim3((3 * nsteps + 255) / 256), dim3(256),
                       0, s, gpgd, 3 * nsteps);
    PMX_LAUNCH_CHECK();
  }
  float gsc = 0.0f;  // stays 0: every PGD step stores gF
  PMX_TRY(adjoint_open(p, N, nsteps, coef, gstate, s));
  CoefGrad cg;
  cg.out = gcoef;
  cg.nsteps = nsteps;
  cg.xb0 = tp.xj;
  PMX_TRY(reverse_steps(p, tp, 0, nsteps, N, f, fac, coef, gstate, nw, gm, s,
                        &gsc, pgd, gpgd, cg));
  AdjUpdate u0{0.f, 0.f, 0.f, 3, 0, 0, 0};
  particle_grid(p, N, f, &u0.gx, &u0.gy, &u0.gz);
  PMX_TRY(force_vjp(p, tp.xj, gstate + 2 * BN3, gstate, gstate + BN3, N, f,
                    fac, nw, gm, tp.fj, u0, s));
  if (gcoef && nsteps > 0)
    PMX_TRY(launch_coef_dot(tp.xj, nullptr, nullptr, gstate, gstate + BN3,
                            reinterpret_cast<const float4*>(tp.fj),
                            MeshDims{p.d.B, p.d.nx, p.d.ny, p.d.nz, N}, f,
                            fac, 0.0f, coef_scratch(nw), gcoef, 2 * nsteps,
                            -1, s));
  return kOk;
}

}  // namespace pmx

// ================================================================ C ABI ==
using namespace pmx;

struct pmx_plan {
  Plan p;
};

static inline const Plan& P(const pmx_plan* h) { return h->p; }

// ------------------------------------------------------ static scratch --
// Workspace of pmx_nbody_run_noalloc: a file-scope device array, so that
// callers that cannot allocate (code under stream capture, serving loops)
// can run the time loop.  Every device the library runs on holds its own
// copy (64 MiB) until the process ends; calls on one device share it, so
// calls that overlap on one device race.  In host code `scratch` names a
// host-side stand-in: its device address comes from hipGetSymbolAddress,
// once per device, and only that address is ever passed to a kernel.
constexpr long long kScratchFloats = 16LL << 20;
static __device__ __align__(256) float scratch[kScratchFloats];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static float* g_scratch_ptr[kMaxDevices] = {};

// the current device's scratch address (looked up on its first use)
static int scratch_ptr(float** out) {
  int dev = -1;
  PMX_HIP_CHECK(hipGetDevice(&dev));
  PMX_REQUIRE(dev >= 0 && dev < kMaxDevices,
              "static scratch: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    PMX_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    PMX_REQUIRE(p != nullptr, "static scratch: no device address");
    g_scratch_ptr[dev] = static_cast<float*>(p);
  }
  *out = g_scratch_ptr[dev];
  return kOk;
}

extern "C" {

int pmx_version(void) { return PMX_VERSION; }

int pmx_profile_start(void) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  for (auto& r : g_prof) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  g_prof.clear();
  g_prof_on = true;
  return 0;
}

int pmx_profile_stop(double* ms, long long* count, int nkinds) {
  g_prof_on = false;
  std::lock_guard<std::mutex> lk(g_prof_mu);
  for (int k = 0; k < nkinds; ++k) {
    if (ms) ms[k] = 0.0;
    if (count) count[k] = 0;
  }
  int status = 0;
  for (auto& r : g_prof) {
    float t = 0.f;
    if (hipEventSynchronize(r.b) != hipSuccess ||
        hipEventElapsedTime(&t, r.a, r.b) != hipSuccess)
      status = kHipError;
    if (r.kind < nkinds) {
      if (ms) ms[r.kind] += t;
      if (count) count[r.kind] += 1;
    }
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  g_prof.clear();
  if (status) set_error("pmx_profile_stop: event query failed");
  return status;
}

int pmx_profile_kinds(void) { return kProfKinds; }
const char* pmx_profile_kind_name(int k) {
  return (k >= 0 && k < kProfKinds) ? kProfNames[k] : "";
}
const char* pmx_last_error(void) { return last_error(); }

int pmx_plan_create(pmx_plan** plan, int B, int nx, int ny, int nz) {
  PMX_REQUIRE(plan != nullptr, "pmx_plan_create: null output");
  PMX_REQUIRE(nz % 2 == 0, "pmx_plan_create: nz must be even");
  Plan* p = nullptr;
  PMX_TRY(plan_create(&p, B, nx, ny, nz));
  pmx_plan* h = new pmx_plan();
  h->p = *p;
  delete p;
  *plan = h;
  return 0;
}

int pmx_plan_destroy(pmx_plan* plan) {
  if (!plan) return 0;
  if (plan->p.dev_mem) PMX_HIP_CHECK(hipFree(plan->p.dev_mem));
  delete plan;
  return 0;
}

int pmx_fft_plan_create(pmx_plan** plan, int B, int nx, int ny, int nz) {
  return pmx_plan_create(plan, B, nx, ny, nz);
}

int pmx_fft_plan_destroy(pmx_plan* plan) { return pmx_plan_destroy(plan); }

long long pmx_plan_spec_elems(const pmx_plan* plan) {
  return plan ? plan->p.d.S : -1;
}

int pmx_plan_set_tiles(pmx_plan* plan, int ty, int tx) {
  PMX_REQUIRE(plan, "pmx_plan_set_tiles: null plan");
  PMX_REQUIRE((ty == 8 || ty == 16) && (tx == 8 || tx == 16),
              "pmx_plan_set_tiles: tile widths must be 8 or 16");
  plan->p.tw_y = ty;
  plan->p.tw_x = tx;
  return 0;
}

#define CHECK_MESH(B, N, nx, ny, nz)                                   \
  PMX_REQUIRE((B) >= 1 && (N) >= 0 && (nx) >= 1 && (ny) >= 1 &&        \
                  (nz) >= 1,                                           \
              "bad shape: need B >= 1, N >= 0 and positive mesh sides")

int pmx_cic_paint(float* mesh, const float* pos, const float* weight, int B,
                  long long N, int nx, int ny, int nz, float ps,
                  void* stream) {
  CHECK_MESH(B, N, nx, ny, nz);
  PMX_REQUIRE(mesh && (pos || N == 0), "pmx_cic_paint: null pointer");
  MeshDims d{B, nx, ny, nz, N};
  return launch_paint(mesh, pos, weight, 1.0f, d, ps, 0, S(stream));
}

// Particles that number one per mesh cell are painted with the brick paint
// (their grid order, when it is the Lagrangian one, keeps every brick's
// box small; any order is painted correctly, scattered bricks spill to
// atomics).
static bool grid_paint(int B, long long N, int nx, int ny, int nz, float ps,
                       BrickGrid* g) {
  if (ps != 1.0f || N != (long long)nx * ny * nz) return false;
  *g = make_brick_grid(nx, ny, nz, 1);
  const MeshDims d{B, nx, ny, nz, N};
  return brick_paint_ok(d, *g);
}

long long pmx_cic_paint_work_floats(int B, long long N, int nx, int ny,
                                    int nz, float ps) {
  BrickGrid g;
  if (B < 1 || nx < 1 || ny < 1 || nz < 1 ||
      !grid_paint(B, N, nx, ny, nz, ps, &g))
    return 0;
  return brick_work_floats(MeshDims{B, nx, ny, nz, N}, g, 1);
}

int pmx_cic_paint_ws(float* mesh, const float* pos, const float* weight,
                     int B, long long N, int nx, int ny, int nz, float ps,
                     float* work, long long work_floats, void* stream) {
  CHECK_MESH(B, N, nx, ny, nz);
  PMX_REQUIRE(mesh && (pos || N == 0), "pmx_cic_paint_ws: null pointer");
  MeshDims d{B, nx, ny, nz, N};
  BrickGrid g;
  if (work && grid_paint(B, N, nx, ny, nz, ps, &g) &&
      work_floats >= brick_work_floats(d, g, 1)) {
    float* m[1] = {mesh};
    return launch_brick_paint(m, 1, pos, weight, 1, 1.0f, ps, d, g, work,
                              S(stream), true);
  }
  return launch_paint(mesh, pos, weight, 1.0f, d, ps, 0, S(stream));
}

int pmx_cic_readout(float* out, const float* mesh, const float* pos, int B,
                    long long N, int nx, int ny, int nz, float ps,
                    void* stream) {
  CHECK_MESH(B, N, nx, ny, nz);
  PMX_REQUIRE(mesh && ((out && pos) || N == 0),
              "pmx_cic_readout: null pointer");
  MeshDims d{B, nx, ny, nz, N};
  return launch_readout(out, mesh, pos, d, ps, S(stream));
}

int pmx_cic_paint_adj(float* dpos, float* dweight, const float* gmesh,
                      const float* pos, const float* weight, int B,
                      long long N, int nx, int ny, int nz, float ps,
                      void* stream) {
  CHECK_MESH(B, N, nx, ny, nz);
  PMX_REQUIRE(gmesh && (pos || N == 0), "pmx_cic_paint_adj: null pointer");
  if (!dpos && !dweight) return kOk;  // no cotangent asked for
  MeshDims d{B, nx, ny, nz, N};
  return launch_paint_adj(dpos, dweight, gmesh, pos, weight, d, ps, 1.0f, 0,
                          S(stream));
}

int pmx_cic_readout_adj(float* gmesh, float* dpos, const float* gout,
                        const float* mesh, const float* pos, int B,
                        long long N, int nx, int ny, int nz, float ps,
                        void* stream) {
  CHECK_MESH(B, N, nx, ny, nz);
  PMX_REQUIRE(gout && (pos || N == 0), "pmx_cic_readout_adj: null pointer");
  PMX_REQUIRE(dpos == nullptr || mesh != nullptr,
              "pmx_cic_readout_adj: position gradient needs the mesh");
  if (!gmesh && !dpos) return kOk;  // no cotangent asked for
  MeshDims d{B, nx, ny, nz, N};
  return launch_readout_adj(gmesh, dpos, gout, mesh, pos, d, ps, 0,
                            S(stream));
}

int pmx_fft_r2c(const pmx_plan* plan, const float* in, float* spec,
                float scale, void* stream) {
  PMX_REQUIRE(plan && in && spec, "pmx_fft_r2c: null pointer");
  const Plan& p = P(plan);
  float2* w = reinterpret_cast<float2*>(spec);
  PMX_TRY(r2c_zy(p, &in, &w, 1, scale, S(stream)));
  XJob xj{};
  xj.nin = xj.nout = 1;
  xj.in[0] = w;
  xj.out[0] = w;
  xj.term[0][0] = term(1.0f, 0, 0, 0, 0);
  xj.do_fwd = 1;
  xj.do_inv = 0;
  return launch_xpass(p, xj, S(stream));
}

int pmx_fft_c2r(const pmx_plan* plan, float* spec, float* out, float scale,
                void* stream) {
  PMX_REQUIRE(plan && out && spec, "pmx_fft_c2r: null pointer");
  const Plan& p = P(plan);
  float2* w = reinterpret_cast<float2*>(spec);
  XJob xj{};
  xj.nin = xj.nout = 1;
  xj.in[0] = w;
  xj.out[0] = w;
  xj.term[0][0] = term(1.0f, 0, 0, 0, 0);
  xj.do_fwd = 0;
  xj.do_inv = 1;
  PMX_TRY(launch_xpass(p, xj, S(stream)));
  YJobs yj{};
  yj.n = 1;
  yj.j[0].in = w;
  yj.j[0].nout = 1;
  yj.j[0].out[0] = w;
  yj.j[0].scale[0] = 1.0f;
  PMX_TRY(launch_ypass(p, yj, 1, S(stream)));
  ZJobs zj{};
  zj.n = 1;
  zj.j[0] = zc2r(w, out, scale, 0);
  return launch_zpass_c2r(p, zj, S(stream));
}

int pmx_spec_to_full(const pmx_plan* plan, const float* spec, float* full,
                     void* stream) {
  PMX_REQUIRE(plan && spec && full, "pmx_spec_to_full: null pointer");
  return launch_spec_to_full(P(plan), reinterpret_cast<const float2*>(spec),
                             reinterpret_cast<float2*>(full), S(stream));
}

int pmx_full_to_spec(const pmx_plan* plan, const float* full, float* spec,
                     void* stream) {
  PMX_REQUIRE(plan && spec && full, "pmx_full_to_spec: null pointer");
  return launch_full_to_spec(P(plan), reinterpret_cast<const float2*>(full),
                             reinterpret_cast<float2*>(spec), S(stream));
}

int pmx_fft_c2r_poisson_grad(const pmx_plan* plan, const float* dk,
                             float* work, float* fx, float* fy, float* fz,
                             float scale, void* stream) {
  PMX_REQUIRE(plan && dk && work && fx && fy && fz,
              "pmx_fft_c2r_poisson_grad: null pointer");
  const Plan& p = P(plan);
  float2* w = reinterpret_cast<float2*>(work);
  return poisson_grad(p, reinterpret_cast<const float2*>(dk), 0, w,
                      w + p.d.S, w + 2 * p.d.S, fx, fy, fz, scale, S(stream));
}

int pmx_fft_c2r_longrange_grad(const pmx_plan* plan, const float* dk,
                               float* work, float* fx, float* fy, float* fz,
                               float scale, double r_split, void* stream) {
  PMX_REQUIRE(plan && dk && work && fx && fy && fz,
              "pmx_fft_c2r_longrange_grad: null pointer");
  const Plan& p = P(plan);
  float2* w = reinterpret_cast<float2*>(work);
  return poisson_grad(p, reinterpret_cast<const float2*>(dk), 0, w,
                      w + p.d.S, w + 2 * p.d.S, fx, fy, fz, scale, S(stream),
                      nullptr, r_split * r_split);
}

int pmx_fft_longrange_grad_adj(const pmx_plan* plan, const float* gfx,
                               const float* gfy, const float* gfz,
                               float* work, float* gdk, float scale,
                               double r_split, void* stream) {
  PMX_REQUIRE(plan && gfx && gfy && gfz && work && gdk,
              "pmx_fft_longrange_grad_adj: null pointer");
  return poisson_grad_adj(P(plan), gfx, gfy, gfz,
                          reinterpret_cast<float2*>(work),
                          reinterpret_cast<float2*>(gdk), scale, S(stream),
                          nullptr, r_split * r_split);
}

int pmx_fft_c2r_pgd_grad(const pmx_plan* plan, const float* dk, float* work,
                         float* fx, float* fy, float* fz, float scale,
                         float kl, float ks, int deriv, void* stream) {
  PMX_REQUIRE(plan && dk && work && fx && fy && fz,
              "pmx_fft_c2r_pgd_grad: null pointer");
  PMX_REQUIRE(kl != 0.0f && ks != 0.0f,
              "pmx_fft_c2r_pgd_grad: kl and ks must be nonzero");
  PMX_REQUIRE(deriv >= 0 && deriv <= 2,
              "pmx_fft_c2r_pgd_grad: deriv must be 0, 1 or 2");
  const Plan& p = P(plan);
  float2* w = reinterpret_cast<float2*>(work);
  const PgdTerm pt{kl, ks, 1 + deriv};
  return poisson_grad(p, reinterpret_cast<const float2*>(dk), 0, w,
                      w + p.d.S, w + 2 * p.d.S, fx, fy, fz, scale, S(stream),
                      &pt);
}

int pmx_fft_pgd_grad_adj(const pmx_plan* plan, const float* gfx,
                         const float* gfy, const float* gfz, float* work,
                         float* gdk, float scale, float kl, float ks,
                         void* stream) {
  PMX_REQUIRE(plan && gfx && gfy && gfz && work && gdk,
              "pmx_fft_pgd_grad_adj: null pointer");
  PMX_REQUIRE(kl != 0.0f && ks != 0.0f,
              "pmx_fft_pgd_grad_adj: kl and ks must be nonzero");
  const PgdTerm pt{kl, ks, 1};
  return poisson_grad_adj(P(plan), gfx, gfy, gfz,
                          reinterpret_cast<float2*>(work),
                          reinterpret_cast<float2*>(gdk), scale, S(stream),
                          &pt);
}

int pmx_mesh_dot(const float* a, const float* b, long long n, double* out,
                 void* stream) {
  PMX_REQUIRE(a && b && out, "pmx_mesh_dot: null pointer");
  PMX_REQUIRE(n >= 0, "pmx_mesh_dot: n < 0");
  hipStream_t s = S(stream);
  hipLaunchKernelGGL(k_zero_dbl, dim3(1), dim3(64), 0, s, out, 1);
  PMX_LAUNCH_CHECK();
  if (n == 0) return kOk;
  hipLaunchKernelGGL(k_dot, dim3(std::max(1u, std::min(gridfor(n), 2048u))),
                     dim3(256), 0, s, a, b, n, out);
  PMX_LAUNCH_CHECK();
  return kOk;
}

int pmx_pgd(const pmx_plan* plan, float* state, long long N, float f,
            float alpha, float kl, float ks, float* work, void* stream) {
  PMX_REQUIRE(plan && state && work, "pmx_pgd: null pointer");
  PMX_REQUIRE(N >= 1, "pmx_pgd: need at least one particle");
  PMX_REQUIRE(kl != 0.0f && ks != 0.0f, "pmx_pgd: kl and ks must be nonzero");
  PMX_REQUIRE(f >= 1.0f, "pmx_pgd: pm_nc_factor must be >= 1");
  const Plan& p = P(plan);
  hipStream_t s = S(stream);
  const NbodyWork nw = nbody_work(p, work);
  const PgdTerm pt{kl, ks, 1};
  MeshDims md{p.d.B, p.d.nx, p.d.ny, p.d.nz, N};
  PMX_TRY(force_at(p, state, N, f, nw, s, &pt));
  return launch_pgd_shift(state, m3(nw.f4), md, f, alpha / f, s);
}

long long pmx_pgd_adj_work_floats(const pmx_plan* plan, long long N) {
  if (!plan || N < 1) return -1;
  const Plan& p = P(plan);
  return nbody_work_floats(p) + 7 * p.d.rows * p.d.nz;
}

int pmx_pgd_adjoint(const pmx_plan* plan, const float* X, long long N,
                    float f, float alpha, float kl, float ks, float* gstate,
                    double* gpgd, float* work, void* stream) {
  PMX_REQUIRE(plan && X && gstate && work, "pmx_pgd_adjoint: null pointer");
  PMX_REQUIRE(N >= 1, "pmx_pgd_adjoint: need at least one particle");
  PMX_REQUIRE(kl != 0.0f && ks != 0.0f,
              "pmx_pgd_adjoint: kl and ks must be nonzero");
  PMX_REQUIRE(f >= 1.0f, "pmx_pgd_adjoint: pm_nc_factor must be >= 1");
  const Plan& p = P(plan);
  hipStream_t s = S(stream);
  int gx, gy, gz;
  particle_grid(p, N, f, &gx, &gy, &gz);
  PMX_REQUIRE(gx > 0, "pmx_pgd_adjoint: the particles must form the "
                      "Lagrangian grid of the mesh / pm_nc_factor");
  const long long M = p.d.rows * p.d.nz;
  const long long BN3 = 3LL * p.d.B * N;
  const NbodyWork nw = nbody_work(p, work);
  float* gm = work + nbody_work_floats(p);  // 3 cotangent meshes
  float* fpm = gm + 3 * M;                  // PGD meshes at X (float4)
  NbodyWork nwp = nw;
  nwp.f4 = fpm;
  const PgdTerm pt{kl, ks, 1};
  PMX_TRY(force_at(p, X, N, f, nwp, s, &pt));
  float* gX = gstate;
  float* gP = gstate + BN3;
  float* gF = gstate + 2 * BN3;
  if (gpgd) {
    hipLaunchKernelGGL(k_zero_dbl, dim3(1), dim3(64), 0, s, gpgd, 3);
    PMX_LAUNCH_CHECK();
    // before gX is updated: gX is the cotangent of the displacement
    PMX_TRY(pgd_param_grads(p, X, gX, fpm, alpha, kl, ks, N, f, nw, gpgd, s));
  }
  AdjUpdate u{};
  u.mode = 0;
  u.gx = gx;
  u.gy = gy;
  u.gz = gz;
  return force_vjp(p, X, gF, gX, gP, N, f, alpha, nw, gm, fpm, u, s, &pt, gX);
}

int pmx_power_spectrum_bins_adj(const pmx_plan* plan, float* spec,
                                const double* kx, const double* ky,
                                const double* kz, const double* edges,
                                int nedges, const double* gbin,
                                void* stream) {
  PMX_REQUIRE(plan && spec && kx && ky && kz && edges && gbin,
              "pmx_power_spectrum_bins_adj: null pointer");
  return launch_power_bins_adj(P(plan), reinterpret_cast<float2*>(spec), kx,
                               ky, kz, edges, nedges, gbin, S(stream));
}

int pmx_force_meshes(const pmx_plan* plan, const float* rho, float* work,
                     float* fx, float* fy, float* fz, float rho_scale,
                     void* stream) {
  PMX_REQUIRE(plan && rho && work && fx && fy && fz,
              "pmx_force_meshes: null pointer");
  return force_meshes(P(plan), rho, reinterpret_cast<float2*>(work), fx, fy,
                      fz, rho_scale, S(stream));
}

int pmx_force_mesh_f4(const pmx_plan* plan, const float* rho, float* work,
                      float* phi, float* f4, float rho_scale, void* stream) {
  PMX_REQUIRE(plan && rho && work && phi && f4,
              "pmx_force_mesh_f4: null pointer");
  PMX_REQUIRE((reinterpret_cast<uintptr_t>(f4) & 15) == 0,
              "pmx_force_mesh_f4: f4 must be 16-byte aligned");
  const Plan& p = P(plan);
  PMX_REQUIRE(grad_stencil_ok(p),
              "pmx_force_mesh_f4: nz must be a multiple of 4, sides >= 4 "
              "(nz >= 8), B * cells < 2^33");
  return force_mesh_f4(p, rho, reinterpret_cast<float2*>(work), phi, f4,
                       rho_scale, S(stream), nullptr);
}

int pmx_force_meshes_adj(const pmx_plan* plan, const float* gfx,
                         const float* gfy, const float* gfz, float* work,
                         float* grho, float rho_scale, void* stream) {
  PMX_REQUIRE(plan && gfx && gfy && gfz && work && grho,
              "pmx_force_meshes_adj: null pointer");
  return force_meshes_adj(P(plan), gfx, gfy, gfz,
                          reinterpret_cast<float2*>(work), grho, rho_scale,
                          S(stream));
}

int pmx_fft_poisson_grad_adj(const pmx_plan* plan, const float* gfx,
                             const float* gfy, const float* gfz, float* work,
                             float* gdk, float scale, void* stream) {
  PMX_REQUIRE(plan && gfx && gfy && gfz && work && gdk,
              "pmx_fft_poisson_grad_adj: null pointer");
  return poisson_grad_adj(P(plan), gfx, gfy, gfz,
                          reinterpret_cast<float2*>(work),
                          reinterpret_cast<float2*>(gdk), scale, S(stream));
}

long long pmx_lpt2_work_floats(const pmx_plan* plan) {
  if (!plan) return -1;
  const Plan& p = P(plan);
  return 6 * 2 * p.d.S + 7 * p.d.rows * p.d.nz;
}

int pmx_lpt2_source(const pmx_plan* plan, const float* dk, float* work,
                    float* src_k, void* stream) {
  PMX_REQUIRE(plan && dk && work && src_k, "pmx_lpt2_source: null pointer");
  const Plan& p = P(plan);
  hipStream_t s = S(stream);
  float2* w = reinterpret_cast<float2*>(work);
  float2* ws[6];
  for (int i = 0; i < 6; ++i) ws[i] = w + i * p.d.S;
  const long long M = p.d.rows * p.d.nz;
  float* r = work + 6 * 2 * p.d.S;
  float* phi[7];
  for (int i = 0; i < 7; ++i) phi[i] = r + i * M;
  // phi_ij = c2r(dk lap (i g_i)(i g_j)); x pass splits by the x exponent
  XJob xj{};
  xj.nin = 1;
  xj.nout = 3;
  xj.in[0] = reinterpret_cast<const float2*>(dk);
  for (int e = 0; e < 3; ++e) {
    xj.out[e] = ws[e];
    xj.term[0][e] = term(1.0f, 1, e, 0, 0);
  }
  xj.do_fwd = 0;
  xj.do_inv = 1;
  PMX_TRY(launch_xpass(p, xj, s));
  YJobs yj{};
  yj.n = 3;
  // from ex=0: yy (ey=2) -> ws3, yz (ey=1) -> ws4, zz (ey=0) -> ws0
  yj.j[0].in = ws[0];
  yj.j[0].nout = 3;
  yj.j[0].out[0] = ws[3];
  yj.j[0].ey[0] = 2;
  yj.j[0].out[1] = ws[4];
  yj.j[0].ey[1] = 1;
  yj.j[0].out[2] = ws[0];
  yj.j[0].ey[2] = 0;
  // from ex=1: xy (ey=1) -> ws5, xz (ey=0) -> ws1
  yj.j[1].in = ws[1];
  yj.j[1].nout = 2;
  yj.j[1].out[0] = ws[5];
  yj.j[1].ey[0] = 1;
  yj.j[1].out[1] = ws[1];
  yj.j[1].ey[1] = 0;
  // from ex=2: xx -> ws2
  yj.j[2].in = ws[2];
  yj.j[2].nout = 1;
  yj.j[2].out[0] = ws[2];
  yj.j[2].ey[0] = 0;
  for (int j = 0; j < 3; ++j)
    for (int o = 0; o < 3; ++o) yj.j[j].scale[o] = 1.0f;
  PMX_TRY(launch_ypass(p, yj, 1, s));
  // phi: 0 xx, 1 yy, 2 zz, 3 xy, 4 xz, 5 yz
  ZJobs zj{};
  zj.n = 3;
  zj.j[0] = zc2r(ws[2], phi[0], 1.0f, 0);
  zj.j[1] = zc2r(ws[3], phi[1], 1.0f, 0);
  zj.j[2] = zc2r(ws[0], phi[2], 1.0f, 2);
  PMX_TRY(launch_zpass_c2r(p, zj, s));
  zj.j[0] = zc2r(ws[5], phi[3], 1.0f, 0);
  zj.j[1] = zc2r(ws[1], phi[4], 1.0f, 1);
  zj.j[2] = zc2r(ws[4], phi[5], 1.0f, 1);
  PMX_TRY(launch_zpass_c2r(p, zj, s));
  hipLaunchKernelGGL(k_lpt2_combine, dim3(gridfor(M)), dim3(256), 0, s,
                     phi[6], phi[0], phi[1], phi[2], phi[3], phi[4], phi[5],
                     M);
  PMX_LAUNCH_CHECK();
  const long long Mb = (long long)p.d.nx * p.d.ny * p.d.nz;
  return pmx_fft_r2c(plan, phi[6], src_k, (float)(1.0 / (double)Mb), stream);
}

int pmx_lpt_assemble(float* state, const float* dx1, const float* dx2,
                     const double* c, int B, int nx, int ny, int nz,
                     void* stream) {
  PMX_REQUIRE(state && dx1 && c, "pmx_lpt_assemble: null pointer");
  CHECK_MESH(B, 0, nx, ny, nz);
  const long long M = (long long)nx * ny * nz;
  const long long BM = (long long)B * M;
  const float* d2x = dx2;
  const float* d2y = dx2 ? dx2 + BM : nullptr;
  const float* d2z = dx2 ? dx2 + 2 * BM : nullptr;
  hipLaunchKernelGGL(k_lpt_assemble, dim3(gridfor(BM)), dim3(256), 0,
                     S(stream), state, dx1, dx1 + BM, dx1 + 2 * BM, d2x, d2y,
                     d2z, B, nx, ny, nz, (float)c[0], (float)c[1],
                     (float)c[2], (float)c[3], (float)c[4], (float)c[5]);
  PMX_LAUNCH_CHECK();
  return 0;
}

int pmx_lpt_assemble_adj(const float* gstate, const double* c, float* g1,
                         float* g2, int B, int nx, int ny, int nz,
                         void* stream) {
  PMX_REQUIRE(gstate && g1 && c, "pmx_lpt_assemble_adj: null pointer");
  CHECK_MESH(B, 0, nx, ny, nz);
  const long long BM = (long long)B * nx * ny * nz;
  hipLaunchKernelGGL(k_lpt_assemble_adj, dim3(gridfor(BM)), dim3(256), 0,
                     S(stream), gstate, g1, g2, BM, (float)c[0], (float)c[1],
                     (float)c[2], (float)c[3], (float)c[4], (float)c[5]);
  PMX_LAUNCH_CHECK();
  return 0;
}

int pmx_lpt2_source_adj(const pmx_plan* plan, const float* work,
                        const float* gsrc, float* gphi, void* stream) {
  PMX_REQUIRE(plan && work && gsrc && gphi,
              "pmx_lpt2_source_adj: null pointer");
  const Plan& p = P(plan);
  const long long M = p.d.rows * p.d.nz;
  const float* phi = work + 6 * 2 * p.d.S;
  hipLaunchKernelGGL(k_lpt2_combine_adj, dim3(gridfor(M)), dim3(256), 0,
                     S(stream), phi, gsrc, gphi, M);
  PMX_LAUNCH_CHECK();
  return 0;
}

int pmx_fourier_apply(const pmx_plan* plan, int nin, const float* const* in,
                      const int* e, const float* scales, int lap,
                      float in_scale, float* work, float* out, int accumulate,
                      void* stream) {
  PMX_REQUIRE(plan && in && e && scales && work && out,
              "pmx_fourier_apply: null pointer");
  PMX_REQUIRE(nin >= 1 && nin <= 3, "pmx_fourier_apply: 1 <= nin <= 3");
  PMX_REQUIRE((lap & ~3) == 0, "pmx_fourier_apply: unknown kernel flag");
  Term t[3] = {};
  for (int q = 0; q < nin; ++q) {
    PMX_REQUIRE(in[q] != nullptr, "pmx_fourier_apply: null input");
    for (int d = 0; d < 3; ++d)
      PMX_REQUIRE(e[3 * q + d] >= 0 && e[3 * q + d] <= 2,
                  "pmx_fourier_apply: exponents must be 0, 1 or 2");
    t[q] = term(scales[q], lap & 1, e[3 * q], e[3 * q + 1], e[3 * q + 2]);
    t[q].cic = (lap >> 1) & 1;
  }
  return fourier_apply(P(plan), nin, in, t, in_scale,
                       reinterpret_cast<float2*>(work), out, accumulate,
                       S(stream));
}

int pmx_white_noise(float* out, long long n, unsigned long long key_lo,
                    unsigned long long key_hi, unsigned long long counter,
                    double scale, void* stream) {
  PMX_REQUIRE(out || n == 0, "pmx_white_noise: null output");
  PMX_REQUIRE(n >= 0, "pmx_white_noise: n < 0");
  return launch_white_noise(out, n, key_lo, key_hi, counter, scale,
                            S(stream));
}

int pmx_power_spectrum_bins(const pmx_plan* plan, const float* spec,
                            const double* kx, const double* ky,
                            const double* kz, const double* edges, int nedges,
                            double* psum, void* stream) {
  PMX_REQUIRE(plan && spec && kx && ky && kz && edges && psum,
              "pmx_power_spectrum_bins: null pointer");
  return launch_power_bins(P(plan), reinterpret_cast<const float2*>(spec), kx,
                           ky, kz, edges, nedges, psum, S(stream));
}

int pmx_power_spectrum_counts(const pmx_plan* plan, const double* kx,
                              const double* ky, const double* kz,
                              const double* edges, int nedges, double* nsum,
                              void* stream) {
  PMX_REQUIRE(plan && kx && ky && kz && edges && nsum,
              "pmx_power_spectrum_counts: null pointer");
  return launch_power_bins(P(plan), nullptr, kx, ky, kz, edges, nedges, nsum,
                           S(stream), 1);
}

int pmx_amp_from_table(const pmx_plan* plan, const float* table,
                       long long ntable, float* amp, void* stream) {
  PMX_REQUIRE(plan && table && amp, "pmx_amp_from_table: null pointer");
  const PlanDev& d = P(plan).d;
  const long long need = (long long)(d.nx / 2) * (d.nx / 2) +
                         (long long)(d.ny / 2) * (d.ny / 2) +
                         (long long)(d.nz / 2) * (d.nz / 2) + 1;
  PMX_REQUIRE(ntable >= need,
              "pmx_amp_from_table: table shorter than the largest mx^2 + my^2 "
              "+ mz^2 + 1");
  const long long tot = (long long)d.nx * d.ny * (d.H + 1);
  hipLaunchKernelGGL(k_amp_from_table, dim3(gridfor(tot)), dim3(256), 0,
                     S(stream), table, ntable, amp, d);
  PMX_LAUNCH_CHECK();
  return 0;
}

int pmx_spec_mul_real(const pmx_plan* plan, float* spec, const float* amp,
                      void* stream) {
  PMX_REQUIRE(plan && spec && amp, "pmx_spec_mul_real: null pointer");
  hipLaunchKernelGGL(k_spec_mul_real, dim3(gridfor(P(plan).d.S)), dim3(256),
                     0, S(stream), reinterpret_cast<float2*>(spec), amp,
                     P(plan).d);
  PMX_LAUNCH_CHECK();
  return 0;
}

int pmx_readout_kick_drift(float* X, float* P, float* F, const float* fx,
                           const float* fy, const float* fz, int B,
                           long long N, int nx, int ny, int nz, float f,
                           float fac, float k2, float k1, float d, int flags,
                           void* stream) {
  CHECK_MESH(B, N, nx, ny, nz);
  PMX_REQUIRE(X && P && F, "pmx_readout_kick_drift: null state pointer");
  MeshDims md{B, nx, ny, nz, N};
  if (flags & 4) {  // kick/drift only, from the stored force
    KickDrift c{};
    c.k1 = k1;
    c.d = d;
    c.k2 = k2;
    c.do_k2 = flags & 1;
    c.do_k1d = (flags >> 1) & 1;
    return launch_kick_drift_only(X, P, F, md, c, S(stream));
  }
  PMX_REQUIRE(fx && fy && fz, "pmx_readout_kick_drift: null mesh pointer");
  KickDrift c{};
  c.fac = fac;
  c.f = f;
  c.k2 = k2;
  c.k1 = k1;
  c.d = d;
  c.do_k2 = flags & 1;
  c.do_k1d = (flags >> 1) & 1;
  return launch_readout_kick_drift(X, P, F, Mesh3{fx, fy, fz, nullptr, nullptr}, md, c,
                                   S(stream));
}

long long pmx_nbody_work_floats(const pmx_plan* plan) {
  return plan ? nbody_work_floats(P(plan)) : -1;
}

int pmx_nbody_run(const pmx_plan* plan, float* state, long long N, float f,
                  float fac, const double* coef, int b0, int b1, int jump,
                  float* work, void* stream) {
  PMX_REQUIRE(plan && state && work, "pmx_nbody_run: null pointer");
  PMX_REQUIRE(N >= 1, "pmx_nbody_run: need at least one particle");
  PMX_REQUIRE(b0 >= 0 && b1 >= b0 && (coef || b1 == b0),
              "pmx_nbody_run: bad step range");
  return nbody_run(P(plan), state, N, f, fac, coef, b0, b1, jump, work,
                   S(stream));
}

int pmx_nbody_run_noalloc(const pmx_plan* plan, float* state, long long N,
                          float f, float fac, const double* coef, int b0,
                          int b1, int jump, void* stream) {
  PMX_REQUIRE(plan && state, "pmx_nbody_run_noalloc: null pointer");
  PMX_REQUIRE(N >= 1, "pmx_nbody_run_noalloc: need at least one particle");
  PMX_REQUIRE(b0 >= 0 && b1 >= b0 && (coef || b1 == b0),
              "pmx_nbody_run_noalloc: bad step range");
  const long long need = nbody_work_floats(P(plan));
  if (need > kScratchFloats) {
    set_error("pmx_nbody_run_noalloc: this plan needs " +
              std::to_string(need) + " floats of scratch; the static scratch "
              "holds " + std::to_string(kScratchFloats) +
              " (use pmx_nbody_run with a workspace)");
    return kBadArg;
  }
  float* work = nullptr;
  PMX_TRY(scratch_ptr(&work));
  return nbody_run(P(plan), state, N, f, fac, coef, b0, b1, jump, work,
                   S(stream));
}

int pmx_nbody_run_pgd(const pmx_plan* plan, float* state, long long N,
                      float f, float fac, const double* coef,
                      const float* pgd, int b0, int b1, int jump, float* work,
                      void* stream) {
  PMX_REQUIRE(plan && state && work && pgd,
              "pmx_nbody_run_pgd: null pointer");
  PMX_REQUIRE(N >= 1, "pmx_nbody_run_pgd: need at least one particle");
  PMX_REQUIRE(b0 >= 0 && b1 >= b0 && (coef || b1 == b0),
              "pmx_nbody_run_pgd: bad step range");
  for (int i = b0; i < b1; ++i)
    PMX_REQUIRE(pgd[3 * i + 1] != 0.0f && pgd[3 * i + 2] != 0.0f,
                "pmx_nbody_run_pgd: kl and ks must be nonzero");
  return nbody_run(P(plan), state, N, f, fac, coef, b0, b1, jump, work,
                   S(stream), nullptr, pgd);
}

long long pmx_nbody_adj_work_floats(const pmx_plan* plan, long long N,
                                    int every) {
  return (plan && every >= 0) ? nbody_adj_work_floats(P(plan), N, every) : -1;
}

int pmx_nbody_batched_step(int ngroups, const pmx_plan* const* plans,
                           float* const* states, const long long* N,
                           const float* f, const float* fac,
                           const double* const* coefs, const int* step_begin,
                           const int* step_end, const int* jump_start,
                           float* const* work, void* stream) {
  PMX_REQUIRE(ngroups >= 0, "pmx_nbody_batched_step: ngroups < 0");
  if (ngroups == 0) return 0;
  PMX_REQUIRE(plans && states && N && f && fac && coefs && step_begin &&
                  step_end && jump_start && work,
              "pmx_nbody_batched_step: null array");
  // validate every group before anything is enqueued
  for (int g = 0; g < ngroups; ++g) {
    PMX_REQUIRE(plans[g] && states[g] && work[g],
                "pmx_nbody_batched_step: null pointer in a group");
    PMX_REQUIRE(N[g] >= 1, "pmx_nbody_batched_step: group without particles");
    PMX_REQUIRE(step_begin[g] >= 0 && step_end[g] >= step_begin[g] &&
                    (coefs[g] || step_end[g] == step_begin[g]),
                "pmx_nbody_batched_step: bad step range");
  }
  for (int g = 0; g < ngroups; ++g)
    PMX_TRY(nbody_run(P(plans[g]), states[g], N[g], f[g], fac[g], coefs[g],
                      step_begin[g], step_end[g], jump_start[g], work[g],
                      S(stream)));
  return 0;
}

long long pmx_nbody_tape_floats(const pmx_plan* plan, long long N,
                                int nsteps) {
  return (plan && N >= 1 && nsteps >= 0) ? tape_floats(P(plan), N, nsteps)
                                         : -1;
}

int pmx_nbody_run_taped(const pmx_plan* plan, float* state, long long N,
                        float f, float fac, const double* coef, int nsteps,
                        float* tape, float* work, void* stream) {
  PMX_REQUIRE(plan && state && tape && work,
              "pmx_nbody_run_taped: null pointer");
  PMX_REQUIRE(N >= 1, "pmx_nbody_run_taped: need at least one particle");
  PMX_REQUIRE(nsteps >= 0 && (coef || nsteps == 0),
              "pmx_nbody_run_taped: bad steps");
  const Tape tp = tape_of(P(plan), tape, N, nsteps);
  return nbody_run(P(plan), state, N, f, fac, coef, 0, nsteps, 1, work,
                   S(stream), &tp);
}

long long pmx_nbody_compact_tape_floats(const pmx_plan* plan, long long N,
                                        int nsteps) {
  return (plan && N >= 1 && nsteps >= 0)
             ? compact_tape_floats(P(plan), N, nsteps)
             : -1;
}

int pmx_nbody_run_compact_taped(const pmx_plan* plan, float* state,
                                long long N, float f, float fac,
                                const double* coef, int nsteps, float* tape,
                                float* work, void* stream) {
  PMX_REQUIRE(plan && state && tape && work,
              "pmx_nbody_run_compact_taped: null pointer");
  PMX_REQUIRE(N >= 1, "pmx_nbody_run_compact_taped: need at least one "
                      "particle");
  PMX_REQUIRE(nsteps >= 0 && (coef || nsteps == 0),
              "pmx_nbody_run_compact_taped: bad steps");
  const Tape tp = compact_tape_of(P(plan), tape, N, nsteps);
  return nbody_run(P(plan), state, N, f, fac, coef, 0, nsteps, 1, work,
                   S(stream), &tp);
}

int pmx_nbody_adjoint_compact_taped(const pmx_plan* plan, float* tape,
                                    int nsteps, long long N, float f,
                                    float fac, const double* coef,
                                    float* gstate, float* work,
                                    double* gcoef, void* stream) {
  PMX_REQUIRE(plan && tape && gstate && work,
              "pmx_nbody_adjoint_compact_taped: null pointer");
  PMX_REQUIRE(nsteps >= 0 && (coef || nsteps == 0),
              "pmx_nbody_adjoint_compact_taped: bad steps");
  return nbody_adjoint_taped(P(plan), tape, nsteps, N, f, fac, coef, gstate,
                             work, S(stream), gcoef, true);
}

long long pmx_nbody_pgd_tape_floats(const pmx_plan* plan, long long N,
                                    int nsteps) {
  return (plan && N >= 1 && nsteps >= 0) ? pgd_tape_floats(P(plan), N, nsteps)
                                         : -1;
}

int pmx_nbody_run_pgd_taped(const pmx_plan* plan, float* state, long long N,
                            float f, float fac, const double* coef,
                            const float* pgd, int nsteps, float* tape,
                            float* work, void* stream) {
  PMX_REQUIRE(plan && state && tape && work && (pgd || nsteps == 0),
              "pmx_nbody_run_pgd_taped: null pointer");
  PMX_REQUIRE(N >= 1, "pmx_nbody_run_pgd_taped: need at least one particle");
  PMX_REQUIRE(nsteps >= 0 && (coef || nsteps == 0),
              "pmx_nbody_run_pgd_taped: bad steps");
  for (int i = 0; i < nsteps; ++i)
    PMX_REQUIRE(pgd[3 * i + 1] != 0.0f && pgd[3 * i + 2] != 0.0f,
                "pmx_nbody_run_pgd_taped: kl and ks must be nonzero");
  const Tape tp = tape_of(P(plan), tape, N, nsteps, true);
  return nbody_run(P(plan), state, N, f, fac, coef, 0, nsteps, 1, work,
                   S(stream), &tp, pgd);
}

int pmx_nbody_adjoint_pgd_taped(const pmx_plan* plan, float* tape,
                                int nsteps, long long N, float f, float fac,
                                const double* coef, const float* pgd,
                                float* gstate, double* gpgd, float* work,
                                double* gcoef, void* stream) {
  PMX_REQUIRE(plan && tape && gstate && work && (pgd || nsteps == 0),
              "pmx_nbody_adjoint_pgd_taped: null pointer");
  PMX_REQUIRE(nsteps >= 0 && (coef || nsteps == 0),
              "pmx_nbody_adjoint_pgd_taped: bad steps");
  return nbody_adjoint_pgd_taped(P(plan), tape, nsteps, N, f, fac, coef, pgd,
                                 gstate, gpgd, work, S(stream), gcoef);
}

int pmx_nbody_adjoint_taped(const pmx_plan* plan, float* tape, int nsteps,
                            long long N, float f, float fac,
                            const double* coef, float* gstate, float* work,
                            double* gcoef, void* stream) {
  PMX_REQUIRE(plan && tape && gstate && work,
              "pmx_nbody_adjoint_taped: null pointer");
  PMX_REQUIRE(nsteps >= 0 && (coef || nsteps == 0),
              "pmx_nbody_adjoint_taped: bad steps");
  return nbody_adjoint_taped(P(plan), tape, nsteps, N, f, fac, coef, gstate,
                             work, S(stream), gcoef);
}

int pmx_nbody_adjoint(const pmx_plan* plan, const float* ckpt, int every,
                      int nsteps, long long N, float f, float fac,
                      const double* coef, float* gstate, float* work,
                      double* gcoef, void* stream) {
  PMX_REQUIRE(plan && ckpt && gstate && work,
              "pmx_nbody_adjoint: null pointer");
  PMX_REQUIRE(nsteps >= 0 && (coef || nsteps == 0),
              "pmx_nbody_adjoint: bad steps");
  return nbody_adjoint(P(plan), ckpt, every, nsteps, N, f, fac, coef, gstate,
                       work, S(stream), false, nullptr, gcoef);
}

int pmx_nbody_run_last_segment(const pmx_plan* plan, float* state,
                               long long N, float f, float fac,
                               const double* coef, int nsteps, int every,
                               float* work, void* stream) {
  PMX_REQUIRE(plan && state && work && (coef || nsteps == 0),
              "pmx_nbody_run_last_segment: null pointer");
  PMX_REQUIRE(N >= 1 && nsteps >= 1 && every >= 1,
              "pmx_nbody_run_last_segment: need particles, steps and "
              "every >= 1");
  const int nseg = (nsteps + every - 1) / every;
  const int b0 = (nseg - 1) * every;
  const Tape tp = segment_tape(P(plan), N, every, work, b0);
  return nbody_run(P(plan), state, N, f, fac, coef, b0, nsteps, nseg == 1,
                   work, S(stream), &tp);
}

int pmx_nbody_adjoint_last_taped(const pmx_plan* plan, const float* ckpt,
                                 int every, int nsteps, long long N, float f,
                                 float fac, const double* coef,
                                 float* gstate, float* work, double* gcoef,
                                 void* stream) {
  PMX_REQUIRE(plan && ckpt && gstate && work,
              "pmx_nbody_adjoint_last_taped: null pointer");
  PMX_REQUIRE(nsteps >= 1 && coef,
              "pmx_nbody_adjoint_last_taped: bad steps");
  return nbody_adjoint(P(plan), ckpt, every, nsteps, N, f, fac, coef, gstate,
                       work, S(stream), true, nullptr, gcoef);
}

long long pmx_nbody_phi_floats(const pmx_plan* plan, int nsteps) {
  if (!plan || nsteps < 0) return -1;
  const Plan& p = P(plan);
  return (long long)(nsteps + 1) * p.d.rows * p.d.nz;
}

int pmx_nbody_run_phi(const pmx_plan* plan, float* state, long long N,
                      float f, float fac, const double* coef, int b0, int b1,
                      int jump, float* work, float* phi, void* stream) {
  PMX_REQUIRE(plan && state && work && phi, "pmx_nbody_run_phi: null pointer");
  PMX_REQUIRE(N >= 1, "pmx_nbody_run_phi: need at least one particle");
  PMX_REQUIRE(b0 >= 0 && b1 >= b0 && (coef || b1 == b0),
              "pmx_nbody_run_phi: bad step range");
  Tape tp;
  tp.phi = phi;
  return nbody_run(P(plan), state, N, f, fac, coef, b0, b1, jump, work,
                   S(stream), &tp);
}

int pmx_nbody_run_checkpointed(const pmx_plan* plan, float* state,
                               const float* src, long long N, float f,
                               float fac, const double* coef, int b0, int b1,
                               int jump, int every, float* ckpt, float* work,
                               float* phi, void* stream) {
  PMX_REQUIRE(plan && state && work && ckpt,
              "pmx_nbody_run_checkpointed: null pointer");
  PMX_REQUIRE(N >= 1, "pmx_nbody_run_checkpointed: need at least one "
              "particle");
  PMX_REQUIRE(b0 >= 0 && b1 >= b0 && (coef || b1 == b0) && every >= 1,
              "pmx_nbody_run_checkpointed: bad step range or every < 1");
  PMX_REQUIRE(!src || b1 > b0,
              "pmx_nbody_run_checkpointed: a separate initial state needs "
              "at least one step");
  Tape tp;
  tp.phi = phi;
  return nbody_run(P(plan), state, N, f, fac, coef, b0, b1, jump, work,
                   S(stream), phi ? &tp : nullptr, nullptr, src, ckpt,
                   every);
}

int pmx_nbody_run_last_segment_phi(const pmx_plan* plan, float* state,
                                   long long N, float f, float fac,
                                   const double* coef, int nsteps, int every,
                                   float* work, float* phi, void* stream) {
  PMX_REQUIRE(plan && state && work && phi && (coef || nsteps == 0),
              "pmx_nbody_run_last_segment_phi: null pointer");
  PMX_REQUIRE(N >= 1 && nsteps >= 1 && every >= 1,
              "pmx_nbody_run_last_segment_phi: need particles, steps and "
              "every >= 1");
  const int nseg = (nsteps + every - 1) / every;
  const int b0 = (nseg - 1) * every;
  Tape tp = segment_tape(P(plan), N, every, work, b0);
  tp.phi = phi;
  return nbody_run(P(plan), state, N, f, fac, coef, b0, nsteps, nseg == 1,
                   work, S(stream), &tp);
}

int pmx_nbody_adjoint_phi(const pmx_plan* plan, const float* ckpt,
                          const float* phi, int every, int nsteps,
                          long long N, float f, float fac, const double* coef,
                          float* gstate, float* work, int last_taped,
                          double* gcoef, void* stream) {
  PMX_REQUIRE(plan && ckpt && phi && gstate && work,
              "pmx_nbody_adjoint_phi: null pointer");
  PMX_REQUIRE(nsteps >= 0 && (coef || nsteps == 0) &&
                  (!last_taped || nsteps >= 1),
              "pmx_nbody_adjoint_phi: bad steps");
  return nbody_adjoint(P(plan), ckpt, every, nsteps, N, f, fac, coef, gstate,
                       work, S(stream), last_taped != 0, phi, gcoef);
}

}  // extern "C"
