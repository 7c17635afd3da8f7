/* pmx -- MI355X (gfx950) particle-mesh step with FlowPM semantics.
 *
 * C ABI of libpmx_hip.so.  Plain pointers and sizes only: every float* is a
 * device pointer (HBM), every `stream` is a hipStream_t passed as void*.
 * Every entry point returns 0 on success or a negative status; the message
 * of the last failure on the calling thread is pmx_last_error().  The
 * library never allocates memory visible to the caller: scratch comes from
 * caller-provided workspaces whose sizes the library reports; a plan owns
 * only its small float64-derived tables.
 *
 * The reference (DifferentiableUniverseInitiative/flowpm) has no native
 * layer or FFI: its path is Python over TensorFlow ops.  Each entry point
 * below names the reference function / TF op it replaces.
 *
 * Layouts: mesh [B][nx][ny][nz] float; particles [B][N][3] float;
 * state [3][B][N][3] float (X, P, F).  "spec" buffers hold the half
 * spectrum in the library's layout: pmx_plan_spec_elems() float2 values
 * (main block [B][nx][ny][nz/2] followed by the kz=nz/2 plane [B][nx][ny]).
 */
#ifndef PMX_H_
#define PMX_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PMX_VERSION 1000

typedef struct pmx_plan pmx_plan;

int pmx_version(void);
const char* pmx_last_error(void);

/* Live kernel timing (no reference counterpart; the reference's scripts
 * time whole runs with nvprof / the TF profiler,
 * scripts/Modi2020/fft_benchmark.py:126-168).  While started, every launch
 * is bracketed by hipEvents on its stream; stop synchronises and returns the
 * summed milliseconds and launch counts per kernel kind
 * (pmx_profile_kind_name). */
int pmx_profile_start(void);
int pmx_profile_stop(double* ms, long long* count, int nkinds);
int pmx_profile_kinds(void);
const char* pmx_profile_kind_name(int kind);

/* Plan for a batch of B meshes of nx*ny*nz cells (each in
 * {8,12,16,24,32,48,64,96,128,192,256,384,512,768,1024}, nz even).
 * Holds the twiddles, (2 pi fftfreq)^2 and 4th-order gradient tables
 * (flowpm/kernels.py:10-25, :28-78) computed in float64. */
int pmx_plan_create(pmx_plan** plan, int B, int nx, int ny, int nz);
int pmx_plan_destroy(pmx_plan* plan);
/* The same plan under the names of the FFT plan pair (TF builds its FFT3D
 * plans inside the op, utils.py:213-272; one plan here serves the FFT, the
 * Fourier multipliers and the nbody loop of one batch shape). */
int pmx_fft_plan_create(pmx_plan** plan, int B, int nx, int ny, int nz);
int pmx_fft_plan_destroy(pmx_plan* plan);
long long pmx_plan_spec_elems(const pmx_plan* plan);
/* Column-tile widths (float2 columns) of the y and x line passes; the
 * autotuner sets them.  Allowed: 8 or 16 (16 only for lines <= 512). */
int pmx_plan_set_tiles(pmx_plan* plan, int tile_y, int tile_x);

/* ---- CIC: flowpm/utils.py:32-91 (cic_paint), :163-210 (cic_readout) ---- */
/* mesh += paint(pos * pos_scale, weight)  (weight may be NULL = ones) */
int pmx_cic_paint(float* mesh, const float* pos, const float* weight, int B,
                  long long N, int nx, int ny, int nz, float pos_scale,
                  void* stream);
/* pmx_cic_paint with a workspace: particles that number one per mesh cell
 * (N = nx*ny*nz, pos_scale 1: the final paint of an nbody run) go through
 * the brick paint of the time loop, which writes the mesh without atomics
 * (mesh += paint, as pmx_cic_paint); others through pmx_cic_paint's
 * scatter.  work: pmx_cic_paint_work_floats() floats (0: not applicable;
 * work may then be NULL). */
long long pmx_cic_paint_work_floats(int B, long long N, int nx, int ny,
                                    int nz, float pos_scale);
int pmx_cic_paint_ws(float* mesh, const float* pos, const float* weight,
                     int B, long long N, int nx, int ny, int nz,
                     float pos_scale, float* work, long long work_floats,
                     void* stream);
/* out[b][p] = sum_corners mesh * w(pos * pos_scale) */
int pmx_cic_readout(float* out, const float* mesh, const float* pos, int B,
                    long long N, int nx, int ny, int nz, float pos_scale,
                    void* stream);
/* VJP of cic_paint w.r.t. positions and weights for mesh cotangent gmesh
 * (dpos or dweight may be NULL). */
int pmx_cic_paint_adj(float* dpos, float* dweight, const float* gmesh,
                      const float* pos, const float* weight, int B,
                      long long N, int nx, int ny, int nz, float pos_scale,
                      void* stream);
/* VJP of cic_readout: gmesh += paint(pos, gout) and dpos (either may be
 * NULL; with both NULL the call does nothing). */
int pmx_cic_readout_adj(float* gmesh, float* dpos, const float* gout,
                        const float* mesh, const float* pos, int B,
                        long long N, int nx, int ny, int nz, float pos_scale,
                        void* stream);

/* ---- FFT: flowpm/utils.py:213-241 (r2c3d), :244-272 (c2r3d) ---- */
/* spec = fft3d(in) * scale, half spectrum (r2c3d with scale = 1/norm) */
int pmx_fft_r2c(const pmx_plan* plan, const float* in, float* spec,
                float scale, void* stream);
/* out = real(ifft3d(spec)) * M * scale (c2r3d with scale = norm/M);
 * spec is overwritten. */
int pmx_fft_c2r(const pmx_plan* plan, float* spec, float* out, float scale,
                void* stream);
/* half spectrum <-> the reference's full (B,nx,ny,nz) complex64 layout;
 * full_to_spec keeps the Hermitian part, so that c2r of the result equals
 * the real part of the full inverse. */
int pmx_spec_to_full(const pmx_plan* plan, const float* spec, float* full,
                     void* stream);
int pmx_full_to_spec(const pmx_plan* plan, const float* full, float* spec,
                     void* stream);

/* ---- Fourier-space Poisson x gradient: kernels.py:28-102 used by
 * tfpm.py:66-95 (lpt1) and :183-219 (apply_longrange) ----
 * out_d = c2r3d(dk * laplace_kernel * gradient_kernel(d) * scale) for
 * d = x, y, z.  dk is a half spectrum (not modified); work holds 3 spec
 * buffers. */
int pmx_fft_c2r_poisson_grad(const pmx_plan* plan, const float* dk,
                             float* work, float* fx, float* fy, float* fz,
                             float scale, void* stream);
/* The same with the long-range filter of apply_longrange(split = r_split)
 * (kernels.py:81-102): the multiplier is laplace_kernel (float64, rounded
 * to float) times exp(-k^2 r_split^2) (float64 exponent, float exp), as the
 * reference multiplies the two after casting each to complex64. */
int pmx_fft_c2r_longrange_grad(const pmx_plan* plan, const float* dk,
                               float* work, float* fx, float* fy, float* fz,
                               float scale, double r_split, void* stream);
/* Adjoint of pmx_fft_c2r_longrange_grad (the gradient of
 * apply_longrange(split = r_split), tfpm.py:183-219, w.r.t. delta_k):
 * gdk = sum_d conj(lap exp(-k^2 r_split^2) i g_d) FFT3(g_d) * scale, half
 * spectrum.  work: 3 spec buffers. */
int pmx_fft_longrange_grad_adj(const pmx_plan* plan, const float* gfx,
                               const float* gfy, const float* gfz,
                               float* work, float* gdk, float scale,
                               double r_split, void* stream);

/* The meshes of apply_pgd (tfpm.py:222-270): out_d = c2r3d(dk *
 * laplace_kernel * PGD_kernel(kl, ks) * gradient_kernel(d) * scale)
 * (kernels.py:105-134, PGD = exp(-kl^2/k^2) exp(-k^4/ks^4) in float32).
 * deriv = 1 / 2 replaces PGD_kernel by its derivative w.r.t. kl / ks (the
 * parameter gradients).  dk: half spectrum (not modified); work: 3 spec. */
int pmx_fft_c2r_pgd_grad(const pmx_plan* plan, const float* dk, float* work,
                         float* fx, float* fy, float* fz, float scale,
                         float kl, float ks, int deriv, void* stream);
/* Its adjoint w.r.t. dk (deriv = 0): gdk = sum_d conj(lap PGD i g_d)
 * FFT3(g_d) * scale.  work: 3 spec. */
int pmx_fft_pgd_grad_adj(const pmx_plan* plan, const float* gfx,
                         const float* gfy, const float* gfz, float* work,
                         float* gdk, float scale, float kl, float ks,
                         void* stream);

/* One PGD correction of a state, tfpm.pgd (tfpm.py:369-422): X += alpha /
 * f * readout(c2r(delta_k lap PGD_kernel(kl, ks) i g_d), f X), delta_k =
 * r2c3d(paint(f X) / nbar).  state (3,B,N,3) in place (P, F untouched); the
 * plan is the force mesh (nc * f).  work: pmx_nbody_work_floats floats. */
int pmx_pgd(const pmx_plan* plan, float* state, long long N, float f,
            float alpha, float kl, float ks, float* work, void* stream);
/* Its reverse: gstate (3,B,N,3) holds the cotangent of the corrected state
 * on entry and of the input state on exit (gX += J^T gX; gP, gF unchanged);
 * X the input positions; gpgd (device float64[3], may be NULL) receives
 * d loss / d (alpha, kl, ks) (overwritten).  The particles must form the
 * Lagrangian grid of the mesh / f.  work: pmx_pgd_adj_work_floats. */
long long pmx_pgd_adj_work_floats(const pmx_plan* plan, long long N);
int pmx_pgd_adjoint(const pmx_plan* plan, const float* X, long long N,
                    float f, float alpha, float kl, float ks, float* gstate,
                    double* gpgd, float* work, void* stream);

/* out[0] = sum_i a[i] b[i] in float64 (device double, overwritten): the
 * contractions of the parameter gradients (no reference counterpart; TF
 * forms them inside its autodiff). */
int pmx_mesh_dot(const float* a, const float* b, long long n, double* out,
                 void* stream);

/* Fused density -> force meshes: r2c3d(rho * rho_scale) followed by the
 * Poisson x gradient inverse above, with the forward x transform, the
 * multiply and the inverse x transform in one pass (tfpm.py:344-354). */
int pmx_force_meshes(const pmx_plan* plan, const float* rho, float* work,
                     float* fx, float* fy, float* fz, float rho_scale,
                     void* stream);
/* The force mesh of the time loop, interleaved (tfpm.py:344-354 with the
 * three meshes in one float4 per cell): phi = c2r3d(r2c3d(rho * rho_scale)
 * * laplace_kernel) into the real mesh phi (which may be rho itself), then
 * f4[cell] = (D_x phi, D_y phi, D_z phi, 0), D the periodic fourth-order
 * central difference 2/3 (x[n+1] - x[n-1]) - 1/12 (x[n+2] - x[n-2]) whose
 * Fourier multiplier is exactly gradient_kernel (kernels.py:51-78).
 * work: 1 spec buffer; nz % 4 == 0, sides >= 4 (nz >= 8), B * cells
 * < 2^33; f4 16-byte aligned. */
int pmx_force_mesh_f4(const pmx_plan* plan, const float* rho, float* work,
                      float* phi, float* f4, float rho_scale, void* stream);
/* Its adjoint: grho = sum_d T_d^T g_d, T_d^T = c2r(conj(lap i g_d) r2c(.)),
 * the conjugate transform (tfpm.py:183-219 reversed). */
int pmx_force_meshes_adj(const pmx_plan* plan, const float* gfx,
                         const float* gfy, const float* gfz, float* work,
                         float* grho, float rho_scale, void* stream);
/* Adjoint of pmx_fft_c2r_poisson_grad: gdk (half spectrum, adjoint in the
 * real-linear sense) from the three mesh cotangents.  work: 3 spec. */
int pmx_fft_poisson_grad_adj(const pmx_plan* plan, const float* gfx,
                             const float* gfy, const float* gfz, float* work,
                             float* gdk, float scale, void* stream);

/* ---- second-order LPT source: tfpm.py:98-145 ----
 * src_k = r2c3d(3/7 [sum phi_ii phi_jj - sum phi_ij^2]) from dk.
 * work: 6 spec buffers followed by 7 real meshes. */
int pmx_lpt2_source(const pmx_plan* plan, const float* dk, float* work,
                    float* src_k, void* stream);
long long pmx_lpt2_work_floats(const pmx_plan* plan);

/* ---- lpt_init assembly: tfpm.py:148-180 ----
 * dx1, dx2: the unscaled displacement meshes (the lpt1 transforms of the
 * linear field and of the 2LPT source) at the particle grid (B, nx, ny, nz),
 * x / y / z meshes one after the other (dx2 may be NULL); the growth
 * factors are applied here, with DX1 = c[0] dx1 and DX2 = c[3] dx2:
 *   state[0] = Q + DX1 + DX2,
 *   state[1] = c[1] DX1 + c[4] DX2   (= c[1] c[0] dx1 + c[4] c[3] dx2),
 *   state[2] = c[2] DX1 + c[5] DX2,
 * Q in np.indices order; coef6 = (D1, a^2 f1 E, a^2 E gf / D1, D2,
 * a^2 f2 E, a^2 E gf2 / D2) as tfpm.py:168-177 forms them.  A caller must
 * not pre-multiply dx1 / dx2 by D1 / D2. */
int pmx_lpt_assemble(float* state, const float* dx1, const float* dx2,
                     const double* coef6, int B, int nx, int ny, int nz,
                     void* stream);

/* Adjoint of pmx_lpt_assemble: g1 (3 meshes, x/y/z) = c[0] (gX + c[1] gP +
 * c[2] gF) per component, and likewise g2 with c[3..5] (g2 may be NULL). */
int pmx_lpt_assemble_adj(const float* gstate, const double* coef6, float* g1,
                         float* g2, int B, int nx, int ny, int nz,
                         void* stream);
/* Adjoint of the pointwise source combination inside pmx_lpt2_source:
 * work is the buffer a pmx_lpt2_source call filled (its phi_ij meshes);
 * gsrc is the cotangent of the real-space source (before its r2c); gphi
 * receives 6 meshes (xx, yy, zz, xy, xz, yz). */
int pmx_lpt2_source_adj(const pmx_plan* plan, const float* work,
                        const float* gsrc, float* gphi, void* stream);

/* Generic Fourier filter used by the adjoints of lpt_init and force, and
 * by compensate_cic (utils.py:11-29):
 * out (+)= c2r3d( sum_q scales[q] [laplace_kernel] [cic_compensation]
 *          prod_d gradient_kernel(d)^e[3q+d] r2c3d(in[q]) * in_scale ),
 * nin <= 3 real meshes, exponents 0..2 (kernels.py:28-78, :137-152).
 * `lap` holds kernel flags: bit 0 laplace_kernel, bit 1 cic_compensation.
 * work: 3 spec buffers. */
int pmx_fourier_apply(const pmx_plan* plan, int nin, const float* const* in,
                      const int* e, const float* scales, int lap,
                      float in_scale, float* work, float* out, int accumulate,
                      void* stream);

/* Seeded white noise in HBM (replaces TF's RNG in utils.py:337-355):
 * out[i] = float32(scale * z_i), z_i the i-th standard normal of the
 * Philox4x64-10 stream with key (key_lo, key_hi) -- the generator of
 * numpy.random.Philox: normals 4q..4q+3 come from block q, i.e. the words
 * philox4x64_10(counter + q + 1, key), by two Box-Muller pairs
 * (u1 = ((w >> 11) + 1) 2^-53, u2 = (w' >> 11) 2^-53, float64). */
int pmx_white_noise(float* out, long long n, unsigned long long key_lo,
                    unsigned long long key_hi, unsigned long long counter,
                    double scale, void* stream);

/* Power-spectrum bins (flowpm/pk.py:7-103): psum[b][i] = sum over the
 * modes of batch member b with numpy.digitize(|k|, edges) == i of
 * W |F|^2, where F is the unnormalised transform held in spec (the half
 * spectrum of pmx_fft_r2c with scale 1) and W the reference's weights
 * (1 for the first and last full-spectrum kz index, 2 otherwise; pk.py:16-19),
 * accounted over the full spectrum.  kx, ky, kz: device float64 tables of
 * the full per-axis wavenumbers (nx, ny, nz values, pk.py:24-27); edges:
 * nedges increasing float64 bin edges (nedges + 1 <= 8192); psum: device
 * float64 [B][nedges + 1], overwritten. */
int pmx_power_spectrum_bins(const pmx_plan* plan, const float* spec,
                            const double* kx, const double* ky,
                            const double* kz, const double* edges, int nedges,
                            double* psum, void* stream);
/* The weighted mode count of the same bins (pk.py:32-35: bincount of the
 * full-spectrum digitize with the weights W), one array nsum[nedges + 1]
 * (float64, exact), without touching any field. */
int pmx_power_spectrum_counts(const pmx_plan* plan, const double* kx,
                              const double* ky, const double* kz,
                              const double* edges, int nedges, double* nsum,
                              void* stream);

/* Adjoint of the binning (the gradient of power_spectrum, pk.py:40-103,
 * w.r.t. the field): spec (the unnormalised half spectrum that
 * pmx_power_spectrum_bins read) is multiplied in place by gbin[b][bin] times
 * the column's weight (W(j) + W(nz - j)) / 2, or W(j) on the kz = 0 and
 * nz/2 planes; pmx_fft_c2r of the result with scale 2 (twice the
 * unnormalised inverse) is the field cotangent when gbin[b][i] =
 * d loss / d psum[b][i]. */
int pmx_power_spectrum_bins_adj(const pmx_plan* plan, float* spec,
                                const double* kx, const double* ky,
                                const double* kz, const double* edges,
                                int nedges, const double* gbin,
                                void* stream);

/* The linear-field amplitude of a box with equal sides (tfpm.py:56-62 with
 * |k| = 2 pi sqrt(mx^2 + my^2 + mz^2) / L): amp[mode] = table[mx^2 + my^2 +
 * mz^2] over the half-spectrum layout below, m_d the signed frequency index
 * (numpy fftfreq order).  table[s] = sqrt(P(2 pi sqrt(s) / L) / V) is
 * evaluated once per distinct |k| on the host; ntable must exceed the
 * largest (nx/2)^2 + (ny/2)^2 + (nz/2)^2.  No counterpart in the reference
 * (which evaluates pk on every mode). */
int pmx_amp_from_table(const pmx_plan* plan, const float* table,
                       long long ntable, float* amp, void* stream);
/* spec *= amp, amp a real half-spectrum array of one mesh (nx*ny*nz/2 main
 * values then nx*ny Nyquist-plane values), applied to every batch member:
 * the sqrt(P(k)/V) weighting of tfpm.py:56-62. */
int pmx_spec_mul_real(const pmx_plan* plan, float* spec, const float* amp,
                      void* stream);

/* ---- fused readout + kick + drift: tfpm.py:273-314, 352-366 ----
 * F = readout(f{x,y,z}, X*f) * fac / f; then, by flags,
 *   bit0: P += k2 F      bit1: P += k1 F ; X += d P
 *   bit2: skip the readout and use F as stored (kick/drift only). */
int pmx_readout_kick_drift(float* X, float* P, float* F, const float* fx,
                           const float* fy, const float* fz, int B,
                           long long N, int nx, int ny, int nz, float f,
                           float fac, float k2, float k1, float d, int flags,
                           void* stream);

/* ---- nbody time loop: tfpm.py:425-503 (force :317-366) ----
 * Runs steps [step_begin, step_end) of the FastPM KDK scheme on state
 * (3,B,N,3) in place.  coef[3*i + {0,1,2}] = (k1, d, k2) of step i: opening
 * kick, drift and closing kick factors (float64, from the growth tables).
 * jump_start != 0 first evaluates force(state) as tfpm.py:470 does.
 * The plan is the force mesh (nc * pm_nc_factor); f = pm_nc_factor.
 * work: pmx_nbody_work_floats() floats. */
long long pmx_nbody_work_floats(const pmx_plan* plan);
int pmx_nbody_run(const pmx_plan* plan, float* state, long long N, float f,
                  float fac, const double* coef, int step_begin, int step_end,
                  int jump_start, float* work, void* stream);

/* pmx_nbody_run without a workspace argument.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP stream/graph capture, tight serving loops).  The workspace is a
 * static device array of 16 Mi floats (64 MiB) that every device the
 * library runs on holds until the process ends; plans whose
 * pmx_nbody_work_floats() exceeds it are rejected (kBadArg, nothing
 * enqueued).  The call allocates nothing (no hipMalloc / hipFree).  All
 * calls on one device share the array: calls that overlap on one device
 * (other streams, other threads) race and must be serialised by the caller. */
int pmx_nbody_run_noalloc(const pmx_plan* plan, float* state, long long N,
                          float f, float fac, const double* coef,
                          int step_begin, int step_end, int jump_start,
                          void* stream);

/* pmx_nbody_run with the PGD correction (tfpm.py:369-422, :487-488): after
 * the drift of step i, X += alpha_i / f * readout(c2r(delta_k lap
 * PGD_kernel(kl_i, ks_i) grad_d), f X) with delta_k the density of the
 * drifted positions (kernels.py:105-134).  pgd[3*i + {0,1,2}] =
 * (alpha, kl, ks) of step i (float32, as pgdparams). */
int pmx_nbody_run_pgd(const pmx_plan* plan, float* state, long long N,
                      float f, float fac, const double* coef,
                      const float* pgd, int step_begin, int step_end,
                      int jump_start, float* work, void* stream);

/* Reverse sweep (hand-written adjoint with recompute).  ckpt holds the
 * states at the start of every segment of `every` steps (segment s starts
 * at step s*every; segment 0 at step 0 *before* the jump-start force);
 * gstate (3,B,N,3) holds the cotangent of the final state on entry and the
 * cotangent of the initial state on exit.  Each segment is recomputed from
 * its checkpoint keeping X and the force meshes of every step, then swept
 * in reverse.  work: pmx_nbody_adj_work_floats(plan, N, every) floats.
 *
 * gcoef (every reverse sweep below takes it; NULL: not wanted): a device
 * array of 2 nsteps + 1 doubles to which the sweep ADDS the scalar
 * cotangents of the time loop's coefficients, from which a caller forms the
 * gradients w.r.t. the coefficients of step_coefficients and fac
 * (tfpm.py:273-314, :352-356; the cosmology-parameter gradients of
 * scripts/check_grads.py:71-98): gcoef[i] = sum_p gP . F_i (F_i the force
 * of step i, gP the momentum cotangent at the two kicks that use it: the
 * cotangent of k2_i and of k1_{i+1}), gcoef[nsteps + i] = sum_p gX . P_i
 * (P_i the momentum of step i's drift: the cotangent of d_i), and
 * gcoef[2 nsteps] = sum_p gP . F_jump (the cotangent of k1_0). */
long long pmx_nbody_adj_work_floats(const pmx_plan* plan, long long N,
                                    int every);
int pmx_nbody_adjoint(const pmx_plan* plan, const float* ckpt, int every,
                      int nsteps, long long N, float f, float fac,
                      const double* coef, float* gstate, float* work,
                      double* gcoef, void* stream);

/* The checkpointed reverse sweep without recomputing the last segment:
 * pmx_nbody_run_last_segment runs the last segment of a run of nsteps
 * steps in segments of `every` (steps [(nseg - 1) every, nsteps), with the
 * jump-start force when it is the only segment) on state, recording its
 * positions and force meshes in work (pmx_nbody_adj_work_floats(plan, N,
 * every) floats) exactly where pmx_nbody_adjoint's recomputation would;
 * pmx_nbody_adjoint_last_taped is pmx_nbody_adjoint on that same work,
 * which then recomputes segments 0 .. nseg - 2 only.  (No reference
 * counterpart: TF keeps every forward intermediate, tfpm.py:425-503.) */
int pmx_nbody_run_last_segment(const pmx_plan* plan, float* state,
                               long long N, float f, float fac,
                               const double* coef, int nsteps, int every,
                               float* work, void* stream);
int pmx_nbody_adjoint_last_taped(const pmx_plan* plan, const float* ckpt,
                                 int every, int nsteps, long long N, float f,
                                 float fac, const double* coef,
                                 float* gstate, float* work, double* gcoef,
                                 void* stream);

/* Checkpointing that also keeps the potential of every force evaluation
 * (no reference counterpart: TF keeps every forward intermediate,
 * tfpm.py:425-503).  phi holds pmx_nbody_phi_floats(plan, nsteps) floats:
 * slot 0 (B*nx*ny*nz floats) the jump-start force's potential, slot 1 + i
 * that of step i.  pmx_nbody_run_phi and pmx_nbody_run_last_segment_phi
 * are pmx_nbody_run and pmx_nbody_run_last_segment that also store the
 * potentials of the forces they evaluate in their slots (same results);
 * pmx_nbody_adjoint_phi is pmx_nbody_adjoint (last_taped = 0) or
 * pmx_nbody_adjoint_last_taped (last_taped = 1) whose segment
 * recomputation takes every force mesh from its kept potential (the
 * fourth-order gradient) instead of painting and transforming: the same
 * gradients for about half the recomputation time.  No PGD. */
long long pmx_nbody_phi_floats(const pmx_plan* plan, int nsteps);
/* pmx_nbody_run (phi NULL) or pmx_nbody_run_phi (phi set) that also writes
 * the checkpoints of the steps it crosses: after step i with
 * (i + 1 - step_begin) % every == 0 and i + 1 < step_end, the state a run
 * ending there would leave (X, P after the closing kick, F; 3*B*N*3
 * floats) goes to slot (i + 1 - step_begin) / every - 1 of ckpt.  The
 * same states as separate pmx_nbody_run calls per segment with copies
 * between them (bit for bit), without the copies and the split kicks: the
 * readout of each such step stores them.  src (NULL: the state in place):
 * the initial state is read from src (not modified; at least one step) and
 * state only receives the result, so the caller's copy of the initial
 * state can serve as both.  (No reference counterpart: TF keeps every
 * forward intermediate, tfpm.py:425-503.)  Not with PGD. */
int pmx_nbody_run_checkpointed(const pmx_plan* plan, float* state,
                               const float* src, long long N, float f,
                               float fac, const double* coef, int step_begin,
                               int step_end, int jump_start, int every,
                               float* ckpt, float* work, float* phi,
                               void* stream);
int pmx_nbody_run_phi(const pmx_plan* plan, float* state, long long N,
                      float f, float fac, const double* coef, int step_begin,
                      int step_end, int jump_start, float* work, float* phi,
                      void* stream);
int pmx_nbody_run_last_segment_phi(const pmx_plan* plan, float* state,
                                   long long N, float f, float fac,
                                   const double* coef, int nsteps, int every,
                                   float* work, float* phi, void* stream);
int pmx_nbody_adjoint_phi(const pmx_plan* plan, const float* ckpt,
                          const float* phi, int every, int nsteps,
                          long long N, float f, float fac, const double* coef,
                          float* gstate, float* work, int last_taped,
                          double* gcoef, void* stream);

/* Several independent groups of simulations of uneven sizes (one plan per
 * size class; the reference batches only equal sizes, tfpm.py:15-63):
 * group g advances its state over steps [step_begin[g], step_end[g]) of
 * its own coefficient table, exactly as pmx_nbody_run would, all enqueued
 * on `stream` in group order (for concurrent groups, capture one stream
 * per group into a graph, as flowpm_amd.batched does).  Arrays have
 * ngroups entries; work[g] holds pmx_nbody_work_floats(plans[g]) floats. */
int pmx_nbody_batched_step(int ngroups, const pmx_plan* const* plans,
                           float* const* states, const long long* N,
                           const float* f, const float* fac,
                           const double* const* coefs, const int* step_begin,
                           const int* step_end, const int* jump_start,
                           float* const* work, void* stream);

/* Taped forward + reverse sweep: the same integration as pmx_nbody_run over
 * steps [0, nsteps) with the jump-start force, recording the positions and
 * the force meshes of every force evaluation in `tape`
 * (pmx_nbody_tape_floats(plan, N, nsteps) floats), so that
 * pmx_nbody_adjoint_taped needs no recomputation.  This replaces TF's
 * stored forward intermediates of tfpm.py:425-503 under autodiff; the
 * adjoint work is pmx_nbody_adj_work_floats(plan, N, 0) floats. */
long long pmx_nbody_tape_floats(const pmx_plan* plan, long long N,
                                int nsteps);
int pmx_nbody_run_taped(const pmx_plan* plan, float* state, long long N,
                        float f, float fac, const double* coef, int nsteps,
                        float* tape, float* work, void* stream);
int pmx_nbody_adjoint_taped(const pmx_plan* plan, float* tape, int nsteps,
                            long long N, float f, float fac,
                            const double* coef, float* gstate, float* work,
                            double* gcoef, void* stream);

/* Compact tape (no reference counterpart: TF keeps every forward
 * intermediate, tfpm.py:425-503): the same run as pmx_nbody_run_taped
 * keeping X and the potential of every force evaluation instead of its
 * float4 force mesh (pmx_nbody_compact_tape_floats(plan, N, nsteps) =
 * (nsteps + 1)(3 B N + B nx ny nz) floats, about half of
 * pmx_nbody_tape_floats); the forward uses the packed meshes of an untaped
 * run, and pmx_nbody_adjoint_compact_taped forms each step's force mesh from
 * its potential with the gradient stencil before the step's VJP: the same
 * gradients as pmx_nbody_adjoint_taped, bit for bit.  work and gcoef as for
 * pmx_nbody_adjoint_taped. */
long long pmx_nbody_compact_tape_floats(const pmx_plan* plan, long long N,
                                        int nsteps);
int pmx_nbody_run_compact_taped(const pmx_plan* plan, float* state,
                                long long N, float f, float fac,
                                const double* coef, int nsteps, float* tape,
                                float* work, void* stream);
int pmx_nbody_adjoint_compact_taped(const pmx_plan* plan, float* tape,
                                    int nsteps, long long N, float f,
                                    float fac, const double* coef,
                                    float* gstate, float* work,
                                    double* gcoef, void* stream);

/* Taped run and reverse sweep with the PGD correction (tfpm.py:369-422,
 * :487-488; pgd as in pmx_nbody_run_pgd).  The tape
 * (pmx_nbody_pgd_tape_floats floats) also keeps, per step, the positions
 * before the correction and the PGD meshes at them.  The reverse sweep
 * returns the state cotangent in gstate (as pmx_nbody_adjoint_taped) and,
 * if gpgd is not NULL, d loss / d (alpha, kl, ks) of every step in
 * gpgd[3*i + {0,1,2}] (device float64, overwritten), the gradients
 * scripts/fit_pgd.py:77-117 fits the parameters with; gcoef as for
 * pmx_nbody_adjoint (the drift momentum of step i is the drift to the
 * positions before its correction).  work:
 * pmx_nbody_adj_work_floats(plan, N, 0) floats. */
long long pmx_nbody_pgd_tape_floats(const pmx_plan* plan, long long N,
                                    int nsteps);
int pmx_nbody_run_pgd_taped(const pmx_plan* plan, float* state, long long N,
                            float f, float fac, const double* coef,
                            const float* pgd, int nsteps, float* tape,
                            float* work, void* stream);
int pmx_nbody_adjoint_pgd_taped(const pmx_plan* plan, float* tape,
                                int nsteps, long long N, float f, float fac,
                                const double* coef, const float* pgd,
                                float* gstate, double* gpgd, float* work,
                                double* gcoef, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PMX_H_ */
