// Host-side check of the C ABI (include/pmx.h) under AddressSanitizer and
// UndefinedBehaviorSanitizer, without a GPU: the library's sources are
// compiled with the sanitizers on their host code (one FFT line length, to
// keep the build short), and
// every entry point is called with the arguments its validation must
// refuse (null pointers and plans, sizes below 1, odd or unsupported mesh
// sides, mismatched workspaces).  Each call must return its negative
// status and leave a message in pmx_last_error(), which is per thread.
// Creating a plan for a valid mesh needs device memory: without a device
// that call must fail cleanly (a HIP status, no plan).  The workspace-size
// functions are evaluated at extreme sizes, where UBSan would flag signed
// overflow.  Prints "ok" and exits 0 when every check holds.
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>

#include "../../include/pmx.h"

static int g_bad = 0;

// rc must be the status `want` (any negative status when want is 0) with a
// message; -1 is the argument check, which runs before any HIP call
static void expect_fail(int rc, int want, const char* what) {
  const char* msg = pmx_last_error();
  const bool code_ok = want ? rc == want : rc < 0;
  if (!code_ok || msg == nullptr || msg[0] == '\0') {
    printf("FAIL %s: rc %d, message '%s'\n", what, rc, msg ? msg : "(null)");
    ++g_bad;
  }
}

#define FAILS(call) expect_fail((call), -1, #call)

static void expect_ok(int rc, const char* what) {
  if (rc != 0) {
    printf("FAIL %s: rc %d, message '%s'\n", what, rc, pmx_last_error());
    ++g_bad;
  }
}

int main() {
  float buf[64] = {};
  double dbuf[16] = {};
  long long lbuf[64] = {};
  const pmx_plan* np = nullptr;

  if (pmx_version() <= 0) {
    printf("FAIL pmx_version\n");
    ++g_bad;
  }
  // plans: null output, sizes, parity, supported sides; a valid plan needs
  // a device
  pmx_plan* plan = nullptr;
  FAILS(pmx_plan_create(nullptr, 1, 16, 16, 16));
  FAILS(pmx_plan_create(&plan, 0, 16, 16, 16));
  FAILS(pmx_plan_create(&plan, 1, 16, 16, 15));
  FAILS(pmx_plan_create(&plan, 1, 7, 16, 16));
  FAILS(pmx_plan_create(&plan, 1, 16, 2048, 16));
  FAILS(pmx_plan_create(&plan, -3, 16, 16, 16));
  FAILS(pmx_fft_plan_create(&plan, 1, 16, 16, 14));
  {
    const int rc = pmx_plan_create(&plan, 1, 16, 16, 16);
    if (rc == 0) {
      // a device is present after all: the plan must work and be freed
      if (pmx_plan_spec_elems(plan) != 16LL * 16 * 9) {
        printf("FAIL spec elems %lld\n", pmx_plan_spec_elems(plan));
        ++g_bad;
      }
      pmx_plan_destroy(plan);
    } else {
      expect_fail(rc, 0, "pmx_plan_create without a device");
      if (plan != nullptr) {
        printf("FAIL plan set on failure\n");
        ++g_bad;
      }
    }
  }
  if (pmx_plan_destroy(nullptr) != 0) {
    printf("FAIL destroy(null)\n");
    ++g_bad;
  }
  FAILS(pmx_plan_set_tiles(nullptr, 16, 16));
  if (pmx_plan_spec_elems(nullptr) >= 0 || pmx_nbody_work_floats(nullptr) >= 0 ||
      pmx_lpt2_work_floats(nullptr) >= 0 ||
      pmx_nbody_adj_work_floats(nullptr, 8, 1) >= 0 ||
      pmx_nbody_tape_floats(nullptr, 8, 2) >= 0 ||
      pmx_nbody_pgd_tape_floats(nullptr, 8, 2) >= 0 ||
      pmx_nbody_phi_floats(nullptr, 2) >= 0 ||
      pmx_nbody_compact_tape_floats(nullptr, 8, 2) >= 0 ||
      pmx_pgd_adj_work_floats(nullptr, 8) >= 0) {
    printf("FAIL a size function accepted a null plan\n");
    ++g_bad;
  }

  // CIC on raw pointers
  FAILS(pmx_cic_paint(nullptr, buf, nullptr, 1, 4, 4, 4, 4, 1.0f, nullptr));
  FAILS(pmx_cic_paint(buf, nullptr, nullptr, 1, 4, 4, 4, 4, 1.0f, nullptr));
  FAILS(pmx_cic_paint(buf, buf, nullptr, 0, 4, 4, 4, 4, 1.0f, nullptr));
  FAILS(pmx_cic_paint(buf, buf, nullptr, 1, -1, 4, 4, 4, 1.0f, nullptr));
  FAILS(pmx_cic_paint(buf, buf, nullptr, 1, 4, 0, 4, 4, 1.0f, nullptr));
  FAILS(pmx_cic_paint_ws(nullptr, buf, nullptr, 1, 64, 4, 4, 4, 1.0f, buf, 64,
                         nullptr));
  FAILS(pmx_cic_paint_ws(buf, buf, nullptr, 1, 64, 4, 4, -4, 1.0f, buf, 64,
                         nullptr));
  FAILS(pmx_cic_readout(nullptr, buf, buf, 1, 4, 4, 4, 4, 1.0f, nullptr));
  FAILS(pmx_cic_readout(buf, nullptr, buf, 1, 4, 4, 4, 4, 1.0f, nullptr));
  FAILS(pmx_cic_readout(buf, buf, buf, 1, 4, 4, 0, 4, 1.0f, nullptr));
  // no cotangent asked for: nothing to do, no HIP call
  expect_ok(pmx_cic_paint_adj(nullptr, nullptr, buf, buf, nullptr, 1, 4, 4, 4,
                              4, 1.0f, nullptr),
            "pmx_cic_paint_adj without outputs");
  expect_ok(pmx_cic_readout_adj(nullptr, nullptr, buf, buf, buf, 1, 4, 4, 4, 4,
                                1.0f, nullptr),
            "pmx_cic_readout_adj without outputs");
  FAILS(pmx_cic_paint_adj(buf, nullptr, nullptr, buf, nullptr, 1, 4, 4, 4, 4,
                          1.0f, nullptr));
  FAILS(pmx_cic_readout_adj(nullptr, buf, buf, nullptr, buf, 1, 4, 4, 4, 4,
                            1.0f, nullptr));
  FAILS(pmx_cic_readout_adj(buf, buf, buf, buf, nullptr, 1, 4, 4, 4, 4, 1.0f,
                            nullptr));
  FAILS(pmx_cic_readout_adj(buf, buf, nullptr, buf, buf, 1, 4, 4, 4, 4, 1.0f,
                            nullptr));
  FAILS(pmx_cic_paint_adj(buf, buf, buf, nullptr, buf, 1, 4, 4, 4, 4, 1.0f,
                          nullptr));
  FAILS(pmx_readout_kick_drift(nullptr, buf, buf, buf, buf, buf, 1, 4, 4, 4, 4,
                               1.0f, 1.0f, 0.f, 0.f, 0.f, 0, nullptr));
  FAILS(pmx_readout_kick_drift(buf, buf, buf, buf, buf, buf, 0, 4, 4, 4, 4,
                               1.0f, 1.0f, 0.f, 0.f, 0.f, 0, nullptr));
  // workspace of the brick paint: 0 where it does not apply, never negative
  // or overflowing
  {
    const long long w = pmx_cic_paint_work_floats(1, 64LL * 64 * 64, 64, 64, 64, 1.0f);
    if (w <= 0) {
      printf("FAIL cic work floats %lld\n", w);
      ++g_bad;
    }
    if (pmx_cic_paint_work_floats(1, 100, 64, 64, 64, 1.0f) != 0 ||
        pmx_cic_paint_work_floats(1, 64LL * 64 * 64, 64, 64, 64, 2.0f) != 0 ||
        pmx_cic_paint_work_floats(0, 64, 4, 4, 4, 1.0f) != 0 ||
        pmx_cic_paint_work_floats(1, 64, 4, 4, 0, 1.0f) != 0) {
      printf("FAIL cic work floats of a non-grid paint\n");
      ++g_bad;
    }
    const long long big = pmx_cic_paint_work_floats(
        64, 1024LL * 1024 * 1024, 1024, 1024, 1024, 1.0f);
    if (big < 0) {
      printf("FAIL cic work floats at 64 x 1024^3: %lld\n", big);
      ++g_bad;
    }
  }

  // every plan-based entry point refuses a null plan
  float* fp = buf;
  const float* cfp = buf;
  FAILS(pmx_fft_r2c(np, cfp, fp, 1.0f, nullptr));
  FAILS(pmx_fft_c2r(np, fp, fp, 1.0f, nullptr));
  FAILS(pmx_spec_to_full(np, cfp, fp, nullptr));
  FAILS(pmx_full_to_spec(np, cfp, fp, nullptr));
  FAILS(pmx_fft_c2r_poisson_grad(np, cfp, fp, fp, fp, fp, 1.0f, nullptr));
  FAILS(pmx_fft_c2r_longrange_grad(np, cfp, fp, fp, fp, fp, 1.0f, 0.5, nullptr));
  FAILS(pmx_fft_longrange_grad_adj(np, cfp, cfp, cfp, fp, fp, 1.0f, 0.5, nullptr));
  FAILS(pmx_fft_c2r_pgd_grad(np, cfp, fp, fp, fp, fp, 1.0f, 0.3f, 10.f, 0, nullptr));
  FAILS(pmx_fft_pgd_grad_adj(np, cfp, cfp, cfp, fp, fp, 1.0f, 0.3f, 10.f, nullptr));
  FAILS(pmx_pgd(np, fp, 8, 1.0f, 0.1f, 0.3f, 10.f, fp, nullptr));
  FAILS(pmx_pgd_adjoint(np, cfp, 8, 1.0f, 0.1f, 0.3f, 10.f, fp, dbuf, fp, nullptr));
  FAILS(pmx_force_meshes(np, cfp, fp, fp, fp, fp, 1.0f, nullptr));
  FAILS(pmx_force_mesh_f4(np, cfp, fp, fp, fp, 1.0f, nullptr));
  FAILS(pmx_force_meshes_adj(np, cfp, cfp, cfp, fp, fp, 1.0f, nullptr));
  FAILS(pmx_fft_poisson_grad_adj(np, cfp, cfp, cfp, fp, fp, 1.0f, nullptr));
  FAILS(pmx_lpt2_source(np, cfp, fp, fp, nullptr));
  FAILS(pmx_lpt2_source_adj(np, cfp, cfp, fp, nullptr));
  {
    const float* ins[1] = {cfp};
    const int e[3] = {0, 0, 0};
    const float sc[1] = {1.0f};
    FAILS(pmx_fourier_apply(np, 1, ins, e, sc, 1, 1.0f, fp, fp, 0, nullptr));
  }
  FAILS(pmx_power_spectrum_bins(np, cfp, dbuf, dbuf, dbuf, dbuf, 4, dbuf, nullptr));
  FAILS(pmx_power_spectrum_counts(np, dbuf, dbuf, dbuf, dbuf, 4, dbuf, nullptr));
  FAILS(pmx_power_spectrum_bins_adj(np, fp, dbuf, dbuf, dbuf, dbuf, 4, dbuf, nullptr));
  FAILS(pmx_amp_from_table(np, cfp, 4, fp, nullptr));
  FAILS(pmx_spec_mul_real(np, fp, cfp, nullptr));
  FAILS(pmx_nbody_run(np, fp, 8, 1.0f, 0.4f, dbuf, 0, 1, 1, fp, nullptr));
  FAILS(pmx_nbody_run_noalloc(np, fp, 8, 1.0f, 0.4f, dbuf, 0, 1, 1, nullptr));
  FAILS(pmx_nbody_run_pgd(np, fp, 8, 1.0f, 0.4f, dbuf, cfp, 0, 1, 1, fp, nullptr));
  FAILS(pmx_nbody_adjoint(np, cfp, 1, 1, 8, 1.0f, 0.4f, dbuf, fp, fp, dbuf,
                          nullptr));
  FAILS(pmx_nbody_run_last_segment(np, fp, 8, 1.0f, 0.4f, dbuf, 2, 1, fp, nullptr));
  FAILS(pmx_nbody_adjoint_last_taped(np, cfp, 1, 2, 8, 1.0f, 0.4f, dbuf, fp, fp,
                                     dbuf, nullptr));
  FAILS(pmx_nbody_run_phi(np, fp, 8, 1.0f, 0.4f, dbuf, 0, 1, 1, fp, fp,
                          nullptr));
  FAILS(pmx_nbody_run_last_segment_phi(np, fp, 8, 1.0f, 0.4f, dbuf, 2, 1, fp,
                                       fp, nullptr));
  FAILS(pmx_nbody_run_checkpointed(np, fp, nullptr, 8, 1.0f, 0.4f, dbuf, 0, 2,
                                   1, 1, fp, fp, nullptr, nullptr));
  FAILS(pmx_nbody_adjoint_phi(np, cfp, cfp, 1, 2, 8, 1.0f, 0.4f, dbuf, fp, fp,
                              1, dbuf, nullptr));
  FAILS(pmx_nbody_run_taped(np, fp, 8, 1.0f, 0.4f, dbuf, 1, fp, fp, nullptr));
  FAILS(pmx_nbody_run_compact_taped(np, fp, 8, 1.0f, 0.4f, dbuf, 1, fp, fp,
                                    nullptr));
  FAILS(pmx_nbody_adjoint_compact_taped(np, fp, 1, 8, 1.0f, 0.4f, dbuf, fp, fp,
                                        dbuf, nullptr));
  FAILS(pmx_nbody_adjoint_taped(np, fp, 1, 8, 1.0f, 0.4f, dbuf, fp, fp, dbuf,
                                nullptr));
  FAILS(pmx_nbody_run_pgd_taped(np, fp, 8, 1.0f, 0.4f, dbuf, cfp, 1, fp, fp,
                                nullptr));
  FAILS(pmx_nbody_adjoint_pgd_taped(np, fp, 1, 8, 1.0f, 0.4f, dbuf, cfp, fp,
                                    dbuf, fp, dbuf, nullptr));
  {
    const pmx_plan* plans[1] = {nullptr};
    float* states[1] = {fp};
    float* works[1] = {fp};
    const double* coefs[1] = {dbuf};
    const float ff[1] = {1.0f}, fac[1] = {0.4f};
    const int b0[1] = {0}, b1[1] = {1}, js[1] = {1};
    FAILS(pmx_nbody_batched_step(1, plans, states, lbuf, ff, fac, coefs, b0, b1,
                                 js, works, nullptr));
    FAILS(pmx_nbody_batched_step(-1, plans, states, lbuf, ff, fac, coefs, b0,
                                 b1, js, works, nullptr));
    FAILS(pmx_nbody_batched_step(1, nullptr, states, lbuf, ff, fac, coefs, b0,
                                 b1, js, works, nullptr));
  }
  // raw-pointer helpers
  FAILS(pmx_mesh_dot(nullptr, cfp, 4, dbuf, nullptr));
  FAILS(pmx_mesh_dot(cfp, cfp, 4, nullptr, nullptr));
  FAILS(pmx_white_noise(nullptr, 16, 1, 2, 0, 1.0, nullptr));
  FAILS(pmx_lpt_assemble(nullptr, cfp, nullptr, dbuf, 1, 4, 4, 4, nullptr));
  FAILS(pmx_lpt_assemble(fp, cfp, nullptr, dbuf, 0, 4, 4, 4, nullptr));
  FAILS(pmx_lpt_assemble_adj(nullptr, dbuf, fp, fp, 1, 4, 4, 4, nullptr));
  if (pmx_profile_kinds() <= 0 || pmx_profile_kind_name(0) == nullptr ||
      pmx_profile_kind_name(-1) == nullptr ||
      pmx_profile_kind_name(1 << 20) == nullptr) {
    printf("FAIL profile kind names\n");
    ++g_bad;
  }

  // the error message is per thread
  pmx_plan_set_tiles(nullptr, 16, 16);
  const std::string mine = pmx_last_error();
  std::string theirs;
  std::thread t([&] {
    pmx_cic_paint(nullptr, nullptr, nullptr, 1, 4, 4, 4, 4, 1.0f, nullptr);
    theirs = pmx_last_error();
  });
  t.join();
  if (mine == theirs || std::string(pmx_last_error()) != mine) {
    printf("FAIL per-thread messages: '%s' / '%s'\n", mine.c_str(),
           theirs.c_str());
    ++g_bad;
  }

  if (g_bad) {
    printf("FAILED %d check(s)\n", g_bad);
    return 1;
  }
  printf("ok\n");
  return 0;
}
