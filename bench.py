#!/usr/bin/env python
"""PM-step benchmark on one MI355X.

Workload: a single simulation with 256^3 particles on a 256^3 force mesh,
box 512 Mpc/h, Planck15 Eisenstein-Hu P(k), seeded Gaussian initial
conditions, 2LPT at a0 = 0.1, 40 linear-in-a KDK steps to a = 1, float32.
A "step" is one PM step (paint, forward FFT, Poisson x gradient, three
inverse FFTs, fused readout + kick + drift).

Reports (one JSON line):
  value / ms_per_step      forward ms per PM step (K timed steps, including
                           the jump-start force of nbody)
  fwd_adjoint_ms_per_step  forward + hand-written adjoint per step, with
                           the automatic trajectory policy (at 256^3 the
                           whole trajectory is taped: no recomputation);
                           fwd_adjoint_recompute8_ms_per_step: checkpoints
                           every 8 steps, segments recomputed from the
                           kept potential of every force (gradient stencil,
                           no paint or transforms);
                           fwd_adjoint_recompute8_states_only_ms_per_step:
                           the same keeping the states only (segments
                           recomputed from scratch);
                           fwd_adjoint_with_omega_c_ms_per_step: the
                           automatic policy with the gradient w.r.t.
                           Omega_c as well;
                           fwd_adjoint_compact_tape_ms_per_step: the
                           compact tape (positions and potentials)
  roofline                 dominant kernel: algorithmic bytes per launch /
                           average launch time from HIP events recorded on
                           the launch stream during a forward run
  cpu_baseline             the CPU oracle (torch, float32, the reference's
                           op graph restated) timed for one PM step at the
                           same size on this host's cores
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("ms/PM step fwd and fwd+adjoint at 256^3; batched sims/sec at "
          "64^3; HBM GB/s")
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(kind, n, N):
  """Bytes each kernel must read + write once, per launch (B = 1)."""
  M = n**3
  S8 = 8 * n * n * (n // 2 + 1)  # one half-spectrum buffer, bytes
  table = {
      "fill": 4 * M,
      # brick paint: deposit reads the particles, gather writes the mesh
      "paint": 12 * N,
      "paint_gather": 4 * M,
      "fft_z_r2c": 4 * M + S8,
      "fft_y_fwd": 2 * S8,
      # the potential: x forward, multiply, x inverse; y inverse; z c2r
      "fft_x_fused": 2 * S8,
      "fft_y_inv": 2 * S8,
      "fft_z_c2r": S8 + 4 * M,
      # potential in, the three force meshes out
      "grad": 4 * M + 3 * 4 * M,
      # X, P read and written; F stays in registers inside fused steps
      "readout_kick_drift": 3 * 4 * M + 24 * N + 24 * N,
  }
  return table.get(kind)


def main():
  ap = argparse.ArgumentParser()
  ap.add_argument("--gpus", type=int, default=1)
  ap.add_argument("--steps", type=int, default=40)
  ap.add_argument("--warmup", type=int, default=2)
  ap.add_argument("--n", type=int, default=256)
  ap.add_argument("--box", type=float, default=512.0)
  ap.add_argument("--lpt-order", type=int, default=2, choices=(1, 2))
  ap.add_argument("--every", type=int, default=None,
                  help="adjoint checkpoint interval (0: taped, no "
                  "recomputation; default: automatic)")
  ap.add_argument("--no-adjoint", action="store_true")
  ap.add_argument("--no-cpu-baseline", action="store_true")
  ap.add_argument("--no-batched", action="store_true")
  ap.add_argument("--copy-calibration-mb", type=int, default=0,
                  help="first run a plain device copy of this many MiB "
                  "(the known-size kernel that scales rocprofv3 "
                  "FETCH_SIZE / WRITE_SIZE counters)")
  args = ap.parse_args()
  if args.gpus != 1:
    print("bench.py measures one GPU: there is no multi-GPU path (batched "
          "simulations would shard by member with no communication, one "
          "large simulation by slabs with all-to-all transposes).  Run with "
          "--gpus 1.", file=sys.stderr)
    sys.exit(2)

  import flowpm_amd as fa
  from flowpm_amd import _lib
  from flowpm_amd import _plans

  dev = torch.device("cuda:0")
  if args.copy_calibration_mb:
    src = torch.ones(args.copy_calibration_mb << 18, device=dev)
    dst = torch.empty_like(src)
    for _ in range(3):
      dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
  n, L, K, W = args.n, args.box, args.steps, args.warmup
  N = n**3
  cosmo = fa.cosmology.Planck15()
  kk = np.logspace(-4, 1, 256)
  pk_tab = fa.tfpower.linear_matter_power(cosmo, kk)
  pk = lambda k: fa.scipy.interpolate.interp_tf(k, kk, pk_tab)
  gen = torch.Generator().manual_seed(100)
  white = (torch.randn(1, n, n, n, generator=gen, dtype=torch.float64) *
           n**1.5).float().to(dev)
  lin = fa.whitenoise_to_linear(n, L, white, pk)
  state0 = fa.lpt_init(cosmo, lin, 0.1, order=args.lpt_order)
  stages = np.linspace(0.1, 1.0, K + 1)
  coef = fa.tfpm.step_coefficients(cosmo, stages)
  plan = _plans.get(dev, 1, n, n, n)
  from flowpm_amd import autotune
  # untimed: the tuner runs this workload (same ICs and stages) per choice
  tiles, tile_ms = autotune.tune(dev, 1, n, n, n, state=state0,
                                 stages=stages, cosmo=cosmo)
  fac = float(np.float32(1.5 * cosmo.Omega_m))
  work = torch.empty(_lib.lib().pmx_nbody_work_floats(plan.handle),
                     dtype=torch.float32, device=dev)
  cptr = _lib.dbl(coef)
  stream = _lib.stream_of(work)

  def run(st, b0, b1, jump):
    _lib.call("pmx_nbody_run", plan.handle, _lib.ptr(st), N, 1.0, fac, cptr,
              b0, b1, jump, _lib.ptr(work), stream)

  # warmup: W steps on a copy
  st = state0.clone()
  run(st, 0, min(W, K), 1)
  torch.cuda.synchronize()

  # timed forward: K steps (jump-start force + K KDK steps)
  st = state0.clone()
  torch.cuda.synchronize()
  e0 = torch.cuda.Event(enable_timing=True)
  e1 = torch.cuda.Event(enable_timing=True)
  e0.record()
  run(st, 0, K, 1)
  e1.record()
  torch.cuda.synchronize()
  fwd_ms = e0.elapsed_time(e1)
  ms_step = fwd_ms / K

  # live per-kernel timing on the launch stream (separate forward run)
  st = state0.clone()
  torch.cuda.synchronize()
  kernels = live_profile(_lib, lambda: run(st, 0, K, 1))
  dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
  alg = algorithmic_bytes(dom, n, N)
  avg_s = kernels[dom]["avg_us"] * 1e-6
  achieved = alg / avg_s / 1e9
  for name, v in kernels.items():
    b = algorithmic_bytes(name, n, N)
    if b:
      v["alg_bytes"] = b
      v["GBps"] = b / (v["avg_us"] * 1e-6) / 1e9
    h = hbm_traffic(name, n, tiles)
    if h:
      v["hbm_bytes"] = h  # measured (committed counter summary)
      v["hbm_GBps"] = h / (v["avg_us"] * 1e-6) / 1e9

  # the final paint of the run (cic_paint drop-in of one particle per cell:
  # the brick paint on a workspace)
  zeros = torch.zeros(1, n, n, n, device=dev)
  fa.cic_paint(zeros, st[0])
  torch.cuda.synchronize()
  e0.record()
  for _ in range(5):
    rho_final = fa.cic_paint(zeros, st[0])
  e1.record()
  torch.cuda.synchronize()
  final_paint_ms = e0.elapsed_time(e1) / 5
  del rho_final, zeros

  # forward + adjoint (checkpoint every k, recompute in backward)
  fa_ms = fa8_ms = fa8s_ms = fac_ms = facompact_ms = every_used = None
  adj_kernels = None
  if not args.no_adjoint:
    x0 = state0.clone().requires_grad_(True)
    gseed = torch.randn(state0.shape, generator=torch.Generator().manual_seed(
        101)).to(dev)
    # warm the adjoint path on the full run, so the caching allocator holds
    # the tape and adjoint workspace before the timed run
    fin = fa.nbody(cosmo, x0, stages, n, checkpoint_every=args.every)
    torch.autograd.grad(fin, x0, grad_outputs=gseed)
    torch.cuda.synchronize()
    e0.record()
    fin = fa.nbody(cosmo, x0, stages, n, checkpoint_every=args.every)
    (gx,) = torch.autograd.grad(fin, x0, grad_outputs=gseed)
    e1.record()
    torch.cuda.synchronize()
    fa_ms = e0.elapsed_time(e1) / K
    assert torch.isfinite(gx).all()
    every_used = fa.tfpm._resolve_every(plan, N, K, args.every, dev)

    # the recomputing adjoint (segments of 8 steps) for comparison: with
    # the potentials of every force kept (the default when they fit), and
    # with the states only
    def ckpt8_ms(keep):
      fa.tfpm.set_keep_potentials(keep)
      try:
        for rep in range(2):  # warm-up, then timed
          e0.record()
          fin = fa.nbody(cosmo, x0, stages, n, checkpoint_every=8)
          (gx8,) = torch.autograd.grad(fin, x0, grad_outputs=gseed)
          e1.record()
          torch.cuda.synchronize()
        assert torch.isfinite(gx8).all()
        return e0.elapsed_time(e1) / K
      finally:
        fa.tfpm.set_keep_potentials(None)
    fa8_ms = ckpt8_ms(True)
    fa8s_ms = ckpt8_ms(False)

    # the compact tape: positions and potentials, force meshes formed again
    # in the backward pass
    fa.tfpm.set_compact_tape(True)
    try:
      for rep in range(2):  # warm-up, then timed
        e0.record()
        fin = fa.nbody(cosmo, x0, stages, n, checkpoint_every=0)
        (gxc,) = torch.autograd.grad(fin, x0, grad_outputs=gseed)
        e1.record()
        torch.cuda.synchronize()
      assert torch.isfinite(gxc).all()
      facompact_ms = e0.elapsed_time(e1) / K
    finally:
      fa.tfpm.set_compact_tape(None)
    del gxc, fin

    # forward + adjoint with the gradient w.r.t. Omega_c as well (the
    # reverse sweep also accumulates the step coefficients' cotangents)
    oc = torch.tensor(cosmo.Omega_c, dtype=torch.float64, requires_grad=True)
    cosmo_t = fa.cosmology.Planck15(Omega_c=oc)
    for rep in range(2):  # warm-up, then timed
      e0.record()
      fin = fa.nbody(cosmo_t, x0, stages, n, checkpoint_every=args.every)
      gxc, goc = torch.autograd.grad(fin, (x0, oc), grad_outputs=gseed)
      e1.record()
      torch.cuda.synchronize()
    assert torch.isfinite(gxc).all() and bool(torch.isfinite(goc))
    fac_ms = e0.elapsed_time(e1) / K
    del gxc, goc, fin

    def fwd_adj():
      fin = fa.nbody(cosmo, x0, stages, n, checkpoint_every=args.every)
      torch.autograd.grad(fin, x0, grad_outputs=gseed)
    adj_kernels = live_profile(_lib, fwd_adj)

  batched = None
  if not args.no_batched:
    batched = batched_runs(fa, cosmo, dev)

  copy_gbps = copy_peak(dev)

  cpu = None
  if not args.no_cpu_baseline:
    cpu = cpu_baseline(state0, n, cosmo)

  out = {
      "metric": METRIC,
      "value": ms_step,
      "unit": "ms/step",
      "n_gpus": 1,
      "steps": K,
      "warmup": W,
      "ms_per_step": ms_step,
      "higher_is_better": False,
      "scaling": "weak",
      "vs_baseline": None,
      "dtype": "f32",
      "data": "synthetic",
      "config": {
          "workload": "single sim, %d^3 particles, %d^3 force mesh, box %g "
                      "Mpc/h, Planck15 EH P(k), %dLPT a0=0.1, %d KDK steps "
                      "to a=1" % (n, n, L, args.lpt_order, K),
          "mesh": n, "particles": N, "box_mpc_h": L,
          "lpt_order": args.lpt_order,
          "a0": 0.1, "checkpoint_every": every_used,
          "fft_tiles": {"y": tiles[0], "x": tiles[1]},
      },
      "autotune_run_ms": {"%dx%d" % k: v for k, v in tile_ms.items()},
      "fwd_adjoint_ms_per_step": fa_ms,
      "adjoint_over_forward": (fa_ms / ms_step) if fa_ms else None,
      "fwd_adjoint_recompute8_ms_per_step": fa8_ms,
      "fwd_adjoint_recompute8_states_only_ms_per_step": fa8s_ms,
      "fwd_adjoint_with_omega_c_ms_per_step": fac_ms,
      "fwd_adjoint_compact_tape_ms_per_step": facompact_ms,
      "final_cic_paint_ms": final_paint_ms,
      "roofline": {
          "bound": "hbm",
          "kernel": dom,
          "achieved": achieved,
          "peak": HBM_PEAK_GBS,
          "unit": "GB/s",
          "frac": achieved / HBM_PEAK_GBS,
          "traffic": hbm_traffic(dom, n, tiles),
          "copy_peak": copy_gbps,
          "frac_of_copy": achieved / copy_gbps,
      },
      "step_alg_bytes": sum(
          (algorithmic_bytes(k, n, N) or 0) * v["launches"] / (K + 1)
          for k, v in kernels.items()),
      "kernels": kernels,
      "fwd_adjoint_kernels": adj_kernels,
      "batched": batched,
      "cpu_baseline": cpu,
  }
  print(json.dumps(out))


# rocprofv3 counter kernel names (regular expressions) of the profiler
# kinds, for the kernels the forward step launches
# ({ty} / {tx}: the column-tile widths of the y / x passes the run used)
_KIND_KERNEL = {
    "readout_kick_drift": r"k_readout_kick_drift<", "paint": r"k_brick_deposit<1>",
    "paint_gather": r"k_brick_gather<1, ", "fft_z_r2c": r"k_zpass_r2c<",
    "fft_y_fwd": r"k_ypass<\d+, {ty}, \d+, -1, false>",
    "fft_x_fused": r"k_xpass<\d+, {tx}, \d+, false>",
    "fft_y_inv": r"k_ypass<\d+, {ty}, \d+, 1, false>",
    "fft_z_c2r": r"k_zpass_c2r<\d+, 16, ", "grad": r"k_grad_mesh<3>", "fill": r"k_fill",
}
TRAFFIC_TABLE = os.path.join(ROOT, "profiles", "pm_step_256_hbm_traffic.md")


def hbm_traffic(kind, n, tiles=(16, 16)):
  """HBM bytes per launch of this kernel kind, from the committed rocprofv3
  FETCH_SIZE / WRITE_SIZE summary of the 256^3 workload (each counter scaled
  by what it reads for a plain 1 GiB copy); None when not measured.
  tiles: the (y, x) column-tile widths of the run."""
  import re
  if n != 256 or kind not in _KIND_KERNEL or not os.path.exists(TRAFFIC_TABLE):
    return None
  pat = re.compile(_KIND_KERNEL[kind].format(ty=tiles[0], tx=tiles[1]))
  for line in open(TRAFFIC_TABLE):
    if line.startswith("| `") and pat.search(line):
      cols = [c.strip() for c in line.strip().strip("|").split("|")]
      return float(cols[-1]) * 1e6
  return None


def live_profile(_lib, fn):
  """Per-kind launch counts and average durations (HIP events recorded by
  the library on its launch stream) over one call of fn."""
  _lib.call("pmx_profile_start")
  fn()
  nk = _lib.lib().pmx_profile_kinds()
  ms = (_lib._c.c_double * nk)()
  cnt = (_lib._ll * nk)()
  _lib.call("pmx_profile_stop", ms, cnt, nk)
  out = {}
  for k in range(nk):
    if cnt[k]:
      name = _lib.lib().pmx_profile_kind_name(k).decode()
      out[name] = {"launches": int(cnt[k]),
                   "avg_us": 1e3 * ms[k] / cnt[k],
                   "total_ms": ms[k]}
  return out


def _ic_states(fa, cosmo, n, L, B, seed, dev):
  kk = np.logspace(-4, 1, 256)
  pk_tab = fa.tfpower.linear_matter_power(cosmo, kk)
  pk = lambda k: fa.scipy.interpolate.interp_tf(k, kk, pk_tab)
  gen = torch.Generator().manual_seed(seed)
  white = (torch.randn(B, n, n, n, generator=gen, dtype=torch.float64) *
           n**1.5).float().to(dev)
  return fa.lpt_init(cosmo, fa.whitenoise_to_linear(n, L, white, pk), 0.1,
                     order=1)


def _time(fn, reps):
  fn()
  torch.cuda.synchronize()
  e0 = torch.cuda.Event(enable_timing=True)
  e1 = torch.cuda.Event(enable_timing=True)
  e0.record()
  for _ in range(reps):
    fn()
  e1.record()
  torch.cuda.synchronize()
  return e0.elapsed_time(e1) / reps


def batched_runs(fa, cosmo, dev):
  """Nbody throughput of many simulations (ICs excluded): 64^3 sims (box
  200, 1LPT a0=0.1, 10 steps) batched in one plan under a captured graph vs
  one at a time; the uneven mix of 8 each of 32/64/96/128^3 (boxes
  100/200/300/400) forward under one graph, and forward + gradient."""
  from flowpm_amd.batched import BatchedNbody
  from flowpm_amd import autotune
  stages = np.linspace(0.1, 1.0, 11)
  B = 64
  st = _ic_states(fa, cosmo, 64, 200.0, B, 1000, dev)
  # tile widths of every plan from the autotuner, on the timed workloads
  # (untimed)
  tiles = {"64x64^3": autotune.tune(dev, B, 64, 64, 64, state=st,
                                    stages=stages, cosmo=cosmo)[0],
           "1x64^3": autotune.tune(dev, 1, 64, 64, 64,
                                   state=st[:, :1].contiguous(),
                                   stages=stages, cosmo=cosmo)[0]}
  runner = BatchedNbody([dict(cosmo=cosmo, state=st, stages=stages, nc=64)])
  t_b = _time(runner.run, 3)
  from flowpm_amd import _lib
  eager = BatchedNbody([dict(cosmo=cosmo, state=st, stages=stages, nc=64)],
                       graph=False)
  eager.run()
  torch.cuda.synchronize()
  kern = live_profile(_lib, eager.run)
  one = BatchedNbody([dict(cosmo=cosmo, state=st[:, :1].contiguous(),
                           stages=stages, nc=64)], graph=False)
  t_1 = _time(one.run, 5)
  groups = []
  for i, (n, L) in enumerate(((32, 100.), (64, 200.), (96, 300.),
                              (128, 400.))):
    gst = _ic_states(fa, cosmo, n, L, 8, 2000 + i, dev)
    tiles["8x%d^3" % n] = autotune.tune(dev, 8, n, n, n, state=gst,
                                        stages=stages, cosmo=cosmo)[0]
    groups.append(dict(cosmo=cosmo, state=gst, stages=stages, nc=n))
  mix = BatchedNbody(groups)
  t_mix = _time(mix.run, 3)

  # forward and gradient of every class (taped run + reverse sweep per
  # class) in one graph replay
  mixg = BatchedNbody(groups, grad=True)
  cts = [torch.ones_like(g["state"]) for g in groups]
  mixg.run_grad(cts)
  t_mix_grad = _time(mixg.run_grad, 2)
  del mixg
  return {
      "sims64_batch": B,
      "sims64_per_s_batched": B / (t_b * 1e-3),
      "sims64_per_s_sequential": 1.0 / (t_1 * 1e-3),
      "speedup_vs_sequential": (B / t_b) / (1.0 / t_1),
      "uneven_mix_ms": t_mix,
      "uneven_mix_sims_per_s": 32 / (t_mix * 1e-3),
      "uneven_mix_fwd_grad_ms": t_mix_grad,
      "sims64_batch_kernels": kern,
      "fft_tiles": {k: "%dx%d" % v for k, v in tiles.items()},
      "uneven_mix_fwd_grad_sims_per_s": 32 / (t_mix_grad * 1e-3),
      "note": "10-step nbody per sim, ICs excluded; forward: one graph "
              "replay; fwd_grad: one graph replay of every class's taped "
              "run and reverse sweep",
  }


def cpu_baseline(state0, n, cosmo, reps=5):
  """PM steps (force, opening kick, drift, closing kick: one force
  evaluation each) of the same workload with the CPU oracle in float32 on
  all host threads: one warm-up step, then the median of ``reps`` steps."""
  from oracle import background as obg
  from oracle import pm as opm
  torch.set_num_threads(os.cpu_count() or 1)
  g = obg.GrowthTable(obg.planck15(), dtype=np.float64)
  st = state0.detach().cpu()
  a = np.linspace(0.1, 0.1 + 0.0225 * (reps + 1), reps + 2)
  times = []
  for i in range(reps + 1):
    a0, a1 = a[i], a[i + 1]
    ah = (a0 * a1)**0.5
    t0 = time.perf_counter()
    st = opm.force(g, st, [n] * 3)
    st = opm.kick(g, st, a0, a0, ah)
    st = opm.drift(g, st, a0, ah, a1)
    st = opm.kick(g, st, ah, a1, a1)
    times.append(time.perf_counter() - t0)
  return {"value": 1e3 * float(np.median(times[1:])), "unit": "ms/step",
          "cores": torch.get_num_threads(), "kind": "port",
          "sample": "%d^3 workload: 1 warm-up + median of %d PM steps (one "
                    "force evaluation + kick + drift + kick each) with the "
                    "float32 torch-CPU oracle" % (n, reps),
          "step_ms_all": [1e3 * t for t in times]}


def copy_peak(dev, mib=1024, reps=5):
  """A plain device-to-device copy of ``mib`` MiB (torch copy_): read +
  write bytes over the median HIP-event time, GB/s."""
  src = torch.ones(mib << 18, device=dev)
  dst = torch.empty_like(src)
  dst.copy_(src)
  e0 = torch.cuda.Event(enable_timing=True)
  e1 = torch.cuda.Event(enable_timing=True)
  ts = []
  for _ in range(reps):
    e0.record()
    dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ts.append(e0.elapsed_time(e1))
  del src, dst
  return 2 * (mib << 20) / (float(np.median(ts)) * 1e-3) / 1e9


if __name__ == "__main__":
  main()
