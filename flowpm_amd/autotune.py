"""Tile-shape autotuner for the FFT line passes.

For a plan (batch, nx, ny, nz) it times runs of the library's time loop
(paint, the force-mesh transforms, fused readout + kick + drift), each
restarted from the same state, with HIP events, for every allowed
(tile_y, tile_x) column width and keeps the fastest: by default two PM
steps on particles near the plan's own grid, or the caller's workload
(its initial state and stages).  The steps are timed
rather than the transforms alone because the tile width that is faster on
cache-resident spectra need not be faster inside the step, where the
particle streams keep the spectra out of the 256 MB last-level cache.
Choices are cached per process and, when ``FLOWPM_AMD_TUNE_CACHE`` names a
file, in that JSON file across processes.
"""
import numpy as np
import json
import os

import torch

from . import _lib
from . import _plans

_WIDTHS = (8, 16)


def _cache_path():
  return os.environ.get("FLOWPM_AMD_TUNE_CACHE")


def _load():
  path = _cache_path()
  if path and os.path.exists(path):
    with open(path) as f:
      return {tuple(json.loads(k)): tuple(v) for k, v in json.load(f).items()}
  return {}


def _save(table):
  path = _cache_path()
  if not path:
    return
  with open(path, "w") as f:
    json.dump({json.dumps(list(k)): list(v) for k, v in table.items()}, f)


def time_force_meshes(plan, reps=5):
  """Mean ms of one fused force-mesh pipeline on this plan."""
  dev = plan.device
  rho = torch.rand((plan.B, plan.nx, plan.ny, plan.nz), device=dev)
  f = torch.empty((3,) + tuple(rho.shape), device=dev)
  work = plan.spec(3)
  st = _lib.stream_of(rho)

  def once():
    _lib.call("pmx_force_meshes", plan.handle, _lib.ptr(rho), _lib.ptr(work),
              _lib.ptr(f[0]), _lib.ptr(f[1]), _lib.ptr(f[2]), 1.0, st)

  once()
  torch.cuda.synchronize(dev)
  e0 = torch.cuda.Event(enable_timing=True)
  e1 = torch.cuda.Event(enable_timing=True)
  e0.record()
  for _ in range(reps):
    once()
  e1.record()
  torch.cuda.synchronize(dev)
  return e0.elapsed_time(e1) / reps


class _StepBench:
  """A run of the library's time loop, restarted from the same state every
  call: by default two KDK steps (early-time coefficients) on particles at
  the plan's grid points displaced by 0.3 cells rms; with ``state``,
  ``stages`` and ``cosmo`` the caller's own workload (particles on the
  plan's grid, pm_nc_factor 1)."""

  def __init__(self, plan, nsteps=2, state=None, stages=None, cosmo=None):
    from . import cosmology, tfpm
    dev = plan.device
    c = cosmo if cosmo is not None else cosmology.Planck15()
    if state is None:
      g = torch.Generator(device="cpu").manual_seed(0)
      q = torch.stack(torch.meshgrid(
          torch.arange(plan.nx, dtype=torch.float32),
          torch.arange(plan.ny, dtype=torch.float32),
          torch.arange(plan.nz, dtype=torch.float32), indexing="ij"),
                      dim=-1).reshape(1, -1, 3)
      st = torch.zeros((3, plan.B, q.shape[1], 3), dtype=torch.float32)
      st[0] = q + 0.3 * torch.randn((plan.B, q.shape[1], 3), generator=g)
      state = st.to(dev)
      stages = np.linspace(0.1, 0.12, nsteps + 1)
    self.init = state.detach().to(device=dev, dtype=torch.float32)
    self.state = torch.empty_like(self.init)
    self.N = int(self.init.shape[2])
    self.coef = _lib.dbl(tfpm.step_coefficients(c, stages))
    self.fac = float(np.float32(1.5 * c.Omega_m))
    self.nsteps = len(stages) - 1
    self.work = torch.empty(_lib.lib().pmx_nbody_work_floats(plan.handle),
                            dtype=torch.float32, device=dev)
    self.plan = plan
    self.stream = _lib.stream_of(self.work)

  def __call__(self):
    # every run starts from the same state: repeated runs would otherwise
    # integrate further and further (the cost of the particle passes grows
    # with the displacements)
    self.state.copy_(self.init)
    _lib.call("pmx_nbody_run", self.plan.handle, _lib.ptr(self.state),
              self.N, 1.0, self.fac, self.coef, 0, self.nsteps, 1,
              _lib.ptr(self.work), self.stream)


def _time(fn, dev, reps):
  fn()
  torch.cuda.synchronize(dev)
  e0 = torch.cuda.Event(enable_timing=True)
  e1 = torch.cuda.Event(enable_timing=True)
  e0.record()
  for _ in range(reps):
    fn()
  e1.record()
  torch.cuda.synchronize(dev)
  return e0.elapsed_time(e1) / reps


def time_steps(plan, reps=3, bench=None):
  """Mean ms of two PM steps of the time loop on this plan."""
  return _time(bench or _StepBench(plan), plan.device, reps)


def tune(device, B, nx, ny, nz, reps=3, rounds=3, verbose=False, state=None,
         stages=None, cosmo=None):
  """Measures every tile choice for this shape, applies and returns the
  fastest (tile_y, tile_x) and the timings (ms per run of the time loop:
  two PM steps, or the caller's ``state`` over ``stages``).

  The choices are timed in turn, ``rounds`` times over, and each keeps its
  best time: the differences between tile widths are a few per cent of a
  step, the size of the drift of the GPU's clocks between consecutive
  measurements."""
  key = (int(B), int(nx), int(ny), int(nz))
  plan = _plans.get(device, *key)
  table = _load()
  results = {}
  bench = _StepBench(plan, state=state, stages=stages, cosmo=cosmo)
  for _ in range(rounds):
    for ty in _WIDTHS:
      for tx in _WIDTHS:
        _lib.call("pmx_plan_set_tiles", plan.handle, ty, tx)
        t = time_steps(plan, reps, bench)
        results[(ty, tx)] = min(results.get((ty, tx), t), t)
  if verbose:
    for (ty, tx), t in sorted(results.items()):
      print("tiles y=%d x=%d: %.4f ms" % (ty, tx, t))
  best = min(results, key=results.get)
  _lib.call("pmx_plan_set_tiles", plan.handle, *best)
  _plans.TILE_CHOICES[key] = best
  table[key] = best
  _save(table)
  return best, results


def apply_cached():
  """Loads cached choices so that new plans pick them up."""
  _plans.TILE_CHOICES.update(_load())
