"""Tile-shape autotuner for the FFT line passes.

For a plan (batch, nx, ny, nz) it times the fused density -> force-mesh
pipeline (z r2c, y fwd, fused x, y inv, z c2r) with HIP events for every
allowed (tile_y, tile_x) column width and keeps the fastest.  Choices are
cached per process and, when ``FLOWPM_AMD_TUNE_CACHE`` names a file, in that
JSON file across processes.
"""
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


def tune(device, B, nx, ny, nz, reps=5, verbose=False):
  """Measures every tile choice for this shape, applies and returns the
  fastest (tile_y, tile_x) and the timings."""
  key = (int(B), int(nx), int(ny), int(nz))
  plan = _plans.get(device, *key)
  table = _load()
  results = {}
  for ty in _WIDTHS:
    for tx in _WIDTHS:
      _lib.call("pmx_plan_set_tiles", plan.handle, ty, tx)
      results[(ty, tx)] = time_force_meshes(plan, reps)
      if verbose:
        print("tiles y=%d x=%d: %.4f ms" % (ty, tx, results[(ty, tx)]))
  best = min(results, key=results.get)
  _lib.call("pmx_plan_set_tiles", plan.handle, *best)
  _plans.TILE_CHOICES[key] = best
  table[key] = best
  _save(table)
  return best, results


def apply_cached():
  """Loads cached choices so that new plans pick them up."""
  _plans.TILE_CHOICES.update(_load())
