"""Many simulations of uneven mesh sizes advanced together.

The reference batches only simulations of one size (a leading batch axis,
flowpm/tfpm.py:15-63); here each size class is one batched plan.  The whole KDK loop of every
class is captured once into a HIP graph (the library's launches are plain
stream work, so torch's graph capture records them) and replayed: one
launch of the graph runs every class for every step with no host
round trip.  Each class runs on its own stream (default priority) forked
from and joined back into the capture stream, so the graph holds one branch
per class and small meshes execute concurrently with large ones instead of
queueing behind them.  Coefficients are per class (each class may have its
own cosmology and stages).
"""
import numpy as np
import torch

from . import _lib
from . import _plans
from . import tfpm


class _Group:

  def __init__(self, cosmo, state, stages, nc, pm_nc_factor):
    self.plan, self.B, self.N, self.f, self.fac = tfpm._force_setup(
        cosmo, state, nc, pm_nc_factor)
    self.state = tfpm._f32(state).clone()
    self.coef = tfpm.step_coefficients(cosmo, stages)
    self.cptr = _lib.dbl(self.coef)
    self.nsteps = len(self.coef) // 3
    self.work = torch.empty(_lib.lib().pmx_nbody_work_floats(
        self.plan.handle), dtype=torch.float32, device=state.device)

  def run_all(self, stream):
    """Every step in one library call: consecutive steps share the fused
    readout + closing kick + next opening kick + drift pass."""
    _lib.call("pmx_nbody_run", self.plan.handle, _lib.ptr(self.state), self.N,
              self.f, self.fac, self.cptr, 0, self.nsteps, 1,
              _lib.ptr(self.work), stream)


class BatchedNbody:
  """Advances several (cosmo, state, stages, nc) groups together.

  ``groups``: list of dicts with keys cosmo, state (3, B, N, 3), stages, nc
  and optionally pm_nc_factor.  ``run()`` integrates every group to its last
  stage and returns the final states; with ``graph=True`` the loop is
  captured once and replayed on later ``run()`` calls from the same initial
  states (``reset(states)`` loads new ones in place)."""

  def __init__(self, groups, graph=True):
    self.groups = [_Group(g["cosmo"], g["state"], g["stages"], g["nc"],
                          g.get("pm_nc_factor", 1)) for g in groups]
    self.init = [g.state.clone() for g in self.groups]
    self.streams = [torch.cuda.Stream() for _ in self.groups]
    self.use_graph = graph
    self.graph = None
    self.nsteps = max(max(g.nsteps for g in self.groups), 1)

  def _loop(self):
    main = torch.cuda.current_stream()
    for st in self.streams:
      st.wait_stream(main)
    for g, st in zip(self.groups, self.streams):
      with torch.cuda.stream(st):
        g.run_all(_lib._vp(st.cuda_stream))
    for st in self.streams:
      main.wait_stream(st)

  def run_one_call(self):
    """Integrates every group with one pmx_nbody_batched_step call (all
    groups enqueued on the current stream, no graph); returns the final
    states."""
    self.reset()
    G = len(self.groups)
    c = _lib._c
    arr = lambda typ, vals: (typ * G)(*vals)
    _lib.call("pmx_nbody_batched_step", G,
              arr(_lib._vp, [g.plan.handle for g in self.groups]),
              arr(_lib._vp, [g.state.data_ptr() for g in self.groups]),
              arr(_lib._ll, [g.N for g in self.groups]),
              arr(_lib._f, [g.f for g in self.groups]),
              arr(_lib._f, [g.fac for g in self.groups]),
              arr(_lib._dp, [c.cast(g.cptr, _lib._dp) for g in self.groups]),
              arr(_lib._i, [0] * G),
              arr(_lib._i, [g.nsteps for g in self.groups]),
              arr(_lib._i, [1] * G),
              arr(_lib._vp, [g.work.data_ptr() for g in self.groups]),
              _lib.stream_of(self.groups[0].state))
    return [g.state for g in self.groups]

  def reset(self, states=None):
    for k, g in enumerate(self.groups):
      src = self.init[k] if states is None else tfpm._f32(states[k])
      g.state.copy_(src)

  def run(self):
    if not self.use_graph:
      self.reset()
      self._loop()
      return [g.state for g in self.groups]
    if self.graph is None:
      # warm once eagerly (loads kernels), then capture
      self.reset()
      self._loop()
      torch.cuda.synchronize()
      self.graph = torch.cuda.CUDAGraph()
      s = torch.cuda.Stream()
      s.wait_stream(torch.cuda.current_stream())
      with torch.cuda.stream(s):
        self.reset()
        with torch.cuda.graph(self.graph, stream=s):
          self._loop()
      torch.cuda.current_stream().wait_stream(s)
      torch.cuda.synchronize()
    self.reset()
    self.graph.replay()
    return [g.state for g in self.groups]
