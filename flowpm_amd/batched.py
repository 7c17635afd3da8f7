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
own cosmology and stages).  With ``grad=True`` each class also keeps a tape,
and ``run_grad`` replays a second graph holding every class's taped run and
hand-written reverse sweep: forward and gradient of the whole uneven batch
in one graph launch.
"""
import numpy as np
import torch

from . import _lib
from . import _plans
from . import tfpm
from .cosmology import grad_params


class _Group:

  def __init__(self, cosmo, state, stages, nc, pm_nc_factor, grad=False):
    if grad_params(cosmo):
      raise NotImplementedError(
          "BatchedNbody: gradients w.r.t. cosmological parameters are "
          "available through flowpm_amd.nbody, not the batched runner")
    self.plan, self.B, self.N, self.f, self.fac = tfpm._force_setup(
        cosmo, state, nc, pm_nc_factor)
    self.state = tfpm._f32(state).clone()
    self.coef = tfpm.step_coefficients(cosmo, stages)
    self.cptr = _lib.dbl(self.coef)
    self.nsteps = len(self.coef) // 3
    L = _lib.lib()
    dev = state.device
    self.work = torch.empty(L.pmx_nbody_work_floats(self.plan.handle),
                            dtype=torch.float32, device=dev)
    if grad:
      # tape of the forward run and the reverse sweep's workspace; gfinal
      # holds the cotangent of the final state, gstate receives the
      # cotangent of the initial state
      h = self.plan.handle
      self.tape = torch.empty(L.pmx_nbody_tape_floats(h, self.N, self.nsteps),
                              dtype=torch.float32, device=dev)
      self.adj_work = torch.empty(L.pmx_nbody_adj_work_floats(h, self.N, 0),
                                  dtype=torch.float32, device=dev)
      self.gfinal = torch.zeros_like(self.state)
      self.gstate = torch.zeros_like(self.state)

  def run_all(self, stream):
    """Every step in one library call: consecutive steps share the fused
    readout + closing kick + next opening kick + drift pass."""
    _lib.call("pmx_nbody_run", self.plan.handle, _lib.ptr(self.state), self.N,
              self.f, self.fac, self.cptr, 0, self.nsteps, 1,
              _lib.ptr(self.work), stream)

  def run_grad(self, stream):
    """Taped forward run, then the hand-written reverse sweep of the
    cotangent in gfinal (copied into gstate, which the sweep turns into the
    initial state's cotangent)."""
    _lib.call("pmx_nbody_run_taped", self.plan.handle, _lib.ptr(self.state),
              self.N, self.f, self.fac, self.cptr, self.nsteps,
              _lib.ptr(self.tape), _lib.ptr(self.work), stream)
    self.gstate.copy_(self.gfinal)
    _lib.call("pmx_nbody_adjoint_taped", self.plan.handle,
              _lib.ptr(self.tape), self.nsteps, self.N, self.f, self.fac,
              self.cptr, _lib.ptr(self.gstate), _lib.ptr(self.adj_work),
              None, stream)


class BatchedNbody:
  """Advances several (cosmo, state, stages, nc) groups together.

  ``groups``: list of dicts with keys cosmo, state (3, B, N, 3), stages, nc
  and optionally pm_nc_factor.  ``run()`` integrates every group to its last
  stage and returns the final states; with ``graph=True`` the loop is
  captured once and replayed on later ``run()`` calls from the initial
  states (``reset(states)`` replaces them in place; the graph reads the
  same buffers)."""

  def __init__(self, groups, graph=True, grad=False):
    self.groups = [_Group(g["cosmo"], g["state"], g["stages"], g["nc"],
                          g.get("pm_nc_factor", 1), grad=grad)
                   for g in groups]
    self.init = [g.state.clone() for g in self.groups]
    self.streams = [torch.cuda.Stream() for _ in self.groups]
    self.use_graph = graph
    self.with_grad = grad
    self.graph = None
    self.graph_grad = None
    self.nsteps = max(max(g.nsteps for g in self.groups), 1)

  def _loop(self, grad=False):
    main = torch.cuda.current_stream()
    for st in self.streams:
      st.wait_stream(main)
    for g, st in zip(self.groups, self.streams):
      with torch.cuda.stream(st):
        if grad:
          g.run_grad(_lib._vp(st.cuda_stream))
        else:
          g.run_all(_lib._vp(st.cuda_stream))
    for st in self.streams:
      main.wait_stream(st)

  def _capture(self, grad):
    # warm once eagerly (loads kernels), then capture
    self.reset()
    self._loop(grad)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
      self.reset()
      with torch.cuda.graph(graph, stream=s):
        self._loop(grad)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    return graph

  def run_grad(self, cotangents=None):
    """Forward and gradient of every group: each group's taped run and its
    reverse sweep, all classes in one graph replay (graph=True).
    ``cotangents``: one tensor per group shaped like its state, the
    cotangent of the final state (None: the last ones given, initially
    zero).  Returns (final states, initial-state cotangents)."""
    if not self.with_grad:
      raise RuntimeError("BatchedNbody was built without grad=True")
    if cotangents is not None:
      if len(cotangents) != len(self.groups):
        raise ValueError("run_grad: one cotangent per group")
      for g, c in zip(self.groups, cotangents):
        if tuple(c.shape) != tuple(g.state.shape):
          raise ValueError("run_grad: cotangent shape %s, expected %s" %
                           (tuple(c.shape), tuple(g.state.shape)))
        g.gfinal.copy_(tfpm._f32(c))
    if not self.use_graph:
      self.reset()
      self._loop(True)
    else:
      if self.graph_grad is None:
        self.graph_grad = self._capture(True)
      self.reset()
      self.graph_grad.replay()
    return ([g.state for g in self.groups], [g.gstate for g in self.groups])

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
    """Copies the initial states into the working states; with ``states``
    (one per group, same shapes) they first become the initial states, so
    later ``run()`` calls (graph replays included) integrate them."""
    if states is not None:
      if len(states) != len(self.groups):
        raise ValueError("reset: one state per group")
      for k, st in enumerate(states):
        if tuple(st.shape) != tuple(self.init[k].shape):
          raise ValueError("reset: state %d has shape %s, expected %s" %
                           (k, tuple(st.shape), tuple(self.init[k].shape)))
        self.init[k].copy_(tfpm._f32(st))
    for k, g in enumerate(self.groups):
      g.state.copy_(self.init[k])

  def run(self):
    if not self.use_graph:
      self.reset()
      self._loop()
      return [g.state for g in self.groups]
    if self.graph is None:
      self.graph = self._capture(False)
    self.reset()
    self.graph.replay()
    return [g.state for g in self.groups]
