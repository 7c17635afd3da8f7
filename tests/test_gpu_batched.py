"""Batched simulations of uneven sizes under a captured HIP graph.

Every replay of the graph is checked against the eager per-group nbody run,
with eager work and synchronisations between replays (graph replays must not
depend on state left by earlier replays)."""
import numpy as np
import pytest
import torch

import flowpm_amd as fa
from flowpm_amd.batched import BatchedNbody

pytestmark = pytest.mark.gpu


def _state(cosmo, n, B, seed, dev):
  gen = torch.Generator().manual_seed(seed)
  white = (torch.randn(B, n, n, n, generator=gen, dtype=torch.float64) *
           n**1.5).float().to(dev)
  k = np.logspace(-4, 1, 256)
  pkt = fa.tfpower.linear_matter_power(cosmo, k)
  pk = lambda x: fa.scipy.interpolate.interp_tf(x, k, pkt)
  lin = fa.whitenoise_to_linear(n, 3.0 * n, white, pk)
  return fa.lpt_init(cosmo, lin, 0.1, order=1)


def _rel(a, b):
  return float((a - b).norm() / b.norm())


def test_batched_uneven_graph_replays(gpu):
  cosmo = fa.cosmology.Planck15()
  stages = np.linspace(0.1, 1.0, 6)
  spec = [(16, 2), (32, 1), (24, 3), (20, 2), (40, 1)]  # radix-5 sides too
  groups = []
  ref = []
  for i, (n, B) in enumerate(spec):
    st = _state(cosmo, n, B, 1000 + i, gpu)
    groups.append(dict(cosmo=cosmo, state=st, stages=stages, nc=n))
    ref.append(fa.nbody(cosmo, st, stages, n))
  runner = BatchedNbody(groups, graph=True)
  for rep in range(3):
    out = runner.run()
    torch.cuda.synchronize()
    for o, r in zip(out, ref):
      e = _rel(o, r)
      assert e < 1e-5, (rep, e)
    # eager work between replays
    junk = fa.cic_paint(torch.zeros(1, 16, 16, 16, device=gpu),
                        torch.rand(1, 100, 3, device=gpu) * 16)
    torch.cuda.synchronize()
    assert float(junk.sum()) == pytest.approx(100.0, rel=1e-5)


def test_same_size_batch_matches_singles(gpu):
  """A batch of B simulations in one plan equals B separate runs."""
  cosmo = fa.cosmology.Planck15()
  stages = np.linspace(0.1, 1.0, 5)
  st = _state(cosmo, 32, 3, 7, gpu)
  out = fa.nbody(cosmo, st, stages, 32)
  for b in range(3):
    one = fa.nbody(cosmo, st[:, b:b + 1].contiguous(), stages, 32)
    assert _rel(out[:, b:b + 1], one) < 1e-5


def test_batched_step_entry_matches_nbody(gpu):
  """pmx_nbody_batched_step over uneven groups = one nbody per group."""
  cosmo = fa.cosmology.Planck15()
  stages = np.linspace(0.1, 1.0, 5)
  spec = [(16, 1), (24, 2), (32, 1)]
  groups, ref = [], []
  for i, (n, B) in enumerate(spec):
    st = _state(cosmo, n, B, 2000 + i, gpu)
    groups.append(dict(cosmo=cosmo, state=st, stages=stages, nc=n))
    ref.append(fa.nbody(cosmo, st, stages, n))
  out = BatchedNbody(groups, graph=False).run_one_call()
  torch.cuda.synchronize()
  for o, r in zip(out, ref):
    assert _rel(o, r) < 1e-6


def test_autotune_picks_a_tile_and_results_do_not_depend_on_it(gpu):
  """The tuner times every (tile_y, tile_x) on two PM steps and applies the
  fastest; every choice gives the same trajectory (the tile width only
  changes which workgroup transforms a line, not the arithmetic)."""
  from flowpm_amd import _plans, autotune
  c = fa.cosmology.Planck15()
  n, B = 32, 2
  st = _state(c, n, B, 41, gpu)
  stages = np.linspace(0.1, 1.0, 4)
  plan = _plans.get(gpu, B, n, n, n)
  outs = {}
  for tiles in ((8, 8), (8, 16), (16, 8), (16, 16)):
    fa._lib.call("pmx_plan_set_tiles", plan.handle, *tiles)
    outs[tiles] = fa.nbody(c, st, stages, n)
  for tiles, o in outs.items():
    assert _rel(o, outs[(16, 16)]) < 1e-6, tiles
  best, ms = autotune.tune(gpu, B, n, n, n)
  assert sorted(ms) == sorted(outs)
  assert ms[best] == min(ms.values()) and all(v > 0 for v in ms.values())
  assert _plans.TILE_CHOICES[(B, n, n, n)] == best


def test_reset_with_new_states_then_run(gpu):
  """reset(new) replaces the initial states: the next run() -- eager and
  graph replay -- integrates them."""
  cosmo = fa.cosmology.Planck15()
  stages = np.linspace(0.1, 1.0, 4)
  spec = [(16, 1), (24, 2)]
  old = [_state(cosmo, n, B, 300 + i, gpu) for i, (n, B) in enumerate(spec)]
  new = [_state(cosmo, n, B, 400 + i, gpu) for i, (n, B) in enumerate(spec)]
  ref = [fa.nbody(cosmo, s, stages, n) for s, (n, _) in zip(new, spec)]
  for graph in (False, True):
    runner = BatchedNbody([dict(cosmo=cosmo, state=s, stages=stages, nc=n)
                           for s, (n, _) in zip(old, spec)], graph=graph)
    runner.run()
    runner.reset(new)
    out = runner.run()
    torch.cuda.synchronize()
    for o, r in zip(out, ref):
      assert _rel(o, r) < 1e-6, graph


def test_batched_forward_and_gradient_graph(gpu):
  """run_grad: every class's taped run and reverse sweep in one graph
  replay equals the per-group nbody forward and its autograd gradient."""
  cosmo = fa.cosmology.Planck15()
  stages = np.linspace(0.1, 1.0, 5)
  spec = [(16, 2), (32, 1), (24, 1)]
  sts = [_state(cosmo, n, B, 500 + i, gpu) for i, (n, B) in enumerate(spec)]
  gen = torch.Generator().manual_seed(7)
  cts = [torch.randn(s.shape, generator=gen).to(gpu) for s in sts]
  ref_f, ref_g = [], []
  for s, (n, _), c in zip(sts, spec, cts):
    x0 = s.detach().clone().requires_grad_(True)
    out = fa.nbody(cosmo, x0, stages, n, checkpoint_every=0)
    (g,) = torch.autograd.grad(out, x0, grad_outputs=c)
    ref_f.append(out.detach())
    ref_g.append(g)
  runner = BatchedNbody([dict(cosmo=cosmo, state=s, stages=stages, nc=n)
                         for s, (n, _) in zip(sts, spec)], grad=True)
  for rep in range(3):
    fin, grads = runner.run_grad(cts if rep == 0 else None)
    torch.cuda.synchronize()
    for o, r in zip(fin, ref_f):
      assert _rel(o, r) < 1e-6, rep
    for o, r in zip(grads, ref_g):
      assert _rel(o, r) < 1e-6, rep
    fa.cic_paint(torch.zeros(1, 16, 16, 16, device=gpu),
                 torch.rand(1, 100, 3, device=gpu) * 16)
    torch.cuda.synchronize()
