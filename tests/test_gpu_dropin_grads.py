"""Gradients of the single-op drop-ins against the float64 oracle's torch
autograd (TF's conventions: d|x|/dx = sign(x), floor has no gradient).

kick / drift (tfpm.py:273-314), force (:317-366), lpt1 (:66-95),
lpt2_source (:98-145), apply_longrange (:183-219), apply_pgd (:222-270),
pgd (:369-422), nbody's intermediate states (:498-503) and a second
backward pass through a taped nbody.  Gradients w.r.t. complex spectra use
torch's convention (dL/dRe + i dL/dIm) on both sides.  Bars: relative RMS
1e-5 for the linear ops, 1e-4 for the PM ops (float32 forward), cosine
>= 0.999 for the gradients of PM ops.
"""
import numpy as np
import pytest
import torch

import flowpm_amd as fa
from oracle import background as obg
from oracle import pm as opm

pytestmark = pytest.mark.gpu


def rel(a, b):
  a = a.detach().cpu().to(torch.complex128 if a.is_complex() else
                          torch.float64)
  b = b.detach().cpu().to(a.dtype)
  return float(((a - b).abs()**2).mean().sqrt() / (b.abs()**2).mean().sqrt())


def cos(a, b):
  if a.is_complex():
    a, b = torch.view_as_real(a), torch.view_as_real(b)
  a = a.detach().double().cpu().reshape(-1)
  b = b.detach().double().cpu().reshape(-1)
  return float(a @ b / (a.norm() * b.norm()))


def _state(n, seed, amp=1.5, B=1):
  gen = torch.Generator().manual_seed(seed)
  x = opm.lagrangian_grid([n] * 3, B, torch.float64)
  x = x + amp * torch.randn(x.shape, generator=gen, dtype=torch.float64)
  p = torch.randn(x.shape, generator=gen, dtype=torch.float64)
  f = torch.randn(x.shape, generator=gen, dtype=torch.float64)
  return torch.stack([x, p, f])


def _lpt_state(n, seed, a=0.1):
  """1LPT state of a Planck15 linear field in a box of 3 n Mpc/h (float64):
  small, smooth displacements, so every brick's box fits its LDS budget
  and the paint never falls back to float atomics (whose summation order
  varies from run to run)."""
  from oracle import power as opw
  gen = torch.Generator().manual_seed(seed)
  white = torch.randn(1, n, n, n, generator=gen, dtype=torch.float64) * n**1.5
  k = np.logspace(-4, 1, 256)
  pkt = opw.linear_matter_power(obg.planck15(), k)
  pk = lambda x: opw.interp_tf(x.reshape(-1), k, pkt).reshape(x.shape)
  lin = opm.whitenoise_to_linear(white, 3.0 * n, pk)
  return opm.lpt_init(obg.GrowthTable(obg.planck15()), lin, a, order=1)


def _spectrum(n, seed):
  gen = torch.Generator().manual_seed(seed)
  x = torch.randn(1, n, n, n, generator=gen, dtype=torch.float64)
  return opm.r2c3d(x)


def _both(t, gpu, dtype=None):
  """(oracle leaf, GPU leaf) copies of t."""
  o = t.clone().requires_grad_(True)
  dt = dtype or (torch.complex64 if t.is_complex() else torch.float32)
  g = t.to(dt).to(gpu).requires_grad_(True)
  return o, g


def test_kick_drift_gradients(gpu):
  g64 = obg.GrowthTable(obg.planck15())
  c = fa.cosmology.Planck15()
  st = _state(8, 1)
  w = torch.randn(st.shape, generator=torch.Generator().manual_seed(2),
                  dtype=torch.float64)
  for fo, fg in ((opm.kick, fa.tfpm.kick), (opm.drift, fa.tfpm.drift)):
    so, sg = _both(st, gpu)
    (fo(g64, so, 0.2, 0.25, 0.3) * w).sum().backward()
    (fg(c, sg, 0.2, 0.25, 0.3) * w.float().to(gpu)).sum().backward()
    assert rel(sg.grad, so.grad) < 1e-6


@pytest.mark.parametrize("n,f", [(16, 1), (16, 2)])
def test_force_gradient(gpu, n, f):
  g64 = obg.GrowthTable(obg.planck15())
  st = _state(n, 3)
  w = torch.randn(st.shape, generator=torch.Generator().manual_seed(4),
                  dtype=torch.float64)
  so, sg = _both(st, gpu)
  (opm.force(g64, so, [n] * 3, pm_nc_factor=f) * w).sum().backward()
  out = fa.tfpm.force(fa.cosmology.Planck15(), sg, n, pm_nc_factor=f)
  (out * w.float().to(gpu)).sum().backward()
  # the incoming F is replaced: no gradient flows to it
  assert float(sg.grad[2].abs().max()) == 0.0
  assert cos(sg.grad, so.grad) > 0.999
  assert rel(sg.grad, so.grad) < 1e-2


def test_lpt1_and_apply_longrange_gradients(gpu):
  n = 16
  dk = _spectrum(n, 5)
  q = opm.lagrangian_grid([n] * 3, 1, torch.float64) + 0.3
  w = torch.randn(1, n**3, 3, generator=torch.Generator().manual_seed(6),
                  dtype=torch.float64)
  cases = [(lambda d, x: opm.lpt1(d, x), lambda d, x: fa.tfpm.lpt1(d, x))]
  for split in (0.0, 0.7):
    cases.append((
        lambda d, x, s=split: opm.apply_longrange(x, d, split=s, factor=0.4),
        lambda d, x, s=split: fa.tfpm.apply_longrange(x, d, split=s,
                                                      factor=0.4)))
  for fo, fg in cases:
    do, dg = _both(dk, gpu)
    xo, xg = _both(q, gpu)
    (fo(do, xo) * w).sum().backward()
    (fg(dg, xg) * w.float().to(gpu)).sum().backward()
    assert rel(dg.grad, do.grad) < 1e-5
    assert rel(xg.grad, xo.grad) < 1e-4


def test_lpt2_source_gradient(gpu):
  n = 16
  dk = _spectrum(n, 7)
  gen = torch.Generator().manual_seed(8)
  wc = torch.complex(torch.randn(1, n, n, n, generator=gen,
                                 dtype=torch.float64),
                     torch.randn(1, n, n, n, generator=gen,
                                 dtype=torch.float64))
  do, dg = _both(dk, gpu)
  (opm.lpt2_source(do) * wc.conj()).real.sum().backward()
  (fa.tfpm.lpt2_source(dg) * wc.to(torch.complex64).to(gpu).conj()
   ).real.sum().backward()
  assert rel(dg.grad, do.grad) < 1e-5


def test_apply_pgd_forward_and_gradients(gpu):
  n = 16
  dk = _spectrum(n, 9)
  q = opm.lagrangian_grid([n] * 3, 1, torch.float64) + 0.4
  w = torch.randn(1, n**3, 3, generator=torch.Generator().manual_seed(10),
                  dtype=torch.float64)
  par = torch.tensor([0.3, 0.9, 4.0], dtype=torch.float64)
  do, dg = _both(dk, gpu)
  xo, xg = _both(q, gpu)
  po, pg = _both(par, gpu)
  fo = opm.apply_pgd(xo, do, po[0], po[1], po[2])
  fg = fa.tfpm.apply_pgd(xg, dg, pg[0], pg[1], pg[2])
  assert rel(fg, fo) < 1e-5
  (fo * w).sum().backward()
  (fg * w.float().to(gpu)).sum().backward()
  assert rel(dg.grad, do.grad) < 1e-5
  assert rel(xg.grad, xo.grad) < 1e-4
  assert rel(pg.grad, po.grad) < 1e-4, (pg.grad, po.grad)


@pytest.mark.parametrize("f", [1, 2])
def test_pgd_forward_and_gradients(gpu, f):
  n = 16
  st = _state(n, 11, amp=0.8)
  w = torch.randn(st.shape, generator=torch.Generator().manual_seed(12),
                  dtype=torch.float64)
  par = torch.tensor([0.3, 0.9 / f, 4.0], dtype=torch.float64)
  so, sg = _both(st, gpu)
  po, pg = _both(par, gpu)
  out_o = opm.pgd(so, po, [n] * 3, pm_nc_factor=f)
  out_g = fa.tfpm.pgd(sg, pg, [n] * 3, pm_nc_factor=f)
  q = opm.lagrangian_grid([n] * 3, 1, torch.float64)
  assert rel(out_g[0] - q.float().to(gpu), out_o[0] - q) < 1e-4
  assert torch.equal(out_g[1:], sg[1:].detach())
  (out_o * w).sum().backward()
  (out_g * w.float().to(gpu)).sum().backward()
  assert cos(sg.grad, so.grad) > 0.999
  assert rel(pg.grad, po.grad) < 2e-2, (pg.grad, po.grad)


def test_nbody_intermediate_states_gradient(gpu):
  """Losses on two intermediate states and the last one, differentiated
  w.r.t. the initial state (tfpm.py:498-503 returns them inside the tape)."""
  n = 32
  g64 = obg.GrowthTable(obg.planck15())
  c = fa.cosmology.Planck15()
  st = _lpt_state(n, 13)
  stages = np.linspace(0.1, 1.0, 5)
  gen = torch.Generator().manual_seed(14)
  ws = [torch.randn(st.shape, generator=gen, dtype=torch.float64)
        for _ in range(4)]
  so, sg = _both(st, gpu)
  oo = opm.nbody(g64, so, stages, [n] * 3, return_intermediate_states=True)
  og = fa.nbody(c, sg, stages, n, return_intermediate_states=True)
  assert [a for a, _ in og] == list(stages[1:])
  for (_, a), (_, b) in zip(og, oo):
    assert rel(a, b) < 1e-4
  sum((s * w).sum() for (_, s), w in zip(oo[1::2], ws)).backward()
  sum((s * w.float().to(gpu)).sum() for (_, s), w in zip(og[1::2], ws)
      ).backward()
  assert cos(sg.grad, so.grad) > 0.999
  # the values equal the single-run path's (to float32 rounding: the
  # taped steps read a float4 force mesh, the plain run a packed one)
  plain = fa.nbody(c, st.float().to(gpu), stages, n,
                   return_intermediate_states=True)
  for (_, a), (_, b) in zip(og, plain):
    assert rel(a, b) < 1e-6


def test_taped_nbody_second_backward(gpu):
  """A second backward pass (retain_graph) reads the same tape: the same
  gradient, bit for bit (smooth ICs: no brick spills to float atomics)."""
  n = 32
  c = fa.cosmology.Planck15()
  st = _lpt_state(n, 15)
  sg = st.float().to(gpu).requires_grad_(True)
  out = fa.nbody(c, sg, np.linspace(0.1, 1.0, 4), n, checkpoint_every=0)
  loss = (out**2).sum()
  g1, = torch.autograd.grad(loss, sg, retain_graph=True)
  g2, = torch.autograd.grad(loss, sg)
  assert torch.equal(g1, g2)


def test_power_spectrum_gradient(gpu):
  from oracle import spectrum as osp
  rng = np.random.default_rng(16)
  shape, L = (2, 16, 16, 24), [40.0, 40.0, 60.0]
  x = torch.tensor(rng.standard_normal(shape))
  kmin, dk = 0.1, 0.25
  xo, xg = _both(x, gpu)
  _, Po = osp.power_spectrum_torch(xo, kmin, dk, L)
  _, Pg = fa.power_spectrum(xg, kmin=kmin, dk=dk, boxsize=L)
  ok = torch.isfinite(Po[0])
  assert rel(Pg[:, ok], Po[:, ok]) < 1e-5
  w = torch.tensor(rng.standard_normal(int(ok.sum())))
  (Po[:, ok] * w).sum().backward()
  (Pg[:, ok.to(gpu)] * w.float().to(gpu)).sum().backward()
  assert rel(xg.grad, xo.grad) < 1e-4


def test_pgd_fit_loss_gradient(gpu):
  """The PGD-fit loss of scripts/fit_pgd.py:44-70 -- pgd, cic_paint,
  compensate_cic, power_spectrum, sum (1 - P / P_target)^2 -- differentiated
  w.r.t. (alpha, kl, ks) against the float64 oracle's autograd."""
  from oracle import spectrum as osp
  n, L = 16, 64.0
  g64 = obg.GrowthTable(obg.planck15())
  st = _state(n, 17, amp=0.0, B=2)
  gen = torch.Generator().manual_seed(18)
  st[0] = st[0] + 0.6 * torch.randn(st[0].shape, generator=gen,
                                    dtype=torch.float64)
  st[2] = 0
  st = opm.nbody(g64, st, np.linspace(0.3, 1.0, 3), [n] * 3)
  kmin, dk = np.pi / L, 2 * np.pi / L
  par = torch.tensor([0.2, 0.4, 6.0], dtype=torch.float64)

  def loss_o(p):
    s = opm.pgd(st, p, [n] * 3)
    f = opm.cic_paint(torch.zeros(2, n, n, n, dtype=torch.float64), s[0])
    _, pk = osp.power_spectrum_torch(osp.compensate_cic_torch(f), kmin, dk,
                                     [L] * 3)
    return pk.mean(0)

  po = par.clone().requires_grad_(True)
  pk_o = loss_o(po)
  ok = torch.isfinite(pk_o.detach())
  target = 0.8 * pk_o.detach()[ok]
  Lo = ((1 - pk_o[ok] / target)**2).sum()
  Lo.backward()
  pg = par.float().to(gpu).requires_grad_(True)
  sg = st.float().to(gpu)
  s = fa.tfpm.pgd(sg, pg, [n] * 3)
  f = fa.cic_paint(torch.zeros(2, n, n, n, device=gpu), s[0])
  _, pk = fa.power_spectrum(fa.compensate_cic(f), kmin=kmin, dk=dk,
                            boxsize=[L] * 3)
  pk = pk.mean(0)
  Lg = ((1 - pk[ok.to(gpu)] / target.float().to(gpu))**2).sum()
  assert abs(float(Lg.detach()) - float(Lo)) < 1e-3 * max(1.0, float(Lo))
  Lg.backward()
  print("fit_pgd loss gradient: gpu %s oracle %s" % (pg.grad.cpu().numpy(),
                                                    po.grad.numpy()))
  assert cos(pg.grad, po.grad) > 0.999
  assert rel(pg.grad, po.grad) < 2e-2


def test_checkpointed_nbody_second_backward(gpu):
  """Checkpoints every 2 steps: the last segment runs taped in the forward
  pass; a second backward pass (retain_graph) recomputes every segment and
  gives the same gradient."""
  n = 32
  c = fa.cosmology.Planck15()
  st = _lpt_state(n, 16)
  sg = st.float().to(gpu).requires_grad_(True)
  stages = np.linspace(0.1, 1.0, 6)
  out = fa.nbody(c, sg, stages, n, checkpoint_every=2)
  loss = (out**2).sum()
  g1, = torch.autograd.grad(loss, sg, retain_graph=True)
  g2, = torch.autograd.grad(loss, sg)
  assert torch.equal(g1, g2)
  ref = fa.nbody(c, sg, stages, n, checkpoint_every=0)
  g0, = torch.autograd.grad((ref**2).sum(), sg)
  assert torch.equal(g1, g0)
