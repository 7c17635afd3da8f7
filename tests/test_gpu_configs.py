"""The full-size workloads on the GPU.

* 256^3 (box 512, 2LPT at a = 0.1, 40 steps): final positions, momenta and
  painted density against the float64 oracle's run of the same seeded ICs
  (subsampled in tests/golden/c4_256.npz, written by
  tests/golden/make_fixtures.py), at the 1e-4 relative RMS bar; the
  gradient of the density chi^2 with checkpoints every 8 steps and
  recomputation equals the taped one.
* 512^3 (box 1024, 1LPT, 10 steps): properties that hold at any size --
  the paint conserves mass, the force sums to zero at every step, the
  gradient with checkpoints every 2 steps equals the taped one, and the
  large-scale power grows as D1(a)^2.
* The uneven batch of 8 each of 32^3, 64^3, 96^3 and 128^3 simulations
  (boxes 100-400, 1LPT, 10 steps) under the captured graph: every member
  equals its class's batched nbody run, and the first member of each class
  matches the float64 oracle fixture (tests/golden/c3_<n>.npz).
"""
import numpy as np
import pytest
import torch

import flowpm_amd as fa
from flowpm_amd.batched import BatchedNbody

pytestmark = pytest.mark.gpu


def _rel(a, b):
  a = np.asarray(a, dtype=np.float64)
  b = np.asarray(b, dtype=np.float64)
  return float(np.sqrt(np.mean((a - b)**2) / np.mean(b**2)))


def _white(n, seeds):
  """Real white noise N(0, sqrt(M)): one torch.Generator stream per member
  (float64, the injection the fixtures were written with)."""
  out = []
  for s in seeds:
    gen = torch.Generator().manual_seed(int(s))
    out.append(torch.randn(n, n, n, generator=gen, dtype=torch.float64) *
               n**1.5)
  return torch.stack(out).float()


def _pk(cosmo):
  kk = np.logspace(-4, 1, 256)
  pkt = fa.tfpower.linear_matter_power(cosmo, kk)
  return lambda k: fa.scipy.interpolate.interp_tf(k, kk, pkt)


def _compare_fixture(fx, X, P, rho, tol=1e-4):
  """X, P: final (N, 3) of member 0 (device), rho its painted (n, n, n)."""
  n = int(fx["n"])
  idx = torch.as_tensor(fx["idx"], device=X.device)
  q = np.stack(np.unravel_index(fx["idx"], (n, n, n)), axis=1)
  Xg = X[idx].double().cpu().numpy()
  Pg = P[idx].double().cpu().numpy()
  Rg = rho[torch.as_tensor(fx["planes"], device=rho.device)]
  Rg = Rg.double().cpu().numpy()
  ex = _rel(Xg - q, fx["X"] - q)
  ep = _rel(Pg, fx["P"])
  er = _rel(Rg - 1, fx["rho"].astype(np.float64) - 1)
  print("%d^3 vs float64 oracle: dX %.3g  P %.3g  delta %.3g" %
        (n, ex, ep, er), flush=True)
  assert ex < tol and ep < tol and er < tol, (ex, ep, er)


def test_c4_256_forward_against_float64_fixture(gpu, golden_dir):
  fx = np.load(golden_dir / "c4_256.npz")
  n, box, nsteps = int(fx["n"]), float(fx["box"]), int(fx["nsteps"])
  assert (n, box, nsteps, int(fx["order"])) == (256, 512.0, 40, 2)
  c = fa.cosmology.Planck15()
  white = _white(n, [int(fx["seed"])]).to(gpu)
  lin = fa.whitenoise_to_linear(n, box, white, _pk(c))
  st = fa.lpt_init(c, lin, 0.1, order=2)
  fin = fa.nbody(c, st, np.linspace(0.1, 1.0, nsteps + 1), n)
  rho = fa.cic_paint(torch.zeros(1, n, n, n, device=gpu), fin[0])
  _compare_fixture(fx, fin[0, 0], fin[1, 0], rho[0])


def _chi2_grad(c, lin, target, stages, n, every, order):
  lin = lin.detach().clone().requires_grad_(True)
  st = fa.lpt_init(c, lin, 0.1, order=order)
  fin = fa.nbody(c, st, stages, n, checkpoint_every=every)
  rho = fa.cic_paint(torch.zeros_like(target), fin[0])
  loss = ((rho - target)**2).sum()
  loss.backward()
  return lin.grad.detach().clone()


def test_c4_256_checkpoint8_gradient_equals_taped(gpu):
  """The chi^2 gradient w.r.t. the linear field with checkpoints every 8
  steps (recomputing each segment) equals the taped gradient.  Equal up to
  the order of the float atomics of the few late-time bricks that overflow
  their LDS box (measured 1e-6 .. 1.5e-6 over runs)."""
  n, box = 256, 512.0
  c = fa.cosmology.Planck15()
  pk = _pk(c)
  lin = fa.whitenoise_to_linear(n, box, _white(n, [100]).to(gpu), pk)
  lin_t = fa.whitenoise_to_linear(n, box, _white(n, [101]).to(gpu), pk)
  target = fa.cic_paint(torch.zeros(1, n, n, n, device=gpu),
                        fa.lpt_init(c, lin_t, 0.1, order=1)[0])
  stages = np.linspace(0.1, 1.0, 41)
  g8 = _chi2_grad(c, lin, target, stages, n, 8, 2)
  g0 = _chi2_grad(c, lin, target, stages, n, 0, 2)
  e = float((g8 - g0).norm() / g0.norm())
  print("256^3 x 40: checkpoint-8 vs taped gradient rel %.3g" % e)
  assert torch.isfinite(g0).all() and e < 1e-5


def test_c5_512_properties(gpu):
  n, box, nsteps = 512, 1024.0, 10
  c = fa.cosmology.Planck15()
  stages = np.linspace(0.1, 1.0, nsteps + 1)
  white = fa.white_noise(n, seed=512, type="real", device=gpu)
  lin = fa.whitenoise_to_linear(n, box, white, _pk(c))
  del white
  st0 = fa.lpt_init(c, lin, 0.1, order=1)
  zeros = torch.zeros(1, n, n, n, device=gpu)
  rho0 = fa.cic_paint(zeros, st0[0])
  # mass: every particle deposits weight 1
  m0 = float(rho0.double().sum())
  assert abs(m0 - n**3) < 1e-6 * n**3, m0
  # the force sums to zero at every step (odd stencil, adjoint paint /
  # readout pair)
  out = fa.nbody(c, st0, stages, n, return_intermediate_states=True)
  assert [a for a, _ in out] == list(stages[1:])
  for a, s in out:
    F = s[2, 0].double()
    r = float(F.sum(0).norm() / F.norm(dim=1).sum())
    assert r < 1e-5, (a, r)
  fin = out[-1][1]
  del out
  rho1 = fa.cic_paint(zeros, fin[0])
  m1 = float(rho1.double().sum())
  assert abs(m1 - n**3) < 1e-6 * n**3, m1
  # large scales grow linearly: P(k, a=1) / P(k, a=0.1) = (D1(1) / D1(0.1))^2
  # on the three lowest bins (k < 0.025 h/Mpc)
  kf = 2 * np.pi / box
  _, p0 = fa.pk.power_spectrum(rho0, kmin=kf / 2, dk=kf, boxsize=[box] * 3)
  _, p1 = fa.pk.power_spectrum(rho1, kmin=kf / 2, dk=kf, boxsize=[box] * 3)
  ratio = (p1[0, :3] / p0[0, :3]).double().cpu().numpy()
  growth = float(fa.tfbackground.D1(c, 1.0) / fa.tfbackground.D1(c, 0.1))**2
  print("512^3: P(k) growth on the lowest bins %s, D1^2 ratio %.4f" %
        (ratio, growth))
  assert np.all(np.abs(ratio / growth - 1) < 0.05), ratio / growth
  del rho0, rho1, fin
  # checkpoint-every-2 gradient = taped gradient (chi^2 against the
  # initial density, w.r.t. the linear field).  Not bit for bit here: at
  # 512^3 a few late-time bricks overflow their LDS box and spill to float
  # atomics, whose summation order differs between the recomputed and the
  # taped run; the chaotic trajectory amplifies those roundings (measured
  # 6e-6 .. 3e-5 over runs).
  target = _target(c, lin, zeros)
  g2 = _chi2_grad(c, lin, target, stages, n, 2, 1)
  g0 = _chi2_grad(c, lin, target, stages, n, 0, 1)
  e = float((g2 - g0).norm() / g0.norm())
  print("512^3 x 10: checkpoint-2 vs taped gradient rel %.3g" % e)
  assert torch.isfinite(g0).all() and e < 1e-4


def _target(c, lin, zeros):
  with torch.no_grad():
    return fa.cic_paint(zeros, fa.lpt_init(c, 1.5 * lin, 0.1, order=1)[0])


def test_c3_uneven_batch_against_classes_and_fixtures(gpu, golden_dir):
  c = fa.cosmology.Planck15()
  pk = _pk(c)
  stages = np.linspace(0.1, 1.0, 11)
  groups, fixtures = [], []
  for j, n in enumerate([32, 64, 96, 128]):
    fx = np.load(golden_dir / ("c3_%d.npz" % n))
    assert (int(fx["n"]), float(fx["box"]), int(fx["seed"])) == \
        (n, 100.0 * n / 32, 1000 + 8 * j)
    white = _white(n, [1000 + 8 * j + b for b in range(8)]).to(gpu)
    lin = fa.whitenoise_to_linear(n, 100.0 * n / 32, white, pk)
    groups.append(dict(cosmo=c, state=fa.lpt_init(c, lin, 0.1, order=1),
                       stages=stages, nc=n))
    fixtures.append(fx)
  runner = BatchedNbody(groups, graph=True)
  for rep in range(2):
    out = [o.clone() for o in runner.run()]
    torch.cuda.synchronize()
    for g, o, fx in zip(groups, out, fixtures):
      n = g["nc"]
      ref = fa.nbody(c, g["state"], stages, n)
      e = float((o - ref).norm() / ref.norm())
      assert e < 1e-6, (rep, n, e)
      rho = fa.cic_paint(torch.zeros(1, n, n, n, device=gpu), o[0, :1])
      _compare_fixture(fx, o[0, 0], o[1, 0], rho[0])
