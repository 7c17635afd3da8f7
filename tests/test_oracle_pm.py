"""Pins oracle/pm.py (flowpm/utils.py, flowpm/kernels.py, flowpm/tfpm.py).

Each check is independent of the oracle's own formula:
  fftk / kernels      known values at n=4 (exact fftfreq, sin at pi/2, ...)
  r2c3d / c2r3d       numpy.fft.fftn (library), brute-force DFT sum at n=4,
                      float64 round trip rtol 1e-9 (tests/test_utils.py:51-61)
  cic_paint/readout   brute-force Python loops over the 8 corners
                      (tests/test_utils.py:12-48 procedure, 16^3, 100 parts),
                      mass conservation, delta deposit at integer positions,
                      periodic wrap, paint/readout adjoint identity
  lpt1                plane-wave closed form psi_d = -A g(k_d)/|k|^2 sin(k.q)
  lpt2_source         two-plane-wave closed form via the invariant
                      1/2[(tr Phi)^2 - tr(Phi^2)] (index bookkeeping pinned)
  force               momentum conservation sum F = 0 (antisymmetric kernel)
  nbody/lpt_init      one small long mode grows as D1(a)/D1(a0) (linear theory)
  autograd            float64 gradient vs central finite differences
"""
import itertools

import numpy as np
import torch

from oracle import background as bg
from oracle import pm


def test_kernels_known_values():
  k = pm.fftk([4, 4, 4], symmetric=False)
  np.testing.assert_allclose(k[0].ravel(), [0, np.pi / 2, -np.pi, -np.pi / 2])
  assert k[0].shape == (4, 1, 1) and k[2].shape == (1, 1, 4)
  ks = pm.fftk([4, 4, 4], symmetric=True)
  assert ks[2].shape == (1, 1, 3)
  g = pm.gradient_kernel(k, 0)
  np.testing.assert_allclose(g[1, 0, 0], 1j * 8 / 6.0, rtol=1e-15)
  assert abs(g[2, 0, 0]) < 1e-15  # Nyquist: -2e-16, not exactly 0
  lap = pm.laplace_kernel(k)
  assert lap[0, 0, 0] == 1.0  # DC: kk forced to 1 before masking
  np.testing.assert_allclose(lap[1, 1, 0], 1 / (2 * (np.pi / 2)**2))
  assert pm.longrange_kernel(k, 0) == 1.0


def test_fft_against_numpy_and_bruteforce():
  rng = np.random.default_rng(0)
  x = rng.standard_normal((2, 4, 6, 8))
  c = pm.r2c3d(torch.tensor(x)).numpy()
  np.testing.assert_allclose(c, np.fft.fftn(x, axes=(1, 2, 3)) / (4 * 6 * 8),
                             atol=1e-14)
  # brute-force DFT of one mode
  kx, ky, kz = 1, 5, 3
  ref = 0
  for i, j, l in itertools.product(range(4), range(6), range(8)):
    ref += x[1, i, j, l] * np.exp(-2j * np.pi * (kx * i / 4 + ky * j / 6 +
                                                  kz * l / 8))
  np.testing.assert_allclose(c[1, kx, ky, kz], ref / 192, atol=1e-14)
  r = pm.c2r3d(pm.r2c3d(torch.tensor(x))).numpy()
  np.testing.assert_allclose(r, x, rtol=1e-9, atol=1e-12)


def _brute_paint(nc, pos, w):
  out = np.zeros(nc)
  for p in range(len(pos)):
    f = np.floor(pos[p])
    for o in itertools.product([0, 1], repeat=3):
      n = f + np.array(o)
      wt = np.prod(1 - np.abs(pos[p] - n)) * w[p]
      out[tuple(np.mod(n.astype(int), nc))] += wt
  return out


def test_cic_paint_readout_bruteforce():
  rng = np.random.default_rng(1)
  nc = [16, 16, 16]
  pos = rng.random((100, 3)) * 16
  pos[:5] -= 16.3  # negative coordinates wrap
  pos[5:10] += 16  # beyond the box wrap
  w = rng.random(100)
  mesh = pm.cic_paint(torch.zeros(1, *nc, dtype=torch.float64),
                      torch.tensor(pos[None]), torch.tensor(w[None]))
  np.testing.assert_allclose(mesh[0].numpy(), _brute_paint(nc, pos, w),
                             atol=1e-12)
  np.testing.assert_allclose(mesh.sum().item(), w.sum(), rtol=1e-12)
  base = rng.random(nc) * 100
  val = pm.cic_readout(torch.tensor(base[None]), torch.tensor(pos[None]))
  ref = []
  for p in range(len(pos)):
    f = np.floor(pos[p])
    s = 0
    for o in itertools.product([0, 1], repeat=3):
      n = f + np.array(o)
      s += base[tuple(np.mod(n.astype(int), nc))] * np.prod(
          1 - np.abs(pos[p] - n))
    ref.append(s)
  np.testing.assert_allclose(val[0].numpy(), ref, rtol=1e-12)


def test_cic_integer_positions_and_adjoint():
  nc = [8, 8, 8]
  q = pm.lagrangian_grid(nc, 1, torch.float64)
  m = torch.rand(1, *nc, dtype=torch.float64)
  # readout at integer grid positions is exact sampling
  np.testing.assert_allclose(pm.cic_readout(m, q)[0].numpy(),
                             m.reshape(-1).numpy(), rtol=1e-15)
  # a uniform grid paints exactly 1 per cell
  np.testing.assert_allclose(
      pm.cic_paint(torch.zeros(1, *nc, dtype=torch.float64), q).numpy(), 1.0)
  x = torch.rand(1, 300, 3, dtype=torch.float64) * 8
  w = torch.rand(1, 300, dtype=torch.float64)
  lhs = (pm.cic_paint(torch.zeros_like(m), x, w) * m).sum()
  rhs = (w * pm.cic_readout(m, x)).sum()
  np.testing.assert_allclose(lhs.item(), rhs.item(), rtol=1e-13)


def test_lpt1_plane_wave():
  n = 16
  A = 0.7
  m = np.array([2, -3, 1])
  q = np.indices([n, n, n]).reshape(3, -1).T
  kv = 2 * np.pi * m / n
  delta = A * np.cos(q @ kv).reshape(n, n, n)
  dk = pm.r2c3d(torch.tensor(delta[None]))
  pos = pm.lagrangian_grid([n, n, n], 1, torch.float64)
  psi = pm.lpt1(dk, pos)[0].numpy()
  # wrap k into the fftfreq range used by the kernel
  kw = np.where(kv > np.pi, kv - 2 * np.pi, kv)
  g = (8 * np.sin(kw) - np.sin(2 * kw)) / 6
  ref = -A * g[None, :] / np.sum(kw**2) * np.sin(q @ kv)[:, None]
  np.testing.assert_allclose(psi, ref, atol=1e-12)


def test_lpt2_source_two_plane_waves():
  n = 16
  q = np.indices([n, n, n]).reshape(3, -1).T
  modes = [(0.5, np.array([1, 2, 0])), (0.3, np.array([0, 1, -3]))]
  delta = sum(A * np.cos(q @ (2 * np.pi * m / n)) for A, m in modes)
  dk = pm.r2c3d(torch.tensor(delta.reshape(1, n, n, n)))
  src = pm.c2r3d(pm.lpt2_source(dk))[0].numpy().ravel()
  # Phi_ij = c2r(dk lap (i g_i)(i g_j)) = -sum A g_i g_j / k^2 cos(k.q)
  phi = np.zeros((len(q), 3, 3))
  for A, m in modes:
    kv = 2 * np.pi * m / n
    kw = np.where(kv > np.pi, kv - 2 * np.pi, kv)
    g = (8 * np.sin(kw) - np.sin(2 * kw)) / 6
    phi += -A * (np.outer(g, g) / np.sum(kw**2))[None] * np.cos(
        q @ kv)[:, None, None]
  tr = np.trace(phi, axis1=1, axis2=2)
  tr2 = np.einsum("pij,pji->p", phi, phi)
  ref = 3.0 / 7.0 * 0.5 * (tr**2 - tr2)
  np.testing.assert_allclose(src, ref, atol=1e-12)


def test_force_momentum_conservation():
  g = bg.GrowthTable(bg.planck15())
  torch.manual_seed(3)
  n = 16
  x = pm.lagrangian_grid([n] * 3, 2, torch.float64)
  x = x + 1.5 * torch.randn_like(x)
  st = torch.stack([x, torch.zeros_like(x), torch.zeros_like(x)])
  F = pm.force(g, st, [n] * 3)[2]
  assert F.abs().max() > 1e-2
  np.testing.assert_allclose(F.sum(1).numpy(), 0.0, atol=1e-10)
  F2 = pm.force(g, st, [n] * 3, pm_nc_factor=2)[2]
  np.testing.assert_allclose(F2.sum(1).numpy(), 0.0, atol=1e-10)


def test_single_mode_linear_growth():
  """A small-amplitude fundamental mode must grow as D1(a)/D1(a0) through
  lpt_init + nbody.  CIC and the 4th-order gradient weaken the force on the
  fundamental of a 32^3 mesh by < 1 %, so the check is 2 %."""
  c = bg.planck15()
  g = bg.GrowthTable(c)
  n = 32
  q = np.indices([n] * 3).reshape(3, -1).T
  A = 1e-3
  lin = A * np.cos(2 * np.pi * q[:, 0] / n).reshape(1, n, n, n)
  st = pm.lpt_init(g, torch.tensor(lin), 0.1, order=1)
  dx0 = (st[0] - pm.lagrangian_grid([n] * 3, 1, torch.float64))[0, :, 0]
  stages = np.linspace(0.1, 1.0, 10)
  fin = pm.nbody(g, st, stages, [n] * 3)
  dx1 = (fin[0] - pm.lagrangian_grid([n] * 3, 1, torch.float64))[0, :, 0]
  ratio = (dx1 @ dx0) / (dx0 @ dx0)
  np.testing.assert_allclose(ratio.item(), g.D1(1.0) / g.D1(0.1), rtol=2e-2)
  # velocity: P = a^2 f1 E dX, so P/dX at a=1 is f1(1) E(1)
  p1 = fin[1][0, :, 0]
  np.testing.assert_allclose((p1 @ dx1 / (dx1 @ dx1)).item(),
                             g.f1(1.0) * g.E(1.0), rtol=3e-2)


def test_oracle_gradient_vs_finite_differences():
  c = bg.planck15()
  g = bg.GrowthTable(c)
  n = 8
  torch.manual_seed(0)
  white = torch.randn(1, n, n, n, dtype=torch.float64) * n**1.5
  pk = lambda k: 2e3 * np.exp(-k)  # smooth positive spectrum
  target = torch.rand(1, n, n, n, dtype=torch.float64)
  stages = np.linspace(0.1, 1.0, 4)

  def loss(w):
    lin = pm.whitenoise_to_linear(w, 50.0, pk)
    st = pm.lpt_init(g, lin, 0.1, order=2)
    st = pm.nbody(g, st, stages, [n] * 3)
    rho = pm.cic_paint(torch.zeros_like(lin), st[0])
    return ((rho - target)**2).sum()

  w = white.clone().requires_grad_(True)
  L = loss(w)
  (gr,) = torch.autograd.grad(L, w)
  rng = torch.Generator().manual_seed(5)
  for _ in range(3):
    v = torch.randn(white.shape, generator=rng, dtype=torch.float64)
    eps = 1e-4
    fd = (loss(white + eps * v) - loss(white - eps * v)) / (2 * eps)
    np.testing.assert_allclose((gr * v).sum().item(), fd.item(), rtol=2e-5)


def test_pgd_plane_wave_closed_form():
  """apply_pgd on delta = A cos(k.q) read out at grid points: the
  displacement is alpha * (-A g(k) / k^2) * PGD(k) * sin(k.q), PGD(k) =
  exp(-kl^2/k^2) exp(-k^4/ks^4) (kernels.py:105-134, tfpm.py:222-270)."""
  n, A, alpha, kl, ks = 16, 0.4, 0.7, 0.3, 1.7
  m = np.array([1, -2, 3])
  q = np.indices([n, n, n]).reshape(3, -1).T
  kv = 2 * np.pi * m / n
  delta = A * np.cos(q @ kv).reshape(n, n, n)
  dk = pm.r2c3d(torch.tensor(delta[None]))
  pos = pm.lagrangian_grid([n, n, n], 1, torch.float64)
  s = pm.apply_pgd(pos, dk, alpha, kl, ks)[0].numpy()
  kw = np.where(kv > np.pi, kv - 2 * np.pi, kv)
  g = (8 * np.sin(kw) - np.sin(2 * kw)) / 6
  kk = np.sum(kw**2)
  w = np.exp(-kl**2 / kk) * np.exp(-kk**2 / ks**4)
  ref = alpha * -A * g[None, :] / kk * w * np.sin(q @ kv)[:, None]
  np.testing.assert_allclose(s, ref, atol=1e-12)


def test_pgd_limits_and_momentum():
  """kl -> 0, ks -> inf: the PGD kernel is 1 and apply_pgd equals
  apply_longrange (same paint, FFT and readout); the displacement of a
  displaced grid sums to zero over the particles (odd kernel, adjoint CIC)."""
  n = 16
  gen = torch.Generator().manual_seed(3)
  q = pm.lagrangian_grid([n] * 3, 1, torch.float64)
  x = q + 0.8 * torch.randn(q.shape, generator=gen, dtype=torch.float64)
  rho = pm.cic_paint(torch.zeros(1, n, n, n, dtype=torch.float64), x)
  dk = pm.r2c3d(rho, norm=n**3)
  a = pm.apply_pgd(x, dk, 0.9, 1e-9, 1e9)
  b = pm.apply_longrange(x, dk, factor=0.9)
  np.testing.assert_allclose(a.numpy(), b.numpy(), rtol=1e-12, atol=1e-14)
  st = torch.stack([x, torch.zeros_like(x), torch.zeros_like(x)])
  out = pm.pgd(st, [0.5, 0.4, 2.0], [n] * 3)
  shift = (out[0] - x)[0]
  assert shift.abs().max() > 1e-3
  assert shift.sum(0).abs().max() < 1e-10 * shift.abs().sum()


def test_gradient_kernel_is_fourth_order_central_difference():
  """gradient_kernel (kernels.py:51-78) multiplies mode k by
  i (8 sin w - sin 2w) / 6; under numpy's forward transform that is the
  transfer function of the periodic stencil
  2/3 (x[n+1] - x[n-1]) - 1/12 (x[n+2] - x[n-2]), which the GPU time loop
  applies in real space."""
  rng = np.random.default_rng(0)
  for n in ([16, 12, 20], [8, 8, 8], [32, 24, 16]):
    phi = rng.normal(size=n)
    kvec = pm.fftk(n, symmetric=False)
    for d in range(3):
      g = pm.gradient_kernel(kvec, d)
      ref = np.real(np.fft.ifftn(np.fft.fftn(phi) * g))
      st = (2 / 3) * (np.roll(phi, -1, d) - np.roll(phi, 1, d)) - (1 / 12) * (
          np.roll(phi, -2, d) - np.roll(phi, 2, d))
      np.testing.assert_allclose(st, ref, atol=1e-13 * np.abs(ref).max())
