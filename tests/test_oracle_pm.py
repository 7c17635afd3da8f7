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
  force               momentum conservation sum F = 0 (antisymmetric kernel);
                      pm_nc_factor = 2: linear growth of a small mode and the
                      first force equal to the linear-theory F (the 1/nbar
                      and 1/f factors)
  apply_longrange     split != 0: exp(-k^2 r^2) times the plane-wave closed
                      form
  whitenoise_to_linear / kmesh
                      sum of cosine modes on a rectangular mesh and box:
                      sqrt(P(|k|)/V) per mode, |k| = 2 pi |m / L|
  lpt_init            P = a^3 E dX/da and F = a^2 E dP/da by central
                      differences in a (orders 1 and 2); the 2LPT term equals
                      the physical second-order displacement
                      +(3/7) D1^2 Omega_m^-1/143 psi2 of a two-plane-wave field
                      (closed form), the 1LPT term D1 psi1
  nbody/lpt_init      one small long mode grows as D1(a)/D1(a0) (linear theory)
  autograd            float64 gradient vs central finite differences
Every one of these fails under a sign or coefficient change of the oracle
function it pins (dropped n/L or sqrt, DX2 sign, D1 for D2, f1 for f2, gf
for gf2, dropped 1/D1, 1/nbar or 1/f, r instead of r^2).
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


def test_whitenoise_to_linear_single_mode_rectangular_box():
  """A white field that is a sum of real cosine modes m on a rectangular
  mesh and box comes out as sqrt(P(|k|)/V) times each mode, with
  |k| = 2 pi sqrt(sum_d (m_d / L_d)^2) in h/Mpc (tfpm.py:15-63 with the
  noise injected, scripts/check_grads.py:39-65).  Pins kmesh (the n/L
  scaling per axis) and the sqrt(P/V) amplitude."""
  nc = [8, 12, 16]
  L = [50.0, 70.0, 90.0]
  V = L[0] * L[1] * L[2]
  pk = lambda k: 1e3 * np.exp(-3.0 * k) + 5.0 * k
  q = np.indices(nc).reshape(3, -1).T.astype(np.float64)
  modes = [(1.3, np.array([1, -2, 3])), (0.6, np.array([3, 0, -7])),
           (2.0, np.array([0, 4, 1]))]
  white = np.zeros(len(q))
  ref = np.zeros(len(q))
  for B, m in modes:
    ph = 2 * np.pi * (q @ (m / np.array(nc)))
    white += B * np.cos(ph)
    k = 2 * np.pi * np.sqrt(np.sum((m / np.array(L))**2))
    ref += np.sqrt(pk(k) / V) * B * np.cos(ph)
  lin = pm.whitenoise_to_linear(torch.tensor(white.reshape(1, *nc)), L, pk)
  np.testing.assert_allclose(lin[0].numpy().ravel(), ref, rtol=1e-12,
                             atol=1e-14)
  # the largest |k| of a mode equals the kmesh value at that mode
  km = pm.kmesh(nc, L)
  np.testing.assert_allclose(
      km[3, 0, 16 - 7], 2 * np.pi * np.sqrt((3 / 50.)**2 + (7 / 90.)**2),
      rtol=1e-14)


def _two_mode_linear(n, A=(0.04, 0.03)):
  q = np.indices([n] * 3).reshape(3, -1).T
  lin = A[0] * np.cos(2 * np.pi * q @ np.array([1, 2, 0]) / n) + \
      A[1] * np.cos(2 * np.pi * q @ np.array([0, 1, -2]) / n)
  return torch.tensor(lin.reshape(1, n, n, n))


def test_lpt_init_velocity_and_force_are_time_derivatives():
  """FastPM's momentum and force variables (tfpm.py:148-180):
  P = a^3 E(a) dX/da and F = a^2 E(a) dP/da.  Central differences of X and
  P in a (same linear field, lpt_init at a +- h) must reproduce the P and F
  lpt_init returns, for order 1 and for order 2 (pins the D1 / f1 / gf and
  D2 / f2 / gf2 coefficients of the assembly)."""
  g = bg.GrowthTable(bg.planck15())
  lin = _two_mode_linear(16, (0.4, 0.3))
  for order in (1, 2):
    for a in (0.1, 0.37):
      st = pm.lpt_init(g, lin, a, order=order)
      # h spans several cells of the 256-point ln-a table, so the
      # difference quotient is not the slope of one interpolation segment
      h = 0.04 * a
      sp = pm.lpt_init(g, lin, a + h, order=order)
      sm = pm.lpt_init(g, lin, a - h, order=order)
      dXda = (sp[0] - sm[0]) / (2 * h)
      dPda = (sp[1] - sm[1]) / (2 * h)
      E = float(g.E(a))
      P_ref = a**3 * E * dXda
      F_ref = a**2 * E * dPda
      sc = st[1].abs().max().item()
      np.testing.assert_allclose(st[1].numpy(), P_ref.numpy(), rtol=0,
                                 atol=2e-3 * sc)
      sc = st[2].abs().max().item()
      np.testing.assert_allclose(st[2].numpy(), F_ref.numpy(), rtol=0,
                                 atol=2e-3 * sc)


def test_lpt_init_second_order_displacement_closed_form():
  """The 2LPT term of lpt_init (tfpm.py:170-177) is the physical
  second-order displacement: X - Q - D1 psi1 = -(3/7) D1^2 Omega_m(a)^-1/143
  grad phi2 / (the same factor at a = 1), with nabla^2 phi2 =
  sum_{i<j} (phi_ii phi_jj - phi_ij^2) (Bouchet et al. 1995; the
  reference normalises D2 to 1 at a = 1 and puts the 3/7 in the source).
  For two plane waves A_m cos(k_m.q) phi2 is a closed form: the source is
  S0 cos(k1.q) cos(k2.q) with S0 = A1 A2 [|g1|^2 |g2|^2 - (g1.g2)^2] /
  (k1^2 k2^2), and each cos(K.q) moves the particles by
  -g(K)/|K|^2 sin(K.q) (the discrete kernels of kernels.py:28-78).  Pins the
  sign of DX2 and that it grows as D1^2 (D2, not D1) to the 1 % the
  Omega_m^-1/143 approximation allows."""
  g = bg.GrowthTable(bg.planck15())
  n = 16
  q = np.indices([n] * 3).reshape(3, -1).T
  modes = [(0.4, np.array([1, 2, 0])), (0.3, np.array([0, 1, -3]))]
  lin = sum(A * np.cos(2 * np.pi * q @ m / n) for A, m in modes)
  lin = torch.tensor(lin.reshape(1, n, n, n))

  def gk(kv):
    kw = np.where(kv > np.pi, kv - 2 * np.pi, kv)
    kw = np.where(kw <= -np.pi, kw + 2 * np.pi, kw)
    return (8 * np.sin(kw) - np.sin(2 * kw)) / 6, np.sum(kw**2)

  (A1, m1), (A2, m2) = modes
  k1, k2 = 2 * np.pi * m1 / n, 2 * np.pi * m2 / n
  g1, kk1 = gk(k1)
  g2, kk2 = gk(k2)
  S0 = A1 * A2 * ((g1 @ g1) * (g2 @ g2) - (g1 @ g2)**2) / (kk1 * kk2)
  psi2 = np.zeros((len(q), 3))
  for K in (k1 + k2, k1 - k2):  # cos a cos b = (cos(a+b) + cos(a-b)) / 2
    gK, kK = gk(K)
    psi2 += -(S0 / 2) * gK[None, :] / kK * np.sin(q @ K)[:, None]
  Q = pm.lagrangian_grid([n] * 3, 1, torch.float64)[0].numpy()
  om = lambda a: float(g.Omega_m_a(a))
  for a in (0.1, 0.5):
    x1 = pm.lpt_init(g, lin, a, order=1)[0][0].numpy()
    x2 = pm.lpt_init(g, lin, a, order=2)[0][0].numpy()
    d1 = float(g.D1(a))
    # x = q - D1 grad phi1 + D2 grad phi2 with D2 = -(3/7) D1^2 (...) and
    # grad phi2 = -psi2 (psi2 the closed-form lpt1 displacement of the
    # source): the second-order term is +(3/7) D1^2 (...) psi2
    ref = (3 / 7) * d1**2 * (om(a) / om(1.0))**(-1 / 143) * psi2
    np.testing.assert_allclose(x2 - x1, ref, rtol=0,
                               atol=1.2e-2 * np.abs(ref).max())
    assert np.abs(x2 - x1).max() > 0.3 * np.abs(ref).max()
  # the 1LPT part is D1 psi1 with psi1 the lpt1 closed form
  psi1 = sum(-A * gk(2 * np.pi * m / n)[0][None, :] /
             gk(2 * np.pi * m / n)[1] * np.sin(q @ (2 * np.pi * m / n))[:, None]
             for A, m in modes)
  x1 = pm.lpt_init(g, lin, 0.5, order=1)[0][0].numpy()
  np.testing.assert_allclose(x1 - Q, float(g.D1(0.5)) * psi1, atol=1e-12)


def test_force_pm_nc_factor_two_linear_growth():
  """pm_nc_factor = 2 (tfpm.py:342-356): positions scaled by f, density
  divided by nbar = f^-3, force divided by f.  A small fundamental mode must
  still grow as D1(a)/D1(a0) (any error in the 1/nbar or 1/f factors
  changes the force and the growth rate)."""
  c = bg.planck15()
  g = bg.GrowthTable(c)
  n = 16
  q = np.indices([n] * 3).reshape(3, -1).T
  lin = 1e-3 * np.cos(2 * np.pi * q[:, 1] / n).reshape(1, n, n, n)
  st = pm.lpt_init(g, torch.tensor(lin), 0.1, order=1)
  Q = pm.lagrangian_grid([n] * 3, 1, torch.float64)
  dx0 = (st[0] - Q)[0, :, 1]
  fin = pm.nbody(g, st, np.linspace(0.1, 1.0, 10), [n] * 3, pm_nc_factor=2)
  dx1 = (fin[0] - Q)[0, :, 1]
  ratio = (dx1 @ dx0) / (dx0 @ dx0)
  np.testing.assert_allclose(ratio.item(), g.D1(1.0) / g.D1(0.1), rtol=2e-2)
  # and the first force itself: F = a^2 E dP/da of the linear solution,
  # i.e. the lpt_init F, to the CIC / gradient transfer of the fine mesh
  F = pm.force(g, st, [n] * 3, pm_nc_factor=2)[2][0, :, 1]
  Fl = st[2][0, :, 1]
  np.testing.assert_allclose((F @ Fl / (Fl @ Fl)).item(), 1.0, rtol=2e-2)


def test_apply_longrange_split_plane_wave():
  """longrange_kernel with r_split != 0 (kernels.py:81-102): for a plane
  wave apply_longrange(split=r) is exp(-k^2 r^2) times the split = 0
  force, and that force is the closed form -A g(k)/k^2 sin(k.q) times the
  factor (tfpm.py:183-219)."""
  n, A, fac = 16, 0.5, 1.7
  q = np.indices([n, n, n]).reshape(3, -1).T
  pos = pm.lagrangian_grid([n, n, n], 1, torch.float64)
  for m in (np.array([1, -2, 3]), np.array([4, 0, 1])):
    kv = 2 * np.pi * m / n
    delta = A * np.cos(q @ kv).reshape(1, n, n, n)
    dk = pm.r2c3d(torch.tensor(delta))
    kw = np.where(kv > np.pi, kv - 2 * np.pi, kv)
    kk = np.sum(kw**2)
    g = (8 * np.sin(kw) - np.sin(2 * kw)) / 6
    base = fac * -A * g[None, :] / kk * np.sin(q @ kv)[:, None]
    for r in (0.0, 0.7, 1.9):
      f = pm.apply_longrange(pos, dk, split=r, factor=fac)[0].numpy()
      np.testing.assert_allclose(f, np.exp(-kk * r**2) * base, atol=1e-12)
