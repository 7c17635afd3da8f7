"""GPU parity of the HIP path against the CPU oracle (float64 unless noted).

Tolerances: FFTs and CIC at float32 rounding (relative RMS 1e-6 .. 1e-5);
force / LPT / nbody final positions, velocities and painted density at the
1e-4 relative RMS that the PM-step parity bar states; gradients by cosine
>= 0.999 against the float64 oracle's autograd.
"""
import numpy as np
import pytest
import torch

import flowpm_amd as fa
from oracle import background as obg
from oracle import pm as opm

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
  a = a.detach().cpu().to(torch.complex128 if a.is_complex() else
                          torch.float64)
  b = b.detach().cpu().to(a.dtype)
  return float(((a - b).abs()**2).mean().sqrt() / (b.abs()**2).mean().sqrt())


def close(a, b, tol, what=""):
  e = rel_rms(a, b)
  assert e < tol, "%s rel RMS %.3g >= %.3g" % (what, e, tol)
  return e


def cosine(a, b):
  a = a.detach().double().cpu().reshape(-1)
  b = b.detach().double().cpu().reshape(-1)
  return float(a @ b / (a.norm() * b.norm()))


@pytest.mark.parametrize("shape", [(1, 16, 16, 16), (2, 32, 32, 32),
                                   (1, 16, 16, 48), (1, 96, 96, 96),
                                   (1, 24, 48, 12), (1, 128, 64, 256),
                                   (1, 256, 256, 8), (2, 512, 16, 8),
                                   (1, 80, 100, 60), (1, 20, 40, 160),
                                   (2, 320, 120, 40), (1, 960, 8, 200),
                                   (1, 800, 24, 240), (1, 400, 480, 20),
                                   (1, 640, 12, 100)])
def test_r2c_c2r_against_numpy(gpu, shape):
  """r2c3d / c2r3d against numpy.fft on every radix (16, 8, 4, 2, 3, 5:
  100 = 4 * 5 * 5, 960 = 16 * 4 * 3 * 5, 800 = 16 * 2 * 5 * 5) and both
  column-tile widths (lines above 512 use 8-column tiles)."""
  rng = np.random.default_rng(sum(shape))
  x = rng.standard_normal(shape)
  M = np.prod(shape[1:])
  c = fa.r2c3d(torch.tensor(x, dtype=torch.float32, device=gpu))
  ref = np.fft.fftn(x, axes=(1, 2, 3)) / M
  assert c.dtype == torch.complex64 and tuple(c.shape) == shape
  err = np.sqrt(np.mean(np.abs(c.cpu().numpy() - ref)**2) /
                np.mean(np.abs(ref)**2))
  assert err < 2e-6, err
  back = fa.c2r3d(c)
  close(back, torch.tensor(x), 2e-6)
  # c2r of a non-Hermitian spectrum = real part of the full inverse
  z = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
  r = fa.c2r3d(torch.tensor(z, dtype=torch.complex64, device=gpu), norm=7.0)
  refr = np.fft.ifftn(z, axes=(1, 2, 3)).real * 7.0
  close(r, torch.tensor(refr), 2e-6)


@pytest.mark.parametrize("tiles", [(8, 8), (8, 16), (16, 8), (16, 16)])
def test_fft_every_column_tile_width(gpu, tiles):
  """The y and x passes on 8- and 16-column tiles (8-column tile rows are
  64 bytes, half a cache line, and neighbouring tiles run on one XCD)
  against numpy.fft, on a batch of two rectangular meshes with a Nyquist
  plane that is not a multiple of the tile width."""
  from flowpm_amd import _plans
  shape = (2, 48, 40, 24)
  rng = np.random.default_rng(5)
  x = rng.standard_normal(shape)
  plan = _plans.get(gpu, *shape)
  fa._lib.call("pmx_plan_set_tiles", plan.handle, *tiles)
  try:
    c = fa.r2c3d(torch.tensor(x, dtype=torch.float32, device=gpu))
    ref = np.fft.fftn(x, axes=(1, 2, 3)) / np.prod(shape[1:])
    err = np.sqrt(np.mean(np.abs(c.cpu().numpy() - ref)**2) /
                  np.mean(np.abs(ref)**2))
    assert err < 2e-6, err
    close(fa.c2r3d(c), torch.tensor(x), 2e-6)
  finally:
    fa._lib.call("pmx_plan_set_tiles", plan.handle, 16, 16)


def test_cic_against_oracle(gpu):
  rng = np.random.default_rng(0)
  nc = [16, 16, 16]
  pos = rng.random((2, 5000, 3)) * 16
  pos[:, :50] -= 17.5  # wrap through negative coordinates
  w = rng.random((2, 5000))
  base = rng.random([2] + nc) * 100
  zero = torch.zeros([2] + nc, dtype=torch.float64)
  ref = opm.cic_paint(zero, torch.tensor(pos), torch.tensor(w))
  out = fa.cic_paint(zero.float().to(gpu), torch.tensor(pos).float().to(gpu),
                     torch.tensor(w).float().to(gpu))
  e = rel_rms(out, ref)
  assert e < 1e-6, e
  # onto a non-zero mesh: float32 sum of base (~100) and increments (~1)
  mesh0 = torch.tensor(base)
  ref = opm.cic_paint(mesh0, torch.tensor(pos), torch.tensor(w))
  out = fa.cic_paint(mesh0.float().to(gpu), torch.tensor(pos).float().to(gpu),
                     torch.tensor(w).float().to(gpu))
  e = rel_rms(out, ref)
  assert e < 1e-7, e
  refv = opm.cic_readout(torch.tensor(base), torch.tensor(pos))
  val = fa.cic_readout(torch.tensor(base).float().to(gpu),
                       torch.tensor(pos).float().to(gpu))
  e = rel_rms(val, refv)
  assert e < 1e-6, e
  # mass conservation
  m = fa.cic_paint(torch.zeros(2, *nc, device=gpu),
                   torch.tensor(pos).float().to(gpu))
  np.testing.assert_allclose(m.sum((1, 2, 3)).cpu().numpy(), 5000, rtol=1e-5)


def test_cic_adjoints_against_oracle_autograd(gpu):
  rng = np.random.default_rng(1)
  nc = [8, 16, 12]
  pos = torch.tensor(rng.random((1, 700, 3)) * np.array(nc) * 1.3 - 2)
  w = torch.tensor(rng.random((1, 700)))
  gm = torch.tensor(rng.standard_normal([1] + nc))
  gv = torch.tensor(rng.standard_normal((1, 700)))
  mesh = torch.tensor(rng.standard_normal([1] + nc))
  # oracle
  p_o = pos.clone().requires_grad_(True)
  w_o = w.clone().requires_grad_(True)
  m_o = mesh.clone().requires_grad_(True)
  L = (opm.cic_paint(m_o, p_o, w_o) * gm).sum() + \
      (opm.cic_readout(m_o, p_o) * gv).sum()
  L.backward()
  # HIP
  p_g = pos.float().to(gpu).requires_grad_(True)
  w_g = w.float().to(gpu).requires_grad_(True)
  m_g = mesh.float().to(gpu).requires_grad_(True)
  Lg = (fa.cic_paint(m_g, p_g, w_g) * gm.float().to(gpu)).sum() + \
      (fa.cic_readout(m_g, p_g) * gv.float().to(gpu)).sum()
  Lg.backward()
  close(p_g.grad, p_o.grad, 1e-5)
  close(w_g.grad, w_o.grad, 1e-6)
  close(m_g.grad, m_o.grad, 1e-6)


def test_fft_autograd_against_oracle(gpu):
  rng = np.random.default_rng(2)
  x = torch.tensor(rng.standard_normal((1, 16, 8, 24)))
  gc = torch.tensor(rng.standard_normal((1, 16, 8, 24)) +
                    1j * rng.standard_normal((1, 16, 8, 24)))
  xo = x.clone().requires_grad_(True)
  Lo = (opm.r2c3d(xo, norm=3.0) * gc.conj()).real.sum()
  Lo = Lo + (opm.c2r3d(opm.r2c3d(xo)) * x).sum()
  Lo.backward()
  xg = x.float().to(gpu).requires_grad_(True)
  Lg = (fa.r2c3d(xg, norm=3.0) * gc.to(torch.complex64).to(gpu).conj()
        ).real.sum()
  Lg = Lg + (fa.c2r3d(fa.r2c3d(xg)) * x.float().to(gpu)).sum()
  Lg.backward()
  close(xg.grad, xo.grad, 2e-6)


def _displaced_state(n, B=1, amp=1.5, seed=3, dtype=torch.float64):
  gen = torch.Generator().manual_seed(seed)
  x = opm.lagrangian_grid([n] * 3, B, dtype)
  x = x + amp * torch.randn(x.shape, generator=gen, dtype=dtype)
  p = torch.randn(x.shape, generator=gen, dtype=dtype)
  return torch.stack([x, p, torch.zeros_like(x)])


@pytest.mark.parametrize("n,f,B,amp", [(32, 1, 1, 1.5), (16, 2, 2, 1.5),
                                       (64, 1, 1, 1.5), (32, 1, 2, 9.0),
                                       (24, 1, 1, 40.0), (128, 2, 1, 3.0)])
def test_force_against_oracle(gpu, n, f, B, amp):
  """Small displacements use the LDS brick paint; amp 9 makes some bricks
  overflow to device atomics; amp 40 wraps particles many boxes away;
  128^3 particles on a 256^3 mesh runs the 512-thread column passes."""
  g = obg.GrowthTable(obg.planck15())
  st = _displaced_state(n, B, amp=amp)
  ref = opm.force(g, st, [n] * 3, pm_nc_factor=f)[2]
  out = fa.tfpm.force(fa.cosmology.Planck15(), st.float().to(gpu), n,
                      pm_nc_factor=f)
  close(out[2], ref, 1e-4)
  assert float((out[0].double().cpu() - st[0].float().double()).abs().max()) == 0.0


def test_kick_drift(gpu):
  c = fa.cosmology.Planck15()
  g = obg.GrowthTable(obg.planck15())
  st = _displaced_state(8)
  st[2] = torch.randn_like(st[2])
  k = fa.tfpm.kick(c, st.float().to(gpu), 0.1, 0.12, 0.15)
  close(k, opm.kick(g, st, 0.1, 0.12, 0.15), 1e-6)
  d = fa.tfpm.drift(c, st.float().to(gpu), 0.1, 0.12, 0.15)
  close(d, opm.drift(g, st, 0.1, 0.12, 0.15), 1e-6)


def _ics(n, L, seed, B=1):
  gen = torch.Generator().manual_seed(seed)
  white = torch.randn(B, n, n, n, generator=gen, dtype=torch.float64) * n**1.5
  from oracle import power as opw
  k = np.logspace(-4, 1, 256)
  pkt = opw.linear_matter_power(obg.planck15(), k)
  pk = lambda x: opw.interp_tf(x.reshape(-1), k, pkt).reshape(x.shape)
  return white, pk


@pytest.mark.parametrize("order", [1, 2])
def test_linear_field_and_lpt_init(gpu, order):
  n, L = 32, 100.0
  white, pk = _ics(n, L, 100)
  g = obg.GrowthTable(obg.planck15())
  lin_o = opm.whitenoise_to_linear(white, L, pk)
  lin_g = fa.whitenoise_to_linear(n, L, white.float().to(gpu), pk)
  close(lin_g, lin_o, 2e-6)
  st_o = opm.lpt_init(g, lin_o, 0.1, order=order)
  st_g = fa.lpt_init(fa.cosmology.Planck15(), lin_g, 0.1, order=order)
  q = opm.lagrangian_grid([n] * 3, 1, torch.float64)
  close(st_g[0] - q.float().to(gpu), st_o[0] - q, 1e-5)
  close(st_g[1], st_o[1], 1e-5)
  close(st_g[2], st_o[2], 1e-5)


def test_lpt_ops_dropins(gpu):
  n = 16
  white, pk = _ics(n, 50.0, 7)
  lin = opm.whitenoise_to_linear(white, 50.0, pk)
  dk_o = opm.r2c3d(lin)
  dk_g = fa.r2c3d(lin.float().to(gpu))
  q = opm.lagrangian_grid([n] * 3, 1, torch.float64) + 0.3
  close(fa.tfpm.lpt1(dk_g, q.float().to(gpu)), opm.lpt1(dk_o, q), 1e-5)
  close(fa.tfpm.lpt2_source(dk_g), opm.lpt2_source(dk_o), 1e-5)
  f_o = opm.apply_longrange(q, dk_o, factor=0.4)
  f_g = fa.tfpm.apply_longrange(q.float().to(gpu), dk_g, factor=0.4)
  close(f_g, f_o, 1e-5)
  # the long-range filter exp(-k^2 split^2) (kernels.py:81-102)
  f_o = opm.apply_longrange(q, dk_o, split=0.7, factor=0.4)
  f_g = fa.tfpm.apply_longrange(q.float().to(gpu), dk_g, split=0.7,
                                factor=0.4)
  close(f_g, f_o, 1e-5)


def test_nbody_readme_workflow_32(gpu):
  """The reference README workflow at 32^3, box 100, 1LPT at 0.1,
  stages linspace(0.1, 1, 10): final X, P and painted density."""
  n, L = 32, 100.0
  white, pk = _ics(n, L, 100)
  g = obg.GrowthTable(obg.planck15())
  stages = np.linspace(0.1, 1.0, 10)
  lin = opm.whitenoise_to_linear(white, L, pk)
  st_o = opm.nbody(g, opm.lpt_init(g, lin, 0.1, order=1), stages, [n] * 3)
  c = fa.cosmology.Planck15()
  st_g = fa.nbody(c, fa.lpt_init(c, lin.float().to(gpu), 0.1, order=1),
                  stages, n)
  q = opm.lagrangian_grid([n] * 3, 1, torch.float64)
  ex = rel_rms(st_g[0] - q.float().to(gpu), st_o[0] - q)
  ep = rel_rms(st_g[1], st_o[1])
  rho_o = opm.cic_paint(torch.zeros(1, n, n, n, dtype=torch.float64), st_o[0])
  rho_g = fa.cic_paint(torch.zeros(1, n, n, n, device=gpu), st_g[0])
  er = rel_rms(rho_g - 1, rho_o - 1)
  print("32^3 nbody rel RMS: dX %.3g P %.3g delta %.3g" % (ex, ep, er))
  assert ex < 1e-4 and ep < 1e-4 and er < 1e-4


def test_nbody_64_batch2_against_oracle(gpu):
  """nbody at 64^3 with two batch members: the fused readout visits 4 x 64
  particle tiles here (grid sides divisible by 64), the brick paint and the
  batched FFT all run on B = 2."""
  n, L = 64, 200.0
  white, pk = _ics(n, L, 300, B=2)
  g = obg.GrowthTable(obg.planck15())
  stages = np.linspace(0.1, 1.0, 6)
  lin = opm.whitenoise_to_linear(white, L, pk)
  st_o = opm.nbody(g, opm.lpt_init(g, lin, 0.1, order=2), stages, [n] * 3)
  c = fa.cosmology.Planck15()
  st_g = fa.nbody(c, fa.lpt_init(c, lin.float().to(gpu), 0.1, order=2),
                  stages, n)
  q = opm.lagrangian_grid([n] * 3, 2, torch.float64)
  ex = rel_rms(st_g[0] - q.float().to(gpu), st_o[0] - q)
  ep = rel_rms(st_g[1], st_o[1])
  ef = rel_rms(st_g[2], st_o[2])
  print("64^3 B=2 nbody rel RMS: dX %.3g P %.3g F %.3g" % (ex, ep, ef))
  assert ex < 1e-4 and ep < 1e-4 and ef < 1e-4


@pytest.mark.parametrize("nc,box", [([16, 16, 48], [50.0, 50.0, 150.0]),
                                    ([32, 64, 128], [100.0, 200.0, 400.0]),
                                    ([20, 40, 60], [62.5, 125.0, 187.5]),
                                    ([80, 100, 40], [160.0, 200.0, 80.0])])
def test_nbody_rectangular_against_oracle(gpu, nc, box):
  """End-to-end nbody on rectangular meshes (tests/test_tfpm.py:186-212 of
  the reference: [nc, nc, 3 nc], box [bs, bs, 3 bs], 1LPT at 0.1, stages
  linspace(0.1, 1, 10)); the second shape runs the tiled particle passes
  (32-row slabs) on a non-cubic grid, the last two sides with radix-5 line
  transforms (20, 40, 60, 80, 100) and bricks cut short at the grid's
  edge."""
  gen = torch.Generator().manual_seed(100)
  white = torch.randn([1] + nc, generator=gen, dtype=torch.float64) * (
      nc[0] * nc[1] * nc[2])**0.5
  _, pk = _ics(8, 1.0, 0)
  g = obg.GrowthTable(obg.planck15())
  stages = np.linspace(0.1, 1.0, 10)
  lin = opm.whitenoise_to_linear(white, box, pk)
  st_o = opm.nbody(g, opm.lpt_init(g, lin, 0.1, order=1), stages, nc)
  c = fa.cosmology.Planck15()
  lin_g = fa.whitenoise_to_linear(nc, box, white.float().to(gpu), pk)
  close(lin_g, lin, 2e-6)
  st_g = fa.nbody(c, fa.lpt_init(c, lin_g, 0.1, order=1), stages, nc)
  q = opm.lagrangian_grid(nc, 1, torch.float64)
  ex = rel_rms(st_g[0] - q.float().to(gpu), st_o[0] - q)
  ep = rel_rms(st_g[1], st_o[1])
  rho_o = opm.cic_paint(torch.zeros([1] + nc, dtype=torch.float64), st_o[0])
  rho_g = fa.cic_paint(torch.zeros([1] + nc, device=gpu), st_g[0])
  er = rel_rms(rho_g - 1, rho_o - 1)
  print("rectangular %s nbody rel RMS: dX %.3g P %.3g delta %.3g" %
        (nc, ex, ep, er))
  assert ex < 1e-4 and ep < 1e-4 and er < 1e-4


@pytest.mark.parametrize("n,L", [(32, 100.0), (20, 62.5)])
def test_nbody_pm_nc_factor2_against_oracle(gpu, n, L):
  """End-to-end nbody with the force mesh twice the particle grid
  (tests/test_tfpm.py:158-183 of the reference: pm_nc_factor = 2): 32^3
  particles on a 64^3 force mesh, and 20^3 on a 40^3 mesh (radix-5 line
  transforms, bricks of 4 particles along x)."""
  white, pk = _ics(n, L, 100)
  g = obg.GrowthTable(obg.planck15())
  stages = np.linspace(0.1, 1.0, 10)
  lin = opm.whitenoise_to_linear(white, L, pk)
  st_o = opm.nbody(g, opm.lpt_init(g, lin, 0.1, order=1), stages, [n] * 3,
                   pm_nc_factor=2)
  c = fa.cosmology.Planck15()
  st_g = fa.nbody(c, fa.lpt_init(c, lin.float().to(gpu), 0.1, order=1),
                  stages, n, pm_nc_factor=2)
  q = opm.lagrangian_grid([n] * 3, 1, torch.float64)
  ex = rel_rms(st_g[0] - q.float().to(gpu), st_o[0] - q)
  ep = rel_rms(st_g[1], st_o[1])
  ef = rel_rms(st_g[2], st_o[2])
  print("pm_nc_factor 2 nbody rel RMS: dX %.3g P %.3g F %.3g" % (ex, ep, ef))
  assert ex < 1e-4 and ep < 1e-4 and ef < 1e-4


def test_nbody_gradient_against_oracle(gpu):
  n, L = 16, 64.0
  white, pk = _ics(n, L, 11)
  tgt = torch.rand(1, n, n, n, dtype=torch.float64,
                   generator=torch.Generator().manual_seed(12)) * 2
  g = obg.GrowthTable(obg.planck15())
  stages = np.linspace(0.1, 1.0, 6)
  wo = white.clone().requires_grad_(True)
  lin = opm.whitenoise_to_linear(wo, L, pk)
  st = opm.nbody(g, opm.lpt_init(g, lin, 0.1, order=2), stages, [n] * 3)
  rho = opm.cic_paint(torch.zeros_like(lin), st[0])
  ((rho - tgt)**2).sum().backward()
  c = fa.cosmology.Planck15()
  for every in (0, 1, 2, 8, None):
    wg = white.float().to(gpu).requires_grad_(True)
    ling = fa.whitenoise_to_linear(n, L, wg, pk)
    stg = fa.nbody(c, fa.lpt_init(c, ling, 0.1, order=2), stages, n,
                   checkpoint_every=every)
    rhog = fa.cic_paint(torch.zeros_like(ling), stg[0])
    ((rhog - tgt.float().to(gpu))**2).sum().backward()
    cs = cosine(wg.grad, wo.grad)
    print("gradient cosine (every=%s): %.7f, rel %.3g" %
          (every, cs, rel_rms(wg.grad, wo.grad)))
    assert cs > 0.999


def test_nbody_gradient_pm_nc_factor2_against_oracle(gpu):
  """Full-chain gradient with the force mesh twice the particle grid (16^3
  particles, 32^3 mesh), taped and with checkpoints every 2 steps, against
  the float64 oracle's autograd: the reverse sweep's force VJPs at f = 2
  (paint and readout at scaled positions, cotangent weights fac / f)."""
  n, L = 16, 64.0
  white, pk = _ics(n, L, 13)
  tgt = torch.rand(1, n, n, n, dtype=torch.float64,
                   generator=torch.Generator().manual_seed(14)) * 2
  g = obg.GrowthTable(obg.planck15())
  stages = np.linspace(0.1, 1.0, 6)
  wo = white.clone().requires_grad_(True)
  lin = opm.whitenoise_to_linear(wo, L, pk)
  st = opm.nbody(g, opm.lpt_init(g, lin, 0.1, order=2), stages, [n] * 3,
                 pm_nc_factor=2)
  rho = opm.cic_paint(torch.zeros_like(lin), st[0])
  ((rho - tgt)**2).sum().backward()
  c = fa.cosmology.Planck15()
  for every in (0, 2):
    wg = white.float().to(gpu).requires_grad_(True)
    ling = fa.whitenoise_to_linear(n, L, wg, pk)
    stg = fa.nbody(c, fa.lpt_init(c, ling, 0.1, order=2), stages, n,
                   pm_nc_factor=2, checkpoint_every=every)
    rhog = fa.cic_paint(torch.zeros_like(ling), stg[0])
    ((rhog - tgt.float().to(gpu))**2).sum().backward()
    cs = cosine(wg.grad, wo.grad)
    print("pm_nc_factor 2 gradient cosine (every=%s): %.7f, rel %.3g" %
          (every, cs, rel_rms(wg.grad, wo.grad)))
    assert cs > 0.999


def test_nbody_gradient_64_tiled(gpu):
  """Full-chain gradient at 64^3 (the particle passes of the adjoint visit
  4 x 64 tiles here), taped adjoint, against the float64 oracle."""
  n, L = 64, 200.0
  white, pk = _ics(n, L, 31)
  tgt = torch.rand(1, n, n, n, dtype=torch.float64,
                   generator=torch.Generator().manual_seed(32)) * 2
  g = obg.GrowthTable(obg.planck15())
  stages = np.linspace(0.1, 1.0, 4)
  wo = white.clone().requires_grad_(True)
  lin = opm.whitenoise_to_linear(wo, L, pk)
  st = opm.nbody(g, opm.lpt_init(g, lin, 0.1, order=1), stages, [n] * 3)
  rho = opm.cic_paint(torch.zeros_like(lin), st[0])
  ((rho - tgt)**2).sum().backward()
  c = fa.cosmology.Planck15()
  wg = white.float().to(gpu).requires_grad_(True)
  ling = fa.whitenoise_to_linear(n, L, wg, pk)
  stg = fa.nbody(c, fa.lpt_init(c, ling, 0.1, order=1), stages, n,
                 checkpoint_every=0)
  rhog = fa.cic_paint(torch.zeros_like(ling), stg[0])
  ((rhog - tgt.float().to(gpu))**2).sum().backward()
  cs = cosine(wg.grad, wo.grad)
  print("64^3 gradient cosine (taped): %.7f, rel %.3g" %
        (cs, rel_rms(wg.grad, wo.grad)))
  assert cs > 0.999


def test_nbody_and_gradient_banded_stencil(gpu):
  """A mesh whose x planes are large enough (256 x 512) that the gradient
  stencil and its adjoint visit the cells in y bands: forward (packed force
  mesh) and taped gradient (float4 force mesh, adjoint stencil) against the
  float64 oracle."""
  nc, box = [8, 256, 512], [32.0, 1024.0, 2048.0]
  gen = torch.Generator().manual_seed(41)
  white = torch.randn([1] + nc, generator=gen, dtype=torch.float64) * (
      nc[0] * nc[1] * nc[2])**0.5
  _, pk = _ics(8, 1.0, 0)
  tgt = torch.rand([1] + nc, dtype=torch.float64,
                   generator=torch.Generator().manual_seed(42)) * 2
  g = obg.GrowthTable(obg.planck15())
  stages = np.linspace(0.1, 1.0, 4)
  wo = white.clone().requires_grad_(True)
  lin = opm.whitenoise_to_linear(wo, box, pk)
  st_o = opm.nbody(g, opm.lpt_init(g, lin, 0.1, order=1), stages, nc)
  rho = opm.cic_paint(torch.zeros_like(lin), st_o[0])
  ((rho - tgt)**2).sum().backward()
  c = fa.cosmology.Planck15()
  with torch.no_grad():
    lin_g = fa.whitenoise_to_linear(nc, box, white.float().to(gpu), pk)
    st_g = fa.nbody(c, fa.lpt_init(c, lin_g, 0.1, order=1), stages, nc)
  q = opm.lagrangian_grid(nc, 1, torch.float64)
  st_od = st_o.detach()
  ex = rel_rms(st_g[0] - q.float().to(gpu), st_od[0] - q)
  ep = rel_rms(st_g[1], st_od[1])
  wg = white.float().to(gpu).requires_grad_(True)
  ling = fa.whitenoise_to_linear(nc, box, wg, pk)
  stg = fa.nbody(c, fa.lpt_init(c, ling, 0.1, order=1), stages, nc,
                 checkpoint_every=0)
  rhog = fa.cic_paint(torch.zeros_like(ling), stg[0])
  ((rhog - tgt.float().to(gpu))**2).sum().backward()
  cs = cosine(wg.grad, wo.grad)
  print("banded %s: dX %.3g P %.3g; gradient cosine %.7f" %
        (nc, ex, ep, cs))
  assert ex < 1e-4 and ep < 1e-4
  assert cs > 0.999


def test_gradient_with_spilled_bricks(gpu):
  """Gradient through two KDK steps from a state whose bricks do not fit
  an LDS box (displacements of 9 cells rms on 32^3: the density paint and
  the three-channel paint of the force VJP add those bricks with device
  atomics after the gather), against float64 autograd."""
  n = 32
  st = _displaced_state(n, 1, amp=9.0, seed=5)
  g = obg.GrowthTable(obg.planck15())
  stages = np.linspace(0.4, 0.6, 3)
  w = torch.randn(st.shape, generator=torch.Generator().manual_seed(6),
                  dtype=torch.float64)
  so = st.clone().requires_grad_(True)
  (opm.nbody(g, so, stages, [n] * 3) * w).sum().backward()
  c = fa.cosmology.Planck15()
  for every in (0, 1):
    sg = st.float().to(gpu).requires_grad_(True)
    out = fa.nbody(c, sg, stages, n, checkpoint_every=every)
    (out * w.float().to(gpu)).sum().backward()
    cs = cosine(sg.grad, so.grad)
    print("spilled-brick gradient cosine (every=%d): %.7f, rel %.3g" %
          (every, cs, rel_rms(sg.grad, so.grad)))
    assert cs > 0.9999


def test_taped_adjoint_matches_recompute(gpu):
  """The taped reverse sweep (no recomputation) and the checkpointed one
  (segments recomputed) evaluate the same kernels on the same inputs."""
  n, L = 32, 96.0
  white, pk = _ics(n, L, 21)
  c = fa.cosmology.Planck15()
  stages = np.linspace(0.1, 1.0, 7)
  gseed = torch.randn(3, 1, n**3, 3, generator=torch.Generator().manual_seed(
      22)).to(gpu)
  lin = fa.whitenoise_to_linear(n, L, white.float().to(gpu), pk)
  st0 = fa.lpt_init(c, lin, 0.1, order=2)
  grads = {}
  for every in (0, 3):
    x0 = st0.clone().requires_grad_(True)
    out = fa.nbody(c, x0, stages, n, checkpoint_every=every)
    (grads[every],) = torch.autograd.grad(out, x0, grad_outputs=gseed)
  close(grads[0], grads[3], 1e-6)


@pytest.mark.parametrize("every,B", [(1, 1), (3, 1), (8, 1), (3, 2)])
def test_checkpointed_adjoint_with_and_without_kept_potentials(gpu, every, B):
  """A checkpointed reverse sweep recomputes its segments' force meshes from
  the potentials the forward run kept (gradient stencil) or by painting and
  transforming again: both give the taped sweep's gradient, and so does a
  second backward pass (retain_graph), which recomputes every segment
  (the last one too).  6 steps: every = 8 is one segment, 3 two, 1 six;
  B = 2 keeps one potential per member in each slot."""
  n, L = 32, 96.0
  white, pk = _ics(n, L, 23, B=B)
  c = fa.cosmology.Planck15()
  stages = np.linspace(0.1, 1.0, 7)
  gseed = torch.randn(3, B, n**3, 3, generator=torch.Generator().manual_seed(
      24)).to(gpu)
  lin = fa.whitenoise_to_linear(n, L, white.float().to(gpu), pk)
  st0 = fa.lpt_init(c, lin, 0.1, order=2)
  x0 = st0.clone().requires_grad_(True)
  out0 = fa.nbody(c, x0, stages, n, checkpoint_every=0)
  (g0,) = torch.autograd.grad(out0, x0, grad_outputs=gseed)
  for keep in (True, False):
    fa.tfpm.set_keep_potentials(keep)
    try:
      x0 = st0.clone().requires_grad_(True)
      out = fa.nbody(c, x0, stages, n, checkpoint_every=every)
      close(out, out0, 1e-6, "final state (keep=%s)" % keep)
      (g1,) = torch.autograd.grad(out, x0, grad_outputs=gseed,
                                  retain_graph=True)
      (g2,) = torch.autograd.grad(out, x0, grad_outputs=gseed)
    finally:
      fa.tfpm.set_keep_potentials(None)
    e1 = close(g1, g0, 1e-6, "gradient (keep=%s)" % keep)
    e2 = close(g2, g0, 1e-6, "second gradient (keep=%s)" % keep)
    print("every=%d B=%d keep=%s: rel RMS %.3g / %.3g against the tape" %
          (every, B, keep, e1, e2))


@pytest.mark.parametrize("keep", [False, True])
def test_checkpointed_run_writes_the_segment_states(gpu, keep):
  """pmx_nbody_run_checkpointed (the checkpoints written by the readout of
  each boundary step, no split kick) against pmx_nbody_run per segment with
  the state copied between segments: the same checkpoints and final state,
  bit for bit, with and without the kept potentials, in place and from a
  separate initial state (7 steps, every 2: checkpoints after steps 2, 4
  and 6)."""
  from flowpm_amd import _lib
  n, L, B = 32, 96.0, 1
  white, pk = _ics(n, L, 31, B=B)
  c = fa.cosmology.Planck15()
  stages = np.linspace(0.1, 1.0, 8)
  lin = fa.whitenoise_to_linear(n, L, white.float().to(gpu), pk)
  st0 = fa.lpt_init(c, lin, 0.1, order=2).contiguous()
  plan, B, N, f, fac = fa.tfpm._force_setup(c, st0, n, 1)
  coef = fa.tfpm.step_coefficients(c, stages)
  cptr = _lib.dbl(coef)
  nsteps, every = len(coef) // 3, 2
  L_ = _lib.lib()
  work = torch.empty(L_.pmx_nbody_work_floats(plan.handle),
                     dtype=torch.float32, device=gpu)
  phi_n = L_.pmx_nbody_phi_floats(plan.handle, nsteps)
  phis = [torch.zeros(phi_n, device=gpu) if keep else None for _ in range(2)]
  pp = [None if q is None else _lib.ptr(q) for q in phis]
  stream = _lib.stream_of(st0)
  # reference: one call per segment, the state copied at each boundary
  st = st0.clone()
  ref = []
  for b0 in range(0, nsteps, every):
    if b0 > 0:
      ref.append(st.clone())
    b1 = min(nsteps, b0 + every)
    if keep:
      _lib.call("pmx_nbody_run_phi", plan.handle, _lib.ptr(st), N, f, fac,
                cptr, b0, b1, 1 if b0 == 0 else 0, _lib.ptr(work), pp[0],
                stream)
    else:
      _lib.call("pmx_nbody_run", plan.handle, _lib.ptr(st), N, f, fac, cptr,
                b0, b1, 1 if b0 == 0 else 0, _lib.ptr(work), stream)
  for from_src in (False, True):
    # in place, or reading the initial state from a separate array (which
    # must stay as it was)
    st2 = torch.full_like(st0, float("nan")) if from_src else st0.clone()
    src = st0.clone()
    ck = torch.full((len(ref),) + tuple(st0.shape), float("nan"), device=gpu)
    if keep:
      phis[1].zero_()
    _lib.call("pmx_nbody_run_checkpointed", plan.handle, _lib.ptr(st2),
              _lib.ptr(src) if from_src else None, N, f, fac, cptr, 0,
              nsteps, 1, every, _lib.ptr(ck), _lib.ptr(work), pp[1], stream)
    torch.cuda.synchronize()
    for j, r in enumerate(ref):
      assert torch.equal(ck[j], r), "checkpoint %d differs (src %s)" % (
          j, from_src)
    assert torch.equal(st2, st), "final state differs (src %s)" % from_src
    assert torch.equal(src, st0), "the initial state was modified"
    if keep:
      assert torch.equal(phis[0], phis[1]), "kept potentials differ"


@pytest.mark.parametrize("B", [1, 2])
def test_compact_tape_matches_full_tape(gpu, B):
  """The compact tape (positions and potential of every force; the reverse
  sweep forms each float4 force mesh with the gradient stencil) gives the
  full tape's final state and gradients bit for bit, w.r.t. the state and
  Omega_c, also on a second backward pass."""
  n, L = 32, 96.0
  white, pk = _ics(n, L, 25, B=B)
  c0 = fa.cosmology.Planck15()
  stages = np.linspace(0.1, 1.0, 6)
  gseed = torch.randn(3, B, n**3, 3, generator=torch.Generator().manual_seed(
      26)).to(gpu)
  lin = fa.whitenoise_to_linear(n, L, white.float().to(gpu), pk)
  st0 = fa.lpt_init(c0, lin, 0.1, order=2)
  res = {}
  for compact in (False, True):
    fa.tfpm.set_compact_tape(compact)
    try:
      oc = torch.tensor(0.2589, dtype=torch.float64, requires_grad=True)
      c = fa.cosmology.Planck15(Omega_c=oc)
      x0 = st0.clone().requires_grad_(True)
      out = fa.nbody(c, x0, stages, n, checkpoint_every=0)
      g1 = torch.autograd.grad(out, (x0, oc), grad_outputs=gseed,
                               retain_graph=True)
      g2 = torch.autograd.grad(out, (x0, oc), grad_outputs=gseed)
    finally:
      fa.tfpm.set_compact_tape(None)
    res[compact] = (out.detach(), g1, g2)
  (o0, a0, b0), (o1, a1, b1) = res[False], res[True]
  assert torch.equal(o0, o1)
  for x, y in ((a0, a1), (b0, b1), (a0, b1)):
    assert torch.equal(x[0], y[0]) and float(x[1]) == float(y[1])
  print("B=%d: dL/dOmega_c %.8g with either tape" % (B, float(a0[1])))


@pytest.mark.parametrize("f", [1, 2])
def test_nbody_pgd_against_oracle(gpu, f):
  """nbody with the PGD correction (tfpm.py:369-422, :487-488): final
  state at 32^3 (f = 1) and 16^3 particles on a 32^3 mesh (f = 2)."""
  n = 32 // f
  L = 100.0
  white, pk = _ics(n, L, 40 + f)
  g = obg.GrowthTable(obg.planck15())
  stages = np.linspace(0.1, 1.0, 6)
  pgdp = np.array([[0.3 + 0.05 * i, 0.6, 3.0 + 0.2 * i] for i in range(5)],
                  dtype=np.float32)
  lin = opm.whitenoise_to_linear(white, L, pk)
  st_o = opm.nbody(g, opm.lpt_init(g, lin, 0.1, order=1), stages, [n] * 3,
                   pm_nc_factor=f, pgdparams=pgdp.astype(np.float64))
  c = fa.cosmology.Planck15()
  st_g = fa.nbody(c, fa.lpt_init(c, lin.float().to(gpu), 0.1, order=1),
                  stages, n, pm_nc_factor=f, pgdparams=pgdp)
  q = opm.lagrangian_grid([n] * 3, 1, torch.float64)
  ex = rel_rms(st_g[0] - q.float().to(gpu), st_o[0] - q)
  ep = rel_rms(st_g[1], st_o[1])
  st_n = opm.nbody(g, opm.lpt_init(g, lin, 0.1, order=1), stages, [n] * 3,
                   pm_nc_factor=f)
  moved = rel_rms(st_o[0] - st_n[0], st_o[0] - q)
  print("PGD f=%d: dX %.3g P %.3g (PGD moved X by %.3g)" % (f, ex, ep, moved))
  assert moved > 1e-3  # the correction matters at these parameters
  assert ex < 1e-4 and ep < 1e-4


def test_nbody_pgd_gradients_against_oracle(gpu):
  """Gradients through nbody with the PGD correction: w.r.t. the white
  noise (full chain) and w.r.t. every step's (alpha, kl, ks), as
  scripts/fit_pgd.py:77-117 uses them, against float64 autograd."""
  n, L = 16, 64.0
  white, pk = _ics(n, L, 51)
  tgt = torch.rand(1, n, n, n, dtype=torch.float64,
                   generator=torch.Generator().manual_seed(52)) * 2
  g = obg.GrowthTable(obg.planck15())
  stages = np.linspace(0.1, 1.0, 5)
  p0 = np.array([[0.4, 0.7, 2.5], [0.5, 0.6, 3.0], [0.3, 0.8, 2.0],
                 [0.45, 0.65, 2.8]])
  wo = white.clone().requires_grad_(True)
  po = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
  lin = opm.whitenoise_to_linear(wo, L, pk)
  st = opm.nbody(g, opm.lpt_init(g, lin, 0.1, order=2), stages, [n] * 3,
                 pgdparams=po)
  rho = opm.cic_paint(torch.zeros_like(lin), st[0])
  ((rho - tgt)**2).sum().backward()
  c = fa.cosmology.Planck15()
  wg = white.float().to(gpu).requires_grad_(True)
  pg = torch.tensor(p0, dtype=torch.float32, device=gpu, requires_grad=True)
  ling = fa.whitenoise_to_linear(n, L, wg, pk)
  stg = fa.nbody(c, fa.lpt_init(c, ling, 0.1, order=2), stages, n,
                 pgdparams=pg)
  rhog = fa.cic_paint(torch.zeros_like(ling), stg[0])
  ((rhog - tgt.float().to(gpu))**2).sum().backward()
  cs = cosine(wg.grad, wo.grad)
  gp, go = pg.grad.double().cpu(), po.grad
  ep = float((gp - go).abs().max() / go.abs().max())
  print("PGD gradients: white cosine %.7f; params max rel err %.3g" % (cs, ep))
  print("params grad gpu", gp.numpy(), "oracle", go.numpy())
  assert cs > 0.999
  assert ep < 2e-2


def test_c2_gradient_128_against_oracle(gpu):
  """C2: 128^3, box 256, 2LPT, 20 steps, chi^2 = sum (delta - delta_t)^2
  w.r.t. the white field: the taped adjoint against float64 autograd of the
  oracle, over the whole gradient (cosine > 0.9999) and along 16 random
  smooth Fourier-mode directions (directional derivatives within 1e-2).
  Central finite differences of the float32 forward are no check here:
  the evolved CIC loss is only piecewise smooth, and the step sizes that
  resolve its derivative are below float32's resolution of the loss."""
  n, L = 128, 256.0
  white, pk = _ics(n, L, 100)
  tgt = torch.rand(1, n, n, n, generator=torch.Generator().manual_seed(101),
                   dtype=torch.float64) * 0.2 - 0.1
  g = obg.GrowthTable(obg.planck15())
  stages = np.linspace(0.1, 1.0, 21)
  wo = white.clone().requires_grad_(True)
  lin = opm.whitenoise_to_linear(wo, L, pk)
  st = opm.nbody(g, opm.lpt_init(g, lin, 0.1, order=2), stages, [n] * 3)
  rho = opm.cic_paint(torch.zeros_like(lin), st[0])
  (((rho - 1.0) - tgt)**2).sum().backward()
  go = wo.grad.reshape(n, n, n)
  c = fa.cosmology.Planck15()
  wg = white.float().to(gpu).requires_grad_(True)
  ling = fa.whitenoise_to_linear(n, L, wg, pk)
  stg = fa.nbody(c, fa.lpt_init(c, ling, 0.1, order=2), stages, n)
  rhog = fa.cic_paint(torch.zeros_like(ling), stg[0])
  (((rhog - 1.0) - tgt.float().to(gpu))**2).sum().backward()
  gg = wg.grad.double().cpu().reshape(n, n, n)
  cs = cosine(gg, go)
  rng = np.random.default_rng(7)
  q = np.indices((n, n, n)).astype(np.float64)
  dd = []
  for _ in range(16):
    m = rng.integers(-4, 5, size=3)
    while not m.any():
      m = rng.integers(-4, 5, size=3)
    ph = rng.uniform(0, 2 * np.pi)
    v = torch.tensor(np.cos(sum(2 * np.pi * m[d] * q[d] / n
                                for d in range(3)) + ph))
    dd.append((float((gg * v).sum()), float((go * v).sum())))
  dd = np.array(dd)
  rel = np.abs(dd[:, 0] - dd[:, 1]) / np.abs(dd[:, 1]).max()
  print("C2 128^3 x 20 steps: gradient cosine %.7f; directional max rel %.3g"
        % (cs, rel.max()))
  assert cs > 0.9999
  assert rel.max() < 1e-2


def test_gradient_64_20steps_against_oracle(gpu):
  """The deeply nonlinear case the finite differences cannot resolve:
  64^3, box 128, 2LPT, 20 steps; taped adjoint vs float64 autograd of the
  oracle (cosine > 0.9999)."""
  n, L = 64, 128.0
  white, pk = _ics(n, L, 100)
  tgt = torch.rand(1, n, n, n, generator=torch.Generator().manual_seed(101),
                   dtype=torch.float64) * 0.2 - 0.1
  g = obg.GrowthTable(obg.planck15())
  stages = np.linspace(0.1, 1.0, 21)
  wo = white.clone().requires_grad_(True)
  lin = opm.whitenoise_to_linear(wo, L, pk)
  st = opm.nbody(g, opm.lpt_init(g, lin, 0.1, order=2), stages, [n] * 3)
  rho = opm.cic_paint(torch.zeros_like(lin), st[0])
  (((rho - 1.0) - tgt)**2).sum().backward()
  c = fa.cosmology.Planck15()
  wg = white.float().to(gpu).requires_grad_(True)
  ling = fa.whitenoise_to_linear(n, L, wg, pk)
  stg = fa.nbody(c, fa.lpt_init(c, ling, 0.1, order=2), stages, n,
                 checkpoint_every=0)
  rhog = fa.cic_paint(torch.zeros_like(ling), stg[0])
  (((rhog - 1.0) - tgt.float().to(gpu))**2).sum().backward()
  cs = cosine(wg.grad, wo.grad)
  print("64^3 x 20 steps gradient cosine %.7f rel %.3g" %
        (cs, rel_rms(wg.grad, wo.grad)))
  assert cs > 0.9999


@pytest.mark.parametrize("shape,tol", [((2, 32, 16, 24), 2e-6),
                                       ((1, 64, 64, 64), 2e-6),
                                       ((1, 40, 100, 80), 2e-6),
                                       ((1, 8, 512, 256), 2e-5)])
def test_force_mesh_f4_matches_fourier_gradient(gpu, shape, tol):
  """The time loop's force mesh (one potential, real-space fourth-order
  gradient) against the three Fourier-space gradient meshes of
  pmx_force_meshes on the same density, and against the float64 difference
  of the potential it returns.  The last shape's 512 x 256 planes make the
  stencil visit the cells in y bands; with only 8 x planes its long y / z
  wavelengths dominate the potential ~50-fold over the x force, so the
  float32 difference of neighbouring potential values keeps fewer digits of
  F_x (the tolerance against the Fourier path is wider there; the stencil
  check, relative to the potential, is not)."""
  from flowpm_amd import _lib, _plans
  B, nx, ny, nz = shape
  plan = _plans.get(gpu, B, nx, ny, nz)
  gen = torch.Generator().manual_seed(sum(shape))
  rho = (torch.rand(shape, generator=gen) * 2).to(gpu)
  sc = 1.0 / (nx * ny * nz)
  f3 = torch.empty((3,) + shape, device=gpu)
  work = plan.spec(3)
  _lib.call("pmx_force_meshes", plan.handle, _lib.ptr(rho), _lib.ptr(work),
            _lib.ptr(f3[0]), _lib.ptr(f3[1]), _lib.ptr(f3[2]), sc,
            _lib.stream_of(rho))
  f4 = torch.empty(shape + (4,), device=gpu)
  phi = torch.empty(shape, device=gpu)
  _lib.call("pmx_force_mesh_f4", plan.handle, _lib.ptr(rho), _lib.ptr(work),
            _lib.ptr(phi), _lib.ptr(f4), sc, _lib.stream_of(rho))
  ph = phi.double().cpu()
  errs = []
  for d in range(3):
    a = d + 1
    dphi = (2.0 / 3.0) * (ph.roll(-1, a) - ph.roll(1, a)) - (1.0 / 12.0) * (
        ph.roll(-2, a) - ph.roll(2, a))
    es = float((f4[..., d].double().cpu() - dphi).pow(2).mean().sqrt() /
               ph.pow(2).mean().sqrt())
    errs.append((rel_rms(f4[..., d], f3[d]), es))
  print(shape, "per component (vs Fourier, stencil vs |phi|):", errs)
  for d, (ef, es) in enumerate(errs):
    assert ef < tol, "component %d vs Fourier %.3g" % (d, ef)
    assert es < 5e-7, "component %d stencil %.3g" % (d, es)
  assert float(f4[..., 3].abs().max()) == 0.0


@pytest.mark.parametrize("nc,box", [([32, 32, 32], 100.0),
                                    ([16, 24, 32], 60.0),
                                    ([64, 64, 64], 200.0)])
def test_linear_amplitude_table_matches_per_mode(gpu, nc, box):
  """The equal-sides shortcut of the linear-field amplitude (P(k) once per
  distinct |k|, expanded on the GPU) against pk evaluated on every mode of
  the half spectrum (tfpm.py:56-62)."""
  _, pk = _ics(8, 1.0, 0)
  amp = fa.tfpm._linear_amp(nc, [box] * 3, pk, gpu).cpu().double().numpy()
  km = fa.tfpm._half_kmesh(nc, [box] * 3)
  ref = (np.asarray(pk(km), dtype=np.float64) / box**3)**0.5
  ref32 = ref.astype(np.float32).astype(np.float64)
  assert amp.shape == ref.shape
  err = np.abs(amp - ref32) / np.abs(ref32).max()
  assert err.max() < 1e-6, err.max()


@pytest.mark.parametrize("n,B,permute", [(32, 1, False), (32, 2, True),
                                         (64, 1, False), (40, 2, False)])
def test_cic_paint_one_particle_per_cell_against_oracle(gpu, n, B, permute):
  """cic_paint of one particle per cell goes through the brick paint
  (mesh += paint, weights, wraps); a shuffled particle order is painted
  correctly too (its bricks are scattered and spill to atomics).  The brick
  paint's fixed point rounds each deposit by at most 2^-30 of its brick's
  sum of |weights| (here <= 2048, i.e. <= 1.9e-6 per deposit, about eight
  deposits per cell), hence the 1e-5 bar."""
  st = _displaced_state(n, B, amp=1.2, seed=21)
  X = st[0]
  if permute:
    X = X[:, torch.randperm(n**3, generator=torch.Generator().manual_seed(3))]
  gen = torch.Generator().manual_seed(22)
  w = torch.rand(B, n**3, generator=gen, dtype=torch.float64)
  mesh0 = torch.randn(B, n, n, n, generator=gen, dtype=torch.float64)
  ref = opm.cic_paint(mesh0, X, w)
  out = fa.cic_paint(mesh0.float().to(gpu), X.float().to(gpu),
                     w.float().to(gpu))
  close(out, ref, 1e-5)
  ref1 = opm.cic_paint(torch.zeros_like(mesh0), X)
  out1 = fa.cic_paint(torch.zeros(B, n, n, n, device=gpu), X.float().to(gpu))
  close(out1, ref1, 1e-5)
