"""power_spectrum and compensate_cic on the GPU against the oracle."""
import numpy as np
import pytest
import torch

import flowpm_amd as fa
from oracle import spectrum as osp

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,box,kmin,dk", [
    ((1, 16, 16, 16), (100.0, 100.0, 100.0), 0.01, 0.05),
    ((2, 32, 32, 32), (200.0, 200.0, 200.0), 0.02, 0.03),
    ((1, 16, 24, 32), (50.0, 60.0, 70.0), 0.1, 0.2),
    ((1, 64, 64, 64), (10.0, 10.0, 10.0), 5, 0.5),
])
def test_power_spectrum_against_oracle(gpu, shape, box, kmin, dk):
  gen = torch.Generator().manual_seed(sum(shape))
  x = torch.randn(shape, generator=gen, dtype=torch.float64) + 1.0
  kb_o, P_o = osp.power_spectrum(x.numpy().astype(np.float32), kmin, dk,
                                 np.array(box))
  kb_g, P_g = fa.power_spectrum(x.float().to(gpu), kmin=kmin, dk=dk,
                                boxsize=np.array(box))
  np.testing.assert_array_equal(kb_g, kb_o)
  P_g = P_g.cpu().numpy().astype(np.float64)
  assert P_g.shape == P_o.shape
  nan = np.isnan(P_o)
  assert (np.isnan(P_g) == nan).all()
  err = np.abs(P_g[~nan] - P_o[~nan]) / np.abs(P_o[~nan]).max()
  assert err.max() < 1e-5, err.max()


def test_power_spectrum_of_painted_density(gpu):
  """P(k) of a 2LPT density at 64^3 (the loss / diagnostic of the
  reference's scripts) against the oracle on the same painted mesh."""
  n, L = 64, 128.0
  c = fa.cosmology.Planck15()
  gen = torch.Generator().manual_seed(5)
  white = (torch.randn(1, n, n, n, generator=gen, dtype=torch.float64) *
           n**1.5).float().to(gpu)
  k = np.logspace(-4, 1, 256)
  pkt = fa.tfpower.linear_matter_power(c, k)
  lin = fa.whitenoise_to_linear(
      n, L, white, lambda x: fa.scipy.interpolate.interp_tf(x, k, pkt))
  st = fa.lpt_init(c, lin, 1.0, order=2)
  rho = fa.cic_paint(torch.zeros(1, n, n, n, device=gpu), st[0])
  kb_g, P_g = fa.power_spectrum(rho, kmin=0.1, dk=0.05, boxsize=[L] * 3)
  kb_o, P_o = osp.power_spectrum(rho.cpu().numpy(), 0.1, 0.05, [L] * 3)
  P_g = P_g.cpu().numpy()
  ok = ~np.isnan(P_o)
  err = np.abs(P_g[ok] - P_o[ok]) / np.abs(P_o[ok])
  assert err.max() < 1e-4, err.max()


@pytest.mark.parametrize("shape", [(1, 16, 16, 16), (2, 32, 24, 16)])
def test_compensate_cic_against_oracle(gpu, shape):
  gen = torch.Generator().manual_seed(7)
  x = torch.randn(shape, generator=gen, dtype=torch.float64)
  out = fa.compensate_cic(x.float().to(gpu)).cpu().double()
  ref = torch.as_tensor(osp.compensate_cic(x.numpy()))
  e = float((out - ref).norm() / ref.norm())
  assert e < 2e-6, e


def test_compensate_cic_adjoint(gpu):
  """The filter is real and symmetric: <C x, y> = <x, C y>, and autograd
  returns C g."""
  gen = torch.Generator().manual_seed(8)
  x = torch.randn(1, 16, 16, 16, generator=gen).to(gpu).requires_grad_(True)
  y = torch.randn(1, 16, 16, 16, generator=gen).to(gpu)
  out = fa.compensate_cic(x)
  (g,) = torch.autograd.grad((out * y).sum(), x)
  ref = fa.compensate_cic(y)
  e = float((g - ref).norm() / ref.norm())
  assert e < 1e-6, e


@pytest.mark.parametrize("shape,seed", [((1, 16, 16, 16), 0),
                                        ((2, 32, 24, 16), 12345),
                                        ((1, 8, 8, 10), (1 << 70) + 11)])
def test_white_noise_philox_against_oracle(gpu, shape, seed):
  """The on-device stream equals numpy.random.Philox's words through the
  same Box-Muller transform (float64 transcendentals may differ in the last
  bit between the GPU and libm: compare to 2 float32 ulps)."""
  from oracle import noise
  M = shape[1] * shape[2] * shape[3]
  w = fa.white_noise(shape[1:], batch_size=shape[0], seed=seed, type="real")
  ref = noise.white_noise(shape, seed, float(M)**0.5)
  w = w.cpu().numpy()
  ulp = np.spacing(np.abs(ref).astype(np.float32))
  assert (np.abs(w - ref) <= 2 * ulp).all()
  # a different seed is a different field
  w2 = fa.white_noise(shape[1:], batch_size=shape[0], seed=seed + 1,
                      type="real").cpu().numpy()
  assert np.abs(w2 - w).max() > 1.0


def test_save_state_of_gpu_state(gpu, tmp_path):
  """io.save_state on a device-resident 2LPT state (tests/test_io.py:16-62
  of the reference: B = 2, 16^3, box 100): the written rows equal the
  state copied to the host, scaled in float32."""
  n, L = 16, 100.0
  c = fa.cosmology.Planck15()
  gen = torch.Generator().manual_seed(0)
  white = (torch.randn(2, n, n, n, generator=gen, dtype=torch.float64) *
           n**1.5).float().to(gpu)
  k = np.logspace(-4, 1, 256)
  pkt = fa.tfpower.linear_matter_power(c, k)
  lin = fa.whitenoise_to_linear(
      [n] * 3, [L] * 3, white,
      lambda x: fa.scipy.interpolate.interp_tf(x, k, pkt))
  st = fa.lpt_init(c, lin, 0.1, order=2)
  fn = str(tmp_path / "snap")
  fa.io.save_state(c, st, 0.1, [n] * 3, [L] * 3, fn)
  host = st.cpu().numpy()
  scale = np.float32(np.float32(1.0) / np.float32(n) * np.float32(L))
  for b in range(2):
    np.testing.assert_array_equal(
        fa.io.read_block(fn, "%d/Position" % b), host[0, b] * scale)
    np.testing.assert_array_equal(
        fa.io.read_block(fn, "%d/Velocity" % b), host[1, b] * scale)
  np.testing.assert_allclose(fa.io.read_attrs(fn, "Header")["Time"], 0.1)


@pytest.mark.parametrize("shape,box,kmin,dk", [
    ((16, 16, 16), (100.0, 100.0, 100.0), 0.01, 0.05),
    ((16, 24, 32), (50.0, 60.0, 70.0), 0.1, 0.2),
    ((64, 64, 64), (200.0, 200.0, 200.0), 0.02, 0.03)])
def test_mode_counts_on_gpu_equal_host_bincount(gpu, shape, box, kmin, dk):
  """The weighted mode counts of the binning pass equal the reference
  procedure's full-mesh digitize + bincount (pk.py:7-37) exactly."""
  from flowpm_amd import _lib, _plans, pk as fpk
  k, kedges, Nsum = fpk._initialize_pk(shape, np.array(box), kmin, dk)
  plan = _plans.get(gpu, 1, *shape)
  tab = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float64),
                                  device=gpu)
  kx, ky, kz, edges = tab(k[0]), tab(k[1]), tab(k[2]), tab(kedges)
  nsum = torch.empty(len(kedges) + 1, dtype=torch.float64, device=gpu)
  _lib.call("pmx_power_spectrum_counts", plan.handle, _lib.ptr(kx),
            _lib.ptr(ky), _lib.ptr(kz), _lib.ptr(edges), len(kedges),
            _lib.ptr(nsum), _lib.stream_of(nsum))
  np.testing.assert_array_equal(nsum.cpu().numpy(), Nsum)
