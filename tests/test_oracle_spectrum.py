"""Pins of oracle/spectrum.py (power spectrum, CIC compensation)."""
import numpy as np
import torch

from oracle import spectrum as osp


def _cos_field(shape, m):
  idx = np.indices(shape).astype(np.float64)
  ph = sum(2 * np.pi * m[d] * idx[d] / shape[d] for d in range(3))
  return np.cos(ph)


def test_single_mode_power_closed_form():
  """cos(k.x) has |F|^2 = (M/2)^2 at +-k: binned power = sum of W at the
  two modes * (M/2)^2 / Nsum_bin * V / M^2."""
  shape, L = (16, 16, 16), np.array([100.0, 100.0, 100.0])
  m = (3, 0, 0)  # along x: both modes have kz index 0, W = 1
  field = _cos_field(shape, m)[None]
  kmin, dk = 0.01, 0.05
  kbins, P = osp.power_spectrum(field, kmin, dk, L)
  dig, Nsum, W, k, kedges = osp.initialize_pk(shape, L, kmin, dk)
  kk = 2 * np.pi * 3 / 100.0
  b = np.digitize([kk], kedges)[0]
  M = np.prod(shape)
  expect = 2 * (M / 2)**2 / Nsum[b] * L.prod() / M**2
  assert np.isclose(P[0, b - 1], expect, rtol=1e-12)
  P2 = P.copy()
  P2[0, b - 1] = 0
  empty = Nsum[1:-1] == 0  # 0 / 0 as in the reference
  assert np.isnan(P2[0, empty]).all()
  assert np.abs(P2[0, ~empty]).max() < 1e-20 * expect


def test_mode_weights_w_quirk():
  """The W weights follow the FULL-spectrum kz index (1 at kz = 0 and at
  the last index, 2 elsewhere): a mode along z at index 1 weighs 2 and its
  mirror at nz - 1 weighs 1."""
  shape, L = (8, 8, 8), np.array([8.0, 8.0, 8.0])
  field = _cos_field(shape, (0, 0, 1))[None]
  kmin, dk = 0.1, 0.5
  dig, Nsum, W, k, kedges = osp.initialize_pk(shape, L, kmin, dk)
  kbins, P = osp.power_spectrum(field, kmin, dk, L)
  b = np.digitize([2 * np.pi / 8.0], kedges)[0]
  M = 512
  assert np.isclose(P[0, b - 1], (2 + 1) * (M / 2)**2 / Nsum[b] * 512 / M**2,
                    rtol=1e-12)


def test_mode_counts_direct():
  shape, L = (8, 12, 16), np.array([50.0, 60.0, 70.0])
  dig, Nsum, W, k, kedges = osp.initialize_pk(shape, L, 0.05, 0.1)
  direct = np.zeros(len(kedges) + 1)
  for i in range(shape[0]):
    for j in range(shape[1]):
      for l in range(shape[2]):
        kmag = np.sqrt(k[0][i]**2 + k[1][j]**2 + k[2][l]**2)
        w = 1.0 if l in (0, shape[2] - 1) else 2.0
        direct[np.searchsorted(kedges, kmag, side="right")] += w
  np.testing.assert_array_equal(direct, Nsum)


def test_constant_field_power_only_at_dc():
  shape, L = (8, 8, 8), np.array([10.0] * 3)
  kbins, P = osp.power_spectrum(np.full((1,) + shape, 3.0), 0.5, 0.5, L)
  assert np.abs(P).max() == 0.0


def test_compensation_single_mode_and_constant():
  shape = (16, 8, 12)
  m = (2, 1, 3)
  f = _cos_field(shape, m)[None]
  out = osp.compensate_cic(f)
  w = 1.0
  for d in range(3):
    x = m[d] / shape[d]
    w *= (np.sin(np.pi * x) / (np.pi * x))**-2
  np.testing.assert_allclose(out, f * w, atol=1e-12)
  c = np.full((1,) + shape, 2.5)
  np.testing.assert_allclose(osp.compensate_cic(c), c, atol=1e-12)


def test_torch_versions_equal_numpy_and_finite_differences():
  rng = np.random.default_rng(5)
  shape, L = (8, 8, 10), np.array([40.0, 40.0, 50.0])
  x = rng.standard_normal((2,) + shape)
  kmin, dk = 0.1, 0.3
  _, P = osp.power_spectrum(x, kmin, dk, L)
  xt = torch.tensor(x, requires_grad=True)
  _, Pt = osp.power_spectrum_torch(xt, kmin, dk, L)
  ok = np.isfinite(P)
  np.testing.assert_allclose(Pt.detach().numpy()[ok], P[ok], rtol=1e-12)
  np.testing.assert_allclose(osp.compensate_cic_torch(xt).detach().numpy(),
                             osp.compensate_cic(x), rtol=1e-12, atol=1e-12)
  # loss of the PGD fit (scripts/fit_pgd.py:61-70) on the compensated field
  target = np.abs(P[:, ok[0]].mean(0)) + 1.0

  def loss(t):
    _, pk = osp.power_spectrum_torch(osp.compensate_cic_torch(t), kmin, dk, L)
    pk = pk[:, torch.as_tensor(ok[0])].mean(0)
    return ((1 - pk / torch.as_tensor(target))**2).sum()

  loss(xt).backward()
  g = xt.grad.numpy()
  for idx in [(0, 1, 2, 3), (1, 7, 0, 9), (0, 4, 4, 0)]:
    e = np.zeros_like(x)
    e[idx] = 1e-6
    fd = (float(loss(torch.tensor(x + e))) -
          float(loss(torch.tensor(x - e)))) / 2e-6
    assert abs(fd - g[idx]) < 1e-6 * max(1.0, abs(fd)), (idx, fd, g[idx])
