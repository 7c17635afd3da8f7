"""Fourier-space kernels (flowpm/kernels.py:10-102), host float64 NumPy.

These keep the reference's API for callers that build kernels themselves.
The GPU path never materialises them: the HIP passes evaluate the same
values per element from float64 per-axis tables (k^2 and the 4th-order
gradient weight), so no M-sized kernel array is read from HBM.
"""
import numpy as np

__all__ = ["fftk", "laplace_kernel", "gradient_kernel", "longrange_kernel",
           "PGD_kernel", "cic_compensation"]


def fftk(shape, symmetric=True, finite=False, dtype=np.float64):
  """2 pi fftfreq per axis, shaped to broadcast (kernels.py:10-25)."""
  out = []
  nd = len(shape)
  for d, n in enumerate(shape):
    kd = 2 * np.pi * np.fft.fftfreq(n)
    if symmetric and d == nd - 1:
      kd = kd[:n // 2 + 1]
    view = [1] * nd
    view[d] = len(kd)
    out.append(kd.reshape(view).astype(dtype))
  return out


def laplace_kernel(kvec):
  """1/|k|^2 with the zero mode equal to 1 (kernels.py:28-48: kk is set to 1
  there before the mask is formed, so the mask keeps it)."""
  kk = sum(ki**2 for ki in kvec)
  kk = np.where(kk == 0, 1.0, kk)
  return 1.0 / kk


def gradient_kernel(kvec, direction, order=1):
  """i (8 sin w - sin 2w)/6 for order=1; i k (Nyquist zeroed) for order=0."""
  w = kvec[direction]
  if order == 0:
    wts = np.squeeze(1j * w).copy()
    wts[len(wts) // 2] = 0
    return wts.reshape(w.shape)
  return 1j * (8 * np.sin(w) - np.sin(2 * w)) / 6.0


def longrange_kernel(kvec, r_split):
  if r_split != 0:
    kk = sum(ki**2 for ki in kvec)
    return np.exp(-kk * r_split**2)
  return 1.0


def PGD_kernel(kvec, kl, ks):
  """exp(-kl^2 / k^2) exp(-k^4 / ks^4) (kernels.py:105-134); as there, the
  zero mode uses k^2 = 1 and is kept (its mask is formed after the
  replacement).  The reference evaluates it in float32 (TF converts k^2 to
  the dtype of kl); so does this host version."""
  kk = sum(ki**2 for ki in kvec)
  kk = np.where(kk == 0, 1.0, kk)
  kl2 = np.float32(kl)**2
  ks4 = np.float32(ks)**4
  return (np.exp(-kl2 / kk.astype(np.float32)) *
          np.exp(-(kk**2).astype(np.float32) / ks4)).astype(np.float32)


def cic_compensation(kvec):
  """(sinc(kx/2pi) sinc(ky/2pi) sinc(kz/2pi))^-2 (kernels.py:137-152)."""
  kwts = [np.sinc(kvec[i] / (2 * np.pi)) for i in range(3)]
  return (kwts[0] * kwts[1] * kwts[2])**(-2)
