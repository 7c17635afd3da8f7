"""Oracle: power spectrum and CIC compensation, NumPy float64.

TEST INFRASTRUCTURE ONLY.  Nothing in ``flowpm_amd`` may import this module;
only ``tests/`` use it, as the checker.  It restates:

  _initialize_pk / power_spectrum  flowpm/pk.py:7-37 / 40-103 (full complex
                                   numpy.fft.fftn of the real field, the
                                   reference's W weights over the FULL
                                   spectrum, numpy.digitize + bincount)
  cic_compensation                 flowpm/kernels.py:137-152
  compensate_cic                   flowpm/utils.py:11-29
  power_spectrum_torch /           the same two in torch float64, for the
  compensate_cic_torch             gradient of the PGD-fit loss
                                   (scripts/fit_pgd.py:44-70) by autograd

Pins (tests/test_oracle_spectrum.py): a single cosine mode puts exactly
(M/2)^2 W per +-k into one bin (closed form for the binned power); the
weighted mode counts equal a direct count over the grid; a constant field
has no power outside the DC bin; compensate_cic of a single cosine mode
multiplies it by the closed-form window (prod_d sinc(k_d/2pi))^-2, and of
a constant field returns it unchanged; the torch versions equal the NumPy
ones and their autograd equals central finite differences.
"""
import numpy as np
import torch


def initialize_pk(shape, boxsize, kmin, dk):
  """pk.py:7-37: returns dig (flat bin index per full-spectrum mode), the
  weighted counts Nsum, W, the per-axis k and the edges."""
  shape = [int(s) for s in shape]
  boxsize = np.asarray(boxsize, dtype=np.float64)
  W = np.empty(shape, dtype="f4")
  W[...] = 2.0
  W[..., 0] = 1.0
  W[..., -1] = 1.0
  kmax = np.pi * np.min(shape) / np.max(boxsize) + dk / 2
  kedges = np.arange(kmin, kmax, dk)
  k = [np.fft.fftfreq(N, 1. / (N * 2 * np.pi / L)) for N, L in
       zip(shape, boxsize)]
  kmag = (k[0][:, None, None]**2 + k[1][None, :, None]**2 +
          k[2][None, None, :]**2)**0.5
  dig = np.digitize(kmag.flat, kedges)
  Nsum = np.bincount(dig, weights=W.flat, minlength=len(kedges) + 1)
  return dig, Nsum, W, k, kedges


def power_spectrum(field, kmin, dk, boxsize):
  """pk.py:40-103 in float64: field (B, nx, ny, nz) -> (kbins, power)."""
  field = np.asarray(field, dtype=np.float64)
  B = field.shape[0]
  shape = field.shape[1:]
  boxsize = np.broadcast_to(np.asarray(boxsize, dtype=np.float64), (3,))
  dig, Nsum, W, k, kedges = initialize_pk(shape, boxsize, kmin, dk)
  out = []
  for b in range(B):
    F = np.fft.fftn(field[b])
    pk = (F * np.conj(F)).real
    Psum = np.bincount(dig, weights=(W.astype(np.float64) * pk).flat,
                       minlength=len(kedges) + 1)
    out.append((Psum / Nsum)[1:-1] * boxsize.prod())
  norm = float(np.prod(shape))**2
  kbins = kedges[:-1] + (kedges[1:] - kedges[:-1]) / 2
  return kbins, np.stack(out) / norm


def cic_compensation(shape):
  """kernels.py:137-152 on the full fftk grid (symmetric=False)."""
  kv = [2 * np.pi * np.fft.fftfreq(n) for n in shape]
  kw = [np.sinc(kv[d] / (2 * np.pi)) for d in range(3)]
  return (kw[0][:, None, None] * kw[1][None, :, None] *
          kw[2][None, None, :])**(-2)


def compensate_cic(field):
  """utils.py:11-29: c2r3d(W * r2c3d(field, norm=M), norm=M)."""
  field = np.asarray(field, dtype=np.float64)
  w = cic_compensation(field.shape[1:])
  return np.stack([np.fft.ifftn(np.fft.fftn(f) * w).real for f in field])


def power_spectrum_torch(field, kmin, dk, boxsize):
  """power_spectrum on a torch float64 (B, nx, ny, nz) tensor,
  differentiable: the bincount is an index_add over the same digitize."""
  B = field.shape[0]
  shape = list(field.shape[1:])
  boxsize = np.broadcast_to(np.asarray(boxsize, dtype=np.float64), (3,))
  dig, Nsum, W, k, kedges = initialize_pk(shape, boxsize, kmin, dk)
  dig_t = torch.as_tensor(dig, dtype=torch.int64)
  Wt = torch.as_tensor(W.astype(np.float64)).reshape(-1)
  out = []
  for b in range(B):
    F = torch.fft.fftn(field[b].to(torch.complex128))
    pk = (F * F.conj()).real.reshape(-1)
    Psum = torch.zeros(len(kedges) + 1, dtype=torch.float64).index_add(
        0, dig_t, Wt * pk)
    out.append((Psum / torch.as_tensor(Nsum))[1:-1] * float(boxsize.prod()))
  norm = float(np.prod(shape))**2
  kbins = kedges[:-1] + (kedges[1:] - kedges[:-1]) / 2
  return kbins, torch.stack(out) / norm


def compensate_cic_torch(field):
  """compensate_cic on a torch float64 tensor (differentiable)."""
  w = torch.as_tensor(cic_compensation(field.shape[1:]))
  F = torch.fft.fftn(field.to(torch.complex128), dim=(-3, -2, -1))
  return torch.fft.ifftn(F * w, dim=(-3, -2, -1)).real
