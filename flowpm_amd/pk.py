"""Power spectrum (flowpm/pk.py:7-103) on the GPU.

``power_spectrum(field, kmin, dk, boxsize)`` keeps the reference's
signature and its arithmetic: the host builds the bin edges and per-axis
wavenumbers as ``_initialize_pk`` does (NumPy float64), the HIP library
transforms the field (unnormalised half spectrum) and sums W |F|^2 per bin
(``pmx_power_spectrum_bins``, the reference's full-spectrum weights
accounted on the half spectrum); the weighted mode counts come from the
same binning pass over W alone (``pmx_power_spectrum_counts``, exact
integers in float64, equal to the reference's bincount) instead of a
full-mesh NumPy digitize; the normalisation follows the reference's float32
steps.  Differentiable w.r.t. the field through the hand-written adjoint
of the binning (``pmx_power_spectrum_bins_adj``), as the PGD fit needs
(scripts/fit_pgd.py:44-70).
"""
import numpy as np
import torch

from . import _lib
from . import _plans
from .utils import _f32, _mesh_dims, fwd_real, inv_real

__all__ = ["power_spectrum"]


def _edges(shape, boxsize, kmin, dk):
  """Per-axis wavenumbers and bin edges (pk.py:7-37), NumPy float64."""
  shape = [int(s) for s in shape]
  boxsize = np.asarray(boxsize, dtype=np.float64)
  kmax = np.pi * np.min(shape) / np.max(boxsize) + dk / 2
  kedges = np.arange(kmin, kmax, dk)
  k = [np.fft.fftfreq(N, 1. / (N * 2 * np.pi / L)) for N, L in
       zip(shape, boxsize)]
  return k, kedges


def _initialize_pk(shape, boxsize, kmin, dk):
  """Bin edges, per-axis wavenumbers and weighted mode counts
  (pk.py:7-37), NumPy float64, on the host over the full mesh (the
  reference's own procedure; power_spectrum counts on the GPU)."""
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
  kmag = (k[0].reshape(-1, 1, 1)**2 + k[1].reshape(1, -1, 1)**2 +
          k[2].reshape(1, 1, -1)**2)**0.5
  dig = np.digitize(kmag.flat, kedges)
  Nsum = np.bincount(dig, weights=W.flat, minlength=len(kedges) + 1)
  return k, kedges, Nsum


class _PowerBins(torch.autograd.Function):
  """psum[b][i] = sum over bin i of W |FFT3(field_b)|^2 (float64,
  pmx_power_spectrum_bins).  Backward: for a bin cotangent g, the field
  cotangent is 2 Re IDFT_unnorm(g_bin(k) W_k F_k) (the derivative of the
  bincount is a gather, of |F|^2 w.r.t. the real field 2 Re conj(F) dF):
  pmx_power_spectrum_bins_adj weights the kept spectrum, a c2r with scale
  2 transforms it back."""

  @staticmethod
  def forward(ctx, field, plan, tabs):
    kx, ky, kz, edges = tabs
    spec = fwd_real(plan, _f32(field), 1.0)  # unnormalised, as fft3d
    psum = torch.empty((plan.B, edges.numel() + 1), dtype=torch.float64,
                       device=field.device)
    _lib.call("pmx_power_spectrum_bins", plan.handle, _lib.ptr(spec),
              _lib.ptr(kx), _lib.ptr(ky), _lib.ptr(kz), _lib.ptr(edges),
              edges.numel(), _lib.ptr(psum), _lib.stream_of(field))
    ctx.plan, ctx.tabs = plan, tabs
    if field.requires_grad:
      ctx.save_for_backward(spec)
    return psum

  @staticmethod
  def backward(ctx, g):
    plan = ctx.plan
    kx, ky, kz, edges = ctx.tabs
    (spec,) = ctx.saved_tensors
    sp = spec.clone()
    g = g.detach().to(torch.float64).contiguous()
    _lib.call("pmx_power_spectrum_bins_adj", plan.handle, _lib.ptr(sp),
              _lib.ptr(kx), _lib.ptr(ky), _lib.ptr(kz), _lib.ptr(edges),
              edges.numel(), _lib.ptr(g), _lib.stream_of(sp))
    return inv_real(plan, sp, 2.0), None, None


def power_spectrum(field, kmin=5, dk=0.5, boxsize=False):
  """Binned power spectrum of a real (B, nx, ny, nz) field.

  Returns (kbins, power): kbins the NumPy bin centres, power a float32
  (B, len(kbins)) tensor on the field's device (pk.py:40-103);
  differentiable w.r.t. ``field`` (the loss of scripts/fit_pgd.py:44-70)."""
  _lib.require_cuda(field)
  B, nx, ny, nz = _mesh_dims(field)
  if boxsize is False or boxsize is None:
    raise ValueError("power_spectrum needs the box size per dimension "
                     "(boxsize.prod() normalises the power, pk.py:90)")
  boxsize = np.broadcast_to(np.asarray(boxsize, dtype=np.float64), (3,))
  k, kedges = _edges((nx, ny, nz), boxsize, kmin, dk)
  dev = field.device
  if len(kedges) < 2:  # no complete bin: empty, as the reference returns
    return kedges[:0].astype(np.float64), torch.zeros(
        (B, 0), dtype=torch.float32, device=dev)
  plan = _plans.get(dev, B, nx, ny, nz)
  tab = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float64),
                                  device=dev)
  tabs = (tab(k[0]), tab(k[1]), tab(k[2]), tab(kedges))
  psum = _PowerBins.apply(field, plan, tabs)
  nsum = torch.empty(len(kedges) + 1, dtype=torch.float64, device=dev)
  _lib.call("pmx_power_spectrum_counts", plan.handle, _lib.ptr(tabs[0]),
            _lib.ptr(tabs[1]), _lib.ptr(tabs[2]), _lib.ptr(tabs[3]),
            len(kedges), _lib.ptr(nsum), _lib.stream_of(field))
  # float32 steps of pk.py:82-103: Psum (complex64) / Nsum (complex64),
  # [1:-1], * boxsize.prod(), / float32(M)**2
  power = psum.to(torch.float32) / nsum.to(torch.float32)
  power = power[:, 1:-1] * np.float32(boxsize.prod())
  norm = np.float32(nx * ny * nz)**2
  kbins = kedges[:-1] + (kedges[1:] - kedges[:-1]) / 2
  return kbins, power / norm
