"""Piecewise-linear interpolant with the reference's quirks
(flowpm/scipy/interpolate.py:6-46), float64 host.

The reference picks the nearest node ``ind`` (clipped to [1, n-2]) and then
``clip_by_value(x, xp[-2], xp[1])`` with min > max, which always evaluates to
xp[-2]: the segment used is therefore always [ind, ind+1] (extrapolated when
x lies left of its nearest node) and the result is NaN when ind = n-2.
"""
import numpy as np
import torch


def interp_tf(x, xp, fp):
  """x and xp: arrays (host); fp: an array, or a torch tensor (the result is
  then a float64 tensor, differentiable w.r.t. fp, as when fp is the power
  spectrum of a cosmology with parameters that require grad)."""
  x = np.asarray(x, np.float64)
  xp = np.asarray(xp, np.float64)
  shape = x.shape
  x = x.reshape(-1)
  # nearest node by bisection (same argmin as the reference's squared
  # distance for sorted xp; ties go to the lower node)
  r = np.clip(np.searchsorted(xp, x), 1, len(xp) - 1)
  ind = np.where(np.abs(x - xp[r - 1]) <= np.abs(xp[r] - x), r - 1, r)
  ind = np.clip(ind, 1, len(xp) - 2)
  right = np.where(ind == len(xp) - 2, 0, 1)  # sign(xp[-2] - xp[ind])
  if isinstance(fp, torch.Tensor):
    fp = fp.to(torch.float64).reshape(-1)
    i0 = torch.from_numpy(ind)
    i1 = torch.from_numpy(ind + right)
    dx = torch.from_numpy(xp[ind + right] - xp[ind]).to(fp.device)
    slope = (fp[i1] - fp[i0]) / dx
    xt = torch.from_numpy(x).to(fp.device)
    xi = torch.from_numpy(xp[ind]).to(fp.device)
    return (slope * xt + (fp[i0] - slope * xi)).reshape(shape)
  fp = np.asarray(fp, np.float64)
  with np.errstate(invalid="ignore", divide="ignore"):
    slope = (fp[ind + right] - fp[ind]) / (xp[ind + right] - xp[ind])
  return (slope * x + (fp[ind] - slope * xp[ind])).reshape(shape)
