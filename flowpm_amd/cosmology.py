"""Cosmological parameter record (flowpm/cosmology.py:5-142).

Parameters are kept as Python floats (float64) on the host: they only feed
the host-side growth / ODE coefficients, which this implementation evaluates
in float64.  ``_workspace`` caches the growth table like the reference
(tfbackground.py:631-647).

A parameter may be given as a one-element torch tensor that requires grad
(``Planck15(Omega_c=torch.tensor(0.2589, requires_grad=True))``, the
pattern of scripts/check_grads.py:71-76): the record keeps its value as a
float, and the quantities computed from the record (the linear power
spectrum, the LPT and time-step coefficients, the force factor) become
differentiable functions of it.  Their derivatives w.r.t. the parameters
are central differences of the float64 host functions (``host_fn``), which
the GPU adjoints chain with the cotangents of those quantities.
"""
from functools import partial

import numpy as np
import torch

__all__ = ["Cosmology", "Planck15", "host_fn", "grad_params"]

_NAMES = ("Omega_c", "Omega_b", "h", "n_s", "sigma8", "Omega_k", "w0", "wa")


class Cosmology:

  def __init__(self, Omega_c, Omega_b, h, n_s, sigma8, Omega_k, w0, wa):
    vals = dict(Omega_c=Omega_c, Omega_b=Omega_b, h=h, n_s=n_s,
                sigma8=sigma8, Omega_k=Omega_k, w0=w0, wa=wa)
    # parameters given as tensors that require grad (name -> tensor)
    self._grad = {}
    for name in _NAMES:
      v = vals[name]
      if isinstance(v, torch.Tensor):
        if v.numel() != 1:
          raise ValueError("cosmological parameter %s must be a scalar" %
                           name)
        if v.requires_grad:
          self._grad[name] = v
        v = v.detach().reshape(()).item()
      setattr(self, "_" + name, float(v))
    self._flags = {}
    self._workspace = {}

  def grad_params(self):
    """The parameters given as tensors that require grad (name -> tensor),
    in the order of the constructor's arguments."""
    return dict(self._grad)

  def replace(self, **kw):
    """A record of plain floats with some parameters changed."""
    d = self.to_dict()
    d.update({k: float(v) for k, v in kw.items()})
    return Cosmology(**d)

  def to_dict(self):
    return {
        "Omega_c": self.Omega_c, "Omega_b": self.Omega_b, "h": self.h,
        "n_s": self.n_s, "sigma8": self.sigma8, "Omega_k": self.Omega_k,
        "w0": self.w0, "wa": self.wa
    }

  def __str__(self):
    return ("Cosmological parameters: \n    h:        %s \n    Omega_b:  %s \n"
            "    Omega_c:  %s \n    Omega_k:  %s \n    w0:       %s \n"
            "    wa:       %s \n    n:        %s \n    sigma8:   %s" %
            (self.h, self.Omega_b, self.Omega_c, self.Omega_k, self.w0,
             self.wa, self.n_s, self.sigma8))

  __repr__ = __str__

  @property
  def Omega(self):
    return 1.0 - self._Omega_k

  @property
  def Omega_b(self):
    return self._Omega_b

  @property
  def Omega_c(self):
    return self._Omega_c

  @property
  def Omega_m(self):
    return self._Omega_b + self._Omega_c

  @property
  def Omega_de(self):
    return self.Omega - self.Omega_m

  @property
  def Omega_k(self):
    return self._Omega_k

  @property
  def k(self):
    if self.Omega > 1.0:
      return 1.0
    if self.Omega == 1.0:
      return 0
    return -1.0

  @property
  def sqrtk(self):
    return abs(self._Omega_k)**0.5

  @property
  def h(self):
    return self._h

  @property
  def w0(self):
    return self._w0

  @property
  def wa(self):
    return self._wa

  @property
  def n_s(self):
    return self._n_s

  @property
  def sigma8(self):
    return self._sigma8


# Planck 2015 paper XII Table 4 final column (cosmology.py:132-142)
Planck15 = partial(Cosmology, Omega_c=0.2589, Omega_b=0.04860, Omega_k=0.0,
                   h=0.6774, n_s=0.9667, sigma8=0.8159, w0=-1.0, wa=0.0)


class _HostFn(torch.autograd.Function):
  """fn(cosmo) (float64, any shape) with its vector-Jacobian product w.r.t.
  the record's grad parameters by central differences of fn."""

  @staticmethod
  def forward(ctx, fn, cosmo, names, *params):
    ctx.fn, ctx.cosmo, ctx.names = fn, cosmo, names
    ctx.meta = [(p.dtype, p.device, p.shape) for p in params]
    return torch.from_numpy(np.array(fn(cosmo), dtype=np.float64))

  @staticmethod
  def backward(ctx, g):
    g = g.detach().to("cpu", torch.float64).numpy()
    out = []
    for name, (dt, dev, shape) in zip(ctx.names, ctx.meta):
      v = getattr(ctx.cosmo, "_" + name)
      h = 1e-6 * max(abs(v), 1e-2)
      fp = np.asarray(ctx.fn(ctx.cosmo.replace(**{name: v + h})), np.float64)
      fm = np.asarray(ctx.fn(ctx.cosmo.replace(**{name: v - h})), np.float64)
      d = float(np.sum(g * (fp - fm)) / (2.0 * h))
      out.append(torch.tensor(d, dtype=dt, device=dev).reshape(shape))
    return (None, None, None) + tuple(out)


def grad_params(cosmo):
  """The parameters a cosmology record holds as tensors that require grad
  (empty for records of plain floats, including other parameter
  objects)."""
  return cosmo.grad_params() if isinstance(cosmo, Cosmology) else {}


def host_fn(fn, cosmo):
  """fn(cosmo), a float64 host computation from the record (numpy array or
  scalar), as a float64 CPU tensor; differentiable w.r.t. the parameters
  the record holds as tensors requiring grad (central differences of fn,
  relative step 1e-6: the host functions are smooth in the parameters)."""
  gp = grad_params(cosmo)
  if not gp or not torch.is_grad_enabled():
    return torch.from_numpy(np.array(fn(cosmo), dtype=np.float64))
  names = tuple(gp)
  return _HostFn.apply(fn, cosmo, names, *[gp[n] for n in names])
