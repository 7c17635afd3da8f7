"""Cosmological parameter record (flowpm/cosmology.py:5-142).

Parameters are kept as Python floats (float64) on the host: they only feed
the host-side growth / ODE coefficients, which this implementation evaluates
in float64.  ``_workspace`` caches the growth table like the reference
(tfbackground.py:631-647).
"""
from functools import partial

__all__ = ["Cosmology", "Planck15"]


class Cosmology:

  def __init__(self, Omega_c, Omega_b, h, n_s, sigma8, Omega_k, w0, wa):
    self._Omega_c = float(Omega_c)
    self._Omega_b = float(Omega_b)
    self._h = float(h)
    self._n_s = float(n_s)
    self._sigma8 = float(sigma8)
    self._Omega_k = float(Omega_k)
    self._w0 = float(w0)
    self._wa = float(wa)
    self._flags = {}
    self._workspace = {}

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
