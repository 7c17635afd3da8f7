"""Oracle: CPU restatement of FlowPM's growth / background coefficients.

TEST INFRASTRUCTURE ONLY.  Nothing in ``flowpm_amd`` may import this module;
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as a checker.

Follows flowpm/tfbackground.py of the reference step for step, but in a
selectable floating-point type (float64 by default, float32 to mimic the
reference's own precision):

* ``fde``            tfbackground.py:10-51   (``log(a - 1e-5)``)
* ``E``              tfbackground.py:85-120
* ``dfde``           tfbackground.py:143-176
* ``dEa``            tfbackground.py:179-212
* ``Omega_m_a``      tfbackground.py:215-242
* ``Omega_de_a``     tfbackground.py:245-274
* ``growth_ode``     tfbackground.py:531-577
* ``rk4``            tfbackground.py:490-528 (RK4 in ``a``)
* ``odesolve``       tfbackground.py:580-628 (matter-dominated y0, normalised
                     to the last grid point, F1p/F2p from the ODE rhs)
* table / interp     tfbackground.py:631-647 (``np.logspace(-2, 0, 256)``) and
                     the ``tfp.math.interp_regular_1d_grid`` calls in ln a
                     (tfbackground.py:650-977), constant extension outside.

Pinned by tests/test_oracle_background.py: the Einstein-de Sitter closed forms
(D1 = a, D2 -> a^2 normalised, f1 = 1, f2 = 2, Gf = a^1.5, gf = 1.5 a^0.5),
D1(a=0.01) against the ratio of the two stored Planck15 P(k) tables, and an
independent scipy ``solve_ivp`` integration in ln a of the same ODE.
"""
import numpy as np

__all__ = ["Cosmo", "planck15", "GrowthTable"]


class Cosmo:
  """Plain parameter record (reference: flowpm/cosmology.py:5-125)."""

  def __init__(self, Omega_c, Omega_b, h, n_s, sigma8, Omega_k=0.0, w0=-1.0,
               wa=0.0):
    self.Omega_c = float(Omega_c)
    self.Omega_b = float(Omega_b)
    self.h = float(h)
    self.n_s = float(n_s)
    self.sigma8 = float(sigma8)
    self.Omega_k = float(Omega_k)
    self.w0 = float(w0)
    self.wa = float(wa)

  @property
  def Omega_m(self):  # cosmology.py:82-83
    return self.Omega_b + self.Omega_c

  @property
  def Omega_de(self):  # cosmology.py:86-87 with Omega = 1 - Omega_k (:66-69)
    return (1.0 - self.Omega_k) - self.Omega_m


def planck15(**over):
  """Planck15 preset, cosmology.py:132-142."""
  p = dict(Omega_c=0.2589, Omega_b=0.04860, Omega_k=0.0, h=0.6774,
           n_s=0.9667, sigma8=0.8159, w0=-1.0, wa=0.0)
  p.update(over)
  return Cosmo(**p)


def _fde(c, a, dt, eps=1e-5):
  a = np.asarray(a, dt)
  return (dt(-3.0) * (dt(1.0) + dt(c.w0)) + dt(3.0) * dt(c.wa) *
          ((a - dt(1.0)) / np.log(a - dt(eps)) - dt(1.0)))


def _E(c, a, dt):
  a = np.asarray(a, dt)
  return np.sqrt(dt(c.Omega_m) / a**3 + dt(c.Omega_k) / a**2 +
                 dt(c.Omega_de) * a**_fde(c, a, dt))


def _dfde(c, a, dt, eps=1e-5):
  a = np.asarray(a, dt)
  la = np.log(a - dt(eps))
  return dt(3.0) * dt(c.wa) * (la - (a - dt(1.0)) / (a - dt(eps))) / la**2


def _dEa(c, a, dt):
  a = np.asarray(a, dt)
  return dt(0.5) * (dt(-3.0) * dt(c.Omega_m) / a**4 - dt(2.0) * dt(c.Omega_k) /
                    a**3 + _dfde(c, a, dt) * dt(c.Omega_de) *
                    a**_fde(c, a, dt)) / _E(c, a, dt)


def _Om_a(c, a, dt):
  a = np.asarray(a, dt)
  return dt(c.Omega_m) * a**dt(-3.0) / _E(c, a, dt)**2


def _Ode_a(c, a, dt):
  a = np.asarray(a, dt)
  return dt(c.Omega_de) * a**_fde(c, a, dt) / _E(c, a, dt)**2


def growth_rhs(c, a, y, dt=np.float64):
  """tfbackground.py:560-577: y = (d1, d1', d2, d2') with ' = d/da."""
  d1, d1f, d2, d2f = y
  om = _Om_a(c, a, dt)
  ode = _Ode_a(c, a, dt)
  a = dt(a)
  h = ode - dt(0.5) * om + dt(2.0)
  r1 = dt(1.5) * om * d1 / a**2 - (d1f / a) * h
  r2 = dt(1.5) * om * d2 / a**2 - (d2f / a) * h - dt(1.5) * (om * d1**2) / a**2
  return np.array([d1f, r1, d2f, r2], dtype=dt)


class GrowthTable:
  """ODE table + interpolators (tfbackground.py:580-977)."""

  def __init__(self, cosmo, dtype=np.float64, log10_amin=-2.0, steps=256):
    dt = self.dt = dtype
    self.c = cosmo
    a = np.logspace(log10_amin, 0.0, steps).astype(dt)  # :641-642
    y0 = np.array([a[0], 1.0, -3.0 / 7 * a[0]**2, -6.0 / 7 * a[0]], dt)  # :601
    ys = np.empty((steps, 4), dt)
    ys[0] = y0
    yt = y0
    for i in range(steps - 1):  # rk4_integration :506-528, RK4 in a
      t = a[i]
      h = a[i + 1] - a[i]
      k1 = growth_rhs(cosmo, t, yt, dt)
      k2 = growth_rhs(cosmo, t + h / dt(2), yt + h * k1 / dt(2), dt)
      k3 = growth_rhs(cosmo, t + h / dt(2), yt + h * k2 / dt(2), dt)
      k4 = growth_rhs(cosmo, t + h, yt + h * k3, dt)
      yt = yt + h * (k1 + dt(2) * k2 + dt(2) * k3 + k4) / dt(6)
      ys[i + 1] = yt
    second = np.stack([growth_rhs(cosmo, a[i], ys[i], dt)
                       for i in range(steps)])  # :609
    n1, n2 = ys[-1, 0], ys[-1, 2]  # normalise to the last point :613-618
    self.a = a
    self.tab = {
        "D1": ys[:, 0] / n1, "D1f": ys[:, 1] / n1, "F1p": second[:, 1] / n1,
        "D2": ys[:, 2] / n2, "D2f": ys[:, 3] / n2, "F2p": second[:, 3] / n2,
    }
    self.x0 = np.log(a[0])
    self.x1 = np.log(a[-1])

  def _interp(self, name, a):
    """tfp.math.interp_regular_1d_grid in ln a, constant extension."""
    y = self.tab[name]
    n = len(y)
    x = np.log(np.asarray(a, self.dt))
    u = (x - self.x0) / (self.x1 - self.x0) * (n - 1)
    u = np.clip(u, 0, n - 1)
    i = np.clip(np.floor(u).astype(np.int64), 0, n - 2)
    w = u - i
    return (1 - w) * y[i] + w * y[i + 1]

  # tfbackground.py:650-977
  def D1(self, a): return self._interp("D1", a)
  def D2(self, a): return self._interp("D2", a)
  def D1f(self, a): return self._interp("D1f", a)
  def D2f(self, a): return self._interp("D2f", a)
  def E(self, a): return _E(self.c, a, self.dt)
  def dEa(self, a): return _dEa(self.c, a, self.dt)
  def Omega_m_a(self, a): return _Om_a(self.c, a, self.dt)

  def f1(self, a):
    return self.D1f(a) * np.asarray(a, self.dt) / self.D1(a)

  def f2(self, a):
    return self.D2f(a) * np.asarray(a, self.dt) / self.D2(a)

  def Gf(self, a):
    a = np.asarray(a, self.dt)
    return self.D1f(a) * a**3 * self.E(a)

  def Gf2(self, a):
    a = np.asarray(a, self.dt)
    return self.D2f(a) * a**3 * self.E(a)

  def gf(self, a):
    a = np.asarray(a, self.dt)
    return (self._interp("F1p", a) * a**3 * self.E(a) +
            self.D1f(a) * a**3 * self.dEa(a) +
            3 * a**2 * self.E(a) * self.D1f(a))

  def gf2(self, a):
    a = np.asarray(a, self.dt)
    return (self._interp("F2p", a) * a**3 * self.E(a) +
            self.D2f(a) * a**3 * self.dEa(a) +
            3 * a**2 * self.E(a) * self.D2f(a))
