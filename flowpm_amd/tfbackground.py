"""Background and growth coefficients, host float64 (flowpm/tfbackground.py).

Same API as the reference (``E(cosmo, a)``, ``D1(cosmo, a)``, ...), same
algorithm -- the growth ODE integrated with RK4 in ``a`` on
``np.logspace(-2, 0, 256)``, normalised to the last node, then linearly
interpolated in ln a -- but evaluated in float64 instead of TensorFlow
float32.  These scalars set the kick / drift / LPT coefficients; the particle
and mesh arithmetic stays float32 on the GPU.

Reference map: fde :10-51, w :54-82, E :85-120, H :123-140, dfde :143-176,
dEa :179-212, Omega_m_a :215-242, Omega_de_a :245-274, rk4_integration
:490-528, growth_ode :531-577, odesolve_func :580-628, maybe_compute_ODE
:631-647, D1/D2/D1f/D2f :650-784, f1/f2 :787-844, Gf/Gf2 :847-899,
gf/gf2 :902-977.
"""
import numpy as np

from . import constants

__all__ = [
    "fde", "w", "E", "H", "dfde", "dEa", "Omega_m_a", "Omega_de_a",
    "rk4_integration", "growth_ode", "odesolve_func", "maybe_compute_ODE",
    "D1", "D2", "D1f", "D2f", "f1", "f2", "Gf", "Gf2", "gf", "gf2"
]


def _a(a):
  return np.asarray(a, dtype=np.float64)


def fde(cosmo, a, epsilon=1e-5):
  a = _a(a)
  return (-3.0 * (1.0 + cosmo.w0) + 3.0 * cosmo.wa *
          ((a - 1.0) / np.log(a - epsilon) - 1.0))


def w(cosmo, a):
  return cosmo.w0 + cosmo.wa * (1.0 - _a(a))


def E(cosmo, a):
  a = _a(a)
  return np.sqrt(cosmo.Omega_m / a**3 + cosmo.Omega_k / a**2 +
                 cosmo.Omega_de * a**fde(cosmo, a))


def H(cosmo, a):
  return constants.H0 * cosmo.h * E(cosmo, a)


def dfde(cosmo, a, epsilon=1e-5):
  a = _a(a)
  la = np.log(a - epsilon)
  return 3 * cosmo.wa * (la - (a - 1) / (a - epsilon)) / la**2


def dEa(cosmo, a):
  a = _a(a)
  return 0.5 * (-3 * cosmo.Omega_m / a**4 - 2 * cosmo.Omega_k / a**3 +
                dfde(cosmo, a) * cosmo.Omega_de * a**fde(cosmo, a)) / E(
                    cosmo, a)


def Omega_m_a(cosmo, a):
  a = _a(a)
  return cosmo.Omega_m * a**-3 / E(cosmo, a)**2


def Omega_de_a(cosmo, a):
  a = _a(a)
  return cosmo.Omega_de * a**fde(cosmo, a) / E(cosmo, a)**2


def growth_ode(a, y, cosmo):
  """d/da of y = (D1, D1', D2, D2') (rows of y may be batched)."""
  y = np.asarray(y, np.float64)
  a = _a(a)
  d1, d1f, d2, d2f = y[..., 0], y[..., 1], y[..., 2], y[..., 3]
  om = Omega_m_a(cosmo, a)
  friction = Omega_de_a(cosmo, a) - 0.5 * om + 2
  dd1 = 1.5 * om * d1 / a**2 - (d1f / a) * friction
  dd2 = 1.5 * om * d2 / a**2 - (d2f / a) * friction - 1.5 * (om *
                                                              d1**2) / a**2
  return np.stack([d1f, dd1, d2f, dd2], axis=-1)


def rk4_integration(dydt, a, y0, cosmo):
  a = _a(a)
  y = np.empty((len(a), len(y0)))
  y[0] = y0
  for i in range(len(a) - 1):
    t, h = a[i], a[i + 1] - a[i]
    k1 = dydt(t, y[i], cosmo)
    k2 = dydt(t + h / 2.0, y[i] + h * k1 / 2.0, cosmo)
    k3 = dydt(t + h / 2.0, y[i] + h * k2 / 2.0, cosmo)
    k4 = dydt(t + h, y[i] + h * k3, cosmo)
    y[i + 1] = y[i] + h * (k1 + 2.0 * k2 + 2.0 * k3 + k4) / 6.0
  return y


def _rk4_growth(a, y0, cosmo):
  """rk4_integration(growth_ode, a, y0, cosmo) with the equation's
  a-dependent coefficients evaluated once for every stage point (the nodes
  and the midpoints) in vectorised float64, the stepping itself on Python
  floats: the same RK4 to float64 rounding, ~10x faster than evaluating the
  background functions per stage (the gradients w.r.t. cosmological
  parameters solve it twice per parameter)."""
  n = len(a)
  mid = a[:-1] + (a[1:] - a[:-1]) / 2.0
  pts = np.concatenate([a, mid])
  om = Omega_m_a(cosmo, pts)
  fr = Omega_de_a(cosmo, pts) - 0.5 * om + 2
  c1 = (1.5 * om / pts**2).tolist()  # d(D')/da = c1 D - c2 D' (- c1 D1^2)
  c2 = (fr / pts).tolist()
  av = a.tolist()
  y = np.empty((n, 4))
  d1, p1, d2, p2 = (float(v) for v in y0)
  y[0] = (d1, p1, d2, p2)
  for i in range(n - 1):
    h = av[i + 1] - av[i]
    j0, jm, j1 = i, n + i, i + 1

    def f(j, d1, p1, d2, p2):
      c, k = c1[j], c2[j]
      return p1, c * d1 - p1 * k, p2, c * d2 - p2 * k - c * d1 * d1

    k1 = f(j0, d1, p1, d2, p2)
    k2 = f(jm, d1 + h * k1[0] / 2.0, p1 + h * k1[1] / 2.0,
           d2 + h * k1[2] / 2.0, p2 + h * k1[3] / 2.0)
    k3 = f(jm, d1 + h * k2[0] / 2.0, p1 + h * k2[1] / 2.0,
           d2 + h * k2[2] / 2.0, p2 + h * k2[3] / 2.0)
    k4 = f(j1, d1 + h * k3[0], p1 + h * k3[1], d2 + h * k3[2],
           p2 + h * k3[3])
    d1 = d1 + h * (k1[0] + 2.0 * k2[0] + 2.0 * k3[0] + k4[0]) / 6.0
    p1 = p1 + h * (k1[1] + 2.0 * k2[1] + 2.0 * k3[1] + k4[1]) / 6.0
    d2 = d2 + h * (k1[2] + 2.0 * k2[2] + 2.0 * k3[2] + k4[2]) / 6.0
    p2 = p2 + h * (k1[3] + 2.0 * k2[3] + 2.0 * k3[3] + k4[3]) / 6.0
    y[i + 1] = (d1, p1, d2, p2)
  return y


def odesolve_func(a, cosmo):
  a = _a(a)
  # matter-dominated initial conditions (tfbackground.py:601)
  y0 = np.array([a[0], 1.0, -3.0 / 7 * a[0]**2, -6.0 / 7 * a[0]])
  y = _rk4_growth(a, y0, cosmo)
  second = growth_ode(a, y, cosmo)
  n1, n2 = y[-1, 0], y[-1, 2]
  return {
      "a": a,
      "D1": y[:, 0] / n1, "D1f": y[:, 1] / n1, "F1p": second[:, 1] / n1,
      "D2": y[:, 2] / n2, "D2f": y[:, 3] / n2, "F2p": second[:, 3] / n2,
  }


def maybe_compute_ODE(cosmo, log10_amin=-2, steps=256):
  cache = cosmo._workspace.get("cache_ODE")
  if cache is None:
    cache = odesolve_func(np.logspace(log10_amin, 0.0, steps), cosmo)
    cosmo._workspace["cache_ODE"] = cache
  return cache


def _interp(cosmo, a, key):
  """tfp.math.interp_regular_1d_grid on ln a (constant extension)."""
  cache = maybe_compute_ODE(cosmo)
  tab = cache[key]
  x0, x1 = np.log(cache["a"][0]), np.log(cache["a"][-1])
  u = (np.log(_a(a)) - x0) / (x1 - x0) * (len(tab) - 1)
  u = np.clip(u, 0.0, len(tab) - 1)
  i = np.minimum(np.floor(u).astype(np.int64), len(tab) - 2)
  t = u - i
  return tab[i] + t * (tab[i + 1] - tab[i])


def D1(cosmo, a):
  return _interp(cosmo, a, "D1")


def D2(cosmo, a):
  return _interp(cosmo, a, "D2")


def D1f(cosmo, a):
  return _interp(cosmo, a, "D1f")


def D2f(cosmo, a):
  return _interp(cosmo, a, "D2f")


def f1(cosmo, a):
  return D1f(cosmo, a) * _a(a) / D1(cosmo, a)


def f2(cosmo, a):
  return D2f(cosmo, a) * _a(a) / D2(cosmo, a)


def Gf(cosmo, a):
  a = _a(a)
  return D1f(cosmo, a) * a**3 * E(cosmo, a)


def Gf2(cosmo, a):
  a = _a(a)
  return D2f(cosmo, a) * a**3 * E(cosmo, a)


def gf(cosmo, a):
  a = _a(a)
  d1f = D1f(cosmo, a)
  return (_interp(cosmo, a, "F1p") * a**3 * E(cosmo, a) +
          d1f * a**3 * dEa(cosmo, a) + 3 * a**2 * E(cosmo, a) * d1f)


def gf2(cosmo, a):
  a = _a(a)
  d2f = D2f(cosmo, a)
  return (_interp(cosmo, a, "F2p") * a**3 * E(cosmo, a) +
          d2f * a**3 * dEa(cosmo, a) + 3 * a**2 * E(cosmo, a) * d2f)
