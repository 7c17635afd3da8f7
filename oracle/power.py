"""Oracle: CPU restatement of FlowPM's linear power spectrum (input side).

TEST INFRASTRUCTURE ONLY (see oracle/background.py header).

* ``primordial``      flowpm/tfpower.py:9-25
* ``eisenstein_hu``   flowpm/tfpower.py:28-159 ("eisenhu_osc", the
                      ``sin(k s)/(k s + 1e-9)`` form at :152-153)
* ``sigmasqr``        flowpm/tfpower.py:193-218 with ``simps`` from
                      flowpm/scipy/integrate.py:4-36.  Quirk kept: the
                      integration variable runs over [log10 kmin, log10 kmax]
                      while the integrand evaluates k = exp(x).
* ``linear_matter_power`` flowpm/tfpower.py:162-190
* ``interp_tf``       flowpm/scipy/interpolate.py:6-46, including the clip with
                      min > max, which always picks the right-hand neighbour
                      and yields NaN when the nearest point is xp[-2].

Pinned by tests/test_oracle_power.py against the reference's stored nbodykit
Eisenstein-Hu tables (flowpm/data/Planck15_a1p00.txt), at the reference's own
tolerance rtol=1.2e-2 (tests/test_tfpower.py:28).
"""
import numpy as np

TCMB = 2.726  # flowpm/constants.py:12


def primordial(c, k):
  return np.asarray(k, np.float64)**c.n_s


def eisenstein_hu(c, k):
  k = np.asarray(k, np.float64)
  T2 = (TCMB / 2.7)**2
  h2 = c.h**2
  w_m = c.Omega_m * h2
  w_b = c.Omega_b * h2
  fb = c.Omega_b / c.Omega_m
  fc = (c.Omega_m - c.Omega_b) / c.Omega_m
  k_eq = 7.46e-2 * w_m / T2 / c.h
  z_eq = 2.50e4 * w_m / T2**2
  b1 = 0.313 * w_m**-0.419 * (1.0 + 0.607 * w_m**0.674)
  b2 = 0.238 * w_m**0.223
  z_d = 1291.0 * w_m**0.251 / (1.0 + 0.659 * w_m**0.828) * (1.0 + b1 * w_b**b2)
  R_d = 31.5 * w_b / T2**2 * (1.0e3 / z_d)
  R_eq = 31.5 * w_b / T2**2 * (1.0e3 / z_eq)
  sh_d = (2.0 / (3.0 * k_eq) * np.sqrt(6.0 / R_eq) *
          np.log((np.sqrt(1.0 + R_d) + np.sqrt(R_eq + R_d)) /
                 (1.0 + np.sqrt(R_eq))))
  k_silk = (1.6 * w_b**0.52 * w_m**0.73 * (1.0 + (10.4 * w_m)**-0.95) / c.h)
  a1 = (46.9 * w_m)**0.670 * (1.0 + (32.1 * w_m)**-0.532)
  a2 = (12.0 * w_m)**0.424 * (1.0 + (45.0 * w_m)**-0.582)
  alpha_c = a1**(-fb) * a2**(-(fb**3))
  bb1 = 0.944 / (1.0 + (458.0 * w_m)**-0.708)
  bb2 = (0.395 * w_m)**-0.0266
  beta_c = 1.0 / (1.0 + bb1 * (fc**bb2 - 1.0))

  def t_tilde(k1, alpha, beta):
    q = k1 / (13.41 * k_eq)
    L = np.log(np.e + 1.8 * beta * q)
    C = 14.2 / alpha + 386.0 / (1.0 + 69.9 * q**1.08)
    return L / (L + C * q * q)

  f = 1.0 / (1.0 + (k * sh_d / 5.4)**4)
  Tc = f * t_tilde(k, 1.0, beta_c) + (1.0 - f) * t_tilde(k, alpha_c, beta_c)
  y = (1.0 + z_eq) / (1.0 + z_d)
  x = np.sqrt(1.0 + y)
  G = y * (-6.0 * x + (2.0 + 3.0 * y) * np.log((x + 1.0) / (x - 1.0)))
  alpha_b = 2.07 * k_eq * sh_d * (1.0 + R_d)**-0.75 * G
  beta_node = 8.41 * w_m**0.435
  tilde_s = sh_d / (1.0 + (beta_node / (k * sh_d))**3)**(1.0 / 3.0)
  beta_b = 0.5 + fb + (3.0 - 2.0 * fb) * np.sqrt((17.2 * w_m)**2 + 1.0)
  Tb = (t_tilde(k, 1.0, 1.0) / (1.0 + (k * sh_d / 5.2)**2) + alpha_b /
        (1.0 + (beta_b / (k * sh_d))**3) * np.exp(-(k / k_silk)**1.4)) * (
            np.sin(k * tilde_s) / (k * tilde_s + 1e-9))
  return fb * Tb + fc * Tc


def simps(f, a, b, N=128):
  if N % 2 == 1:
    raise ValueError("N must be an even integer.")
  dx = (b - a) / N
  x = np.linspace(a, b, N + 1)
  y = f(x)
  return dx / 3 * np.sum(y[0:-1:2] + 4 * y[1::2] + y[2::2], axis=0)


def sigmasqr(c, R, kmin=0.0001, kmax=1000.0):
  def integrand(logk):
    k = np.exp(logk)
    x = k * R
    w = 3.0 * (np.sin(x) - x * np.cos(x)) / (x * x * x)
    pk = eisenstein_hu(c, k)**2 * primordial(c, k)
    return k * (k * w)**2 * pk
  y = simps(integrand, np.log(kmin) / np.log(10), np.log(kmax) / np.log(10),
            256)
  return 1.0 / (2.0 * np.pi**2.0) * y


def linear_matter_power(c, k, growth=None, a=1.0):
  """P(k, a) = k^ns T(k)^2 D1(a)^2 sigma8^2 / sigma^2(8)."""
  g = 1.0 if growth is None else float(growth.D1(a))
  t = eisenstein_hu(c, k)
  norm = c.sigma8**2 / sigmasqr(c, 8.0)
  return primordial(c, k) * t**2 * g**2 * norm


def interp_tf(x, xp, fp):
  x = np.asarray(x, np.float64)
  xp = np.asarray(xp, np.float64)
  fp = np.asarray(fp, np.float64)
  ind = np.argmin((x[:, None] - xp[None, :])**2, axis=-1)
  ind = np.clip(ind, 1, len(xp) - 2)
  xi = xp[ind]
  # tf.clip_by_value(x, xp[-2], xp[1]) == maximum(minimum(x, xp[1]), xp[-2])
  s = np.sign(np.maximum(np.minimum(x, xp[1]), xp[-2]) - xi).astype(np.int64)
  fp0 = fp[ind]
  xp0 = xp[ind]
  with np.errstate(invalid="ignore", divide="ignore"):
    fp1 = fp[ind + np.sign(s)] - fp0
    xp1 = xp[ind + np.sign(s)] - xp0
    a = fp1 / xp1
  b = fp0 - a * xp0
  return a * x + b
