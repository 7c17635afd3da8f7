"""Linear matter power spectrum, host float64 (flowpm/tfpower.py).

Eisenstein & Hu (1998) transfer function in its "eisenhu_osc" form
(tfpower.py:28-159), primordial k^n_s (:9-25), sigma_8 normalisation via the
Simpson rule over x in [log10 kmin, log10 kmax] with k = e^x (:193-218, the
reference's integration range is kept as is), growth D1(a)^2 (:162-190).
These feed the initial-condition amplitude sqrt(P(k)/V) of linear_field.
"""
import numpy as np

from . import constants as const
from . import tfbackground as bkgrd
from .cosmology import grad_params, host_fn
from .scipy.integrate import simps

__all__ = ["primordial_matter_power", "Eisenstein_Hu", "linear_matter_power",
           "sigmasqr"]


def primordial_matter_power(cosmo, k):
  return np.asarray(k, np.float64)**cosmo.n_s


def Eisenstein_Hu(cosmo, k, type="eisenhu_osc"):
  k = np.asarray(k, np.float64)
  T_2_7_sqr = (const.tcmb / 2.7)**2
  h2 = cosmo.h**2
  w_m = cosmo.Omega_m * h2
  w_b = cosmo.Omega_b * h2
  fb = cosmo.Omega_b / cosmo.Omega_m
  fc = (cosmo.Omega_m - cosmo.Omega_b) / cosmo.Omega_m
  k_eq = 7.46e-2 * w_m / T_2_7_sqr / cosmo.h  # Eq. (3) [h/Mpc]
  z_eq = 2.50e4 * w_m / T_2_7_sqr**2  # Eq. (2)
  b1 = 0.313 * w_m**-0.419 * (1.0 + 0.607 * w_m**0.674)  # Eq. (4)
  b2 = 0.238 * w_m**0.223
  z_d = (1291.0 * w_m**0.251 / (1.0 + 0.659 * w_m**0.828) *
         (1.0 + b1 * w_b**b2))
  R_d = 31.5 * w_b / T_2_7_sqr**2 * (1.0e3 / z_d)  # Eq. (5)
  R_eq = 31.5 * w_b / T_2_7_sqr**2 * (1.0e3 / z_eq)
  sh_d = (2.0 / (3.0 * k_eq) * np.sqrt(6.0 / R_eq) *  # Eq. (6)
          np.log((np.sqrt(1.0 + R_d) + np.sqrt(R_eq + R_d)) /
                 (1.0 + np.sqrt(R_eq))))
  k_silk = (1.6 * w_b**0.52 * w_m**0.73 * (1.0 + (10.4 * w_m)**-0.95) /
            cosmo.h)  # Eq. (7)
  if type == "eisenhu":
    alpha_gamma = (1.0 - 0.328 * np.log(431.0 * w_m) * w_b / w_m +
                   0.38 * np.log(22.3 * w_m) * fb**2)
    gamma_eff = cosmo.Omega_m * cosmo.h * (alpha_gamma + (1.0 - alpha_gamma) /
                                           (1.0 + (0.43 * k * sh_d)**4))
    q = k * (const.tcmb / 2.7)**2 / gamma_eff
    L = np.log(2.0 * np.e + 1.8 * q)
    C = 14.2 + 731.0 / (1.0 + 62.5 * q)
    return L / (L + C * q * q)
  if type != "eisenhu_osc":
    raise NotImplementedError(type)
  a1 = (46.9 * w_m)**0.670 * (1.0 + (32.1 * w_m)**-0.532)  # Eq. (11)
  a2 = (12.0 * w_m)**0.424 * (1.0 + (45.0 * w_m)**-0.582)
  alpha_c = a1**(-fb) * a2**(-(fb**3))
  bc1 = 0.944 / (1.0 + (458.0 * w_m)**-0.708)  # Eq. (12)
  bc2 = (0.395 * w_m)**-0.0266
  beta_c = 1.0 / (1.0 + bc1 * (fc**bc2 - 1.0))

  def T_tilde(k1, alpha, beta):  # Eq. (10), (19)
    q = k1 / (13.41 * k_eq)
    L = np.log(np.e + 1.8 * beta * q)
    C = 14.2 / alpha + 386.0 / (1.0 + 69.9 * q**1.08)
    return L / (L + C * q * q)

  f = 1.0 / (1.0 + (k * sh_d / 5.4)**4)  # Eq. (17), (18)
  Tc = f * T_tilde(k, 1.0, beta_c) + (1.0 - f) * T_tilde(k, alpha_c, beta_c)
  y = (1.0 + z_eq) / (1.0 + z_d)  # Eq. (14), (15)
  x = np.sqrt(1.0 + y)
  G_EH98 = y * (-6.0 * x + (2.0 + 3.0 * y) * np.log((x + 1.0) / (x - 1.0)))
  alpha_b = 2.07 * k_eq * sh_d * (1.0 + R_d)**-0.75 * G_EH98
  beta_node = 8.41 * w_m**0.435  # Eq. (22), (23)
  tilde_s = sh_d / (1.0 + (beta_node / (k * sh_d))**3)**(1.0 / 3.0)
  beta_b = 0.5 + fb + (3.0 - 2.0 * fb) * np.sqrt((17.2 * w_m)**2 + 1.0)
  Tb = (T_tilde(k, 1.0, 1.0) / (1.0 + (k * sh_d / 5.2)**2) + alpha_b /
        (1.0 + (beta_b / (k * sh_d))**3) * np.exp(-(k / k_silk)**1.4)) * (
            np.sin(k * tilde_s) / (k * tilde_s + 1e-9))  # Eq. (21)
  return fb * Tb + fc * Tc


def sigmasqr(cosmo, R, transfer_fn, kmin=0.0001, kmax=1000.0, ksteps=5,
             **kwargs):

  def int_sigma(logk):
    k = np.exp(logk)
    x = k * R
    w = 3.0 * (np.sin(x) - x * np.cos(x)) / (x * x * x)
    pk = transfer_fn(cosmo, k, **kwargs)**2 * primordial_matter_power(
        cosmo, k)
    return k * (k * w)**2 * pk

  y = simps(int_sigma, np.log(kmin) / np.log(10), np.log(kmax) / np.log(10),
            256)
  return 1.0 / (2.0 * np.pi**2.0) * y


def _linear_matter_power(cosmo, k, a, transfer_fn, **kwargs):
  g = bkgrd.D1(cosmo, a)
  t = transfer_fn(cosmo, k, **kwargs)
  pknorm = cosmo.sigma8**2 / sigmasqr(cosmo, 8.0, transfer_fn, **kwargs)
  pk = primordial_matter_power(cosmo, k) * t**2 * g**2
  return np.squeeze(pk * pknorm)


def linear_matter_power(cosmo, k, a=1.0, transfer_fn=Eisenstein_Hu, **kwargs):
  """P(k) at a (tfpower.py:162-190): a float64 numpy array, or, when the
  cosmology holds parameters as tensors requiring grad, a float64 CPU
  tensor differentiable w.r.t. them."""
  k = np.asarray(k, np.float64)
  if grad_params(cosmo):
    return host_fn(
        lambda c: _linear_matter_power(c, k, a, transfer_fn, **kwargs), cosmo)
  return _linear_matter_power(cosmo, k, a, transfer_fn, **kwargs)
