"""Pins oracle/background.py (growth ODE + interpolators, flowpm/tfbackground.py).

Checks that do not restate the oracle's formulas:
* Einstein-de Sitter (Omega_m = 1) closed forms: D1 = a, normalised D2 = a^2,
  f1 = 1, f2 = 2, E = a^-1.5, Gf = a^1.5, gf = 1.5 a^0.5.  Between grid points
  the reference's linear interpolation in ln a has a known error
  h^2/8 (h = ln(100)/255) for D1 and 4x that for D2, which sets the tolerance.
* An independent adaptive integration (scipy solve_ivp, in ln a, rtol 1e-10)
  of the growth equation written in the textbook form
  D'' + (3/a + E'/E) D' = 1.5 Om a^-5 E^-2 D, checked at the table nodes.
* The reference's tolerances against the legacy matter-dominated solver are
  rtol 1e-2 (tests/test_tfbackground.py:55-175); the checks here are tighter.
"""
import numpy as np
from scipy.integrate import solve_ivp

from oracle import background as bg

H = np.log(100.0) / 255


def test_eds_closed_forms():
  c = bg.Cosmo(Omega_c=1.0, Omega_b=0.0, h=0.7, n_s=1.0, sigma8=0.8)
  g = bg.GrowthTable(c)
  a = np.logspace(-2, 0, 97)
  tol1 = H**2 / 8 * 1.05
  np.testing.assert_allclose(g.D1(a), a, rtol=tol1)
  np.testing.assert_allclose(g.D2(a), a**2, rtol=4 * tol1)
  np.testing.assert_allclose(g.f1(a), 1.0, rtol=tol1)
  np.testing.assert_allclose(g.f2(a), 2.0, rtol=4 * tol1)
  np.testing.assert_allclose(g.E(a), a**-1.5, rtol=1e-12)
  np.testing.assert_allclose(g.Gf(a), a**1.5, rtol=tol1)
  np.testing.assert_allclose(g.gf(a), 1.5 * a**0.5, rtol=4 * tol1)
  # exactly on the grid nodes there is no interpolation error
  nodes = np.logspace(-2, 0, 256)
  np.testing.assert_allclose(g.D1(nodes), nodes, rtol=1e-9)


def test_lcdm_against_independent_integration():
  c = bg.planck15()
  g = bg.GrowthTable(c)
  om, ol = c.Omega_m, c.Omega_de

  def E(a):
    return np.sqrt(om / a**3 + ol)

  def dlnE(a):  # dlnE/dlna
    return -1.5 * om / a**3 / E(a)**2

  def rhs(lna, y):
    a = np.exp(lna)
    D, Dp = y  # Dp = dD/dlna
    Omz = om / a**3 / E(a)**2
    return [Dp, -(2 + dlnE(a)) * Dp + 1.5 * Omz * D]

  a0 = 1e-2
  # matter-dominated start, as the reference: D = a, dD/da = 1
  sol = solve_ivp(rhs, [np.log(a0), 0.0], [a0, a0], rtol=1e-11, atol=1e-14,
                  dense_output=True)
  nodes = np.logspace(-2, 0, 256)
  D = sol.sol(np.log(nodes))[0]
  Dn = D / D[-1]
  fn = sol.sol(np.log(nodes))[1] / D
  np.testing.assert_allclose(g.D1(nodes), Dn, rtol=2e-6)
  np.testing.assert_allclose(g.f1(nodes), fn, rtol=2e-5)
  # gf = d(Gf)/da with Gf = a^3 E dD/da (tfbackground.py:847-977)
  Gf = lambda a: a**3 * E(a) * sol.sol(np.log(a))[1] / a / D[-1]
  a = np.array([0.1, 0.3, 0.55, 0.9])
  eps = 1e-6
  gfd = (Gf(a * (1 + eps)) - Gf(a * (1 - eps))) / (2 * eps * a)
  np.testing.assert_allclose(g.gf(a), gfd, rtol=2e-4)


def test_matter_dominated_start_of_d2():
  """Second-order growth in LCDM: -3/7 D1^2 Om^{-1/143} approximation
  (Bouchet et al. 1995) holds to ~1% for D2/D1^2 relative to its a->0 value."""
  c = bg.planck15()
  g = bg.GrowthTable(c)
  a = np.array([0.1, 0.5, 1.0])
  ratio = g.D2(a) / g.D1(a)**2
  # normalised: D2/D1^2 -> 1 at early times; at a=1 it is Om^{-1/143}-ish
  om_a = c.Omega_m / a**3 / (c.Omega_m / a**3 + c.Omega_de)
  approx = om_a**(-1 / 143.) / (om_a[-1]**(-1 / 143.))
  np.testing.assert_allclose(ratio / ratio[-1], approx / approx[-1], rtol=1e-2)


def test_d1_against_reference_pk_tables(golden_dir):
  """D1(0.01)^2 = P(k, 0.01)/P(k, 1) from the reference's stored nbodykit
  tables (flowpm/data).  nbodykit's growth includes radiation, so the tables
  only pin D1 at the 3 % level."""
  k1, p1 = np.loadtxt(golden_dir / "Planck15_a1p00.txt").T
  k0, p0 = np.loadtxt(golden_dir / "Planck15_a0p01.txt").T
  assert np.allclose(k0, k1)
  d_tab = np.sqrt(p0 / p1)
  assert d_tab.std() / d_tab.mean() < 1e-10
  g = bg.GrowthTable(bg.planck15())
  np.testing.assert_allclose(g.D1(0.01), d_tab.mean(), rtol=3e-2)
