"""The product's host-side float64 scalars: flowpm_amd.tfbackground,
flowpm_amd.tfpower and flowpm_amd.scipy (flowpm/tfbackground.py,
flowpm/tfpower.py, flowpm/scipy/*).  They set every kick, drift and LPT
coefficient and the initial P(k), so they are checked directly here, on the
CPU, against the pinned oracle, the Einstein-de Sitter closed forms and the
reference's stored P(k) table (tests/test_tfbackground.py and
tests/test_tfpower.py of the reference check the same functions against
astropy / nbodykit at rtol 1e-2 .. 1e-4)."""
import numpy as np
import pytest

from flowpm_amd import tfbackground as tb
from flowpm_amd import tfpower as tp
from flowpm_amd.cosmology import Cosmology, Planck15
from flowpm_amd.scipy import integrate as fint
from flowpm_amd.scipy import interpolate as fip
from oracle import background as bg
from oracle import power as opw

A = np.array([0.01, 0.0137, 0.05, 0.1, 0.2345, 0.5, 0.77, 0.999, 1.0])


@pytest.mark.parametrize("name", ["E", "dEa", "Omega_m_a", "D1", "D2", "D1f",
                                  "D2f", "f1", "f2", "Gf", "Gf2", "gf", "gf2"])
def test_growth_matches_oracle(name):
  ours = getattr(tb, name)(Planck15(), A)
  ref = getattr(bg.GrowthTable(bg.planck15()), name)(A)
  np.testing.assert_allclose(ours, ref, rtol=1e-10, atol=1e-300)


def test_growth_eds_closed_forms():
  """Omega_m = 1: D1 = a, D2 = a^2 (normalised), f1 = 1, f2 = 2,
  E = a^-1.5, Gf = a^1.5; tolerance from the linear interpolation in ln a
  between the 256 nodes (h^2 / 8 for D1, 4x for D2)."""
  c = Cosmology(Omega_c=1.0, Omega_b=0.0, h=0.7, n_s=1.0, sigma8=0.8,
                Omega_k=0.0, w0=-1.0, wa=0.0)
  h = np.log(100.0) / 255
  tol = h**2 / 8 * 1.05
  a = np.logspace(-2, 0, 97)
  np.testing.assert_allclose(tb.D1(c, a), a, rtol=tol)
  np.testing.assert_allclose(tb.D2(c, a), a**2, rtol=4 * tol)
  np.testing.assert_allclose(tb.f1(c, a), 1.0, rtol=tol)
  np.testing.assert_allclose(tb.f2(c, a), 2.0, rtol=4 * tol)
  np.testing.assert_allclose(tb.E(c, a), a**-1.5, rtol=1e-12)
  np.testing.assert_allclose(tb.Gf(c, a), a**1.5, rtol=tol)


def test_linear_power_against_reference_table(golden_dir):
  """flowpm/data/Planck15_a1p00.txt (nbodykit EH, its own Planck15 with
  massive neutrinos: up to 2.6 % tilt), and the oracle to 1e-12."""
  k1, p1 = np.loadtxt(golden_dir / "Planck15_a1p00.txt").T
  k = np.logspace(-3, 1, 512)
  ours = np.asarray(tp.linear_matter_power(Planck15(), k))
  ref = np.exp(np.interp(np.log(k), np.log(k1), np.log(p1)))
  np.testing.assert_allclose(ours, ref, rtol=3e-2)
  np.testing.assert_allclose(
      ours, opw.linear_matter_power(bg.planck15(), k), rtol=1e-12)


def test_linear_power_growth_scaling():
  """P(k, a) = D1(a)^2 P(k, 1) (tfpower.py:183-190)."""
  k = np.logspace(-3, 0, 64)
  c = Planck15()
  p1 = np.asarray(tp.linear_matter_power(c, k))
  pa = np.asarray(tp.linear_matter_power(c, k, a=0.25))
  np.testing.assert_allclose(pa, tb.D1(c, 0.25)**2 * p1, rtol=1e-12)


def test_scipy_helpers_match_oracle():
  rng = np.random.default_rng(3)
  xp = np.sort(rng.uniform(0, 10, 50))
  fp = rng.normal(size=50)
  x = rng.uniform(xp[0], xp[-3], 200)
  np.testing.assert_array_equal(fip.interp_tf(x, xp, fp),
                                opw.interp_tf(x, xp, fp))
  # Simpson's rule is exact for cubics
  np.testing.assert_allclose(fint.simps(lambda t: t**3 - 2 * t, 0.0, 2.0),
                             0.0, atol=1e-12)
  np.testing.assert_allclose(fint.simps(np.exp, 0.0, 1.0), np.e - 1,
                             rtol=1e-9)


@pytest.mark.parametrize("stages", [np.linspace(0.1, 1.0, 41),
                                    np.logspace(-2, 0, 17), [0.1, 0.5]])
def test_step_coefficients_equal_scalar_factors(stages):
  """The KDK coefficients of a whole run, evaluated on all steps at once,
  against the scalar kick / drift factors (tfpm.py:273-314, 472-498) that
  kick() and drift() use, step by step."""
  from flowpm_amd import tfpm
  c = Planck15()
  ours = tfpm.step_coefficients(c, stages)
  ref = []
  for i in range(len(stages) - 1):
    a0, a1 = float(stages[i]), float(stages[i + 1])
    ah = (a0 * a1)**0.5
    ref += [tfpm._kick_fac(c, a0, a0, ah), tfpm._drift_fac(c, a0, ah, a1),
            tfpm._kick_fac(c, ah, a1, a1)]
  np.testing.assert_allclose(ours, ref, rtol=1e-14, atol=0)
  assert tfpm.step_coefficients(c, [0.3]) == []


def test_parameter_tensors_and_host_differences():
  """A parameter given as a tensor that requires grad: the record keeps its
  value; host_fn's central differences reproduce closed-form derivatives:
  d(1.5 Omega_m)/dOmega_c = 1.5 and, for flat LCDM,
  dE(a)/dOmega_c = (a^-3 - 1) / (2 E(a))."""
  import torch
  from flowpm_amd.cosmology import host_fn
  oc = torch.tensor(0.2589, dtype=torch.float64, requires_grad=True)
  c = Planck15(Omega_c=oc)
  assert c.Omega_c == 0.2589 and set(c.grad_params()) == {"Omega_c"}
  assert not Planck15().grad_params()
  a = np.array([0.1, 0.37, 1.0])
  y = host_fn(lambda cc: np.append(1.5 * cc.Omega_m, tb.E(cc, a)), c)
  w = torch.tensor([1.0, 0.3, -2.0, 0.7], dtype=torch.float64)
  (y * w).sum().backward()
  E = tb.E(Planck15(), a)
  ref = 1.5 * 1.0 + float(np.dot([0.3, -2.0, 0.7], (a**-3 - 1) / (2 * E)))
  assert abs(float(oc.grad) - ref) < 1e-8 * abs(ref)


def test_linear_power_tensor_path():
  """With a parameter tensor, linear_matter_power returns the same values
  as a tensor, interp_tf on it equals the numpy path, and its gradient
  matches a central difference with a 100x larger step."""
  import torch
  k = np.logspace(-4, 1, 256)
  x = np.array([3e-3, 0.02, 0.1, 0.6])
  oc = torch.tensor(0.2589, dtype=torch.float64, requires_grad=True)
  pk_t = tp.linear_matter_power(Planck15(Omega_c=oc), k)
  pk = tp.linear_matter_power(Planck15(), k)
  assert isinstance(pk_t, torch.Tensor) and isinstance(pk, np.ndarray)
  np.testing.assert_array_equal(pk_t.detach().numpy(), pk)
  v = fip.interp_tf(x, k, pk_t)
  np.testing.assert_array_equal(v.detach().numpy(), fip.interp_tf(x, k, pk))
  v.sum().backward()
  h = 1e-4 * 0.2589
  fd = (fip.interp_tf(x, k, tp.linear_matter_power(
      Planck15(Omega_c=0.2589 + h), k)).sum() -
        fip.interp_tf(x, k, tp.linear_matter_power(
            Planck15(Omega_c=0.2589 - h), k)).sum()) / (2 * h)
  assert abs(float(oc.grad) - fd) < 1e-6 * abs(fd)
