"""Gradients w.r.t. cosmological parameters: the scripts/check_grads.py:71-119
workflow (Planck15(Omega_c=tensor) -> linear_matter_power -> interp_tf ->
whitenoise_to_linear -> lpt_init -> nbody -> cic_paint -> mean(field^2)),
differentiated by the hand-written adjoints with the coefficients'
cotangents, against central differences of the float64 oracle's forward
(the reference's own check is a central difference of its forward,
check_grads.py:106-119), and the single-op drop-ins (kick, drift, force)
against their exact linear dependence on the coefficients.
"""
import numpy as np
import pytest
import torch

import flowpm_amd as fa
from oracle import background as obg
from oracle import pm as opm
from oracle import power as opw

pytestmark = pytest.mark.gpu

K = np.logspace(-4, 1, 256)


def _white(n, seed):
  gen = torch.Generator().manual_seed(seed)
  return torch.randn(1, n, n, n, generator=gen, dtype=torch.float64) * n**1.5


def _loss_gpu(omega_c, white, n, L, stages, order, every, f=1):
  cosmo = fa.cosmology.Planck15(Omega_c=omega_c)
  pk_tab = fa.tfpower.linear_matter_power(cosmo, K)
  pk = lambda x: fa.scipy.interpolate.interp_tf(x, K, pk_tab)
  lin = fa.whitenoise_to_linear(n, L, white, pk)
  st = fa.lpt_init(cosmo, lin, 0.1, order=order)
  fin = fa.nbody(cosmo, st, stages, n, pm_nc_factor=f,
                 checkpoint_every=every)
  field = fa.cic_paint(torch.zeros_like(lin), fin[0])
  return (field.double()**2).mean()


def _loss_oracle(omega_c, white, n, L, stages, order, f=1):
  c = obg.planck15(Omega_c=omega_c)
  g = obg.GrowthTable(c)
  pkt = opw.linear_matter_power(c, K)
  pk = lambda x: opw.interp_tf(x.reshape(-1), K, pkt).reshape(x.shape)
  lin = opm.whitenoise_to_linear(white, L, pk)
  st = opm.nbody(g, opm.lpt_init(g, lin, 0.1, order=order), stages, [n] * 3,
                 pm_nc_factor=f)
  rho = opm.cic_paint(torch.zeros_like(lin), st[0])
  return float((rho**2).mean())


@pytest.mark.parametrize("order,every,keep", [(2, 0, None), (1, 0, None),
                                              (2, 2, True), (2, 2, False)])
def test_omega_c_gradient_against_oracle_differences(gpu, order, every,
                                                     keep):
  """d mean(field^2) / d Omega_c at 32^3 (box 100, 4 steps from a = 0.1):
  the GPU gradient (taped, or checkpoints every 2 steps with and without
  kept potentials) against a central difference of the float64 oracle
  (step 1e-6 Omega_c: the CIC loss's derivative jumps where particles cross
  cell edges, and wider differences average over those jumps, 0.6 % at
  1e-4 with the PGD correction below), within 5e-4."""
  n, L = 32, 100.0
  stages = np.linspace(0.1, 1.0, 5)
  white = _white(n, 100)
  oc0 = 0.2589
  h = 1e-6 * oc0
  fd = (_loss_oracle(oc0 + h, white, n, L, stages, order) -
        _loss_oracle(oc0 - h, white, n, L, stages, order)) / (2 * h)
  fa.tfpm.set_keep_potentials(keep)
  try:
    oc = torch.tensor(oc0, dtype=torch.float64, requires_grad=True)
    loss = _loss_gpu(oc, white.float().to(gpu), n, L, stages, order, every)
    loss.backward()
  finally:
    fa.tfpm.set_keep_potentials(None)
  lo = _loss_oracle(oc0, white, n, L, stages, order)
  e = abs(float(oc.grad) - fd) / abs(fd)
  print("order %d every %d keep %s: loss %.6g (oracle %.6g), dL/dOmega_c "
        "%.6g, oracle central difference %.6g, rel %.3g" %
        (order, every, keep, float(loss), lo, float(oc.grad), fd, e))
  assert abs(float(loss) - lo) < 1e-4 * abs(lo)
  assert e < 5e-4


@pytest.mark.parametrize("f,B", [(2, 1), (1, 2)])
def test_omega_c_gradient_force_mesh_factor_and_batch(gpu, f, B):
  """The same check with the force on a mesh twice as fine (pm_nc_factor
  2: F = (1.5 Omega_m / f) readout at f X) and on a batch of two
  simulations (the loss averages both; the coefficients are shared).
  Meshes whose brick boxes fit (no float-atomic spill), so the GPU runs
  repeat bit for bit."""
  n, L = (16, 64.0) if f == 2 else (32, 100.0)
  stages = np.linspace(0.1, 1.0, 4)
  gen = torch.Generator().manual_seed(31)
  white = torch.randn(B, n, n, n, generator=gen, dtype=torch.float64) * \
      n**1.5
  oc0 = 0.2589
  h = 1e-6 * oc0
  fd = (_loss_oracle(oc0 + h, white, n, L, stages, 2, f) -
        _loss_oracle(oc0 - h, white, n, L, stages, 2, f)) / (2 * h)
  oc = torch.tensor(oc0, dtype=torch.float64, requires_grad=True)
  _loss_gpu(oc, white.float().to(gpu), n, L, stages, 2, None, f).backward()
  e = abs(float(oc.grad) - fd) / abs(fd)
  print("f %d B %d: dL/dOmega_c %.6g, oracle central difference %.6g, rel "
        "%.3g" % (f, B, float(oc.grad), fd, e))
  assert e < 5e-4


def test_omega_c_gradient_against_gpu_differences(gpu):
  """The same gradient at 64^3 (box 200, 2LPT, 10 steps) against the
  central difference of the GPU forward itself at +-0.1 % of Omega_c, the
  reference's own check (check_grads.py:106-119).  That difference spans
  particles crossing cell edges, where the CIC loss's derivative jumps, so
  it agrees to about 1 % (0.8 % measured); the oracle comparisons above are
  the tight check."""
  n, L = 64, 200.0
  stages = np.linspace(0.1, 1.0, 11)
  white = _white(n, 7).float().to(gpu)
  oc0 = 0.2589
  oc = torch.tensor(oc0, dtype=torch.float64, requires_grad=True)
  _loss_gpu(oc, white, n, L, stages, 2, None).backward()
  with torch.no_grad():
    v1 = float(_loss_gpu(oc0 * 1.001, white, n, L, stages, 2, None))
    v0 = float(_loss_gpu(oc0 * 0.999, white, n, L, stages, 2, None))
  fd = (v1 - v0) / (0.002 * oc0)
  e = abs(float(oc.grad) - fd) / abs(fd)
  print("64^3 x 10: dL/dOmega_c %.6g, GPU central difference %.6g, rel %.3g"
        % (float(oc.grad), fd, e))
  assert e < 2e-2


def test_gradients_of_several_parameters(gpu):
  """Omega_b, h, n_s and sigma8 as tensors too (one backward pass gives
  all five gradients), each against the oracle's central difference.  (At
  16^3 the brick boxes span the mesh and spill to float atomics, whose
  order varies from run to run: d/dh, a sum with much cancellation, then
  moves by ~5e-4 between runs; at 32^3 every box fits.)"""
  n, L = 32, 100.0
  stages = np.linspace(0.1, 1.0, 4)
  white = _white(n, 3)
  base = dict(Omega_c=0.2589, Omega_b=0.0486, h=0.6774, n_s=0.9667,
              sigma8=0.8159)
  params = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True)
            for k, v in base.items()}
  cosmo = fa.cosmology.Planck15(**params)
  pk_tab = fa.tfpower.linear_matter_power(cosmo, K)
  pk = lambda x: fa.scipy.interpolate.interp_tf(x, K, pk_tab)
  lin = fa.whitenoise_to_linear(n, L, white.float().to(gpu), pk)
  fin = fa.nbody(cosmo, fa.lpt_init(cosmo, lin, 0.1), stages, n)
  field = fa.cic_paint(torch.zeros_like(lin), fin[0])
  (field.double()**2).mean().backward()

  def loss_o(**over):
    c = obg.planck15(**dict(base, **over))
    g = obg.GrowthTable(c)
    pkt = opw.linear_matter_power(c, K)
    pko = lambda x: opw.interp_tf(x.reshape(-1), K, pkt).reshape(x.shape)
    lo = opm.whitenoise_to_linear(white, L, pko)
    st = opm.nbody(g, opm.lpt_init(g, lo, 0.1, order=2), stages, [n] * 3)
    return float((opm.cic_paint(torch.zeros_like(lo), st[0])**2).mean())

  for k, v in base.items():
    h = 1e-6 * v
    fd = (loss_o(**{k: v + h}) - loss_o(**{k: v - h})) / (2 * h)
    e = abs(float(params[k].grad) - fd) / abs(fd)
    print("d/d%s: %.6g, oracle central difference %.6g, rel %.3g" %
          (k, float(params[k].grad), fd, e))
    assert e < 5e-4, k


def test_kick_drift_force_coefficient_gradients(gpu):
  """kick / drift (P += c F, X += c P) and force (F = 1.5 Omega_m G(X)):
  the Omega_c gradient of <w, op(state)> equals <w, d op / d Omega_c>,
  which is exactly w_dst . state_src times dc / dOmega_c (central
  difference of the float64 coefficient) for kick and drift, and
  1.5 <w_F, F> / fac for the force."""
  n = 16
  gen = torch.Generator().manual_seed(4)
  st = opm.lagrangian_grid([n] * 3, 1, torch.float64)
  st = torch.stack([st + torch.randn(st.shape, generator=gen,
                                     dtype=torch.float64),
                    torch.randn(st.shape, generator=gen, dtype=torch.float64),
                    torch.randn(st.shape, generator=gen,
                                dtype=torch.float64)]).float().to(gpu)
  w = torch.randn(st.shape, generator=gen).to(gpu)
  c0 = fa.cosmology.Planck15()
  oc0 = c0.Omega_c
  h = 1e-6 * oc0
  ai, ac, af = 0.3, 0.35, 0.4
  for name, op, dst, src, cf in (
      ("kick", fa.tfpm.kick, 1, 2, fa.tfpm._kick_fac),
      ("drift", fa.tfpm.drift, 0, 1, fa.tfpm._drift_fac)):
    oc = torch.tensor(oc0, dtype=torch.float64, requires_grad=True)
    out = op(fa.cosmology.Planck15(Omega_c=oc), st, ai, ac, af)
    (out.double() * w.double()).sum().backward()
    dc = (cf(c0.replace(Omega_c=oc0 + h), ai, ac, af) -
          cf(c0.replace(Omega_c=oc0 - h), ai, ac, af)) / (2 * h)
    ref = float((w[dst].double() * st[src].double()).sum()) * dc
    print("%s: %.8g against %.8g" % (name, float(oc.grad), ref))
    assert abs(float(oc.grad) - ref) < 1e-6 * abs(ref)
  oc = torch.tensor(oc0, dtype=torch.float64, requires_grad=True)
  out = fa.tfpm.force(fa.cosmology.Planck15(Omega_c=oc), st, n)
  (out.double() * w.double()).sum().backward()
  fac = float(np.float32(1.5 * c0.Omega_m))
  ref = 1.5 * float((w[2].double() * out[2].detach().double()).sum()) / fac
  print("force: %.8g against %.8g" % (float(oc.grad), ref))
  assert abs(float(oc.grad) - ref) < 1e-6 * abs(ref)


def test_intermediate_states_give_the_single_run_gradient(gpu):
  """nbody(return_intermediate_states=True) runs one differentiable step at
  a time (each with its own row of coefficients and the jump-start force
  at its input): the Omega_c gradient of a loss on the last state equals
  the single run's (to float rounding: the two sweeps add the same terms
  in different orders and places)."""
  n, L = 32, 100.0
  stages = np.linspace(0.1, 1.0, 5)
  white = _white(n, 100).float().to(gpu)
  grads = []
  for inter in (False, True):
    oc = torch.tensor(0.2589, dtype=torch.float64, requires_grad=True)
    cosmo = fa.cosmology.Planck15(Omega_c=oc)
    pk_tab = fa.tfpower.linear_matter_power(cosmo, K)
    pk = lambda x: fa.scipy.interpolate.interp_tf(x, K, pk_tab)
    lin = fa.whitenoise_to_linear(n, L, white, pk)
    st = fa.lpt_init(cosmo, lin, 0.1)
    out = fa.nbody(cosmo, st, stages, n, return_intermediate_states=inter)
    fin = out[-1][1] if inter else out
    (fa.cic_paint(torch.zeros_like(lin), fin[0]).double()**2).mean().backward()
    grads.append(float(oc.grad))
  e = abs(grads[1] - grads[0]) / abs(grads[0])
  print("single run %.8g, step by step %.8g, rel %.3g" % (grads[0], grads[1],
                                                         e))
  assert e < 1e-4


@pytest.mark.parametrize("inter", [False, True])
def test_omega_c_gradient_with_pgd(gpu, inter):
  """nbody with the PGD correction (tfpm.py:487-488): the force acts at the
  corrected positions and each drift ends at the positions before the
  correction; d mean(field^2) / d Omega_c against the float64 oracle's
  central difference (one run, and one run per step)."""
  n, L = 32, 100.0
  stages = np.linspace(0.1, 1.0, 5)
  white = _white(n, 41)
  pgdp = np.array([[0.3 + 0.05 * i, 0.6, 3.0 + 0.2 * i] for i in range(4)],
                  dtype=np.float32)
  oc0 = 0.2589

  def loss_o(oc):
    c = obg.planck15(Omega_c=oc)
    g = obg.GrowthTable(c)
    pkt = opw.linear_matter_power(c, K)
    pk = lambda x: opw.interp_tf(x.reshape(-1), K, pkt).reshape(x.shape)
    lin = opm.whitenoise_to_linear(white, L, pk)
    st = opm.nbody(g, opm.lpt_init(g, lin, 0.1, order=1), stages, [n] * 3,
                   pgdparams=pgdp.astype(np.float64))
    return float((opm.cic_paint(torch.zeros_like(lin), st[0])**2).mean())

  h = 1e-6 * oc0
  fd = (loss_o(oc0 + h) - loss_o(oc0 - h)) / (2 * h)
  oc = torch.tensor(oc0, dtype=torch.float64, requires_grad=True)
  cosmo = fa.cosmology.Planck15(Omega_c=oc)
  pk_tab = fa.tfpower.linear_matter_power(cosmo, K)
  pk = lambda x: fa.scipy.interpolate.interp_tf(x, K, pk_tab)
  lin = fa.whitenoise_to_linear(n, L, white.float().to(gpu), pk)
  st = fa.lpt_init(cosmo, lin, 0.1, order=1)
  out = fa.nbody(cosmo, st, stages, n, pgdparams=pgdp,
                 return_intermediate_states=inter)
  fin = out[-1][1] if inter else out
  (fa.cic_paint(torch.zeros_like(lin), fin[0]).double()**2).mean().backward()
  e = abs(float(oc.grad) - fd) / abs(fd)
  print("PGD (per step %s): dL/dOmega_c %.6g, oracle central difference "
        "%.6g, rel %.3g" % (inter, float(oc.grad), fd, e))
  assert e < 5e-4
