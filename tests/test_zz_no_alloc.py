"""pmx_nbody_run_noalloc (the time loop on the library's static scratch)
against pmx_nbody_run with a caller workspace, through the C ABI.

Both run the same kernels on the same inputs; every kernel of the loop is
deterministic (integer paint accumulation, fixed summation orders), so the
states must agree bit for bit.  Two calls in a row, the second on a smaller
mesh, so the second starts from the scratch contents the first left.
"""
import ctypes

import numpy as np
import pytest
import torch

import flowpm_amd as fa
from flowpm_amd import _lib, _plans

pytestmark = pytest.mark.gpu


def _state(n, L, seed, dev):
  gen = torch.Generator().manual_seed(seed)
  white = torch.randn(1, n, n, n, generator=gen, dtype=torch.float64) * n**1.5
  c = fa.cosmology.Planck15()
  kk = np.logspace(-4, 1, 256)
  pkt = fa.tfpower.linear_matter_power(c, kk)
  pk = lambda k: fa.scipy.interpolate.interp_tf(k, kk, pkt)
  lin = fa.whitenoise_to_linear(n, L, white.float().to(dev), pk)
  return fa.lpt_init(c, lin, 0.1, order=1).detach().contiguous()


def _both(n, L, seed, nsteps, dev):
  c = fa.cosmology.Planck15()
  st0 = _state(n, L, seed, dev)
  plan = _plans.get(dev, 1, n, n, n)
  N = n**3
  fac = float(np.float32(1.5 * c.Omega_m))
  coef = _lib.dbl(fa.tfpm.step_coefficients(
      c, np.linspace(0.1, 1.0, nsteps + 1)))
  L_ = _lib.lib()
  a = st0.clone()
  work = torch.empty(L_.pmx_nbody_work_floats(plan.handle),
                     dtype=torch.float32, device=dev)
  _lib.call("pmx_nbody_run", plan.handle, _lib.ptr(a), N, 1.0, fac, coef, 0,
            nsteps, 1, _lib.ptr(work), _lib.stream_of(a))
  b = st0.clone()
  _lib.call("pmx_nbody_run_noalloc", plan.handle, _lib.ptr(b), N, 1.0, fac,
            coef, 0, nsteps, 1, _lib.stream_of(b))
  torch.cuda.synchronize()
  return a, b


def test_noalloc_matches_workspace_run(gpu):
  for n, L, seed, nsteps in ((64, 200.0, 11, 5), (32, 100.0, 12, 3)):
    print("noalloc vs workspace: n=%d steps=%d" % (n, nsteps), flush=True)
    a, b = _both(n, L, seed, nsteps, gpu)
    assert torch.isfinite(a).all()
    assert torch.equal(a, b), float((a - b).abs().max())


def test_noalloc_rejects_plans_beyond_the_scratch(gpu):
  plan = _plans.get(gpu, 1, 128, 128, 128)
  need = _lib.lib().pmx_nbody_work_floats(plan.handle)
  assert need > 16 * 2**20
  st = torch.zeros((3, 1, 128**3, 3), dtype=torch.float32, device=gpu)
  rc = _lib.lib().pmx_nbody_run_noalloc(plan.handle, _lib.ptr(st), 128**3,
                                        1.0, 0.5, None, 0, 0, 1,
                                        _lib.stream_of(st))
  assert rc == -1
  assert b"static scratch holds" in _lib.lib().pmx_last_error()
