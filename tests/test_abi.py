"""CPU checks of the drop-in boundary: the C-ABI library loads and exports
every entry point that include/pmx.h declares, and the host-side API mirrors
the reference's signatures.  No compute calls (no GPU here)."""
import inspect
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
  src = open(os.path.join(ROOT, "include", "pmx.h")).read()
  src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
  return sorted(set(re.findall(r"\b(pmx_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
  from flowpm_amd import _lib
  L = _lib.lib()
  decl = _declared()
  assert len(decl) >= 25
  for name in decl:
    assert hasattr(L, name), name
  assert L.pmx_version() == 1000
  # the ctypes table covers every declaration, with no stale entries
  assert sorted(_lib.exported_symbols()) == decl


def test_error_reporting_without_gpu():
  """Argument validation happens on the host before any HIP call."""
  from flowpm_amd import _lib
  L = _lib.lib()
  rc = L.pmx_cic_paint(None, None, None, 0, 10, 4, 4, 4, 1.0, None)
  assert rc == -1
  assert b"B >= 1" in L.pmx_last_error()
  with pytest.raises(ValueError):
    _lib.check(rc, "pmx_cic_paint")


def test_reference_signatures():
  import flowpm_amd as fa
  sig = lambda f: list(inspect.signature(f).parameters)
  assert sig(fa.utils.cic_paint)[:3] == ["mesh", "part", "weight"]
  assert sig(fa.utils.cic_readout)[:2] == ["mesh", "part"]
  assert sig(fa.utils.r2c3d)[:3] == ["rfield", "norm", "dtype"]
  assert sig(fa.utils.c2r3d)[:3] == ["cfield", "norm", "dtype"]
  assert sig(fa.tfpm.linear_field)[:6] == ["nc", "boxsize", "pk", "kvec",
                                            "batch_size", "seed"]
  assert sig(fa.tfpm.lpt_init)[:5] == ["cosmo", "linear", "a", "order",
                                        "kvec"]
  assert sig(fa.tfpm.nbody)[:7] == ["cosmo", "state", "stages", "nc",
                                     "pm_nc_factor",
                                     "return_intermediate_states",
                                     "pgdparams"]
  assert sig(fa.tfpm.force)[:4] == ["cosmo", "state", "nc", "pm_nc_factor"]
  assert sig(fa.tfpm.kick)[:5] == ["cosmo", "state", "ai", "ac", "af"]
  assert sig(fa.tfpm.apply_longrange)[:4] == ["x", "delta_k", "split",
                                               "factor"]


def test_ops_refuse_cpu_tensors():
  import torch
  import flowpm_amd as fa
  with pytest.raises(ValueError):
    fa.cic_paint(torch.zeros(1, 4, 4, 4), torch.zeros(1, 3, 3))


def test_host_kernels_match_reference_conventions():
  import numpy as np
  from flowpm_amd import kernels as K
  k = K.fftk([4, 4, 4], symmetric=False)
  assert K.laplace_kernel(k)[0, 0, 0] == 1.0
  np.testing.assert_allclose(K.gradient_kernel(k, 1)[0, 1, 0], 8j / 6)
  assert K.fftk([4, 4, 6])[2].shape == (1, 1, 4)
