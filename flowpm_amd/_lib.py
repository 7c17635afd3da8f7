"""ctypes binding of libpmx_hip.so (the C ABI declared in include/pmx.h).

The HIP library is the only compute path: if it is missing, every op raises.
There is no CPU or PyTorch fallback.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpmx_hip.so")

_c = ctypes
_vp = _c.c_void_p
_i = _c.c_int
_ll = _c.c_longlong
_f = _c.c_float
_dp = _c.POINTER(_c.c_double)

# name: (restype, argtypes)
_SIGS = {
    "pmx_version": (_i, []),
    "pmx_last_error": (_c.c_char_p, []),
    "pmx_profile_start": (_i, []),
    "pmx_profile_stop": (_i, [_c.POINTER(_c.c_double), _c.POINTER(_ll), _i]),
    "pmx_profile_kinds": (_i, []),
    "pmx_profile_kind_name": (_c.c_char_p, [_i]),
    "pmx_plan_create": (_i, [_c.POINTER(_vp), _i, _i, _i, _i]),
    "pmx_plan_destroy": (_i, [_vp]),
    "pmx_fft_plan_create": (_i, [_c.POINTER(_vp), _i, _i, _i, _i]),
    "pmx_fft_plan_destroy": (_i, [_vp]),
    "pmx_plan_spec_elems": (_ll, [_vp]),
    "pmx_plan_set_tiles": (_i, [_vp, _i, _i]),
    "pmx_cic_paint": (_i, [_vp, _vp, _vp, _i, _ll, _i, _i, _i, _f, _vp]),
    "pmx_cic_paint_work_floats": (_ll, [_i, _ll, _i, _i, _i, _f]),
    "pmx_cic_paint_ws": (_i, [_vp, _vp, _vp, _i, _ll, _i, _i, _i, _f, _vp,
                              _ll, _vp]),
    "pmx_cic_readout": (_i, [_vp, _vp, _vp, _i, _ll, _i, _i, _i, _f, _vp]),
    "pmx_cic_paint_adj": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _ll, _i, _i, _i,
                               _f, _vp]),
    "pmx_cic_readout_adj": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _ll, _i, _i, _i,
                                 _f, _vp]),
    "pmx_fft_r2c": (_i, [_vp, _vp, _vp, _f, _vp]),
    "pmx_fft_c2r": (_i, [_vp, _vp, _vp, _f, _vp]),
    "pmx_spec_to_full": (_i, [_vp, _vp, _vp, _vp]),
    "pmx_full_to_spec": (_i, [_vp, _vp, _vp, _vp]),
    "pmx_fft_c2r_poisson_grad": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _f, _vp]),
    "pmx_fft_c2r_longrange_grad": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _f,
                                        _c.c_double, _vp]),
    "pmx_fft_longrange_grad_adj": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _f,
                                        _c.c_double, _vp]),
    "pmx_fft_c2r_pgd_grad": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _f, _f, _f,
                                  _i, _vp]),
    "pmx_fft_pgd_grad_adj": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _f, _f, _f,
                                  _vp]),
    "pmx_pgd": (_i, [_vp, _vp, _ll, _f, _f, _f, _f, _vp, _vp]),
    "pmx_pgd_adj_work_floats": (_ll, [_vp, _ll]),
    "pmx_pgd_adjoint": (_i, [_vp, _vp, _ll, _f, _f, _f, _f, _vp, _vp, _vp,
                             _vp]),
    "pmx_mesh_dot": (_i, [_vp, _vp, _ll, _vp, _vp]),
    "pmx_power_spectrum_bins_adj": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _i,
                                         _vp, _vp]),
    "pmx_force_meshes": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _f, _vp]),
    "pmx_force_meshes_adj": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _f, _vp]),
    "pmx_force_mesh_f4": (_i, [_vp, _vp, _vp, _vp, _vp, _f, _vp]),
    "pmx_amp_from_table": (_i, [_vp, _vp, _ll, _vp, _vp]),
    "pmx_power_spectrum_counts": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _vp, _vp]),
    "pmx_fft_poisson_grad_adj": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _f, _vp]),
    "pmx_lpt2_source": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "pmx_lpt2_work_floats": (_ll, [_vp]),
    "pmx_lpt_assemble": (_i, [_vp, _vp, _vp, _dp, _i, _i, _i, _i, _vp]),
    "pmx_lpt_assemble_adj": (_i, [_vp, _dp, _vp, _vp, _i, _i, _i, _i, _vp]),
    "pmx_lpt2_source_adj": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "pmx_fourier_apply": (_i, [_vp, _i, _c.POINTER(_vp), _c.POINTER(_i),
                               _c.POINTER(_f), _i, _f, _vp, _vp, _i, _vp]),
    "pmx_spec_mul_real": (_i, [_vp, _vp, _vp, _vp]),
    "pmx_white_noise": (_i, [_vp, _ll, _c.c_ulonglong, _c.c_ulonglong,
                             _c.c_ulonglong, _c.c_double, _vp]),
    "pmx_power_spectrum_bins": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _i, _vp,
                                     _vp]),
    "pmx_readout_kick_drift": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _i, _ll, _i,
                                    _i, _i, _f, _f, _f, _f, _f, _i, _vp]),
    "pmx_nbody_work_floats": (_ll, [_vp]),
    "pmx_nbody_run": (_i, [_vp, _vp, _ll, _f, _f, _dp, _i, _i, _i, _vp, _vp]),
    "pmx_nbody_run_noalloc": (_i, [_vp, _vp, _ll, _f, _f, _dp, _i, _i, _i,
                                   _vp]),
    "pmx_nbody_run_pgd": (_i, [_vp, _vp, _ll, _f, _f, _dp,
                               _c.POINTER(_f), _i, _i, _i, _vp, _vp]),
    "pmx_nbody_pgd_tape_floats": (_ll, [_vp, _ll, _i]),
    "pmx_nbody_run_pgd_taped": (_i, [_vp, _vp, _ll, _f, _f, _dp,
                                     _c.POINTER(_f), _i, _vp, _vp, _vp]),
    "pmx_nbody_adjoint_pgd_taped": (_i, [_vp, _vp, _i, _ll, _f, _f, _dp,
                                         _c.POINTER(_f), _vp, _vp, _vp,
                                         _vp, _vp]),
    "pmx_nbody_adj_work_floats": (_ll, [_vp, _ll, _i]),
    "pmx_nbody_adjoint": (_i, [_vp, _vp, _i, _i, _ll, _f, _f, _dp, _vp, _vp,
                               _vp, _vp]),
    "pmx_nbody_run_last_segment": (_i, [_vp, _vp, _ll, _f, _f, _dp, _i, _i,
                                        _vp, _vp]),
    "pmx_nbody_adjoint_last_taped": (_i, [_vp, _vp, _i, _i, _ll, _f, _f, _dp,
                                          _vp, _vp, _vp, _vp]),
    "pmx_nbody_phi_floats": (_ll, [_vp, _i]),
    "pmx_nbody_run_phi": (_i, [_vp, _vp, _ll, _f, _f, _dp, _i, _i, _i, _vp,
                               _vp, _vp]),
    "pmx_nbody_run_last_segment_phi": (_i, [_vp, _vp, _ll, _f, _f, _dp, _i,
                                            _i, _vp, _vp, _vp]),
    "pmx_nbody_run_checkpointed": (_i, [_vp, _vp, _vp, _ll, _f, _f, _dp, _i,
                                        _i, _i, _i, _vp, _vp, _vp, _vp]),
    "pmx_nbody_adjoint_phi": (_i, [_vp, _vp, _vp, _i, _i, _ll, _f, _f, _dp,
                                   _vp, _vp, _i, _vp, _vp]),
    "pmx_nbody_batched_step": (_i, [_i, _c.POINTER(_vp), _c.POINTER(_vp),
                                    _c.POINTER(_ll), _c.POINTER(_f),
                                    _c.POINTER(_f), _c.POINTER(_dp),
                                    _c.POINTER(_i), _c.POINTER(_i),
                                    _c.POINTER(_i), _c.POINTER(_vp), _vp]),
    "pmx_nbody_tape_floats": (_ll, [_vp, _ll, _i]),
    "pmx_nbody_compact_tape_floats": (_ll, [_vp, _ll, _i]),
    "pmx_nbody_run_compact_taped": (_i, [_vp, _vp, _ll, _f, _f, _dp, _i, _vp,
                                         _vp, _vp]),
    "pmx_nbody_adjoint_compact_taped": (_i, [_vp, _vp, _i, _ll, _f, _f, _dp,
                                             _vp, _vp, _vp, _vp]),
    "pmx_nbody_run_taped": (_i, [_vp, _vp, _ll, _f, _f, _dp, _i, _vp, _vp,
                                 _vp]),
    "pmx_nbody_adjoint_taped": (_i, [_vp, _vp, _i, _ll, _f, _f, _dp, _vp,
                                     _vp, _vp, _vp]),
}

_LIB = None


def lib():
  """Loads the HIP library (raises if it has not been built)."""
  global _LIB
  if _LIB is None:
    if not os.path.exists(LIB_PATH):
      raise RuntimeError(
          "flowpm_amd: %s is missing; build it with __graft_entry__.build() "
          "(hipcc, gfx950).  There is no CPU fallback." % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
      fn = getattr(L, name)
      fn.restype = res
      fn.argtypes = args
    _LIB = L
  return _LIB


def exported_symbols():
  return list(_SIGS.keys())


def check(status, what=""):
  if status != 0:
    msg = lib().pmx_last_error().decode("utf-8", "replace")
    if status == -1:
      raise ValueError("%s: %s" % (what, msg))
    raise RuntimeError("%s failed (%d): %s" % (what, status, msg))


def call(name, *args):
  check(getattr(lib(), name)(*args), name)


def ptr(t):
  """Device pointer of a tensor (None -> NULL)."""
  if t is None:
    return None
  return _vp(t.data_ptr())


def stream_of(t=None):
  dev = t.device if t is not None else torch.device("cuda")
  return _vp(torch.cuda.current_stream(dev).cuda_stream)


def dbl(values):
  arr = (_c.c_double * len(values))(*[float(v) for v in values])
  return arr


def require_cuda(*tensors):
  for t in tensors:
    if t is None:
      continue
    if not t.is_cuda:
      raise ValueError("flowpm_amd ops run on the GPU only; got a tensor on "
                       "%s" % t.device)
