"""CIC and FFT drop-ins for flowpm/utils.py, on torch tensors in HBM.

Same call signatures as the reference (``name=`` is accepted and ignored),
same results up to float32 rounding, pure-functional (inputs are never
modified).  Each op is a ``torch.autograd.Function`` whose backward is a
hand-written HIP adjoint:

* cic_paint   (utils.py:32-91)   backward: readout-type gather for the
                                 weights and the position gradient
* cic_readout (utils.py:163-210) backward: weighted paint + position gradient
* r2c3d       (utils.py:213-241) backward: the conjugate transform (c2r)
* c2r3d       (utils.py:244-272) backward: the conjugate transform (r2c)
* white_noise (utils.py:337-355) N(0, sqrt(M)) drawn in HBM from the
                                 seeded Philox4x64-10 stream (TF's RNG
                                 stream cannot be reproduced)
* compensate_cic (utils.py:11-29) the CIC window deconvolution
"""
import numpy as np
import torch

from . import _lib
from . import _plans

__all__ = ["cic_paint", "cic_readout", "r2c3d", "c2r3d", "white_noise",
           "compensate_cic"]


def _f32(t):
  return t.detach().to(torch.float32).contiguous()


def _mesh_dims(mesh):
  if mesh.dim() != 4:
    raise ValueError("mesh must be (batch, nx, ny, nz), got %s" %
                     (tuple(mesh.shape),))
  return [int(s) for s in mesh.shape]


def _flat_part(part, B):
  if part.shape[-1] != 3:
    raise ValueError("particle coordinates must end with a dimension of 3")
  return _f32(part).reshape(B, -1, 3)


class _CicPaint(torch.autograd.Function):

  @staticmethod
  def forward(ctx, mesh, part, weight):
    B, nx, ny, nz = _mesh_dims(mesh)
    _lib.require_cuda(mesh, part, weight)
    pos = _flat_part(part, B)
    N = pos.shape[1]
    w = None if weight is None else _f32(weight).reshape(B, N)
    out = _f32(mesh).clone()
    # one particle per cell (the final paint of an nbody run): the time
    # loop's atomic-free brick paint, on a workspace; else the scatter
    nw = _lib.lib().pmx_cic_paint_work_floats(B, N, nx, ny, nz, 1.0)
    work = torch.empty(nw, dtype=torch.float32, device=out.device) \
        if nw > 0 else None
    _lib.call("pmx_cic_paint_ws", _lib.ptr(out), _lib.ptr(pos), _lib.ptr(w),
              B, N, nx, ny, nz, 1.0, _lib.ptr(work), nw, _lib.stream_of(out))
    ctx.save_for_backward(pos, w)
    ctx.dims = (B, N, nx, ny, nz)
    ctx.part_shape = part.shape
    ctx.has_w = weight is not None
    return out

  @staticmethod
  def backward(ctx, g):
    pos, w = ctx.saved_tensors
    B, N, nx, ny, nz = ctx.dims
    g = _f32(g)
    dpos = dw = None
    need_pos = ctx.needs_input_grad[1]
    need_w = ctx.has_w and ctx.needs_input_grad[2]
    if need_pos or need_w:
      dpos = torch.empty_like(pos) if need_pos else None
      dw = torch.empty((B, N), dtype=torch.float32, device=g.device) \
          if need_w else None
      _lib.call("pmx_cic_paint_adj", _lib.ptr(dpos), _lib.ptr(dw),
                _lib.ptr(g), _lib.ptr(pos), _lib.ptr(w), B, N, nx, ny, nz,
                1.0, _lib.stream_of(g))
    if dpos is not None:
      dpos = dpos.reshape(ctx.part_shape)
    gm = g if ctx.needs_input_grad[0] else None
    return gm, dpos, dw


def cic_paint(mesh, part, weight=None, name="CiCPaint"):
  """Paints particles on a 3-D mesh (flowpm/utils.py:32-91).

  mesh (batch, nx, ny, nz); part (batch, npart, 3) in mesh units (higher
  rank is flattened); weight (batch, npart) or None.  Returns mesh + paint.
  """
  return _CicPaint.apply(mesh, part, weight)


class _CicReadout(torch.autograd.Function):

  @staticmethod
  def forward(ctx, mesh, part):
    B, nx, ny, nz = _mesh_dims(mesh)
    _lib.require_cuda(mesh, part)
    m = _f32(mesh)
    pos = _flat_part(part, B)
    N = pos.shape[1]
    out = torch.empty((B, N), dtype=torch.float32, device=m.device)
    _lib.call("pmx_cic_readout", _lib.ptr(out), _lib.ptr(m), _lib.ptr(pos), B,
              N, nx, ny, nz, 1.0, _lib.stream_of(m))
    ctx.save_for_backward(m, pos)
    ctx.dims = (B, N, nx, ny, nz)
    ctx.part_shape = part.shape
    return out

  @staticmethod
  def backward(ctx, g):
    m, pos = ctx.saved_tensors
    B, N, nx, ny, nz = ctx.dims
    g = _f32(g)
    gm = torch.zeros_like(m) if ctx.needs_input_grad[0] else None
    dpos = torch.empty_like(pos) if ctx.needs_input_grad[1] else None
    if gm is not None or dpos is not None:
      _lib.call("pmx_cic_readout_adj", _lib.ptr(gm), _lib.ptr(dpos),
                _lib.ptr(g), _lib.ptr(m), _lib.ptr(pos), B, N, nx, ny, nz,
                1.0, _lib.stream_of(g))
    if dpos is not None:
      dpos = dpos.reshape(ctx.part_shape)
    return gm, dpos


def cic_readout(mesh, part, name="CiCReadout"):
  """Reads the mesh out at particle positions (flowpm/utils.py:163-210).

  Returns (batch, npart)."""
  return _CicReadout.apply(mesh, part)


# ------------------------------------------------------------------ FFTs --
def _full_to_real_half(plan, c):
  """Hermitian part of a full complex spectrum, in the library layout."""
  cr = torch.view_as_real(c.to(torch.complex64).contiguous())
  spec = plan.spec()
  _lib.call("pmx_full_to_spec", plan.handle, _lib.ptr(cr), _lib.ptr(spec),
            _lib.stream_of(spec))
  return spec


def _half_to_full(plan, spec):
  full = torch.empty((plan.B, plan.nx, plan.ny, plan.nz),
                     dtype=torch.complex64, device=spec.device)
  _lib.call("pmx_spec_to_full", plan.handle, _lib.ptr(spec),
            _lib.ptr(torch.view_as_real(full)), _lib.stream_of(spec))
  return full


def fwd_real(plan, x, scale):
  """spec = FFT3(x) * scale, library layout."""
  spec = plan.spec()
  _lib.call("pmx_fft_r2c", plan.handle, _lib.ptr(x), _lib.ptr(spec),
            float(scale), _lib.stream_of(x))
  return spec


def inv_real(plan, spec, scale):
  """x = IDFT3(spec) * scale (spec is consumed)."""
  out = torch.empty((plan.B, plan.nx, plan.ny, plan.nz), dtype=torch.float32,
                    device=spec.device)
  _lib.call("pmx_fft_c2r", plan.handle, _lib.ptr(spec), _lib.ptr(out),
            float(scale), _lib.stream_of(spec))
  return out


class _R2C(torch.autograd.Function):

  @staticmethod
  def forward(ctx, x, norm):
    B, nx, ny, nz = _mesh_dims(x)
    _lib.require_cuda(x)
    plan = _plans.get(x.device, B, nx, ny, nz)
    spec = fwd_real(plan, _f32(x), 1.0 / norm)
    ctx.plan, ctx.norm = plan, norm
    return _half_to_full(plan, spec)

  @staticmethod
  def backward(ctx, g):
    plan = ctx.plan
    spec = _full_to_real_half(plan, g)
    return inv_real(plan, spec, 1.0 / ctx.norm), None


class _C2R(torch.autograd.Function):

  @staticmethod
  def forward(ctx, c, norm):
    if c.dim() != 4:
      raise ValueError("cfield must be (batch, nx, ny, nz)")
    _lib.require_cuda(c)
    B, nx, ny, nz = [int(s) for s in c.shape]
    plan = _plans.get(c.device, B, nx, ny, nz)
    spec = _full_to_real_half(plan, c)
    ctx.plan, ctx.norm = plan, norm
    return inv_real(plan, spec, norm / plan.M)

  @staticmethod
  def backward(ctx, g):
    plan = ctx.plan
    spec = fwd_real(plan, _f32(g), ctx.norm / plan.M)
    return _half_to_full(plan, spec), None


def r2c3d(rfield, norm=None, dtype=torch.complex64, name="R2C3D"):
  """Forward 3-D FFT of a real field divided by ``norm`` (default: number of
  cells), full complex64 spectrum (flowpm/utils.py:213-241)."""
  if dtype not in (torch.complex64, None):
    raise ValueError("the HIP FFT computes in complex64")
  if norm is None:
    norm = int(np.prod(rfield.shape[1:]))
  return _R2C.apply(rfield, float(norm))


def c2r3d(cfield, norm=None, dtype=torch.float32, name="C2R3D"):
  """Real part of the inverse 3-D FFT times ``norm`` (default: number of
  cells), float32 (flowpm/utils.py:244-272)."""
  if dtype not in (torch.float32, None):
    raise ValueError("the HIP FFT computes in float32")
  if norm is None:
    norm = int(np.prod(cfield.shape[1:]))
  return _C2R.apply(cfield, float(norm))


def white_noise(nc, batch_size=1, seed=None, type="complex",
                name="WhiteNoise", device=None, rng="philox"):
  """N(0, sqrt(M)) real white noise (flowpm/utils.py:337-355).

  TF's RNG stream cannot be reproduced, so the stream is this package's:
  ``rng='philox'`` (default) draws it in HBM from the counter-based
  Philox4x64-10 generator keyed by ``seed`` -- the stream of
  ``numpy.random.Philox(key=seed)``, so the host can regenerate any
  element (``pmx_white_noise``); ``rng='torch'`` draws float64 normals from
  ``torch.Generator().manual_seed(seed)`` on the host.  ``type='complex'``
  returns r2c3d of the field (norm M)."""
  if isinstance(nc, int):
    nc = [nc, nc, nc]
  M = nc[0] * nc[1] * nc[2]
  device = device if device is not None else torch.device("cuda")
  shape = [batch_size] + list(nc)
  if rng == "philox":
    if seed is None:
      seed = int.from_bytes(np.random.bytes(16), "little")
    seed = int(seed)
    if seed < 0 or seed >= 1 << 128:
      raise ValueError("philox seed must be in [0, 2^128)")
    white = torch.empty(shape, dtype=torch.float32, device=device)
    _lib.require_cuda(white)
    _lib.call("pmx_white_noise", _lib.ptr(white), white.numel(),
              seed & ((1 << 64) - 1), seed >> 64, 0, float(M)**0.5,
              _lib.stream_of(white))
  elif rng == "torch":
    gen = torch.Generator()
    if seed is not None:
      gen.manual_seed(int(seed))
    else:
      gen.seed()
    white = (torch.randn(shape, generator=gen, dtype=torch.float64) *
             M**0.5).to(torch.float32).to(device)
  else:
    raise ValueError("rng must be 'philox' or 'torch'")
  if type == "real":
    return white
  if type == "complex":
    return r2c3d(white, norm=M)
  raise ValueError("type must be 'real' or 'complex'")


class _CompensateCic(torch.autograd.Function):
  """c2r3d(cic_compensation * r2c3d(field, norm=M), norm=M): a real,
  symmetric Fourier filter, so its adjoint is itself."""

  @staticmethod
  def apply_op(x):
    B, nx, ny, nz = _mesh_dims(x)
    plan = _plans.get(x.device, B, nx, ny, nz)
    xin = _f32(x)
    out = torch.empty_like(xin)
    work = plan.spec(3)
    ins = (_lib._vp * 1)(_lib._vp(xin.data_ptr()))
    e = (_lib._i * 3)(0, 0, 0)
    sc = (_lib._f * 1)(1.0)
    _lib.call("pmx_fourier_apply", plan.handle, 1, ins, e, sc, 2,
              1.0 / (nx * ny * nz), _lib.ptr(work), _lib.ptr(out), 0,
              _lib.stream_of(x))
    return out

  @staticmethod
  def forward(ctx, x):
    return _CompensateCic.apply_op(x)

  @staticmethod
  def backward(ctx, g):
    return _CompensateCic.apply_op(g)


def compensate_cic(field, name="CompensateCiC"):
  """Deconvolves the CIC window (utils.py:11-29): multiplies the spectrum by
  (sinc(kx/2pi) sinc(ky/2pi) sinc(kz/2pi))^-2 (kernels.py:137-152, Jing
  2005 with p = 2), evaluated per element from float64 per-axis tables."""
  _lib.require_cuda(field)
  return _CompensateCic.apply(field)
