"""PM solver drop-ins for flowpm/tfpm.py on the HIP path.

Signatures follow the reference (``name=`` accepted and ignored, ``nc`` an
int or a list).  State tensors are (3, batch, npart, 3) float32 on the GPU:
X (mesh units, never wrapped), P, F, as tfpm.py:179-180.  The growth and ODE
coefficients are host float64 (tfbackground); all particle and mesh work runs
in the HIP library (libpmx_hip.so):

  linear_field   tfpm.py:15-63    r2c, * sqrt(P(k)/V), c2r (3 HIP passes each)
  lpt1           tfpm.py:66-95    fused Poisson x gradient inverse + readout
  lpt2_source    tfpm.py:98-145   six phi_ij by one fan-out inverse, source,
                                  r2c
  lpt_init       tfpm.py:148-180  differentiable (hand-written adjoint)
  apply_longrange tfpm.py:183-219 (split = 0 on the path; split != 0 too)
  kick / drift   tfpm.py:273-314
  force          tfpm.py:317-366  paint -> fused r2c/Poisson/gradient/c2r ->
                                  readout
  nbody          tfpm.py:425-503  C++ time loop; one fused readout + kick +
                                  kick + drift kernel per step; differentiable
                                  with checkpoints every k steps and
                                  recomputation in the backward pass
"""
import numpy as np
import torch

from . import _lib
from . import _plans
from . import tfbackground as bg
from .cosmology import grad_params, host_fn
from .utils import cic_readout, white_noise, _f32, _half_to_full, \
    _full_to_real_half, fwd_real, inv_real

__all__ = ["linear_field", "whitenoise_to_linear", "lpt1", "lpt2_source",
           "lpt_init", "apply_longrange", "kick", "drift", "force", "nbody",
           "step_coefficients", "set_checkpoint_every",
           "set_keep_potentials", "set_compact_tape"]

_CKPT_EVERY = [None]
# automatic policy: record the whole trajectory when the tape fits in this
# fraction of the free device memory, else recompute segments of this size
_TAPE_FRACTION = 0.5
_AUTO_SEGMENT = 8


def set_checkpoint_every(k):
  """How differentiable nbody calls keep the trajectory for the adjoint.

  k >= 1: keep the state every k steps and recompute each segment in the
  backward pass; k = 0: record the positions and force meshes of every step
  in the forward pass (no recomputation; or the positions and potentials,
  see set_compact_tape); None: automatic (k = 0 when the tape, full or
  compact, fits in half of the free device memory, else k = 8)."""
  if k is not None and int(k) < 0:
    raise ValueError("checkpoint interval must be >= 0 (or None)")
  _CKPT_EVERY[0] = None if k is None else int(k)


_KEEP_PHI = [None]


def set_keep_potentials(keep):
  """Whether checkpointed differentiable nbody runs (checkpoint interval
  k >= 1) also keep the potential of every force evaluation (one float per
  mesh cell and step), so that the backward pass recomputes each segment's
  force meshes with the gradient stencil instead of painting and
  transforming again (same gradients, about half the recomputation time):
  True, False, or None (automatic: when they fit in half of the free device
  memory)."""
  _KEEP_PHI[0] = None if keep is None else bool(keep)


def _keep_phi(plan, nsteps, device):
  if _KEEP_PHI[0] is not None:
    return _KEEP_PHI[0]
  need = _lib.lib().pmx_nbody_phi_floats(plan.handle, nsteps) * 4
  return need <= _TAPE_FRACTION * torch.cuda.mem_get_info(device)[0]


_COMPACT = [None]


def set_compact_tape(compact):
  """Whether a taped differentiable nbody run (checkpoint interval 0) keeps
  the potential of every force evaluation instead of its float4 force mesh
  (about half the memory: at 256^3 x 40 steps 11 GB against 18.8 GB; the
  backward pass forms each force mesh again with the gradient stencil, the
  same gradients): True, False, or None (automatic: compact only when the
  full tape does not fit in half of the free device memory)."""
  _COMPACT[0] = None if compact is None else bool(compact)


def _tape_fits(plan, N, nsteps, device, name):
  need = getattr(_lib.lib(), name)(plan.handle, N, nsteps) * 4
  return need <= _TAPE_FRACTION * torch.cuda.mem_get_info(device)[0]


def _compact(plan, N, nsteps, device):
  if _COMPACT[0] is not None:
    return _COMPACT[0]
  return not _tape_fits(plan, N, nsteps, device, "pmx_nbody_tape_floats")


def _resolve_every(plan, N, nsteps, every, device):
  """Checkpoint interval of the automatic policy: 0 (a tape: the full one,
  or the compact one when only that fits in half of the free device
  memory), else segments of 8 steps."""
  if every is not None:
    return int(every)
  if _tape_fits(plan, N, nsteps, device, "pmx_nbody_tape_floats"):
    return 0
  if _COMPACT[0] is not False and _tape_fits(
      plan, N, nsteps, device, "pmx_nbody_compact_tape_floats"):
    return 0
  return _AUTO_SEGMENT


def _nc3(nc):
  if isinstance(nc, (int, np.integer)):
    return [int(nc)] * 3
  return [int(n) for n in nc]


# -------------------------------------------------------- initial field --
def _half_kmesh(nc, boxsize):
  """|k| (h/Mpc) on the library's half-spectrum layout of one mesh:
  nx*ny*(nz/2) main values then the nx*ny values of kz = nz/2."""
  kx = 2 * np.pi * np.fft.fftfreq(nc[0]) / boxsize[0] * nc[0]
  ky = 2 * np.pi * np.fft.fftfreq(nc[1]) / boxsize[1] * nc[1]
  kz = 2 * np.pi * np.fft.fftfreq(nc[2]) / boxsize[2] * nc[2]
  H = nc[2] // 2
  kxy2 = kx[:, None]**2 + ky[None, :]**2
  main = (kxy2[:, :, None] + kz[None, None, :H]**2)**0.5
  nyq = (kxy2 + kz[H]**2)**0.5
  return np.concatenate([main.reshape(-1), nyq.reshape(-1)])


class _WhiteToLinear(torch.autograd.Function):
  """linear = c2r(r2c(white) * amp); amp real and even, so the map is
  symmetric and its adjoint is itself (tfpm.py:56-62).  The cotangent of
  amp (when P(k) depends on cosmological parameters that require grad) is
  Re(W conj(G)) / M per half-spectrum mode, W and G the transforms of the
  white field and of the output cotangent, counted twice for the modes whose
  conjugate partner the half spectrum leaves out (0 < kz < nz/2)."""

  @staticmethod
  def forward(ctx, white, amp):
    B, nx, ny, nz = [int(s) for s in white.shape]
    plan = _plans.get(white.device, B, nx, ny, nz)
    ctx.plan = plan
    w = _f32(white)
    if ctx.needs_input_grad[1]:
      ctx.save_for_backward(amp, w)
    else:
      ctx.save_for_backward(amp)
    return _WhiteToLinear.apply_op(plan, w, amp.detach())

  @staticmethod
  def apply_op(plan, x, amp):
    spec = fwd_real(plan, x, 1.0 / plan.M)
    _lib.call("pmx_spec_mul_real", plan.handle, _lib.ptr(spec), _lib.ptr(amp),
              _lib.stream_of(spec))
    return inv_real(plan, spec, 1.0)

  @staticmethod
  def backward(ctx, g):
    plan = ctx.plan
    saved = ctx.saved_tensors
    amp = saved[0].detach()
    g = _f32(g)
    gw = _WhiteToLinear.apply_op(plan, g, amp) if ctx.needs_input_grad[0] \
        else None
    gamp = None
    if ctx.needs_input_grad[1]:
      B, nx, ny, H = plan.B, plan.nx, plan.ny, plan.nz // 2
      W = torch.view_as_complex(fwd_real(plan, saved[1], 1.0).view(-1, 2))
      G = torch.view_as_complex(fwd_real(plan, g, 1.0).view(-1, 2))
      prod = (W.real.double() * G.real.double() +
              W.imag.double() * G.imag.double())
      nm = B * nx * ny * H
      main = prod[:nm].view(B, nx, ny, H).sum(0)
      main[:, :, 1:] *= 2.0
      nyq = prod[nm:].view(B, nx, ny).sum(0)
      gamp = (torch.cat([main.reshape(-1), nyq.reshape(-1)]) /
              plan.M).to(amp.dtype)
    return gw, gamp


class _AmpTable(torch.autograd.Function):
  """The amplitude over the half spectrum of one mesh from its table per
  s = mx^2 + my^2 + mz^2 (pmx_amp_from_table); the cotangent of a table
  entry is the sum of its modes' cotangents."""

  @staticmethod
  def forward(ctx, tab, plan1, modes):
    dev = plan1.device
    tab32 = tab.detach().to(device=dev, dtype=torch.float32).contiguous()
    amp = torch.empty(modes, dtype=torch.float32, device=dev)
    _lib.call("pmx_amp_from_table", plan1.handle, _lib.ptr(tab32),
              tab32.numel(), _lib.ptr(amp), _lib.stream_of(amp))
    ctx.plan1, ctx.n, ctx.meta = plan1, tab.numel(), (tab.dtype, tab.device)
    return amp

  @staticmethod
  def backward(ctx, gamp):
    p = ctx.plan1
    dev = gamp.device
    H = p.nz // 2

    def m2(n):
      i = torch.arange(n, device=dev, dtype=torch.int64)
      m = torch.where(i < (n + 1) // 2, i, i - n)
      return m * m
    sx, sy, sz = m2(p.nx), m2(p.ny), m2(p.nz)
    sxy = sx[:, None] + sy[None, :]
    idx = torch.cat([(sxy[:, :, None] + sz[None, None, :H]).reshape(-1),
                     (sxy + sz[H]).reshape(-1)])
    gt = torch.zeros(ctx.n, dtype=torch.float64, device=dev)
    gt.index_add_(0, idx, gamp.double())
    return gt.to(dtype=ctx.meta[0], device=ctx.meta[1]), None, None


def whitenoise_to_linear(nc, boxsize, white, pk, kvec=None, name=None):
  """Linear field from an injected real white-noise field
  (scripts/check_grads.py:39-65 with the noise given in real space).

  white: (batch, nx, ny, nz) N(0, sqrt(M)); pk: callable on a float64 numpy
  |k| array (h/Mpc).  Differentiable w.r.t. ``white``."""
  nc = _nc3(nc)
  if np.isscalar(boxsize):
    boxsize = [float(boxsize)] * 3
  _lib.require_cuda(white)
  amp = _linear_amp(nc, [float(b) for b in boxsize], pk, white.device)
  return _WhiteToLinear.apply(white, amp)


def _linear_amp(nc, boxsize, pk, device):
  """sqrt(pk(|k|) / V) over the half spectrum of one mesh, float32 on the
  device (tfpm.py:56-62).  With equal box sides |k| = 2 pi sqrt(s) / L,
  s = mx^2 + my^2 + mz^2, so pk is evaluated once per s (at most
  3 (n/2)^2 + 1 values instead of n^3 / 2 modes: 49153 against 8.4 M at
  256^3) and the table is expanded over the modes on the GPU; other boxes
  evaluate pk on every mode on the host."""
  V = boxsize[0] * boxsize[1] * boxsize[2]
  smax = sum((n // 2)**2 for n in nc)
  modes = nc[0] * nc[1] * (nc[2] // 2 + 1)
  if boxsize[0] == boxsize[1] == boxsize[2] and smax + 1 < modes:
    s = np.arange(smax + 1, dtype=np.float64)
    k = 2 * np.pi * np.sqrt(s) / boxsize[0]
    pkv = pk(k)
    plan = _plans.get(device, 1, *nc)
    if isinstance(pkv, torch.Tensor) and pkv.requires_grad:
      # P(k) depends on parameters that require grad
      tab = (pkv.to(torch.float64).reshape(-1) / V)**0.5
      return _AmpTable.apply(tab, plan, modes)
    tab = (np.asarray(_np(pkv), dtype=np.float64).reshape(-1) / V)**0.5
    tab = torch.as_tensor(tab.astype(np.float32), device=device)
    amp = torch.empty(modes, dtype=torch.float32, device=device)
    _lib.call("pmx_amp_from_table", plan.handle, _lib.ptr(tab), smax + 1,
              _lib.ptr(amp), _lib.stream_of(amp))
    return amp
  pkv = pk(_half_kmesh(nc, boxsize))
  if isinstance(pkv, torch.Tensor) and pkv.requires_grad:
    return ((pkv.to(torch.float64).reshape(-1) / V)**0.5).to(
        device=device, dtype=torch.float32)
  amp = (np.asarray(_np(pkv), dtype=np.float64).reshape(-1) / V)**0.5
  return torch.as_tensor(amp.astype(np.float32), device=device)


def _np(v):
  return v.detach().cpu().numpy() if isinstance(v, torch.Tensor) else v


def linear_field(nc, boxsize, pk, kvec=None, batch_size=1, seed=None,
                 dtype=torch.float32, name="LinearField", device=None):
  """Gaussian linear field with power spectrum pk (tfpm.py:15-63)."""
  nc = _nc3(nc)
  white = white_noise(nc, batch_size=batch_size, seed=seed, type="real",
                      device=device)
  return whitenoise_to_linear(nc, boxsize, white, pk)


# -------------------------------------------------------------------- LPT --
def _poisson_grad_meshes(plan, spec, split=0.0):
  """3 meshes c2r(dk * lap [* exp(-k^2 split^2)] * i g_d), as one
  (3, B, nx, ny, nz) tensor."""
  out = torch.empty((3, plan.B, plan.nx, plan.ny, plan.nz),
                    dtype=torch.float32, device=spec.device)
  work = plan.spec(3)
  if split == 0:
    _lib.call("pmx_fft_c2r_poisson_grad", plan.handle, _lib.ptr(spec),
              _lib.ptr(work), _lib.ptr(out[0]), _lib.ptr(out[1]),
              _lib.ptr(out[2]), 1.0, _lib.stream_of(spec))
  else:
    _lib.call("pmx_fft_c2r_longrange_grad", plan.handle, _lib.ptr(spec),
              _lib.ptr(work), _lib.ptr(out[0]), _lib.ptr(out[1]),
              _lib.ptr(out[2]), 1.0, float(split), _lib.stream_of(spec))
  return out


class _PoissonGradMeshes(torch.autograd.Function):
  """(3, B, nx, ny, nz) meshes c2r3d(dk * lap [* exp(-k^2 split^2)] * i g_d)
  of a full complex spectrum dk (tfpm.py:66-95, 183-219).  The gradient of
  a real loss w.r.t. dk (torch's convention dL/dRe + i dL/dIm) is
  sum_d conj(K_d) FFT3(g_d), K_d the mode's multiplier, a Hermitian
  spectrum (pmx_fft_poisson_grad_adj / pmx_fft_longrange_grad_adj)."""

  @staticmethod
  def forward(ctx, dk, split):
    _lib.require_cuda(dk)
    B, nx, ny, nz = [int(s) for s in dk.shape]
    plan = _plans.get(dk.device, B, nx, ny, nz)
    ctx.plan, ctx.split = plan, float(split)
    return _poisson_grad_meshes(plan, _full_to_real_half(plan, dk),
                                split=float(split))

  @staticmethod
  def backward(ctx, g):
    plan = ctx.plan
    g = _f32(g)
    gdk = plan.spec()
    work = plan.spec(3)
    st = _lib.stream_of(g)
    if ctx.split == 0:
      _lib.call("pmx_fft_poisson_grad_adj", plan.handle, _lib.ptr(g[0]),
                _lib.ptr(g[1]), _lib.ptr(g[2]), _lib.ptr(work), _lib.ptr(gdk),
                1.0, st)
    else:
      _lib.call("pmx_fft_longrange_grad_adj", plan.handle, _lib.ptr(g[0]),
                _lib.ptr(g[1]), _lib.ptr(g[2]), _lib.ptr(work), _lib.ptr(gdk),
                1.0, ctx.split, st)
    return _half_to_full(plan, gdk), None


def _dot(a, b):
  """sum(a * b) of two float32 device tensors, in float64 on the GPU."""
  out = torch.empty(1, dtype=torch.float64, device=a.device)
  _lib.call("pmx_mesh_dot", _lib.ptr(a), _lib.ptr(b), a.numel(),
            _lib.ptr(out), _lib.stream_of(a))
  return out[0]


class _PgdMeshes(torch.autograd.Function):
  """The three meshes of apply_pgd (tfpm.py:222-270): c2r3d(dk * lap *
  PGD_kernel(kl, ks) * i g_d).  Backward: the conjugate multiplier for dk
  (pmx_fft_pgd_grad_adj) and, for kl and ks, the contraction of the mesh
  cotangents with the meshes of dPGD/dkl and dPGD/dks."""

  @staticmethod
  def forward(ctx, dk, kl, ks):
    _lib.require_cuda(dk)
    B, nx, ny, nz = [int(s) for s in dk.shape]
    plan = _plans.get(dk.device, B, nx, ny, nz)
    klf, ksf = float(kl), float(ks)
    if klf == 0 or ksf == 0:
      raise ValueError("PGD scales kl and ks must be nonzero")
    spec = _full_to_real_half(plan, dk)
    out = _pgd_meshes(plan, spec, klf, ksf, 0)
    ctx.plan, ctx.kl, ctx.ks = plan, klf, ksf
    ctx.save_for_backward(spec)
    ctx.ptype = [(p.dtype, p.device) if isinstance(p, torch.Tensor) else None
                 for p in (kl, ks)]
    return out

  @staticmethod
  def backward(ctx, g):
    plan = ctx.plan
    (spec,) = ctx.saved_tensors
    g = _f32(g)
    gdk = None
    if ctx.needs_input_grad[0]:
      sp = plan.spec()
      work = plan.spec(3)
      _lib.call("pmx_fft_pgd_grad_adj", plan.handle, _lib.ptr(g[0]),
                _lib.ptr(g[1]), _lib.ptr(g[2]), _lib.ptr(work), _lib.ptr(sp),
                1.0, ctx.kl, ctx.ks, _lib.stream_of(g))
      gdk = _half_to_full(plan, sp)
    gp = []
    for q in (1, 2):
      if ctx.needs_input_grad[q] and ctx.ptype[q - 1] is not None:
        dt, dev = ctx.ptype[q - 1]
        v = _dot(g, _pgd_meshes(plan, spec, ctx.kl, ctx.ks, q))
        gp.append(v.to(dtype=dt, device=dev))
      else:
        gp.append(None)
    return gdk, gp[0], gp[1]


def _pgd_meshes(plan, spec, kl, ks, deriv):
  out = torch.empty((3, plan.B, plan.nx, plan.ny, plan.nz),
                    dtype=torch.float32, device=spec.device)
  work = plan.spec(3)
  _lib.call("pmx_fft_c2r_pgd_grad", plan.handle, _lib.ptr(spec),
            _lib.ptr(work), _lib.ptr(out[0]), _lib.ptr(out[1]),
            _lib.ptr(out[2]), 1.0, float(kl), float(ks), int(deriv),
            _lib.stream_of(spec))
  return out


def _readout3(meshes, pos):
  return torch.stack([cic_readout(meshes[d], pos) for d in range(3)], dim=2)


def lpt1(dlin_k, pos, kvec=None, name="LTP1"):
  """First-order LPT displacement read out at ``pos`` (tfpm.py:66-95);
  differentiable w.r.t. ``dlin_k`` and ``pos``."""
  return _readout3(_PoissonGradMeshes.apply(dlin_k, 0.0), pos)


def _lpt2_src(plan, spec):
  work = torch.empty(_lib.lib().pmx_lpt2_work_floats(plan.handle),
                     dtype=torch.float32, device=spec.device)
  src = plan.spec()
  _lib.call("pmx_lpt2_source", plan.handle, _lib.ptr(spec), _lib.ptr(work),
            _lib.ptr(src), _lib.stream_of(spec))
  return src, work


# exponents of (i g_x, i g_y, i g_z) of the six phi_ij transforms (xx, yy,
# zz, then xy, xz, yz), the order pmx_lpt2_source_adj returns them in
_PHI_EXPS = (((0, 1, 2), (2, 0, 0, 0, 2, 0, 0, 0, 2)),
             ((3, 4, 5), (1, 1, 0, 1, 0, 1, 0, 1, 1)))


def _phi_adjoint(plan, gphi, work, out, accumulate):
  """out (+)= sum_ij IFFT3(lap (i g_i)(i g_j) FFT3(gphi_ij)) (normalised
  inverse): the adjoint of the six phi_ij = c2r(dk lap (i g_i)(i g_j))
  maps, in real space (the multipliers are real and even)."""
  st = _lib.stream_of(gphi)
  for k, (idx, exps) in enumerate(_PHI_EXPS):
    ins = (_lib._vp * 3)(*[gphi[i].data_ptr() for i in idx])
    e = (_lib._i * 9)(*exps)
    sc = (_lib._f * 3)(1.0, 1.0, 1.0)
    _lib.call("pmx_fourier_apply", plan.handle, 3, ins, e, sc, 1,
              1.0 / plan.M, _lib.ptr(work), _lib.ptr(out),
              1 if (accumulate or k > 0) else 0, st)


class _Lpt2Source(torch.autograd.Function):
  """lpt2_source (tfpm.py:98-145) with its hand-written adjoint: r2c3d's
  conjugate transform, the pointwise adjoint of the source combination
  (pmx_lpt2_source_adj), then the six phi_ij multipliers."""

  @staticmethod
  def forward(ctx, dk):
    _lib.require_cuda(dk)
    B, nx, ny, nz = [int(s) for s in dk.shape]
    plan = _plans.get(dk.device, B, nx, ny, nz)
    src, work2 = _lpt2_src(plan, _full_to_real_half(plan, dk))
    ctx.plan = plan
    ctx.save_for_backward(work2)
    return _half_to_full(plan, src)

  @staticmethod
  def backward(ctx, g):
    plan = ctx.plan
    (work2,) = ctx.saved_tensors
    M = plan.M
    # cotangent of the real source: the adjoint of r2c3d(., norm = M)
    gsrc = inv_real(plan, _full_to_real_half(plan, g), 1.0 / M)
    dev = gsrc.device
    gphi = torch.empty((6, plan.B, plan.nx, plan.ny, plan.nz),
                       dtype=torch.float32, device=dev)
    st = _lib.stream_of(gsrc)
    _lib.call("pmx_lpt2_source_adj", plan.handle, _lib.ptr(work2),
              _lib.ptr(gsrc), _lib.ptr(gphi), st)
    # sum_ij K_ij FFT3(gphi_ij) as a spectrum: the real-space sum, then r2c
    G = torch.empty_like(gsrc)
    _phi_adjoint(plan, gphi, plan.spec(3), G, 0)
    return _half_to_full(plan, fwd_real(plan, G, 1.0))


def lpt2_source(dlin_k, kvec=None, name="LPT2Source"):
  """Second-order LPT source term, full complex spectrum (tfpm.py:98-145);
  differentiable w.r.t. ``dlin_k``."""
  return _Lpt2Source.apply(dlin_k)


def lpt_coefficients(cosmo, a, order):
  """(D1, a^2 f1 E, a^2 E gf / D1, D2, a^2 f2 E, a^2 E gf2 / D2) in float64
  (tfpm.py:168-177)."""
  a = float(a)
  E = float(bg.E(cosmo, a))
  d1 = float(bg.D1(cosmo, a))
  c = [d1, a**2 * float(bg.f1(cosmo, a)) * E,
       a**2 * E * float(bg.gf(cosmo, a)) / d1, 0.0, 0.0, 0.0]
  if order == 2:
    d2 = float(bg.D2(cosmo, a))
    c[3:] = [d2, a**2 * float(bg.f2(cosmo, a)) * E,
             a**2 * E * float(bg.gf2(cosmo, a)) / d2]
  return c


class _LptInit(torch.autograd.Function):
  """coef_t (optional): the six coefficients as a float64 tensor that
  requires grad (a cosmology with parameters that require grad); their
  cotangents are the dot products of the state cotangent with the
  displacement meshes."""

  @staticmethod
  def forward(ctx, linear, coef, order, coef_t=None):
    B, nx, ny, nz = [int(s) for s in linear.shape]
    plan = _plans.get(linear.device, B, nx, ny, nz)
    M = plan.M
    dk = fwd_real(plan, _f32(linear), 1.0 / M)
    dx1 = _poisson_grad_meshes(plan, dk)
    dx2 = None
    work2 = None
    if order == 2:
      src, work2 = _lpt2_src(plan, dk)
      dx2 = _poisson_grad_meshes(plan, src)
    del dk
    state = torch.empty((3, B, M, 3), dtype=torch.float32,
                        device=linear.device)
    _lib.call("pmx_lpt_assemble", _lib.ptr(state), _lib.ptr(dx1),
              _lib.ptr(dx2), _lib.dbl(coef), B, nx, ny, nz,
              _lib.stream_of(state))
    ctx.plan, ctx.coef, ctx.order = plan, coef, order
    if order == 2:
      ctx.save_for_backward(work2)
    ctx.dx = (dx1, dx2) if (len(ctx.needs_input_grad) > 3 and
                            ctx.needs_input_grad[3]) else None
    return state

  @staticmethod
  def _coef_grad(ctx, g):
    """d L / d coef from the state cotangent g (3, B, N, 3): the state is
    X = Q + c0 dx1 + c3 dx2, P = c1 c0 dx1 + c4 c3 dx2,
    F = c2 c0 dx1 + c5 c3 dx2 (pmx_lpt_assemble)."""
    c = ctx.coef
    B, M = ctx.plan.B, ctx.plan.M
    gd = [g[i].permute(2, 0, 1).reshape(3, B, M).double() for i in range(3)]
    out = [0.0] * 6
    for j, dx in enumerate(ctx.dx):
      if dx is None:
        continue
      d = dx.reshape(3, B, M).double()
      dots = [float((gd[i] * d).sum()) for i in range(3)]
      c0, c1, c2 = c[3 * j:3 * j + 3]
      out[3 * j] = dots[0] + c1 * dots[1] + c2 * dots[2]
      out[3 * j + 1] = c0 * dots[1]
      out[3 * j + 2] = c0 * dots[2]
    return torch.tensor(out, dtype=torch.float64)

  @staticmethod
  def backward(ctx, g):
    plan, coef, order = ctx.plan, ctx.coef, ctx.order
    B, nx, ny, nz, M = plan.B, plan.nx, plan.ny, plan.nz, plan.M
    g = _f32(g)
    gcoef = _LptInit._coef_grad(ctx, g) if ctx.dx is not None else None
    if not ctx.needs_input_grad[0]:
      return None, None, None, gcoef
    dev = g.device
    st = _lib.stream_of(g)
    g1 = torch.empty((3, B, nx, ny, nz), dtype=torch.float32, device=dev)
    g2 = torch.empty_like(g1) if order == 2 else None
    _lib.call("pmx_lpt_assemble_adj", _lib.ptr(g), _lib.dbl(coef),
              _lib.ptr(g1), _lib.ptr(g2), B, nx, ny, nz, st)
    work = plan.spec(3)
    glin = torch.empty((B, nx, ny, nz), dtype=torch.float32, device=dev)
    _lib.call("pmx_force_meshes_adj", plan.handle, _lib.ptr(g1[0]),
              _lib.ptr(g1[1]), _lib.ptr(g1[2]), _lib.ptr(work),
              _lib.ptr(glin), 1.0 / M, st)
    if order == 2:
      (work2,) = ctx.saved_tensors
      gsrc = torch.empty_like(glin)
      _lib.call("pmx_force_meshes_adj", plan.handle, _lib.ptr(g2[0]),
                _lib.ptr(g2[1]), _lib.ptr(g2[2]), _lib.ptr(work),
                _lib.ptr(gsrc), 1.0 / M, st)
      gphi = torch.empty((6, B, nx, ny, nz), dtype=torch.float32, device=dev)
      _lib.call("pmx_lpt2_source_adj", plan.handle, _lib.ptr(work2),
                _lib.ptr(gsrc), _lib.ptr(gphi), st)
      # phi_ij^T = c2r(lap (i g_i)(i g_j) r2c(.)/M), added to glin
      _phi_adjoint(plan, gphi, work, glin, 1)
    return glin, None, None, gcoef


def lpt_init(cosmo, linear, a, order=2, kvec=None, name="LPTInit"):
  """LPT initial state (3, batch, npart, 3) from a linear field
  (tfpm.py:148-180); differentiable w.r.t. ``linear`` and w.r.t. the
  cosmology's parameters given as tensors that require grad."""
  assert order in (1, 2)
  _lib.require_cuda(linear)
  coef = lpt_coefficients(cosmo, a, order)
  coef_t = None
  if grad_params(cosmo) and torch.is_grad_enabled():
    coef_t = host_fn(lambda c: lpt_coefficients(c, a, order), cosmo)
  return _LptInit.apply(linear, coef, order, coef_t)


# ------------------------------------------------------------------ nbody --
def apply_longrange(x, delta_k, split=0, factor=1, kvec=None,
                    name="ApplyLongrange"):
  """Long-range force at positions x from delta_k (tfpm.py:183-219); with
  split != 0 the Laplacian is filtered by exp(-k^2 split^2)
  (kernels.py:81-102).  Differentiable w.r.t. ``x`` and ``delta_k``."""
  return _readout3(_PoissonGradMeshes.apply(delta_k, float(split)), x) * factor


def apply_pgd(x, delta_k, alpha, kl, ks, kvec=None, name="ApplyPGD"):
  """PGD displacement alpha * readout(c2r(delta_k lap PGD_kernel(kl, ks)
  i g_d), x) (tfpm.py:222-270, kernels.py:105-134); differentiable w.r.t.
  ``x``, ``delta_k`` and (as tensors) ``alpha``, ``kl``, ``ks``."""
  return _readout3(_PgdMeshes.apply(delta_k, kl, ks), x) * alpha


class _Pgd(torch.autograd.Function):
  """One PGD correction of the state (tfpm.py:369-422), pmx_pgd forward and
  pmx_pgd_adjoint backward (state cotangent and d/d(alpha, kl, ks))."""

  @staticmethod
  def forward(ctx, state, pgdparams, plan, N, f):
    st = _f32(state).clone()
    a, kl, ks = [float(v) for v in _pgd_row(pgdparams)]
    X = st[0].clone()
    _lib.call("pmx_pgd", plan.handle, _lib.ptr(st), N, f, a, kl, ks,
              _lib.ptr(_work(plan)), _lib.stream_of(st))
    ctx.save_for_backward(X)
    ctx.args = (plan, N, f, a, kl, ks)
    ctx.pmeta = (pgdparams.dtype, pgdparams.device, tuple(pgdparams.shape)) \
        if isinstance(pgdparams, torch.Tensor) else None
    return st

  @staticmethod
  def backward(ctx, g):
    (X,) = ctx.saved_tensors
    plan, N, f, a, kl, ks = ctx.args
    gs = _f32(g).clone()
    want = ctx.pmeta is not None and ctx.needs_input_grad[1]
    gp = torch.empty(3, dtype=torch.float64, device=gs.device) \
        if want else None
    work = torch.empty(_lib.lib().pmx_pgd_adj_work_floats(plan.handle, N),
                       dtype=torch.float32, device=gs.device)
    _lib.call("pmx_pgd_adjoint", plan.handle, _lib.ptr(X), N, f, a, kl, ks,
              _lib.ptr(gs), _lib.ptr(gp), _lib.ptr(work), _lib.stream_of(gs))
    gpgd = None
    if want:
      dt, dev, shape = ctx.pmeta
      gpgd = gp.reshape(shape).to(dtype=dt, device=dev)
    return gs, gpgd, None, None, None


def _pgd_row(pgdparams):
  if isinstance(pgdparams, torch.Tensor):
    pgdparams = pgdparams.detach().cpu().numpy()
  arr = np.asarray(pgdparams, dtype=np.float32).reshape(-1)
  if arr.size != 3:
    raise ValueError("pgdparams must hold (alpha, kl, ks)")
  if arr[1] == 0 or arr[2] == 0:
    raise ValueError("PGD scales kl and ks must be nonzero")
  return arr


def pgd(state, pgdparams, nc, pm_nc_factor=1, name="PGD", **kwargs):
  """One PGD correction X += apply_pgd(f X, delta_k(X), alpha, kl, ks) / f
  (tfpm.py:369-422); differentiable w.r.t. ``state`` and (as a tensor)
  ``pgdparams`` = (alpha, kl, ks).  The gradient needs the particles on the
  Lagrangian grid of nc (the nbody state)."""
  _lib.require_cuda(state)
  plan, B, N, f, _ = _force_setup(None, state, nc, pm_nc_factor)
  return _Pgd.apply(state, pgdparams, plan, N, f)


def _kick_fac(cosmo, ai, ac, af):
  return float(1 / (ac**2 * bg.E(cosmo, ac)) *
               (bg.Gf(cosmo, af) - bg.Gf(cosmo, ai)) / bg.gf(cosmo, ac))


def _drift_fac(cosmo, ai, ac, af):
  return float(1. / (ac**3 * bg.E(cosmo, ac)) *
               (bg.D1(cosmo, af) - bg.D1(cosmo, ai)) / bg.D1f(cosmo, ac))


def _state_dims(state):
  if state.dim() != 4 or state.shape[0] != 3 or state.shape[3] != 3:
    raise ValueError("state must be (3, batch, npart, 3), got %s" %
                     (tuple(state.shape),))
  return int(state.shape[1]), int(state.shape[2])


def _axpy(st, dst, src, c):
  """st[dst] += c * st[src] on a (3, B, N, 3) tensor, in place (the
  kick-only path of pmx_readout_kick_drift)."""
  B, N = int(st.shape[1]), int(st.shape[2])
  _lib.call("pmx_readout_kick_drift", _lib.ptr(st[0]), _lib.ptr(st[dst]),
            _lib.ptr(st[src]), None, None, None, B, N, 1, 1, 1, 1.0, 0.0,
            float(c), 0.0, 0.0, 1 | 4, _lib.stream_of(st))


class _Axpy(torch.autograd.Function):
  """kick (P += c F, dst 1, src 2) and drift (X += c P, dst 0, src 1):
  linear, so the cotangent update is g[src] += c g[dst]; with c_t (c as a
  tensor that requires grad) also d L / d c = g[dst] . state[src]."""

  @staticmethod
  def forward(ctx, state, dst, src, c, c_t=None):
    _lib.require_cuda(state)
    _state_dims(state)
    st = _f32(state).clone()
    ctx.src_val = st[src].clone() if (len(ctx.needs_input_grad) > 4 and
                                      ctx.needs_input_grad[4]) else None
    _axpy(st, dst, src, c)
    ctx.args = (dst, src, c)
    return st

  @staticmethod
  def backward(ctx, g):
    dst, src, c = ctx.args
    g = _f32(g)
    gc = None
    if ctx.src_val is not None:
      gc = torch.tensor(float((g[dst].double() * ctx.src_val.double()).sum()),
                        dtype=torch.float64)
    gs = g.clone()
    _axpy(gs, src, dst, c)
    return (gs, None, None, None, gc)[:len(ctx.needs_input_grad)]


def _coef_tensor(cosmo, fn):
  """fn(cosmo) as a tensor differentiable w.r.t. the cosmology's grad
  parameters, or None when it has none."""
  if not grad_params(cosmo) or not torch.is_grad_enabled():
    return None
  return host_fn(fn, cosmo)


def kick(cosmo, state, ai, ac, af, dtype=torch.float32, name="Kick", **kw):
  """P += fac F (tfpm.py:273-292); differentiable w.r.t. ``state`` and the
  cosmology's parameters given as tensors that require grad."""
  c_t = _coef_tensor(cosmo, lambda c: _kick_fac(c, ai, ac, af))
  return _Axpy.apply(state, 1, 2, _kick_fac(cosmo, ai, ac, af), c_t)


def drift(cosmo, state, ai, ac, af, dtype=torch.float32, name="Drift", **kw):
  """X += fac P (tfpm.py:295-314); differentiable w.r.t. ``state`` and the
  cosmology's parameters given as tensors that require grad."""
  c_t = _coef_tensor(cosmo, lambda c: _drift_fac(c, ai, ac, af))
  return _Axpy.apply(state, 0, 1, _drift_fac(cosmo, ai, ac, af), c_t)


def _force_setup(cosmo, state, nc, pm_nc_factor):
  B, N = _state_dims(state)
  nc = _nc3(nc)
  if nc[0] * nc[1] * nc[2] != N:
    raise ValueError("npart (%d) must equal prod(nc) (%d)" %
                     (N, nc[0] * nc[1] * nc[2]))
  f = int(pm_nc_factor)
  plan = _plans.get(state.device, B, nc[0] * f, nc[1] * f, nc[2] * f)
  fac = float(np.float32(1.5 * cosmo.Omega_m)) if cosmo is not None else 0.0
  return plan, B, N, float(f), fac


def _work(plan):
  return torch.empty(_lib.lib().pmx_nbody_work_floats(plan.handle),
                     dtype=torch.float32, device=plan.device)


class _Force(torch.autograd.Function):
  """F = force(X) (tfpm.py:317-366) as a taped run of zero steps: the
  forward keeps X and the force mesh, the backward is the jump-start force
  VJP of the nbody reverse sweep (gX += J^T gF, gF = 0).  The reverse
  sweep rewrites only the fourth (density cotangent) component of the
  taped force mesh, so the tape also serves a second backward pass
  (retain_graph=True); it is released with the graph otherwise."""

  @staticmethod
  def forward(ctx, state, plan, N, f, fac, fac_t=None):
    st = _f32(state).clone()
    tape = torch.empty(_lib.lib().pmx_nbody_tape_floats(plan.handle, N, 0),
                       dtype=torch.float32, device=st.device)
    _lib.call("pmx_nbody_run_taped", plan.handle, _lib.ptr(st), N, f, fac,
              None, 0, _lib.ptr(tape), _lib.ptr(_work(plan)),
              _lib.stream_of(st))
    ctx.save_for_backward(tape)
    ctx.args = (plan, N, f, fac)
    # F = fac G(X): d L / d fac = gF . F / fac
    ctx.f_out = st[2].clone() if (len(ctx.needs_input_grad) > 5 and
                                  ctx.needs_input_grad[5]) else None
    return st

  @staticmethod
  def backward(ctx, g):
    plan, N, f, fac = ctx.args
    (tape,) = ctx.saved_tensors
    g = _f32(g)
    gfac = None
    if ctx.f_out is not None:
      gfac = torch.tensor(
          float((g[2].double() * ctx.f_out.double()).sum()) / fac,
          dtype=torch.float64)
    gs = g.clone()
    work = torch.empty(_lib.lib().pmx_nbody_adj_work_floats(plan.handle, N,
                                                            0),
                       dtype=torch.float32, device=gs.device)
    _lib.call("pmx_nbody_adjoint_taped", plan.handle, _lib.ptr(tape), 0, N,
              f, fac, None, _lib.ptr(gs), _lib.ptr(work), None,
              _lib.stream_of(gs))
    return (gs, None, None, None, None, gfac)[:len(ctx.needs_input_grad)]


def force(cosmo, state, nc, pm_nc_factor=1, dtype=torch.float32, name="Force",
          **kw):
  """Replaces F by the PM force at X (tfpm.py:317-366); differentiable
  w.r.t. ``state`` and the cosmology's parameters given as tensors that
  require grad (through the factor 1.5 Omega_m)."""
  _lib.require_cuda(state)
  plan, B, N, f, fac = _force_setup(cosmo, state, nc, pm_nc_factor)
  fac_t = _coef_tensor(cosmo, lambda c: 1.5 * c.Omega_m)
  if torch.is_grad_enabled() and (state.requires_grad or fac_t is not None):
    return _Force.apply(state, plan, N, f, fac, fac_t)
  st = _f32(state).clone()
  _lib.call("pmx_nbody_run", plan.handle, _lib.ptr(st), N, f, fac, None, 0, 0,
            1, _lib.ptr(_work(plan)), _lib.stream_of(st))
  return st


def step_coefficients(cosmo, stages):
  """Per-step (k1, d, k2) of the KDK scheme of tfpm.py:472-498 (float64):
  opening kick (a0 -> ah), drift (a0 -> a1 at ah), closing kick
  (ah -> a1 at a1), ah = sqrt(a0 a1)."""
  # evaluated on all steps at once: the background functions are
  # elementwise, so each value equals the scalar _kick_fac / _drift_fac
  # (the per-call overhead of 3 x nsteps scalar evaluations was host time
  # with the GPU idle in front of every run)
  st = np.asarray(stages, dtype=np.float64)
  if len(st) < 2:
    return []
  a0, a1 = st[:-1], st[1:]
  ah = (a0 * a1)**0.5
  k1 = 1 / (a0**2 * bg.E(cosmo, a0)) * (
      bg.Gf(cosmo, ah) - bg.Gf(cosmo, a0)) / bg.gf(cosmo, a0)
  d = 1. / (ah**3 * bg.E(cosmo, ah)) * (
      bg.D1(cosmo, a1) - bg.D1(cosmo, a0)) / bg.D1f(cosmo, ah)
  k2 = 1 / (a1**2 * bg.E(cosmo, a1)) * (
      bg.Gf(cosmo, a1) - bg.Gf(cosmo, ah)) / bg.gf(cosmo, a1)
  return [float(v) for v in np.stack([k1, d, k2], axis=1).reshape(-1)]


def _coef_cotangents(coef, fac, gc, gfin):
  """Cotangents of (k1_i, d_i, k2_i per step, fac) from the reverse sweep's
  scalar sums gc (pmx_nbody_adjoint's gcoef: a_i = gP . F_i, b_i = gX . P_i,
  a_J = gP . F_jump) and gfin = gF_final . F_final.  F_i feeds the closing
  kick of step i and the opening kick of step i + 1, the jump-start force
  the opening kick of step 0; every force is fac times a function of the
  positions, so d L / d fac = sum over forces of gF . F / fac."""
  n = len(coef) // 3
  a, b, aj = gc[:n], gc[n:2 * n], gc[2 * n]
  out = np.zeros(3 * n + 1)
  if n:
    out[0] = aj
    out[3:3 * n:3] = a[:n - 1]
    out[1::3][:n] = b
    out[2::3][:n] = a
  k1 = np.asarray(coef[0::3], np.float64)
  k2 = np.asarray(coef[2::3], np.float64)
  wa = k2 + np.append(k1[1:], 0.0)  # kicks that use F_i
  out[3 * n] = ((float(np.dot(wa, a)) + (k1[0] * aj if n else 0.0) + gfin) /
                fac)
  return torch.tensor(out, dtype=torch.float64)


class _Nbody(torch.autograd.Function):
  """cpar (optional): the step coefficients and the force factor,
  (k1, d, k2) per step then fac, as a float64 tensor that requires grad (a
  cosmology with parameters that require grad); the reverse sweep then
  also accumulates their scalar cotangents (gcoef of pmx_nbody_adjoint)."""

  @staticmethod
  def forward(ctx, state, plan, N, f, fac, coef, every, need_grad,
              cpar=None):
    B = plan.B
    nsteps = len(coef) // 3
    x0 = _f32(state)  # read only
    work = _work(plan)
    cptr = _lib.dbl(coef)
    stream = _lib.stream_of(x0)
    if not need_grad:
      st = x0.clone()
      _lib.call("pmx_nbody_run", plan.handle, _lib.ptr(st), N, f, fac, cptr,
                0, nsteps, 1, _lib.ptr(work), stream)
      return st
    ctx.want_c = len(ctx.needs_input_grad) > 8 and ctx.needs_input_grad[8]
    st = _Nbody._run_for_grad(ctx, x0, plan, N, f, fac, coef, every, work,
                              cptr, stream)
    if ctx.want_c:
      ctx.f_final = st[2].clone()
    return st

  @staticmethod
  def _run_for_grad(ctx, x0, plan, N, f, fac, coef, every, work, cptr,
                    stream):
    """x0: the initial state (not modified); returns the final state."""
    B = plan.B
    nsteps = len(coef) // 3
    if every == 0:
      st = x0.clone()
      # the full tape (X and the float4 force mesh of every force), or the
      # compact one (X and the potential)
      ctx.compact = _compact(plan, N, nsteps, x0.device)
      kind = "compact_" if ctx.compact else ""
      tape = torch.empty(getattr(_lib.lib(), "pmx_nbody_%stape_floats" %
                                 kind)(plan.handle, N, nsteps),
                         dtype=torch.float32, device=x0.device)
      _lib.call("pmx_nbody_run_%staped" % kind, plan.handle, _lib.ptr(st), N,
                f, fac, cptr, nsteps, _lib.ptr(tape), _lib.ptr(work), stream)
      # released with the graph; the reverse sweep rewrites only the fourth
      # (density cotangent) component of its force meshes, so a second
      # backward pass (retain_graph=True) reads an intact tape
      ctx.save_for_backward(tape)
      ctx.args = (plan, N, f, fac, coef, every)
      return st
    nseg = max(1, (nsteps + every - 1) // every)
    ckpt = torch.empty((nseg, 3, B, N, 3), dtype=torch.float32,
                       device=x0.device)
    # the reverse sweep's workspace holds the recomputation tape of one
    # segment; the last segment is run into it here, so the backward pass
    # recomputes the other segments only
    k = min(every, max(nsteps, 1))
    adj = torch.empty(_lib.lib().pmx_nbody_adj_work_floats(plan.handle, N, k),
                      dtype=torch.float32, device=x0.device)
    # the potential of every force evaluation, when kept: the backward pass
    # then recomputes a segment's force meshes with the gradient stencil
    phi = None
    if _keep_phi(plan, nsteps, x0.device):
      phi = torch.empty(_lib.lib().pmx_nbody_phi_floats(plan.handle, nsteps),
                        dtype=torch.float32, device=x0.device)
    ckpt[0].copy_(x0)
    s0 = 0
    if nseg >= 3:
      # every segment but the last in one run from the first checkpoint,
      # whose readouts write the checkpoints they cross (no copies, no split
      # kicks)
      st = torch.empty_like(x0)
      _lib.call("pmx_nbody_run_checkpointed", plan.handle, _lib.ptr(st),
                _lib.ptr(ckpt[0]), N, f, fac, cptr, 0, (nseg - 1) * every, 1,
                every, _lib.ptr(ckpt[1]), _lib.ptr(work),
                None if phi is None else _lib.ptr(phi), stream)
      s0 = nseg - 1
    else:
      st = x0.clone()
    for s in range(s0, nseg):
      if s > 0:
        ckpt[s].copy_(st)
      b0, b1 = s * every, min(nsteps, (s + 1) * every)
      if s < nseg - 1 or nsteps == 0:
        if phi is None:
          _lib.call("pmx_nbody_run", plan.handle, _lib.ptr(st), N, f, fac,
                    cptr, b0, b1, 1 if s == 0 else 0, _lib.ptr(work), stream)
        else:
          _lib.call("pmx_nbody_run_phi", plan.handle, _lib.ptr(st), N, f,
                    fac, cptr, b0, b1, 1 if s == 0 else 0, _lib.ptr(work),
                    _lib.ptr(phi), stream)
      elif phi is None:
        _lib.call("pmx_nbody_run_last_segment", plan.handle, _lib.ptr(st), N,
                  f, fac, cptr, nsteps, k, _lib.ptr(adj), stream)
      else:
        _lib.call("pmx_nbody_run_last_segment_phi", plan.handle,
                  _lib.ptr(st), N, f, fac, cptr, nsteps, k, _lib.ptr(adj),
                  _lib.ptr(phi), stream)
    ctx.save_for_backward(ckpt, adj, phi)
    ctx.args = (plan, N, f, fac, coef, every)
    return st

  @staticmethod
  def backward(ctx, g):
    plan, N, f, fac, coef, every = ctx.args
    nsteps = len(coef) // 3
    g = _f32(g)
    gs = g.clone()
    gcoef = gfin = None
    if ctx.want_c:
      gcoef = torch.zeros(2 * nsteps + 1, dtype=torch.float64,
                          device=gs.device)
      # the final force's own cotangent
      gfin = float((g[2].double() * ctx.f_final.double()).sum())
    gptr = _lib.ptr(gcoef) if gcoef is not None else None
    if every == 0:
      (tape,) = ctx.saved_tensors
      work = torch.empty(_lib.lib().pmx_nbody_adj_work_floats(plan.handle, N,
                                                              0),
                         dtype=torch.float32, device=gs.device)
      kind = "compact_" if ctx.compact else ""
      _lib.call("pmx_nbody_adjoint_%staped" % kind, plan.handle,
                _lib.ptr(tape), nsteps, N, f, fac, _lib.dbl(coef),
                _lib.ptr(gs), _lib.ptr(work), gptr, _lib.stream_of(gs))
    else:
      ckpt, adj, phi = ctx.saved_tensors
      k = min(every, max(nsteps, 1))
      if phi is not None:
        # the last segment's tape is used once; a second backward pass
        # (retain_graph) recomputes every segment from the potentials
        last = nsteps > 0 and not getattr(ctx, "tape_used", False)
        _lib.call("pmx_nbody_adjoint_phi", plan.handle, _lib.ptr(ckpt),
                  _lib.ptr(phi), k, nsteps, N, f, fac, _lib.dbl(coef),
                  _lib.ptr(gs), _lib.ptr(adj), 1 if last else 0, gptr,
                  _lib.stream_of(gs))
      elif nsteps > 0 and not getattr(ctx, "tape_used", False):
        _lib.call("pmx_nbody_adjoint_last_taped", plan.handle,
                  _lib.ptr(ckpt), k, nsteps, N, f, fac, _lib.dbl(coef),
                  _lib.ptr(gs), _lib.ptr(adj), gptr, _lib.stream_of(gs))
      else:
        # a second backward pass (retain_graph): the sweep has since
        # recomputed the earlier segments into the same workspace, so every
        # segment is recomputed
        _lib.call("pmx_nbody_adjoint", plan.handle, _lib.ptr(ckpt), k,
                  nsteps, N, f, fac, _lib.dbl(coef), _lib.ptr(gs),
                  _lib.ptr(adj), gptr, _lib.stream_of(gs))
      ctx.tape_used = True
    gcpar = None
    if gcoef is not None:
      gcpar = _coef_cotangents(coef, fac, gcoef.cpu().numpy(), gfin)
    out = (gs, None, None, None, None, None, None, None, gcpar)
    return out[:len(ctx.needs_input_grad)]


def _pgd_table(pgdparams, nsteps):
  """(alpha, kl, ks) per step as a float32 C array (pgdparams: tfpm.py:451,
  one row per step)."""
  if isinstance(pgdparams, torch.Tensor):
    pgdparams = pgdparams.detach().cpu().numpy()
  arr = np.asarray(pgdparams, dtype=np.float32).reshape(-1, 3)
  if len(arr) != nsteps:
    raise ValueError("pgdparams needs one (alpha, kl, ks) row per step: "
                     "%d rows for %d steps" % (len(arr), nsteps))
  if (arr[:, 1] == 0).any() or (arr[:, 2] == 0).any():
    raise ValueError("PGD scales kl and ks must be nonzero")
  return (_lib._f * arr.size)(*arr.ravel().tolist())


def _run(plan, st, N, f, fac, cptr, pgd, b0, b1, jump, work):
  if pgd is None:
    _lib.call("pmx_nbody_run", plan.handle, _lib.ptr(st), N, f, fac, cptr,
              b0, b1, jump, _lib.ptr(work), _lib.stream_of(st))
  else:
    _lib.call("pmx_nbody_run_pgd", plan.handle, _lib.ptr(st), N, f, fac,
              cptr, pgd, b0, b1, jump, _lib.ptr(work), _lib.stream_of(st))


class _NbodyPgd(torch.autograd.Function):
  """nbody with the PGD correction, differentiable w.r.t. the state and the
  PGD parameters: the forward run records a tape (positions before and
  after each correction, PGD and force meshes) and the reverse sweep runs
  the hand-written adjoint kernels, including d loss / d (alpha, kl, ks)."""

  @staticmethod
  def forward(ctx, state, pgdparams, plan, N, f, fac, coef, cpar=None):
    st = _f32(state).clone()
    nsteps = len(coef) // 3
    pgd = _pgd_table(pgdparams, nsteps)
    tape = torch.empty(_lib.lib().pmx_nbody_pgd_tape_floats(plan.handle, N,
                                                            nsteps),
                       dtype=torch.float32, device=st.device)
    work = _work(plan)
    _lib.call("pmx_nbody_run_pgd_taped", plan.handle, _lib.ptr(st), N, f,
              fac, _lib.dbl(coef), pgd, nsteps, _lib.ptr(tape),
              _lib.ptr(work), _lib.stream_of(st))
    ctx.save_for_backward(tape)
    ctx.pgd = pgd
    ctx.args = (plan, N, f, fac, coef)
    ctx.pgd_is_tensor = isinstance(pgdparams, torch.Tensor)
    ctx.pgd_meta = (pgdparams.dtype, pgdparams.device,
                    tuple(pgdparams.shape)) if ctx.pgd_is_tensor else None
    ctx.want_c = len(ctx.needs_input_grad) > 7 and ctx.needs_input_grad[7]
    if ctx.want_c:
      ctx.f_final = st[2].clone()
    return st

  @staticmethod
  def backward(ctx, g):
    plan, N, f, fac, coef = ctx.args
    nsteps = len(coef) // 3
    g = _f32(g)
    gs = g.clone()
    gcoef = gfin = None
    if ctx.want_c:
      gcoef = torch.zeros(2 * nsteps + 1, dtype=torch.float64,
                          device=gs.device)
      gfin = float((g[2].double() * ctx.f_final.double()).sum())
    work = torch.empty(_lib.lib().pmx_nbody_adj_work_floats(plan.handle, N,
                                                            0),
                       dtype=torch.float32, device=gs.device)
    want = ctx.pgd_is_tensor and ctx.needs_input_grad[1]
    gp = torch.empty(max(3 * nsteps, 1), dtype=torch.float64,
                     device=gs.device) if want else None
    (tape,) = ctx.saved_tensors
    _lib.call("pmx_nbody_adjoint_pgd_taped", plan.handle, _lib.ptr(tape),
              nsteps, N, f, fac, _lib.dbl(coef), ctx.pgd, _lib.ptr(gs),
              _lib.ptr(gp), _lib.ptr(work), _lib.ptr(gcoef),
              _lib.stream_of(gs))
    gpgd = None
    if want:
      dt, dev, shape = ctx.pgd_meta
      gpgd = gp[:3 * nsteps].reshape(shape).to(dtype=dt, device=dev)
    gcpar = None
    if gcoef is not None:
      gcpar = _coef_cotangents(coef, fac, gcoef.cpu().numpy(), gfin)
    out = (gs, gpgd, None, None, None, None, None, gcpar)
    return out[:len(ctx.needs_input_grad)]


def nbody(cosmo, state, stages, nc, pm_nc_factor=1,
          return_intermediate_states=False, pgdparams=None, name="NBody",
          checkpoint_every=None):
  """Integrates the state across ``stages`` with the FastPM KDK scheme
  (tfpm.py:425-503).  Differentiable w.r.t. ``state`` through hand-written
  adjoint kernels run in reverse: with ``checkpoint_every`` = k >= 1 the
  backward pass recomputes each segment of k steps from a stored
  checkpoint; with 0 the forward pass records the positions and force
  meshes of every step and nothing is recomputed; None uses the setting of
  ``set_checkpoint_every`` (automatic by default)."""
  _lib.require_cuda(state)
  if len(stages) == 0:
    return state
  plan, B, N, f, fac = _force_setup(cosmo, state, nc, pm_nc_factor)
  coef = step_coefficients(cosmo, stages)
  nsteps = len(stages) - 1
  cpar = None
  if grad_params(cosmo) and torch.is_grad_enabled():
    cpar = host_fn(lambda c: np.append(step_coefficients(c, stages),
                                       1.5 * c.Omega_m), cosmo)
  pgd = None
  if pgdparams is not None:
    pgd = _pgd_table(pgdparams, nsteps)
  grad = torch.is_grad_enabled() and (
      state.requires_grad or getattr(pgdparams, "requires_grad", False) or
      cpar is not None)
  if return_intermediate_states:
    if grad:
      return _nbody_states_differentiable(state, stages, plan, N, f, fac,
                                          coef, pgdparams, checkpoint_every,
                                          cpar)
    st = _f32(state).clone()
    work = _work(plan)
    cptr = _lib.dbl(coef)
    out = []
    for i in range(nsteps):
      _run(plan, st, N, f, fac, cptr, pgd, i, i + 1, 1 if i == 0 else 0,
           work)
      out.append((stages[i + 1], st.clone()))
    return out
  if pgd is not None:
    if grad:
      return _NbodyPgd.apply(state, pgdparams, plan, N, f, fac, coef, cpar)
    st = _f32(state).clone()
    _run(plan, st, N, f, fac, _lib.dbl(coef), pgd, 0, nsteps, 1, _work(plan))
    return st
  need_grad = torch.is_grad_enabled() and (state.requires_grad or
                                           cpar is not None)
  every = checkpoint_every if checkpoint_every is not None else _CKPT_EVERY[0]
  every = _resolve_every(plan, N, len(coef) // 3, every,
                         state.device) if need_grad else 1
  return _Nbody.apply(state, plan, N, f, fac, coef, every, need_grad, cpar)


def _nbody_states_differentiable(state, stages, plan, N, f, fac, coef,
                                 pgdparams, checkpoint_every, cpar=None):
  """[(a_i, state_i)] with gradients: one differentiable nbody run per step
  (row i of the run's coefficient table).  Every run starts with the
  jump-start force at its input positions, which recomputes the force the
  previous step stored (the same kernels on the same positions), so each
  intermediate state equals the single-run one and the composition is the
  same function of the initial state (its F is discarded either way,
  tfpm.py:470).  cpar: the step coefficients and fac as a tensor that
  requires grad (see _Nbody); each run takes its step's row and fac."""
  out = []
  st = state
  for i in range(len(stages) - 1):
    c = coef[3 * i:3 * i + 3]
    cp = None if cpar is None else torch.cat([cpar[3 * i:3 * i + 3],
                                              cpar[-1:]])
    if pgdparams is not None:
      row = pgdparams[i:i + 1] if isinstance(pgdparams, torch.Tensor) else \
          np.asarray(pgdparams, dtype=np.float32).reshape(-1, 3)[i:i + 1]
      if torch.is_grad_enabled() and (st.requires_grad or cp is not None or
                                      getattr(row, "requires_grad", False)):
        st = _NbodyPgd.apply(st, row, plan, N, f, fac, c, cp)
      else:
        s2 = _f32(st).clone()
        _run(plan, s2, N, f, fac, _lib.dbl(c), _pgd_table(row, 1), 0, 1, 1,
             _work(plan))
        st = s2
    else:
      need = torch.is_grad_enabled() and (st.requires_grad or
                                          cp is not None)
      every = checkpoint_every if checkpoint_every is not None else \
          _CKPT_EVERY[0]
      every = _resolve_every(plan, N, 1, every, st.device) if need else 1
      st = _Nbody.apply(st, plan, N, f, fac, c, every, need, cp)
    out.append((stages[i + 1], st))
  return out
