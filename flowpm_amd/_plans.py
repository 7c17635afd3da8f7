"""Cache of pmx plans, one per (device, batch, mesh shape)."""
import ctypes
import threading

import torch

from . import _lib

_PLANS = {}
_LOCK = threading.Lock()

# 2^a, 3 * 2^a, 5 * 2^a, 15 * 2^a and 25 * 2^a from 8 to 1024, multiples of
# 4 (the radix 16/8/4/2/3/5 line transforms; the z stencil needs nz % 4 == 0)
SUPPORTED_SIDES = (8, 12, 16, 20, 24, 32, 40, 48, 60, 64, 80, 96, 100, 120,
                   128, 160, 192, 200, 240, 256, 320, 384, 400, 480, 512, 640,
                   768, 800, 960, 1024)

# tile widths chosen by flowpm_amd.autotune, keyed like the plans
TILE_CHOICES = {}


class Plan:

  def __init__(self, device, B, nx, ny, nz):
    for n in (nx, ny, nz):
      if n not in SUPPORTED_SIDES:
        raise ValueError("mesh side %d not supported; sides must be one of "
                         "%s" % (n, SUPPORTED_SIDES))
    if nz % 2:
      raise ValueError("nz must be even")
    self.device = device
    self.B, self.nx, self.ny, self.nz = B, nx, ny, nz
    h = ctypes.c_void_p()
    with torch.cuda.device(device):
      _lib.call("pmx_plan_create", ctypes.byref(h), B, nx, ny, nz)
    self.handle = h
    self.S = _lib.lib().pmx_plan_spec_elems(h)  # float2 per spectral buffer
    self.M = nx * ny * nz

  def spec(self, n=1):
    """n spectral buffers as one float32 tensor (2*S floats each)."""
    return torch.empty(n * 2 * self.S, dtype=torch.float32,
                       device=self.device)


def get(device, B, nx, ny, nz):
  if isinstance(device, torch.device):
    idx = device.index if device.index is not None else \
        torch.cuda.current_device()
  else:
    idx = int(device)
  key = (idx, int(B), int(nx), int(ny), int(nz))
  with _LOCK:
    p = _PLANS.get(key)
    if p is None:
      p = Plan(torch.device("cuda", idx), *key[1:])
      tiles = TILE_CHOICES.get(key[1:])
      if tiles is not None:
        _lib.call("pmx_plan_set_tiles", p.handle, *tiles)
      _PLANS[key] = p
  return p
