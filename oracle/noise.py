"""Oracle: the seeded white-noise stream (TEST INFRASTRUCTURE ONLY).

Nothing in ``flowpm_amd`` may import this module; only ``tests/`` use it.
The reference draws N(0, sqrt(M)) with TF's RNG (flowpm/utils.py:337-355),
which no other code can reproduce; the build replaces it by a counter-based
stream defined here with numpy's own Philox4x64-10 bit generator
(``numpy.random.Philox``, a library routine, so the integer stream is
pinned by numpy itself): normals 4q..4q+3 come from the four words of
block q by two Box-Muller pairs, u1 = ((w0 >> 11) + 1) 2^-53,
u2 = (w1 >> 11) 2^-53, r = sqrt(-2 ln u1), (r cos 2 pi u2, r sin 2 pi u2).

Pins (tests/test_oracle_noise.py): the raw words equal an independent
pure-Python Philox4x64-10 (Salmon et al. 2011) at several keys; every
pair satisfies z0^2 + z1^2 = -2 ln u1; a million draws have mean 0 and
variance 1 within their sampling error.
"""
import numpy as np


def standard_normals(n, key, counter=0):
  """The first n normals of the stream (float64)."""
  nb = (n + 3) // 4
  g = np.random.Philox(key=key, counter=counter)
  w = g.random_raw(4 * nb).reshape(nb, 4)
  u1 = ((w[:, 0::2] >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * \
      (1.0 / 9007199254740992.0)
  u2 = (w[:, 1::2] >> np.uint64(11)).astype(np.float64) * \
      (1.0 / 9007199254740992.0)
  r = np.sqrt(-2.0 * np.log(u1))
  two_pi = 6.283185307179586
  z = np.empty((nb, 4))
  z[:, 0::2] = r * np.cos(two_pi * u2)
  z[:, 1::2] = r * np.sin(two_pi * u2)
  return z.reshape(-1)[:n]


def white_noise(shape, key, scale):
  """float32(scale * z) in C order over ``shape``."""
  n = int(np.prod(shape))
  return (standard_normals(n, key) * scale).astype(np.float32).reshape(shape)
