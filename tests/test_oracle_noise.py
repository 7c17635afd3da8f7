"""Pins of oracle/noise.py."""
import numpy as np

from oracle import noise

_MASK = (1 << 64) - 1


def _philox4x64_10(ctr, key):
  """Philox4x64-10 written out from its definition (Salmon et al. 2011)."""
  M0, M1 = 0xD2E7470EE14C6C93, 0xCA5A826395121157
  W0, W1 = 0x9E3779B97F4A7C15, 0xBB67AE8584CAA73B
  c, k = list(ctr), list(key)
  for r in range(10):
    if r > 0:
      k = [(k[0] + W0) & _MASK, (k[1] + W1) & _MASK]
    p0, p1 = M0 * c[0], M1 * c[2]
    c = [(p1 >> 64) ^ c[1] ^ k[0], p1 & _MASK, (p0 >> 64) ^ c[3] ^ k[1],
         p0 & _MASK]
  return c


def test_raw_stream_is_philox4x64_10():
  for key in (0, 1, 12345, (1 << 64) + 7, (1 << 127) + 3):
    raw = np.random.Philox(key=key).random_raw(12)
    mine = []
    for q in range(3):
      mine += _philox4x64_10([q + 1, 0, 0, 0], [key & _MASK, key >> 64])
    assert [int(x) for x in raw] == mine


def test_box_muller_pairs():
  z = noise.standard_normals(4000, 99)
  w = np.random.Philox(key=99).random_raw(4000).reshape(-1, 4)
  u1 = ((w[:, 0::2] >> np.uint64(11)) + np.uint64(1)).astype(float) / 2.0**53
  zz = z.reshape(-1, 4)
  r2 = zz[:, 0::2]**2 + zz[:, 1::2]**2
  np.testing.assert_allclose(r2, -2 * np.log(u1), rtol=1e-12, atol=1e-14)


def test_moments():
  z = noise.standard_normals(1 << 20, 2024)
  assert abs(z.mean()) < 5 / 1024
  assert abs(z.var() - 1) < 5 * np.sqrt(2) / 1024
