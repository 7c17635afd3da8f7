"""Snapshot export in the BigFile format (flowpm/io.py:5-60).

Mirrors tests/test_io.py:16-62 of the reference (save a B = 2, 16^3 state,
reload the Header attributes and the Position / Velocity of member 1), with
the file read back by a parser written here from the published layout
instead of the absent ``bigfile`` package.  What pins the bytes:
  * the data files hold exactly the float32 rows state * f32(1/nc) *
    f32(boxsize), bit for bit;
  * attribute hex is the IEEE-754 little-endian encoding of known values
    (0.1 as float64 is 0x3FB999999999999A);
  * the header's checksum column equals what coreutils ``sum -s`` (the SysV
    checksum) prints for the data file, and its byte-sum column the plain
    32-bit sum of the file's bytes;
  * the split over data files follows rows*i/k boundaries with hex names.
The exact ``DTYPE``/``NMEMB`` the bigfile writer records for the
attribute-only Header block, the attribute order and the ``#HUMANE`` text
are parity unpinned (no bigfile output exists here to compare with).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest
import torch
from numpy.testing import assert_allclose, assert_array_equal

import flowpm_amd
from flowpm_amd import io as pio
from flowpm_amd.cosmology import Planck15


def _parse_header(path):
  with open(os.path.join(path, "header")) as f:
    lines = f.read().splitlines()
  assert lines[0].startswith("DTYPE: ")
  assert lines[1].startswith("NMEMB: ")
  assert lines[2].startswith("NFILE: ")
  nfile = int(lines[2][7:])
  files = []
  for line in lines[3:]:
    name, rows, bsum, sysv = [x.strip() for x in line.split(":")]
    files.append((name, int(rows), int(bsum), int(sysv)))
  assert len(files) == nfile
  return lines[0][7:], int(lines[1][7:]), files


def _parse_attrs(path):
  out = {}
  with open(os.path.join(path, "attr-v2")) as f:
    for line in f:
      head, humane = line.rstrip("\n").split(" #HUMANE [ ")
      assert humane.endswith(" ]")
      name, dtype, nmemb, hexdata = head.split(" ")
      out[name] = (dtype, int(nmemb), hexdata)
  return out


def _state(B=2, n=16, seed=0):
  g = torch.Generator().manual_seed(seed)
  return torch.rand((3, B, n**3, 3), generator=g) * n * 1.3 - 0.1 * n


def test_save_state_roundtrip(tmp_path):
  """tests/test_io.py:16-62 with the file read back here."""
  nc = [16, 16, 16]
  boxsize = [100., 100., 100.]
  a0 = 0.1
  cosmo = Planck15()
  state = _state()
  fn = str(tmp_path / "testsave")
  pio.save_state(cosmo, state, a0, nc, boxsize, fn)

  attrs = pio.read_attrs(fn, "Header")
  assert_allclose(attrs["NC"], nc)
  assert_allclose(attrs["BoxSize"], boxsize)
  assert_allclose(attrs["OmegaCDM"], np.float32(cosmo.Omega_c))
  assert_allclose(attrs["OmegaB"], np.float32(cosmo.Omega_b))
  assert_allclose(attrs["OmegaK"], np.float32(cosmo.Omega_k))
  assert_allclose(attrs["h"], np.float32(cosmo.h))
  assert_allclose(attrs["Sigma8"], np.float32(cosmo.sigma8))
  assert_allclose(attrs["w0"], np.float32(cosmo.w0))
  assert_allclose(attrs["wa"], np.float32(cosmo.wa))
  assert_allclose(attrs["Time"], a0)
  assert_allclose(attrs["M0"], [1.])

  pos = pio.read_block(fn, "1/Position")
  vel = pio.read_block(fn, "1/Velocity")
  assert_allclose(pos, state[0, 1].numpy() / nc[0] * boxsize[0], rtol=1e-6)
  assert_allclose(vel, state[1, 1].numpy() / nc[0] * boxsize[0], rtol=1e-6)
  # bit-exact: float32(state * float32(float32(1/nc) * float32(boxsize)))
  scale = np.float32(np.float32(1.) / np.float32(16.) * np.float32(100.))
  for b in range(2):
    for which, blk in ((0, "Position"), (1, "Velocity")):
      got = pio.read_block(fn, "%d/%s" % (b, blk))
      want = state[which, b].numpy() * scale
      assert got.dtype == np.float32
      assert_array_equal(got, want)
  # no force block is written
  assert sorted(os.listdir(os.path.join(fn, "0"))) == ["Position", "Velocity"]


def test_attribute_encoding(tmp_path):
  """Attribute hex is the raw little-endian bytes of the stored values."""
  fn = str(tmp_path / "snap")
  pio.save_state(Planck15(), _state(B=1, n=4), 0.1, [16, 8, 4],
                 [100., 50., 25.], fn,
                 attrs={"Label": "run1", "Seed": 7, "Bad": object()})
  at = _parse_attrs(os.path.join(fn, "Header"))
  assert at["Time"] == ("<f8", 1, "9A9999999999B93F")
  assert at["NC"] == ("<i8", 3, "1000000000000000"
                      "0800000000000000"
                      "0400000000000000")
  assert at["BoxSize"][:2] == ("<f8", 3)
  assert at["BoxSize"][2] == np.array([100., 50., 25.], "<f8").tobytes().hex(
  ).upper()
  # cosmological parameters as float32, like the reference's tensors
  assert at["h"] == ("<f4", 1, np.float32(0.6774).tobytes().hex().upper())
  assert at["M0"] == ("<f8", 1, "000000000000F03F")
  assert at["Label"] == ("|S1", 4, b"run1".hex().upper())
  assert at["Seed"] == ("<i8", 1, "0700000000000000")
  assert "Bad" not in at  # skipped, as the reference's best effort
  assert list(at) == sorted(at)


def test_header_and_checksums(tmp_path):
  fn = str(tmp_path / "snap")
  state = _state(B=1, n=8, seed=3)
  pio.save_state(Planck15(), state, 0.5, [8, 8, 8], [64., 64., 64.], fn)
  blk = os.path.join(fn, "0", "Position")
  dtype, nmemb, files = _parse_header(blk)
  assert (dtype, nmemb) == ("<f4", 3)
  assert [f[0] for f in files] == ["000000"]
  name, rows, bsum, sysv = files[0]
  assert rows == 8**3
  raw = open(os.path.join(blk, name), "rb").read()
  assert len(raw) == rows * 3 * 4
  assert bsum == sum(raw) & 0xffffffff
  if shutil.which("sum"):
    out = subprocess.run(["sum", "-s", os.path.join(blk, name)],
                         capture_output=True, text=True, check=True).stdout
    assert int(out.split()[0]) == sysv
  # the attribute-only Header block has no data files
  hdr = os.path.join(fn, "Header")
  assert _parse_header(hdr)[2] == []
  assert sorted(os.listdir(hdr)) == ["attr-v2", "header"]


def test_sysv_fold_against_coreutils(tmp_path):
  """The checksum fold on inputs whose byte sum passes 2^16 and 2^32."""
  if not shutil.which("sum"):
    pytest.skip("coreutils sum not available")
  rng = np.random.default_rng(5)
  for n in (0, 1, 300, 70000, 5_000_000):
    p = tmp_path / ("f%d" % n)
    data = rng.integers(200, 256, size=n, dtype=np.uint8).tobytes()
    p.write_bytes(data)
    out = subprocess.run(["sum", "-s", str(p)], capture_output=True,
                         text=True, check=True).stdout
    assert pio._sysv(sum(data)) == int(out.split()[0])


def test_multi_file_split(tmp_path, monkeypatch):
  """Rows split as rows*i/k over hex-named files; header rows add up."""
  monkeypatch.setattr(pio, "ROWS_PER_FILE", 100)
  rows = 1037
  data = np.arange(rows * 3, dtype=np.float32).reshape(rows, 3)
  fn = str(tmp_path / "split")
  chunks = (data[i:i + 77] for i in range(0, rows, 77))
  pio.write_block(fn, "blk", rows, chunks, attrs={"x": 1.5})
  dtype, nmemb, files = _parse_header(os.path.join(fn, "blk"))
  k = 11
  assert [f[0] for f in files] == ["%06X" % i for i in range(k)]
  assert files[10][0] == "00000A"
  assert [f[1] for f in files] == [rows * (i + 1) // k - rows * i // k
                                   for i in range(k)]
  assert_array_equal(pio.read_block(fn, "blk"), data)
  assert pio.read_attrs(fn, "blk")["x"].tolist() == [1.5]
  with pytest.raises(ValueError):
    pio.write_block(fn, "short", rows, iter([data[:10]]))


def test_nc_scalar_and_bad_state(tmp_path):
  fn = str(tmp_path / "s")
  st = _state(B=1, n=4)
  pio.save_state(Planck15(), st, 1.0, 4, 10., fn)
  assert_array_equal(pio.read_block(fn, "0/Position"),
                     st[0, 0].numpy() * np.float32(np.float32(0.25) *
                                                   np.float32(10.)))
  with pytest.raises(ValueError):
    pio.save_state(Planck15(), st[0], 1.0, 4, 10., fn)
  assert flowpm_amd.io is pio
