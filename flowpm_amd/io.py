"""Snapshot export in the BigFile on-disk format (flowpm/io.py:5-60).

The reference hands the state to the third-party ``bigfile`` package
(rainwoodman/bigfile, unpinned in setup.py:19-22), which is not installed
here; this module writes the same directory layout itself, from the format
bigfile publishes:

* a file is a directory and every block ``name`` (``/`` nests) is a
  subdirectory holding
  - ``header``: ``DTYPE: <dtype>``, ``NMEMB: <n>``, ``NFILE: <k>`` and one
    line per data file, ``%06X: <items> : <byte sum> : <SysV checksum>``;
  - ``attr-v2``: one attribute per line, ``name dtype nmemb HEX``, HEX being
    the raw little-endian bytes of the values (two upper-case hex digits per
    byte), followed by a ``#HUMANE [ ... ]`` rendering that readers ignore;
  - data files ``000000``, ``000001``, ... (hex index): the rows in C order,
    split evenly over the files (file i holds rows
    [size*i/k, size*(i+1)/k)), at most 32 Mi rows per file;
* the byte sum is the 32-bit wrapping sum of the file's bytes, the SysV
  checksum its fold ``r = (s & 0xffff) + (s >> 16); (r & 0xffff) + (r >> 16)``
  (what ``sum -s`` prints).

``save_state`` writes, like the reference, a ``Header`` block whose
attributes are NC, BoxSize, the float32 cosmological parameters, Time and
M0 (plus the caller's ``attrs``, best effort), then per batch member ``i``
the blocks ``i/Position`` and ``i/Velocity``: X and P times
float32(1/nc) * float32(boxsize), in float32 (io.py:46-60).  The state may
live on the GPU; it is copied to the host one batch member at a time.
"""
import os

import numpy as np
import torch

__all__ = ["save_state", "write_block", "write_attrs", "read_attrs",
           "read_block"]

# rows per data file at most (the format's default split)
ROWS_PER_FILE = 32 * 1024 * 1024


def _sysv(byte_sum):
  s = byte_sum & 0xffffffff
  r = (s & 0xffff) + (s >> 16)
  return (r & 0xffff) + (r >> 16)


def _attr_array(value):
  """The array an attribute value is stored as (numpy's own conversion)."""
  if isinstance(value, str):
    value = value.encode()
  if isinstance(value, bytes):
    return np.frombuffer(value, dtype="S1").copy()
  if isinstance(value, torch.Tensor):
    value = value.detach().cpu().numpy()
  arr = np.asarray(value)
  if arr.dtype.hasobject or arr.dtype.kind == "U":
    raise ValueError("attribute value of type %s is not supported" %
                     arr.dtype)
  return np.ascontiguousarray(arr.reshape(-1)).astype(
      arr.dtype.newbyteorder("<"), copy=False)


def _humane(arr):
  if arr.dtype.kind == "S":
    return repr(b"".join(arr.tolist()).decode("latin-1"))
  return " ".join(repr(v) for v in arr.tolist())


def write_attrs(blockdir, attrs):
  """Write ``attrs`` (name -> value) as the block's attr-v2 file."""
  lines = []
  for name in sorted(attrs):
    if not name or any(c.isspace() for c in name):
      raise ValueError("attribute name %r must be non-empty without spaces" %
                       name)
    arr = attrs[name]
    lines.append("%s %s %d %s #HUMANE [ %s ]\n" %
                 (name, arr.dtype.str, arr.size, arr.tobytes().hex().upper(),
                  _humane(arr)))
  with open(os.path.join(blockdir, "attr-v2"), "w") as f:
    f.writelines(lines)


def _write_header(blockdir, dtype, nmemb, files):
  with open(os.path.join(blockdir, "header"), "w") as f:
    f.write("DTYPE: %s\nNMEMB: %d\nNFILE: %d\n" % (dtype, nmemb, len(files)))
    for i, (rows, bsum) in enumerate(files):
      f.write("%06X: %d : %d : %d\n" % (i, rows, bsum & 0xffffffff,
                                        _sysv(bsum)))


def write_block(filename, block, rows, chunks, attrs=None):
  """Write one data block of ``rows`` float32 rows.

  ``chunks`` yields C-ordered numpy arrays whose first axes together hold
  the ``rows`` rows, in order; every row has the same shape (its element
  count is NMEMB).  The rows are split over ceil(rows / ROWS_PER_FILE) data
  files.
  """
  blockdir = os.path.join(filename, block)
  os.makedirs(blockdir, exist_ok=True)
  nfile = max(1, -(-rows // ROWS_PER_FILE))
  bounds = [rows * i // nfile for i in range(nfile + 1)]
  files, dtype, nmemb = [], None, None
  fid, fh, left, bsum = -1, None, 0, 0
  done = 0
  for chunk in chunks:
    chunk = np.ascontiguousarray(chunk)
    if dtype is None:
      dtype = chunk.dtype.newbyteorder("<").str
      nmemb = int(np.prod(chunk.shape[1:], dtype=np.int64))
    flat = chunk.reshape(chunk.shape[0], -1).astype(dtype, copy=False)
    pos = 0
    while pos < flat.shape[0]:
      if left == 0:
        if fh is not None:
          fh.close()
          files.append((bounds[fid + 1] - bounds[fid], bsum))
        fid += 1
        fh = open(os.path.join(blockdir, "%06X" % fid), "wb")
        left, bsum = bounds[fid + 1] - bounds[fid], 0
        continue
      take = min(left, flat.shape[0] - pos)
      part = flat[pos:pos + take]
      raw = part.tobytes()
      bsum += int(np.frombuffer(raw, np.uint8).sum(dtype=np.uint64))
      fh.write(raw)
      pos += take
      left -= take
      done += take
  if done != rows:
    if fh is not None:
      fh.close()
    raise ValueError("block %s: %d rows written, %d declared" %
                     (block, done, rows))
  if fh is not None:
    fh.close()
    files.append((bounds[fid + 1] - bounds[fid], bsum))
  while len(files) < nfile:  # zero-row files (rows < nfile cannot happen)
    open(os.path.join(blockdir, "%06X" % len(files)), "wb").close()
    files.append((0, 0))
  _write_header(blockdir, dtype or "<f4", nmemb or 1, files)
  write_attrs(blockdir, {k: _attr_array(v) for k, v in (attrs or {}).items()})


def _write_attr_block(filename, block, attrs):
  """An attribute-only block: no data files."""
  blockdir = os.path.join(filename, block)
  os.makedirs(blockdir, exist_ok=True)
  _write_header(blockdir, "<i8", 1, [])
  write_attrs(blockdir, attrs)


def _per_axis(v, name):
  arr = np.asarray(v)
  if arr.ndim == 0:
    arr = np.full(3, arr)
  if arr.shape != (3,):
    raise ValueError("%s must have three entries, got %r" % (name, v))
  return arr


def save_state(cosmo, state, a, nc, boxsize, filename, attrs={}):
  """Save a simulation snapshot (flowpm/io.py:5-60).

  state: (3, B, N, 3) float32 tensor [X; P; F] in mesh units (host or GPU).
  a: scale factor of the snapshot, stored as the ``Time`` attribute.
  nc, boxsize: mesh cells and box length per axis (lists of three; a scalar
  is taken for all three axes).
  filename: the snapshot directory, created if needed.
  attrs: extra Header attributes; values that cannot be stored are skipped,
  as in the reference (io.py:44-49).

  Positions and velocities are written per batch member ``i`` to the blocks
  ``i/Position`` and ``i/Velocity`` in Mpc/h units: the state times
  float32(1/nc) * float32(boxsize), rounded to float32 (io.py:51-60).
  """
  if not isinstance(state, torch.Tensor):
    state = torch.as_tensor(np.asarray(state))
  if state.dim() != 4 or state.shape[0] < 2 or state.shape[3] != 3:
    raise ValueError("state must be (3, B, N, 3), got %s" %
                     (tuple(state.shape),))
  ncv = _per_axis(nc, "nc")
  boxv = _per_axis(boxsize, "boxsize")
  os.makedirs(filename, exist_ok=True)
  header = {
      "NC": _attr_array(nc),
      "BoxSize": _attr_array(boxsize),
      "OmegaCDM": _attr_array(np.float32(cosmo.Omega_c)),
      "OmegaB": _attr_array(np.float32(cosmo.Omega_b)),
      "OmegaK": _attr_array(np.float32(cosmo.Omega_k)),
      "h": _attr_array(np.float32(cosmo.h)),
      "Sigma8": _attr_array(np.float32(cosmo.sigma8)),
      "w0": _attr_array(np.float32(cosmo.w0)),
      "wa": _attr_array(np.float32(cosmo.wa)),
      "Time": _attr_array(a),
      "M0": _attr_array([1.]),
  }
  for key in attrs:
    try:
      header[key] = _attr_array(attrs[key])
    except Exception:  # best effort, as the reference
      pass
  _write_attr_block(filename, "Header", header)

  scale = (np.float32(1.) / ncv.astype(np.float32)) * boxv.astype(np.float32)
  scale = scale.astype(np.float32).reshape(1, 3)
  n = int(state.shape[2])
  step = max(1, ROWS_PER_FILE // 4)

  def rows_of(which, b):
    for lo in range(0, n, step):
      part = state[which, b, lo:lo + step].detach()
      if part.dtype != torch.float32:
        part = part.float()
      yield part.cpu().numpy() * scale

  for b in range(int(state.shape[1])):
    write_block(filename, "%d/Position" % b, n, rows_of(0, b))
    write_block(filename, "%d/Velocity" % b, n, rows_of(1, b))


def read_attrs(filename, block):
  """The attributes of a block, as numpy arrays (attr-v2 file)."""
  out = {}
  path = os.path.join(filename, block, "attr-v2")
  if not os.path.exists(path):
    return out
  with open(path) as f:
    for line in f:
      line = line.split("#HUMANE", 1)[0].split()
      if not line:
        continue
      name, dtype, nmemb, data = line[0], line[1], int(line[2]), (
          line[3] if len(line) > 3 else "")
      arr = np.frombuffer(bytes.fromhex(data), dtype=np.dtype(dtype))
      if arr.size != nmemb:
        raise ValueError("attribute %s: %d values, %d declared" %
                         (name, arr.size, nmemb))
      out[name] = arr
  return out


def read_block(filename, block):
  """All rows of a data block as a (rows, NMEMB) array (NMEMB 1: (rows,))."""
  blockdir = os.path.join(filename, block)
  with open(os.path.join(blockdir, "header")) as f:
    lines = f.read().splitlines()
  meta = dict(l.split(":", 1) for l in lines[:3])
  dtype = np.dtype(meta["DTYPE"].strip())
  nmemb = int(meta["NMEMB"])
  nfile = int(meta["NFILE"])
  parts = []
  for line in lines[3:3 + nfile]:
    fields = [x.strip() for x in line.split(":")]
    fid, rows, bsum, sysv = int(fields[0], 16), int(fields[1]), int(
        fields[2]), int(fields[3])
    raw = open(os.path.join(blockdir, "%06X" % fid), "rb").read()
    if len(raw) != rows * nmemb * dtype.itemsize:
      raise ValueError("%s/%06X: %d bytes, header says %d rows" %
                       (block, fid, len(raw), rows))
    s = int(np.frombuffer(raw, np.uint8).sum(dtype=np.uint64)) & 0xffffffff
    if s != bsum or _sysv(s) != sysv:
      raise ValueError("%s/%06X: checksum mismatch" % (block, fid))
    parts.append(np.frombuffer(raw, dtype=dtype))
  data = np.concatenate(parts) if parts else np.zeros(0, dtype)
  return data.reshape(-1, nmemb) if nmemb > 1 else data
