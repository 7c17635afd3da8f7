"""CPU check of the brick paint's index and wrap logic.

tests/cpu/brick_paint_emu.cpp emulates the deposit / gather / spill phases
of flowpm_amd/csrc/paint.hip on the CPU with the kernels' own index code
(flowpm_amd/csrc/brick_index.h) and compares the mesh with a direct
periodic CIC (flowpm/utils.py:32-91): displaced grids with wrapping boxes,
pm_nc_factor 2 and 4, partial bricks, bulk shifts in both directions,
boxes deposited in x slabs through an accumulation window (as the kernel
does for boxes larger than its LDS window), bricks over the arena slot and
far jumps that spill.  It is built with
AddressSanitizer and UndefinedBehaviorSanitizer (host code only), so an
out-of-range box index aborts the run.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_brick_paint_index_logic_matches_direct_cic(tmp_path):
  exe = str(tmp_path / "brick_emu")
  subprocess.run(["g++", "-std=c++17", "-O1", "-g",
                  "-fsanitize=address,undefined",
                  "-fno-sanitize-recover=all",
                  os.path.join(HERE, "cpu", "brick_paint_emu.cpp"), "-o", exe],
                 check=True)
  r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
  print(r.stdout)
  assert r.returncode == 0, r.stdout + r.stderr
  assert r.stdout.strip().endswith("ok")
  # the cases reach every path: boxes that wrap, bricks that spill and boxes
  # deposited in several x slabs of the accumulation window
  import re
  counts = [tuple(int(v) for v in m) for m in re.findall(
      r"bricks (\d+) \(wrapping (\d+), spilled (\d+), in slabs (\d+)\)",
      r.stdout)]
  assert counts
  assert any(w > 0 for _, w, _, _ in counts)
  assert any(s > 0 for _, _, s, _ in counts)
  assert any(k > 0 for _, _, _, k in counts)  # boxes of several slabs
  assert any(n - s - k > 0 for n, _, s, k in counts)  # boxes of one slab
