"""The C ABI's host code under AddressSanitizer and UBSan (no GPU needed).

tests/cpu/abi_host_check.cpp calls every entry point of include/pmx.h with
arguments its validation must refuse (null pointers and plans, sizes below
1, odd or unsupported sides) and checks the negative status, the per-thread
message, a clean failure of plan creation without a device, and the brick
paint's workspace size at extreme sizes.  The library's sources are built
for it with the sanitizers on the host code only (-Xarch_host), one FFT line
length (PMX_FFT_ONLY_16) to keep the build short; leak detection is off,
since the HIP runtime keeps allocations for the life of the process.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "flowpm_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"
SRCS = ["fft", "cic", "paint", "spectrum", "noise", "stencil", "pm"]
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host",
       "-fsanitize=undefined", "-Xarch_host", "-fno-sanitize-recover=all"]


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="needs hipcc")
def test_abi_host_code_under_asan_ubsan(tmp_path):
  flags = ["-O1", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
           "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-DPMX_FFT_ONLY_16"] + SAN
  procs, objs = [], []
  for s in SRCS:
    o = str(tmp_path / (s + ".o"))
    objs.append(o)
    procs.append(subprocess.Popen(
        [HIPCC] + flags + ["-c", os.path.join(CSRC, s + ".hip"), "-o", o],
        stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
  drv = str(tmp_path / "drv.o")
  procs.append(subprocess.Popen(
      [HIPCC, "-O1", "-std=c++17", "-x", "c++"] + SAN +
      ["-c", os.path.join(HERE, "cpu", "abi_host_check.cpp"), "-o", drv],
      stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
  for p in procs:
    out, _ = p.communicate(timeout=600)
    assert p.returncode == 0, out
  exe = str(tmp_path / "abi_check")
  subprocess.run([HIPCC, "-Xarch_host", "-fsanitize=address", "-Xarch_host",
                  "-fsanitize=undefined", drv] + objs + ["-o", exe],
                 check=True, capture_output=True)
  env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0",
             UBSAN_OPTIONS="print_stacktrace=1")
  r = subprocess.run([exe], capture_output=True, text=True, timeout=300,
                     env=env)
  print(r.stdout[-3000:])
  assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
  assert r.stdout.strip().endswith("ok")
