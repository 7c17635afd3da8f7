"""Relative RMS of the GPU run against the float64 oracle run.

usage: compare.py n fp64_dir gpu_dir
"""
import os
import sys

import numpy as np


def main():
  n, d64, dg = int(sys.argv[1]), sys.argv[2], sys.argv[3]
  idx = np.arange(0, n**3, 16)
  q = np.stack(np.unravel_index(idx, (n, n, n)), axis=1).astype(np.float64)
  load = lambda d, k: np.load(os.path.join(d, "%s_%d.npy" % (k, n))).astype(
      np.float64)
  rr = lambda a, b: np.sqrt(np.mean((a - b)**2) / np.mean(b**2))
  X64, P64, R64 = load(d64, "X"), load(d64, "P"), load(d64, "rho")
  Xg, Pg, Rg = load(dg, "X"), load(dg, "P"), load(dg, "rho")
  print("n=%d: displacement rel RMS %.3g, momentum rel RMS %.3g, density "
        "contrast rel RMS %.3g; max |dX| %.3g cells" %
        (n, rr(Xg - q, X64 - q), rr(Pg, P64), rr(Rg - 1, R64 - 1),
         np.abs(Xg - X64).max()))


if __name__ == "__main__":
  main()
