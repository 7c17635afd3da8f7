"""Float64 oracle run of the PM workload (n^3 particles and mesh, box L,
2LPT at a = 0.1, K steps to a = 1) on the seeded ICs that bench.py uses;
keeps every 16th particle and every 8th x-plane of the painted density.

usage: oracle_run.py n L K outdir
"""
import os
import sys

import numpy as np
import torch

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(
    os.path.abspath(__file__)))))
from oracle import background as obg  # noqa: E402
from oracle import pm as opm  # noqa: E402
from oracle import power as opw  # noqa: E402


def main():
  n, L, K, out = int(sys.argv[1]), float(sys.argv[2]), int(sys.argv[3]), \
      sys.argv[4]
  os.makedirs(out, exist_ok=True)
  gen = torch.Generator().manual_seed(100)
  white = torch.randn(1, n, n, n, generator=gen, dtype=torch.float64) * n**1.5
  k = np.logspace(-4, 1, 256)
  pkt = opw.linear_matter_power(obg.planck15(), k)
  pk = lambda x: opw.interp_tf(x.reshape(-1), k, pkt).reshape(x.shape)
  g = obg.GrowthTable(obg.planck15())
  lin = opm.whitenoise_to_linear(white, L, pk)
  st = opm.nbody(g, opm.lpt_init(g, lin, 0.1, order=2),
                 np.linspace(0.1, 1.0, K + 1), [n] * 3)
  idx = torch.arange(0, n**3, 16)
  np.save(os.path.join(out, "X_%d.npy" % n), st[0, 0, idx].numpy())
  np.save(os.path.join(out, "P_%d.npy" % n), st[1, 0, idx].numpy())
  rho = opm.cic_paint(torch.zeros(1, n, n, n, dtype=torch.float64), st[0])
  np.save(os.path.join(out, "rho_%d.npy" % n), rho[0, ::8].numpy())


if __name__ == "__main__":
  main()
