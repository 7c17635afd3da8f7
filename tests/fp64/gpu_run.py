"""The same workload as oracle_run.py on the GPU (flowpm_amd, float32);
keeps the same particle subset and density planes.

usage: gpu_run.py n L K outdir
"""
import os
import sys

import numpy as np
import torch

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(
    os.path.abspath(__file__)))))
import flowpm_amd as fa  # noqa: E402


def main():
  n, L, K, out = int(sys.argv[1]), float(sys.argv[2]), int(sys.argv[3]), \
      sys.argv[4]
  os.makedirs(out, exist_ok=True)
  dev = torch.device("cuda", 0)
  c = fa.cosmology.Planck15()
  kk = np.logspace(-4, 1, 256)
  pkt = fa.tfpower.linear_matter_power(c, kk)
  pk = lambda k: fa.scipy.interpolate.interp_tf(k, kk, pkt)
  gen = torch.Generator().manual_seed(100)
  white = (torch.randn(1, n, n, n, generator=gen, dtype=torch.float64) *
           n**1.5).float().to(dev)
  lin = fa.whitenoise_to_linear(n, L, white, pk)
  st = fa.nbody(c, fa.lpt_init(c, lin, 0.1, order=2),
                np.linspace(0.1, 1.0, K + 1), n)
  idx = torch.arange(0, n**3, 16, device=dev)
  np.save(os.path.join(out, "X_%d.npy" % n), st[0, 0, idx].cpu().numpy())
  np.save(os.path.join(out, "P_%d.npy" % n), st[1, 0, idx].cpu().numpy())
  rho = fa.cic_paint(torch.zeros(1, n, n, n, device=dev), st[0])
  np.save(os.path.join(out, "rho_%d.npy" % n), rho[0, ::8].cpu().numpy())
  print("wrote", out, flush=True)


if __name__ == "__main__":
  main()
