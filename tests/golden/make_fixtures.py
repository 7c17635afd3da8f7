"""Writes the float64-oracle fixtures of the full-size workloads.

Runs the CPU oracle (float64) once, in the build container, and keeps a
small subsample of its output so that the GPU tests can compare the HIP path
with it without running the oracle:

* ``c4_256.npz``: 256^3 particles on a 256^3 mesh, box 512 Mpc/h, Planck15
  Eisenstein-Hu P(k), white noise seed 100, 2LPT at a = 0.1, 40 KDK steps
  to a = 1 (linspace(0.1, 1, 41)).  About 15 minutes and 45 GB of host
  memory.
* ``c3_<n>.npz`` for n = 32, 64, 96, 128: the first member of each size
  class of the uneven batch (box 100 n / 32 Mpc/h, seed 1000 + 8 j for class
  j), 1LPT at a = 0.1, 10 steps (linspace(0.1, 1, 11)).

Each file holds ``idx`` (the particle indices kept: ``n^3 / 4096`` stride,
Lagrangian C order), ``X`` and ``P`` of those particles (float64), the
painted density of the final positions on the x planes ``planes`` (float32),
and the workload parameters.  White noise is ``torch.randn`` of a
``torch.Generator().manual_seed(seed)`` in float64 times n^1.5 (the
scripts/check_grads.py:39-65 injection), which the GPU tests regenerate.

usage: python tests/golden/make_fixtures.py [c4] [c3]
"""
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(
    os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import background as obg  # noqa: E402
from oracle import pm as opm  # noqa: E402
from oracle import power as opw  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
KEEP = 4096  # particles kept per fixture


def white_noise(n, seed):
  gen = torch.Generator().manual_seed(seed)
  return torch.randn(1, n, n, n, generator=gen, dtype=torch.float64) * n**1.5


def planck15_pk():
  k = np.logspace(-4, 1, 256)
  pkt = opw.linear_matter_power(obg.planck15(), k)
  return lambda x: opw.interp_tf(x.reshape(-1), k, pkt).reshape(x.shape)


def run(name, n, box, seed, order, nsteps, planes):
  t0 = time.time()
  g = obg.GrowthTable(obg.planck15())
  with torch.no_grad():
    lin = opm.whitenoise_to_linear(white_noise(n, seed), box, planck15_pk())
    st = opm.lpt_init(g, lin, 0.1, order=order)
    del lin
    st = opm.nbody(g, st, np.linspace(0.1, 1.0, nsteps + 1), [n] * 3)
    idx = np.arange(0, n**3, n**3 // KEEP)
    X = st[0, 0, idx].numpy()
    P = st[1, 0, idx].numpy()
    rho = opm.cic_paint(torch.zeros(1, n, n, n, dtype=torch.float64), st[0])
    rho = rho[0, list(planes)].numpy().astype(np.float32)
  np.savez(os.path.join(HERE, name), idx=idx, X=X, P=P, rho=rho,
           planes=np.array(planes), n=n, box=box, seed=seed, order=order,
           nsteps=nsteps)
  print("%s: %.0f s" % (name, time.time() - t0), flush=True)


def main(which):
  if "c3" in which:
    for j, n in enumerate([32, 64, 96, 128]):
      run("c3_%d.npz" % n, n, 100.0 * n / 32, 1000 + 8 * j, 1, 10,
          (0, n // 2))
  if "c4" in which:
    run("c4_256.npz", 256, 512.0, 100, 2, 40, (0, 64, 128, 192))


if __name__ == "__main__":
  main(sys.argv[1:] or ["c3", "c4"])
