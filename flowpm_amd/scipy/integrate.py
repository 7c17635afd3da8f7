"""Composite Simpson rule (flowpm/scipy/integrate.py:4-36), float64 host."""
import numpy as np


def simps(f, a, b, N=128):
  if N % 2 == 1:
    raise ValueError("N must be an even integer.")
  x = np.linspace(a, b, N + 1)
  y = f(x)
  w = np.ones(N + 1)
  w[1:-1:2] = 4.0
  w[2:-1:2] = 2.0
  return (b - a) / N / 3.0 * np.tensordot(w, y, axes=(0, 0))
