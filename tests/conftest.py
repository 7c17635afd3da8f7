import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
  sys.path.insert(0, ROOT)


def pytest_configure(config):
  config.addinivalue_line(
      "markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def gpu():
  import torch
  if not torch.cuda.is_available():
    pytest.skip("no GPU")
  from flowpm_amd import _lib
  _lib.lib()  # raises if the HIP library is missing
  return torch.device("cuda:0")


@pytest.fixture(scope="session")
def golden_dir():
  import pathlib
  return pathlib.Path(ROOT) / "tests" / "golden"
