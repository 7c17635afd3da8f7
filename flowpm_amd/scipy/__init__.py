"""Host-side numerical helpers (flowpm/scipy/*)."""
from . import integrate
from . import interpolate
