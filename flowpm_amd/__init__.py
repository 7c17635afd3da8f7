"""flowpm_amd: FlowPM's particle-mesh step on MI355X (gfx950).

Drop-in for the reference's PM path (flowpm/__init__.py:1-7): the same
functions and signatures on torch tensors in HBM, computed by hand-written
HIP kernels in ``libpmx_hip.so`` (see include/pmx.h).  The PM path and its
immediate callers are provided (power_spectrum, compensate_cic and the
snapshot export io.save_state among them); raytracing and the
mesh-TensorFlow solver are not.
"""
from . import constants
from . import cosmology
from . import kernels
from . import tfbackground
from . import tfbackground as background
from . import tfpower
from . import utils
from . import tfpm
from . import pk
from . import io
from .pk import power_spectrum
from .utils import cic_paint, cic_readout, r2c3d, c2r3d, white_noise
from .utils import compensate_cic
from .tfpm import nbody, linear_field, lpt_init, whitenoise_to_linear

__version__ = "0.1.0"
