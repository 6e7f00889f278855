"""MI355X (gfx950) solver for the EffPlast2D pseudo-transient elastoplastic
RVE problem: hand-written HIP kernels behind a C ABI (libep2d.so), with a
Python mirror of the reference's EffPlast2D class."""
from ._lib import (FIELDS, MODE_FP64, MODE_PLAIN, MODE_SPLIT, MODES, Ep2dError,  # noqa: F401
                   LIB_PATH)
from .params import derive_params, read_pa, write_pa  # noqa: F401
from .solver import EffPlast2D, device_count  # noqa: F401
from .batch import Batch, request_bytes  # noqa: F401,E402
