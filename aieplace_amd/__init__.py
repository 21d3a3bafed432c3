"""aieplace_amd -- MI355X-native ePlace global placement hot path.

The product is libeplace.so (hand-written HIP kernels for gfx950 behind the C
ABI in include/eplace.h).  Importing this package loads it and fails loudly if
it has not been built (`python -m aieplace_amd.build`).
"""
from . import api  # noqa: F401  (raises ImportError if libeplace.so is missing)
from .netlist import Netlist, generate  # noqa: F401
from .placer import Placer  # noqa: F401
from .configs import CONFIGS, placer_params  # noqa: F401
