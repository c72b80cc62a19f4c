"""MI355X-native drop-in for the 2D DUGKS time step of MPDUGKS ("Ark").

The compute path is libmpdugks_hip.so (HIP kernels for gfx950 behind the C ABI
of include/mpdugks.h); this package holds its Python front end.
"""
from . import abi  # noqa: F401
from .config import default_config  # noqa: F401
