"""waveforge_amd: MI355X-native 2D TM FDTD + DAS engine.

Hand-written HIP kernels for gfx950 behind a C ABI (``include/waveforge_hip.h``,
``libwaveforge_hip.so``), driven from Python through the reference's ``core``
API (``waveforge_amd.core``). The HIP library is the only compute path.
"""

__version__ = "0.1.0"

from . import _native  # noqa: F401

LIB_PATH = _native.LIB_PATH
