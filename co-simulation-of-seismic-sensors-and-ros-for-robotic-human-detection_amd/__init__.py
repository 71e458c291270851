"""MI355X-native axisymmetric viscoelastic FDTD (drop-in for the reference's
internal/fdtd behind the C ABI in include/geofdtd.h).

    fdtd      the reference's fdtd package API over libgeofdtd.so (ctypes)
    fk        the reference's fk package (frequency-wavenumber solver) API
    bank      bank.Bank's interpolation on the device (batched lookups)
    stream    stream.Convolver for many receivers (chunked convolution)
    sensing   sensing.Engine over a bank, for many receivers (runtime synthesis)
    csrc/     HIP kernels for gfx950 and the C ABI

The directory name is not a Python identifier; load it with
importlib (see tests/conftest.py) under any alias.
"""

from . import fdtd  # noqa: F401
from . import bank  # noqa: F401
from . import fk  # noqa: F401
from . import stream  # noqa: F401
from . import sensing  # noqa: F401

__all__ = ["fdtd", "fk", "bank", "stream", "sensing"]
