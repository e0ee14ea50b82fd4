"""amcafe_amd -- MI355X-native CAFE time step (AMCAFE) behind a C ABI.

The product is the native library ``amcafe_amd/_lib/libamcafe.so`` (HIP kernels for
gfx950 plus host C++), declared in ``include/amcafe.h``.  This package only binds it
with ctypes; it has no compute path of its own and raises if the library is absent.
"""
from ._lib import lib, LibraryMissing, AmcafeError  # noqa: F401
from .sim import Sim, Domain, read_deck  # noqa: F401
from .heat import HeatBatch, HeatField, calibrate, track_desc  # noqa: F401
from .meltpool import production_deck, semi_axes  # noqa: F401
