"""sphhip -- host-side Python mirror of the reference's simulator interface over
libsph_hip.so (the MI355X grid-SPH step, C ABI in include/sph_hip.h)."""
from . import _lib  # noqa: F401
from ._lib import SphError, load, LIB_PATH  # noqa: F401
from .simulator import HIPGridParticleSimulator, DAM_BREAK, FOUNTAIN  # noqa: F401
from . import scenes  # noqa: F401
from .batch import SceneBatch, variant_shape  # noqa: F401
from . import serving  # noqa: F401
from .serving import Server  # noqa: F401
