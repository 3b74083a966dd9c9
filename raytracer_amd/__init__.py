"""MI355X-native MIS path tracer (wavefront HIP kernels for gfx950) behind a C ABI (include/rtg.h)."""
from . import rtg, scenes  # noqa: F401
from .rtg import Context, Scene, resolve  # noqa: F401
