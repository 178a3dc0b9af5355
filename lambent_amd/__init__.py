"""lambent_amd: MI355X (gfx950) implementation of Lambent's path-tracing hot path.

The product is the C-ABI library liblambent_hip.so (include/lambent.h); this
package mirrors the reference's host API on top of it (see lambent_amd.api).
"""
from .api import *  # noqa: F401,F403
from .api import LambentError, CameraError, TextureError, version, library_path  # noqa: F401
