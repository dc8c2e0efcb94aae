"""polymage_amd — PolyMage's fused, overlapped-tiled stencil groups on
AMD Instinct MI355X (gfx950).

The DSL and ``buildPipeline`` keep the reference's API (sandbox/
constructs.py, sandbox/compiler.py); each group becomes one hand-emitted
HIP kernel, compiled with hiprtc and called through the C ABI of a small
generated host stub.
"""

from .expr_types import *  # noqa: F401,F403
from .constructs import *  # noqa: F401,F403
from .compiler import buildPipeline  # noqa: F401

__version__ = '0.1.0'
