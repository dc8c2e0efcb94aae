"""``from compiler import *`` -> polymage_amd.compiler (sandbox/compiler.py)."""
from polymage_amd.compiler import *  # noqa: F401,F403
from polymage_amd.compiler import buildPipeline  # noqa: F401
