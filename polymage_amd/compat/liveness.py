"""``from liveness import compute_liveness`` (sandbox/liveness.py)."""
from polymage_amd.liveness import *  # noqa: F401,F403
