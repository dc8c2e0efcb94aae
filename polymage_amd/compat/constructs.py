"""``from constructs import *`` -> polymage_amd.constructs (sandbox/constructs.py)."""
from polymage_amd.constructs import *  # noqa: F401,F403
