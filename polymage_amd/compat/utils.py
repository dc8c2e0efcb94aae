"""``from utils import *`` -> polymage_amd.utils (sandbox/utils.py)."""
from polymage_amd.utils import *  # noqa: F401,F403
