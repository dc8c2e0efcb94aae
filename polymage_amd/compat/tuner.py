"""``import tuner`` -> polymage_amd.tuner (sandbox/tuner.py generate/execute)."""
from polymage_amd.tuner import *  # noqa: F401,F403
from polymage_amd.tuner import generate, execute  # noqa: F401
