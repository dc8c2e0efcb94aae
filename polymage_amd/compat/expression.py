"""``from expression import *`` -> polymage_amd.expression (sandbox/expression.py)."""
from polymage_amd.expression import *  # noqa: F401,F403
