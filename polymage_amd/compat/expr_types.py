"""``from expr_types import *`` -> polymage_amd.expr_types (sandbox/expr_types.py)."""
from polymage_amd.expr_types import *  # noqa: F401,F403
