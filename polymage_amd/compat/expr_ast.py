"""``from expr_ast import *`` -> polymage_amd.expr_ast (sandbox/expr_ast.py)."""
from polymage_amd.expr_ast import *  # noqa: F401,F403
