"""Import-name shims: put this directory on ``sys.path`` and application
files written for the reference (``from compiler import *``,
``from constructs import *``, ``import tuner``, ``from utils import *``)
run unchanged against polymage_amd."""
import os
PATH = os.path.dirname(os.path.abspath(__file__))
