"""MI355X-native GENESIS 1.3 v4 time-dependent FEL step.

The compute path is the gfx950 library libg4hip.so (C ABI in include/g4hip.h);
this package only binds it for Python callers (tests, benchmarks, scripts).
"""
from .api import *  # noqa: F401,F403
from .api import __all__  # noqa: F401

__version__ = "0.1.0"
