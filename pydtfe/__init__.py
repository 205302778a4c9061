"""``import pydtfe`` for the reference's Python callers: the package layout of
python/pydtfe/__init__.py:1 (the package re-exports its ``pydtfe`` module),
here over the GPU mirror in ``sdtfe_amd.pydtfe``."""
from .pydtfe import *  # noqa: F401,F403
