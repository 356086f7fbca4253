"""CPU oracle for the FLIP hot path (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package.  The product path never does.
"""
from .binding import OracleSim, build, lib  # noqa: F401
