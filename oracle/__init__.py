"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the CPU oracle (liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this
package.  The product library never loads it.
"""
from .oracle import *  # noqa: F401,F403
