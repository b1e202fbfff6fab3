"""Helpers for the GPU tests: load the product library through its C ABI."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "rigid-ipc_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

import rigidipc_amd  # noqa: E402


def context():
    """A rigc context on device 0; fails loudly without the library or GPU."""
    return rigidipc_amd.Context(0)
